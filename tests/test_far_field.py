"""The far-field driver (mma_amd.far_field, replacing Hankel_long_medium_parallel_cluster.py):
harmonic-window bisection against its definition, the Python unit set against CODATA, and on
the GPU the whole chain field -> task farm -> far field against the FP64 oracle."""
import math

import numpy as np
import pytest

from conftest import rel_err
from mma_amd import far_field as FF


def test_find_interval_definition():
    """k with x_k <= x0 < x_k+1, the last interval closed; LookupError outside (mynumerics.py:35-56).
    Checked against that definition, which a bisection off by one (a wrong midpoint or bound)
    breaks on the random grids."""
    rng = np.random.Generator(np.random.PCG64(3))
    for n in (2, 3, 7, 64, 1000):
        x = np.sort(rng.uniform(-5, 5, n))
        probes = np.concatenate([x, rng.uniform(x[0], x[-1], 50)])
        for x0 in probes:
            k = FF.find_interval(x, x0)
            if x0 == x[-1]:
                # the last point closes the last interval, except on a two-point grid, where the
                # reference's bisection returns 1 (its length-2 test, mynumerics.py:50-51)
                assert k == (1 if n == 2 else n - 2)
            else:
                assert x[k] <= x0 < x[k + 1], (n, x0, k)
        with pytest.raises(LookupError):
            FF.find_interval(x, x[-1] + 1e-9)
        with pytest.raises(LookupError):
            FF.find_interval(x, x[0] - 1e-9)
    np.testing.assert_array_equal(FF.find_interval(np.arange(5.0), [0.0, 2.5, 4.0]), [0, 2, 3])


def test_python_unit_set():
    """units.py's atomic unit of time (CODATA 2018: 2.4188843265857e-17 s; its constants are
    older, 4e-9 off) and ConvertPhoton's 800 nm frequency (0.0569542 a.u.)."""
    assert abs(FF.TIMEAU / 2.4188843265857e-17 - 1) < 1e-8
    assert abs(FF.OMEGA_AU2SI * FF.TIMEAU - 1) < 1e-15
    w = FF.omega_au_of_wavelength(800e-9)
    assert abs(w - 0.0569542) < 5e-7
    assert abs(w / (2 * math.pi * FF.C_LIGHT / 800e-9 / FF.OMEGA_AU2SI) - 1) < 1e-12  # omega = 2 pi c / lambda


@pytest.mark.gpu
def test_field_to_far_field_matches_oracle():
    """A CUPRAD-like field through the task farm (merged spectra) and the far-field driver
    (harmonics 3-9 of 800 nm on a 40-radius screen at 1 m) against the FP64 oracle of
    Hankel_long on the same selected planes."""
    from mma_amd import farm, workloads as W
    from oracle import hankel as OH
    from test_farm import _cuprad_like
    field, tgrid, r, z = _cuprad_like(nz=5, nr=6, n_cycles=4)
    res = farm.as_host(farm.run(field, tgrid, r, z, kz_step=1, nz_max=5, kr_step=1, nr_max=6, num_r=1600,
                                dx=0.1, dt=0.25, a=W.GRID["a"]))
    P = farm.OUTPUTS
    w0 = FF.omega_au_of_wavelength(800e-9)
    # vacuum propagation: the reference needs the tables switched off for preset_gas 'vacuum'
    # (its pre-factor looks up 'vacuum_<tables>' otherwise, Hankel_transform.py:409-427)
    out = FF.screen(res[P + "/omegagrid"], res[P + "/rgrid_coarse"], res[P + "/zgrid_coarse"], res, w0, [3.0, 9.0],
                    rmax_FF=4e-3, Nr_FF=40, distance_FF=1.0, include_absorption=False, include_dispersion=False)
    og = res[P + "/omegagrid"]
    k0, k1 = FF.find_interval(og / w0, 3.0), FF.find_interval(og / w0, 9.0)
    assert out.FF_integrated.shape == (40, k1 - k0)
    np.testing.assert_array_equal(out.ogrid, FF.OMEGA_AU2SI * og[k0:k1])
    FS = res[P + "/FSourceTerm"]
    fsource = np.transpose(FS[..., 0] + 1j * FS[..., 1], (0, 2, 1))  # (z, omega, r)
    zs, rs, os_, planes = OH.provider_planes(res[P + "/zgrid_coarse"], res[P + "/rgrid_coarse"],
                                             FF.OMEGA_AU2SI * og, fsource, ko_min=k0, ko_max=k1)
    ref = OH.hankel_long(zs, rs, os_, planes, 1.0, np.linspace(0, 4e-3, 40), include_absorption=False,
                         include_dispersion=False)
    for key in ("FF_integrated", "entry_plane_transform", "exit_plane_transform"):
        e = rel_err(getattr(out, key), ref[key])
        print(key, "%.2e" % e, flush=True)
        assert e < 1e-11, key
    d = FF.outputs(out)
    assert d[FF.OUTPUTS + "/FF_integrated"].shape == (40, k1 - k0, 2)
    # the farm's own result (device tensors) feeds the driver directly, with the same bits
    dev_res = farm.run(field, tgrid, r, z, kz_step=1, nz_max=5, kr_step=1, nr_max=6, num_r=1600, dx=0.1, dt=0.25,
                       a=W.GRID["a"])
    out_dev = FF.screen(dev_res[P + "/omegagrid"], dev_res[P + "/rgrid_coarse"], dev_res[P + "/zgrid_coarse"],
                        dev_res, w0, [3.0, 9.0], rmax_FF=4e-3, Nr_FF=40, distance_FF=1.0, include_absorption=False,
                        include_dispersion=False)
    for key in ("FF_integrated", "entry_plane_transform", "exit_plane_transform"):
        np.testing.assert_array_equal(getattr(out_dev, key), getattr(out, key))
