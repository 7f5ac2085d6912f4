"""The task-farm replacement (mma_amd.farm): CUPRAD (z, t, r) field -> coarsened instances in
MPI-TDSE's simulation order -> one batched launch -> outputs in merge.py's layout.

CPU: units against CODATA atomic units, the instance order and coarsening of MPI-TDSE.c, the
device-side FourInterp and spectra (run on CPU tensors) against the oracle and closed forms.
GPU: every instance of a small field, k1 = 1 and k1 = 5, against the oracle's call1DTDSE, and the
merged spectra read back through the Hankel provider's 'dynamic' source."""
import math

import numpy as np
import pytest
import torch

from conftest import rel_err
from mma_amd import farm, workloads as W


def test_units_are_atomic_units():
    """The C unit set (constants.c:26-43) gives the atomic units of time and field (CODATA
    2018: 2.4188843265857e-17 s, 5.14220674763e11 V/m) to the precision of its older constants
    (e = 1.602176565e-19 is the 2010 value): 1.4e-7 and 1.8e-7 off."""
    assert abs(farm.TIMEAU / 2.4188843265857e-17 - 1) < 3e-7
    assert abs(farm.EFIELDAU / 5.14220674763e11 - 1) < 3e-7
    assert abs(farm.TIMEAU / 2.4188843265857e-17 - 1) > 1e-8  # not the CODATA 2018 set


def test_instance_order_and_coarsening():
    nz, nt, nr = 7, 5, 9
    z, t, r = np.meshgrid(np.arange(nz), np.arange(nt), np.arange(nr), indexing="ij")
    field = (1000.0 * z + r + 0.001 * t) * farm.EFIELDAU  # (z, t, r), value encodes the indices
    f, nz_c, nr_c = farm.coarse_fields(field.astype(np.float64), kz_step=2, nz_max=7, kr_step=3, nr_max=8)
    assert (nz_c, nr_c) == (3, 2) and f.shape == (6, nt)
    for n in range(6):
        kz, kr = farm.instance_index(n, nr_c)
        np.testing.assert_allclose(f[n], 1000.0 * 2 * kz + 3 * kr + 0.001 * np.arange(nt), rtol=1e-15)
    np.testing.assert_array_equal(farm.coarsen_grid(np.arange(10.0), 3, 8), [0.0, 3.0])
    with pytest.raises(ValueError):
        farm.coarse_fields(field, 1, 7, 1, nr + 1)


def test_four_interp_matches_oracle_and_band_limited_signal():
    from oracle import ctdse as O
    n, k1 = 64, 5
    t = np.arange(n)
    sig = np.cos(2 * np.pi * 3 * t / n) + 0.5 * np.sin(2 * np.pi * 7 * t / n)
    got = farm.four_interp(torch.tensor(sig[None, :]), k1)[0].numpy()
    ref = O.four_interp(k1, sig)[:k1 * n]
    assert np.max(np.abs(got - ref)) < 1e-13
    tt = np.arange(k1 * n) / k1  # a band-limited signal is reproduced at the refined points
    exact = np.cos(2 * np.pi * 3 * tt / n) + 0.5 * np.sin(2 * np.pi * 7 * tt / n)
    assert np.max(np.abs(got - exact)) < 1e-13


def test_spectra_closed_form():
    """A cosine at bin m: F_m = conj(N/2) dt / sqrt(2 pi), omega_m = m 2 pi / tmax."""
    n, m, dt, tmax = 128, 5, 0.25, 31.0
    sig = np.cos(2 * np.pi * m * np.arange(n) / n)
    omega, F, M2 = farm.spectra(torch.tensor(sig), dt, tmax)
    assert abs(omega[m].item() - m * 2 * np.pi / tmax) < 1e-14
    np.testing.assert_allclose(F[m].numpy(), [n / 2 * dt / math.sqrt(2 * math.pi), 0.0], atol=1e-12)
    assert abs(M2[m].item() - dt * dt / (2 * math.pi) * (n / 2) ** 2) < 1e-10
    assert np.max(np.abs(F.numpy()[np.arange(F.shape[0]) != m])) < 1e-12


def _cuprad_like(nz=5, nr=7, dt_field_au=0.25, n_cycles=2, e0_au=0.06):
    """A Gaussian-beam field in SI units on a (z, t, r) grid, float32 like CUPRAD's output."""
    w = W.omega_au(800e-9)
    T = n_cycles * 2 * math.pi / w
    n_e = int(math.ceil(T / dt_field_au)) + 1
    t_au = np.arange(n_e) * dt_field_au
    z = np.linspace(0, 1e-3, nz)
    r = np.linspace(0, 150e-6, nr)
    amp = np.exp(-(r[None, :] / 100e-6) ** 2) * (1 - 0.1 * z[:, None] / 1e-3)
    env = W.sin2_pulse(t_au, 1.0, w, n_cycles)
    field = e0_au * farm.EFIELDAU * amp[:, None, :] * env[None, :, None]
    return field.astype(np.float32), t_au * farm.TIMEAU, r, z


@pytest.mark.gpu
@pytest.mark.parametrize("dt_field", [0.25, 1.0])
def test_farm_instances_match_oracle(dt_field):
    """Every instance of the coarsened field against the oracle's call1DTDSE on the same
    (float32 -> float64, divided by EFIELDau) field: k1 = 1 (field dt = dt) and k1 = 5 (field dt
    1.0, dt 0.25 -> refined dt 0.2, FourInterp on the device)."""
    from oracle import ctdse as O
    field, tgrid, r, z = _cuprad_like(dt_field_au=dt_field)
    num_r, dx, a = 2000, 0.1, W.GRID["a"]
    res = farm.run(field, tgrid, r, z, kz_step=2, nz_max=5, kr_step=3, nr_max=7, num_r=num_r, dx=dx, dt=0.25, a=a,
                   e_guess=W.GRID["e_guess"], cv=1e-25, x_int=2.0)
    host = farm.as_host(res)
    P = farm.OUTPUTS
    assert host[P + "/SourceTerm"].shape[:2] == (2, 2)
    x = host[P + "/xgrid_micro"]
    psi0 = host[P + "/ground_state"][:, 0] + 1j * host[P + "/ground_state"][:, 1]
    f, nz_c, nr_c = farm.coarse_fields(field, 2, 5, 3, 7)
    t_au = tgrid / farm.TIMEAU
    for n in range(nz_c * nr_c):
        kz, kr = farm.instance_index(n, nr_c)
        ref = O.call1dtdse(num_r, dx, a, 2.0, 0.25, psi0, x, t_au, f[n])
        for key, name in [("sourceterm", "SourceTerm"), ("PopTot", "PopTot"), ("PopInt", "PopInt"),
                          ("expval", "expval_x"), ("Efield", "Efield")]:
            e = rel_err(host[P + "/" + name][kz, kr], ref[key])
            print(dt_field, n, name, "%.2e" % e, flush=True)
            assert e < 1e-8, (n, name)
        fs = host[P + "/FSourceTerm"][kz, kr]
        assert rel_err(fs[:, 0] + 1j * fs[:, 1], ref["Fsourceterm"]) < 1e-8
        fe = host[P + "/FEfield"][kz, kr]
        assert rel_err(fe[:, 0] + 1j * fe[:, 1], ref["FEfield"]) < 1e-8
    assert rel_err(host[P + "/omegagrid"], ref["omegagrid"]) < 1e-14
    assert rel_err(host[P + "/tgrid"], ref["tgrid"]) < 1e-14
    np.testing.assert_array_equal(host[P + "/zgrid_coarse"], z[[0, 2]])
    np.testing.assert_array_equal(host[P + "/rgrid_coarse"], r[[0, 3]])


@pytest.mark.gpu
def test_farm_output_feeds_the_hankel_provider():
    """The merged FSourceTerm is read by FSources_provider(data_source='dynamic') exactly as the
    Hankel driver reads the merged HDF5 file (Hankel_long_medium_parallel_cluster.py:203-215)."""
    from mma_amd import hankel as H
    field, tgrid, r, z = _cuprad_like(nz=6, nr=4)
    res = farm.as_host(farm.run(field, tgrid, r, z, kz_step=1, nz_max=6, kr_step=1, nr_max=4, num_r=1600, dx=0.1,
                                dt=0.25, a=W.GRID["a"]))
    P = farm.OUTPUTS
    og = res[P + "/omegagrid"]
    prov = H.FSources_provider(res[P + "/zgrid_coarse"], res[P + "/rgrid_coarse"], og, h5_handle=res,
                               h5_path=P + "/FSourceTerm", data_source="dynamic", ko_min=3, ko_max=40)
    planes = list(prov.Fsource_plane)
    FS = res[P + "/FSourceTerm"]
    assert len(planes) == 5  # the provider drops the last plane (Hankel_transform.py:62)
    for k, pl in enumerate(planes):
        np.testing.assert_array_equal(pl, (FS[k, :, 3:40, 0] + 1j * FS[k, :, 3:40, 1]).T)


@pytest.mark.gpu
def test_archive_run_writes_the_merged_layout(tmp_path):
    """The farm on an MMA archive: CUPRAD/outputs (float32 (z, t, r) field and its grids), CTDSE
    inputs with a gas preset and Ntinterp refinement (InterpByDTorNT = 1, Ntinterp = 1: k1 = 2),
    run_archive writing CTDSE/outputs/* into the archive itself; the written outputs equal
    farm.run's on the same inputs bit for bit, carry their units, keep the archive's other
    datasets, and feed the Hankel provider as the reference's merged file does; the far field
    of those sources is written to a results file and read back."""
    from mma_amd import far_field as FF, h5lite
    field, tgrid, r, z = _cuprad_like(nz=4, nr=3, dt_field_au=0.5)
    p = str(tmp_path / "results.h5")
    d = {"/CUPRAD/outputs/output_field": field, "/CUPRAD/outputs/tgrid": tgrid, "/CUPRAD/outputs/rgrid": r,
         "/CUPRAD/outputs/zgrid": z, "/global_inputs/gas_preset": b"Ar",
         "/CTDSE/inputs/Nx_max": np.int32(800), "/CTDSE/inputs/dx": 0.1, "/CTDSE/inputs/InterpByDTorNT": np.int32(1),
         "/CTDSE/inputs/dt": 0.25, "/CTDSE/inputs/Ntinterp": np.int32(1), "/CTDSE/inputs/x_int": 2.0,
         "/CTDSE/inputs/CV_criterion_of_GS": 1e-25, "/CTDSE/inputs/kz_step": np.int32(1),
         "/CTDSE/inputs/kr_step": np.int32(1), "/CTDSE/inputs/Nz_max": np.int32(4), "/CTDSE/inputs/Nr_max": np.int32(3)}
    h5lite.write(p, d, {"/CUPRAD/outputs/output_field": {"units": b"[V/m]"}})
    res = farm.as_host(farm.run_archive(p))
    ref = farm.as_host(farm.run(field, tgrid, r, z, 1, 4, 1, 3, 1600, 0.1, 0.25, W.GRID["a"], e_guess=W.GRID["e_guess"],
                                cv=1e-25, x_int=2.0, interp_by_nt=1, ntinterp=1))
    f = h5lite.H5File(p)
    P = farm.OUTPUTS
    for key in ref:
        np.testing.assert_array_equal(res[key], ref[key])
        np.testing.assert_array_equal(f["/" + key], ref[key])
    assert f.attrs("/" + P + "/PopInt")["units"] == b"[-]" and f.attrs("/" + P + "/SourceTerm")["units"] == b"[a.u.]"
    np.testing.assert_array_equal(f["/CUPRAD/outputs/output_field"], field)
    assert f.attrs("/CUPRAD/outputs/output_field")["units"] == b"[V/m]"
    n_e = len(tgrid)
    assert f["/" + P + "/SourceTerm"].shape == (4, 3, 2 * n_e + 2)  # k1 = 2: outputs.Nt = k1 N_E + 2
    # the far field of the written sources (vacuum), read through the file as the driver does
    w0 = FF.omega_au_of_wavelength(800e-9)
    out = FF.screen(f["/" + P + "/omegagrid"], f["/" + P + "/rgrid_coarse"], f["/" + P + "/zgrid_coarse"], f, w0,
                    [3.0, 9.0], rmax_FF=4e-3, Nr_FF=16, distance_FF=1.0, include_absorption=False,
                    include_dispersion=False)
    mem = FF.screen(ref[P + "/omegagrid"], ref[P + "/rgrid_coarse"], ref[P + "/zgrid_coarse"], ref, w0, [3.0, 9.0],
                    rmax_FF=4e-3, Nr_FF=16, distance_FF=1.0, include_absorption=False, include_dispersion=False)
    np.testing.assert_array_equal(out.FF_integrated, mem.FF_integrated)
    q = str(tmp_path / "results_Hankel.h5")
    FF.save(q, out)
    back = FF.load(q)
    np.testing.assert_array_equal(back[FF.OUTPUTS + "/FF_integrated"][..., 0], out.FF_integrated.real)
    np.testing.assert_array_equal(back[FF.OUTPUTS + "/ogrid"], out.ogrid)
    assert h5lite.H5File(q).attrs("/" + FF.OUTPUTS + "/rgrid")["units"] == b"[SI]"
