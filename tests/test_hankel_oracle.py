"""The Hankel oracle (oracle/hankel.py) against what mathematics and physics fix: the Gaussian
Hankel pair (with and without the near-field phase), rho = 0, linearity, the longitudinal
trapezoid over two planes, the provider's plane selection, the scalar pre-factor's closed form
and renormalisation invariant, the z-modulated branch reducing to the scalar branch for a
constant pressure, photon conversions, and the high-energy limit of the XUV index (plasma
dispersion, built from e, eps0 and m instead of the classical electron radius)."""
import math

import numpy as np
import pytest

from oracle import hankel as H

OMEGA0 = 2.0 * math.pi * H.C_LIGHT / 800e-9


def _gauss_case(n_r=16001, r_max=2e-4, w=3e-5):
    r = np.linspace(0.0, r_max, n_r)
    o = OMEGA0 * np.array([11.0, 25.0, 61.0])
    a = 1.0 / w ** 2
    F = np.exp(-a * r ** 2)[None, :] * np.ones((len(o), 1))
    rho = np.linspace(0.0, 7e-3, 17)
    return r, o, a, F, rho


@pytest.mark.parametrize("near", [False, True])
def test_gaussian_pair(near):
    """int_0^inf r exp(-a r^2) J0(b r) dr = exp(-b^2 / 4a) / 2a, a complex with the near-field
    phase exp(-i k r^2 / 2D).  The trapezoid's own error is h^2 a / 6 ~ 3e-8 relative here."""
    r, o, a, F, rho = _gauss_case()
    D = 1.0
    got = H.hankel_plane(o, r, F, D, rho, near_field_factor=near)
    for i, om in enumerate(o):
        k = om / H.C_LIGHT
        aa = a + (1j * k / (2 * D) if near else 0.0)
        b = k * rho / D
        ref = np.exp(-b ** 2 / (4 * aa)) / (2 * aa)
        assert np.max(np.abs(got[i] - ref)) <= 1e-7 * np.max(np.abs(ref))


def test_rho_zero_is_trapezoid_and_linear():
    rng = np.random.default_rng(0)
    r = np.sort(rng.uniform(0, 1e-4, 300))  # non-uniform radial grid
    o = OMEGA0 * np.array([15.0, 33.0])
    F = rng.standard_normal((2, 300)) + 1j * rng.standard_normal((2, 300))
    G = rng.standard_normal((2, 300)) + 1j * rng.standard_normal((2, 300))
    rho = np.array([0.0, 1e-3, 3e-3])
    D = 0.9
    pf = H.hankel_plane(o, r, F, D, rho)
    for i, om in enumerate(o):
        y = np.exp(-1j * om / H.C_LIGHT * r ** 2 / (2 * D)) * r * F[i]
        assert abs(pf[i, 0] - np.trapezoid(y, r)) <= 1e-14 * np.sum(np.abs(y)) * r[-1]
    pg = H.hankel_plane(o, r, G, D, rho)
    pfg = H.hankel_plane(o, r, 2.0 * F - 3j * G, D, rho)
    assert np.max(np.abs(pfg - (2.0 * pf - 3j * pg))) <= 1e-13 * np.max(np.abs(pfg))


def test_provider_and_two_plane_trapezoid():
    rng = np.random.default_rng(1)
    z = np.array([0.0, 1e-3, 2.5e-3])  # the provider keeps z[0:-1]: two planes
    r = np.linspace(0, 1e-4, 64)
    o = OMEGA0 * np.linspace(11, 21, 6)
    S = rng.standard_normal((3, 6, 64)) + 1j * rng.standard_normal((3, 6, 64))
    zz, rr, oo, planes = H.provider_planes(z, r, o, S)
    assert len(zz) == 2 and len(planes) == 2 and np.array_equal(planes[1], S[1])
    rho = np.linspace(0, 5e-3, 9)
    res = H.hankel_long(zz, rr, oo, planes, 1.0, rho, include_absorption=False, include_dispersion=False)
    p0 = H.hankel_plane(oo, rr, planes[0], 1.0 - zz[0], rho).T
    p1 = H.hankel_plane(oo, rr, planes[1], 1.0 - zz[1], rho).T * np.exp(1j * oo * zz[1] * 0.0)[None, :]
    ref = 0.5 * (zz[1] - zz[0]) * (p0 + p1)
    assert np.max(np.abs(res["FF_integrated"] - ref)) <= 1e-14 * np.max(np.abs(ref))
    assert np.array_equal(res["entry_plane_transform"], p0)


def test_photon_conversions():
    om = OMEGA0 * 21
    assert abs(H.omega_si_to_ev(om) - H.HBAR * om / H.ELCHARGE) <= 1e-12 * H.omega_si_to_ev(om)
    assert abs(H.omega_si_to_lambda(om) - 2 * math.pi * H.C_LIGHT / om) <= 1e-12 * 800e-9 / 21


def test_xuv_index_is_plasma_dispersion_with_f1_electrons():
    """n_XUV = 1 - N f1 e^2 / (2 eps0 m omega^2) (the index written with e, eps0 and m instead of
    the classical electron radius and the wavelength); far above the edges f1 -> Z."""
    eps0 = 1.0 / (4e-7 * math.pi * H.C_LIGHT ** 2)
    for gas, Z in (("Ar_NIST", 18), ("He_NIST", 2), ("Kr_NIST", 36)):
        om = 3.0e5 * H.ELCHARGE / H.HBAR
        assert abs(float(H.getf1(gas, H.omega_si_to_ev(om))) - Z) < 0.05 * Z
    for gas in ("Ar_NIST", "Ar_Henke", "Ne_Henke", "Xe_NIST"):
        for harmonic in (25, 45, 61):  # Henke lists f1 from 29.3 eV (harmonic 19 of 800 nm)
            om = OMEGA0 * harmonic
            f1 = float(H.getf1(gas, H.omega_si_to_ev(om)))
            p = 0.8
            d = H.dispersion_function(om, p, gas)  # (1/v_IR) - (1/v_XUV) = (1 - n_XUV)/c for n_IR = 1
            ref = p * H.N_REF * f1 * H.ELCHARGE ** 2 / (2 * eps0 * H.ELMASS * om ** 2) / H.C_LIGHT
            # 1/v_IR - 1/v_XUV cancels to (1 - n_XUV)/c: rounding of 1/c relative to the difference
            tol = 1e-7 * abs(ref) + 4e-16 / H.C_LIGHT
            assert abs(d - ref) <= tol, (gas, harmonic, d, ref)
            if gas != "Xe_NIST":  # NIST lists f2 of Xe from 64 eV
                # beta and delta = 1 - n_XUV share N r_e lambda^2 / 2 pi: beta / delta = f2 / f1
                f2 = float(H.getf2(gas, H.omega_si_to_ev(om)))
                delta = p * H.N_REF * f1 * H.ELCHARGE ** 2 / (2 * eps0 * H.ELMASS * om ** 2)
                assert abs(p * H.beta_factor_ref(om, gas) / delta - f2 / f1) <= 1e-7 * abs(f2 / f1)


def test_scalar_prefactor_closed_form_and_renorm():
    z = np.linspace(0, 5e-3, 11)
    r = np.linspace(0, 1e-4, 4)
    o = OMEGA0 * np.linspace(15, 41, 7)
    pre, ren = H.prefactor_tables(z, r, o, "vacuum", 1.0, include_absorption=False, include_dispersion=False,
                                  n_ir=1.0003)
    ref = np.exp(1j * o[None, :] * z[:, None] * (1.0003 - 1.0) / H.C_LIGHT)  # phases up to ~480 rad
    assert np.max(np.abs(pre - ref)) <= 1e-12
    assert np.all(ren == 1.0)
    pre, ren = H.prefactor_tables(z, r, o, "Ar", 0.5, absorption_tables="Henke", dispersion_tables="NIST")
    assert np.allclose(np.abs(pre) * ren, 0.5, rtol=1e-13, atol=0)
    assert np.all(np.abs(pre[-1]) == pytest.approx(0.5, rel=1e-15))


def test_z_modulated_constant_pressure_equals_scalar():
    z = np.linspace(0, 5e-3, 9)
    r = np.linspace(0, 1e-4, 4)
    o = OMEGA0 * np.linspace(17, 45, 5)
    zp = np.linspace(0, 6e-3, 41)
    prof = {"zgrid": zp, "value": np.full(41, 0.7)}
    a, ra = H.prefactor_tables(z, r, o, "Ar", prof, absorption_tables="Henke", dispersion_tables="NIST")
    b, rb = H.prefactor_tables(z, r, o, "Ar", 0.7, absorption_tables="Henke", dispersion_tables="NIST")
    assert np.max(np.abs(a - b)) <= 1e-11 * np.max(np.abs(b))
    assert np.max(np.abs(ra - rb)) <= 1e-11 * np.max(np.abs(rb))


def test_z_and_r_modulated_separable_pressure_matches_z_branch():
    """The z-and-r branch (Hankel_transform.py:244-399) on a separable table p(z) s(r) given at
    the r nodes: at each radius it is the z-modulated branch (itself pinned to the scalar one
    above) for the table s(r) p(z), without that branch's exit-plane renormalisation (:215 has
    no counterpart there), i.e. times renorm[0] when the first plane sits at the table's start;
    beta comes from the dispersion tables there (:373-376), so both tables are NIST here.
    A transposed (z, r) table, a wrong radius or a renormalised exit plane fails."""
    z = np.linspace(0, 5e-3, 9)
    r = np.linspace(0, 1e-4, 4)
    o = OMEGA0 * np.linspace(17, 45, 5)
    zp = np.linspace(0, 6e-3, 41)
    pz = 0.5 + 0.4 * np.sin(zp / 6e-3 * 3.0)
    s = np.array([1.0, 0.8, 0.55, 0.3])
    table = {"zgrid": zp, "rgrid": r, "value": pz[:, None] * s[None, :]}
    pre, ren = H.prefactor_tables(z, r, o, "Ar", table, absorption_tables="NIST", dispersion_tables="NIST")
    assert ren is None and pre.shape == (9, 5, 4)
    for j in range(4):
        a, ra = H.prefactor_tables(z, r, o, "Ar", {"zgrid": zp, "value": pz * s[j]}, absorption_tables="NIST",
                                   dispersion_tables="NIST")
        want = a * ra[0][None, :]
        assert np.max(np.abs(pre[:, :, j] - want)) <= 1e-11 * np.max(np.abs(want)), j
    assert np.max(np.abs(ra[0] - 1.0)) > 1e-3  # the absorption term is not negligible here


def test_unlisted_f1_raises():
    with pytest.raises(ValueError):
        H.dispersion_function(OMEGA0 * 7, 1.0, "Ar_Henke")  # 10.8 eV: f1 = -9999 in the Henke table
