"""The Hankel oracle (oracle/hankel.py) against what mathematics and physics fix: the Gaussian
Hankel pair (with and without the near-field phase), rho = 0, linearity, the longitudinal
trapezoid over two planes, the provider's plane selection, the scalar pre-factor's closed form
and renormalisation invariant, the z-modulated branch reducing to the scalar branch for a
constant pressure, the z-modulated branch against its closed form on a piecewise-linear
pressure table (with planes outside the table) and against three misreadings, the cumulative
renormalisation against the impulse response of the longitudinal trapezoid, photon
conversions, and the high-energy limit of the XUV index (plasma dispersion, built from e, eps0
and m instead of the classical electron radius).  Also the product's input checks that need no
device: ascending grids for a tabulated pressure, trapezoidal radial integrators only."""
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


# ---------------------------------------------------------------------------------------------
# z-modulated pre-factor (Hankel_transform.py:146-239) against its closed form on a piecewise
# linear pressure profile, and the cumulative renormalisation (:663-667) against the impulse
# response of the longitudinal integral.

EPS0 = 1.0 / (4e-7 * math.pi * H.C_LIGHT ** 2)
RAMP_ZP = np.linspace(-1e-3, 6e-3, 29)  # pressure table: 0.25 mm nodes starting below z = 0
RAMP_KINK = RAMP_ZP[16]  # 3 mm, a table node that is not a plane
# planes: one before the table, table nodes on both sides of the kink (not the kink itself),
# the table's last node and two planes past it (edge fill)
RAMP_Z = np.array([-1.5e-3, RAMP_ZP[0], RAMP_ZP[3], RAMP_ZP[7], RAMP_ZP[12], RAMP_ZP[19], RAMP_ZP[28], 6.4e-3,
                   7e-3])
RAMP_O = OMEGA0 * np.array([21.0, 33.0, 45.0, 61.0])
RAMP_NIR = 1.0003
RAMP_SCALE = 0.25


def _ramp_p(z):
    """Pressure S (0.02 + 12 (z - zp0)) up to the kink, then S (0.068 - 5 (z - z_kink))."""
    z = np.asarray(z, dtype=np.float64)
    up = 0.02 + 12.0 * (z - RAMP_ZP[0])
    pk = 0.02 + 12.0 * (RAMP_KINK - RAMP_ZP[0])
    return RAMP_SCALE * np.where(z <= RAMP_KINK, up, pk - 5.0 * (z - RAMP_KINK))


def _ramp_P(z):
    """Exact integral of the piecewise-linear pressure from the table's first node to z."""
    z = np.asarray(z, dtype=np.float64)
    s = z - RAMP_ZP[0]
    up = 0.02 * s + 6.0 * s ** 2
    sk = RAMP_KINK - RAMP_ZP[0]
    pk = 0.02 + 12.0 * sk
    d = z - RAMP_KINK
    down = 0.02 * sk + 6.0 * sk ** 2 + pk * d - 2.5 * d ** 2
    return RAMP_SCALE * np.where(z <= RAMP_KINK, up, down)


def _plasma_coef(omega, f):
    """N_ref f e^2 / (2 eps0 m omega^2): delta of the XUV index per unit pressure (f = f1) or beta
    (f = f2), written with e, eps0 and m instead of the classical electron radius and lambda."""
    return H.N_REF * f * H.ELCHARGE ** 2 / (2.0 * EPS0 * H.ELMASS * omega ** 2)


def _ramp_closed_form(dgas="Ar_NIST", agas="Ar_Henke", lost_offset=False, local_beta_p=False,
                      integrate_on_planes=False, drop_f1=False):
    """pre(z, w) = p(z) exp(w [i((n_IR - 1)(z - zp0) + A P(z)) + B (P(z) - P(z_last))] / c) and
    renorm(z, w) = exp(w B (P(z_last) - P(z)) / c), with P = int_{zp0}^{z} p, z clamped to the
    table (edge fill), A and B the plasma coefficients of f1 and f2.  The keyword switches give
    the misreadings this closed form has to tell apart from the reference."""
    zc = np.clip(RAMP_Z, RAMP_ZP[0], RAMP_ZP[-1])
    P = _ramp_P(zc)
    if integrate_on_planes:  # pressure interpolated onto the planes, then integrated over them
        pz = _ramp_p(zc)
        P = np.concatenate([[0.0], np.cumsum(0.5 * np.diff(zc) * (pz[1:] + pz[:-1]))])
    pre = np.empty((len(RAMP_Z), len(RAMP_O)), dtype=np.complex128)
    ren = np.empty((len(RAMP_Z), len(RAMP_O)))
    for i, w in enumerate(RAMP_O):
        A = _plasma_coef(w, float(H.getf1(dgas, H.omega_si_to_ev(w))))
        B = _plasma_coef(w, float(H.getf2(agas, H.omega_si_to_ev(w))))
        offset = zc if lost_offset else zc - RAMP_ZP[0]
        phase = (RAMP_NIR - 1.0) * offset + (0.0 if drop_f1 else A) * P
        absorb = B * (_ramp_p(zc) * (zc - RAMP_ZP[0]) if local_beta_p else P)
        pre[:, i] = _ramp_p(zc) * np.exp(w * (1j * phase + absorb - absorb[-1]) / H.C_LIGHT)
        ren[:, i] = np.exp(w * (absorb[-1] - absorb) / H.C_LIGHT)
    return pre, ren


def test_z_modulated_prefactor_closed_form():
    """The z-modulated branch on a piecewise-linear pressure table, planes on table nodes and
    outside the table: cumulative_trapezoid and linear interp1d are exact there, so the branch
    must equal the closed form of _ramp_closed_form.  Phases reach ~400 rad (n_IR) and a few
    rad (f1), absorption up to ~2.5 nepers; errors are taken per frequency, relative to that
    frequency's largest value; the tolerance covers the 7e-8 difference between the two unit
    sets' electron radius."""
    prof = {"zgrid": RAMP_ZP, "value": _ramp_p(RAMP_ZP)}
    pre, ren = H.prefactor_tables(RAMP_Z, np.zeros(2), RAMP_O, "Ar", prof, absorption_tables="Henke",
                                  dispersion_tables="NIST", n_ir=RAMP_NIR)
    want_pre, want_ren = _ramp_closed_form()
    e_pre = np.max(np.abs(pre - want_pre) / np.max(np.abs(want_pre), axis=0))
    e_ren = np.max(np.abs(ren - want_ren) / want_ren)
    print("z-modulated closed form: pre %.2e renorm %.2e" % (e_pre, e_ren))
    assert e_pre <= 1e-6 and e_ren <= 1e-6, (e_pre, e_ren)
    # the case is not degenerate: absorption and the f1 phase term matter at every frequency
    assert np.all(np.max(want_ren, axis=0) > 1.08)
    no_f1 = _ramp_closed_form(dgas="Ar_NIST", drop_f1=True)[0]
    assert np.all(np.max(np.abs(no_f1 - want_pre), axis=0) > 0.5 * np.max(np.abs(want_pre), axis=0))


@pytest.mark.parametrize("mistake", ["lost_offset", "local_beta_p", "integrate_on_planes"])
def test_z_modulated_closed_form_rejects_misreadings(mistake):
    """Each misreading of the z-modulated branch (zp[0] offset of the n_IR term dropped; the
    local pressure times the path length for the pressure integral in the absorption; the
    pressure integrated over the planes after interpolation, which misses the kink between
    them) moves the tables by far more than the closed-form test's tolerance."""
    good_pre, good_ren = _ramp_closed_form()
    bad_pre, bad_ren = _ramp_closed_form(**{mistake: True})
    e = max(np.max(np.abs(bad_pre - good_pre) / np.max(np.abs(good_pre), axis=0)),
            np.max(np.abs(bad_ren - good_ren) / good_ren))
    assert e > 1e-2, (mistake, e)


def _impulse_case(j, nz=5, zstep=1e-4):
    """Planes on a 2-point radial grid [0, R] with the source zero except at plane j, screen
    radius 0 and no near-field phase: plane k's transform is 1/2 R^2 F_k(w, R) pre_k(w)."""
    z = np.array([0.0, 1.0, 2.5, 3.0, 4.5])[:nz] * zstep
    r = np.array([0.0, 2e-5])
    o = OMEGA0 * np.array([21.0, 29.0])
    planes = [np.zeros((2, 2), dtype=np.complex128) for _ in range(nz)]
    planes[j][:, 1] = 2.0 / r[1] ** 2 * np.array([1.0, 0.5 - 0.25j])
    return z, r, o, planes


@pytest.mark.parametrize("j", [0, 2, 4])
def test_cumulative_renormalisation_impulse_response(j):
    """Scalar pressure, absorption only: pre_k = p exp(a (z_k - z_last)) and renorm_k =
    exp(a (z_last - z_k)) with a = w p B / c.  With a source only in plane j, the longitudinal
    trapezoid through plane k+1 holds plane j with the weight W_k = dz_{j-1}/2 [k + 1 >= j] +
    dz_j/2 [k >= j], so FF after plane k+1 = W_k p c_j exp(a (z_j - z_last)), and the reference
    stores cumulative_field[k] = renorm_k times that sum (renorm of plane k, applied after plane
    k+1 was added, :663-667): cumulative_field[k] = W_k p c_j exp(a (z_j - z_k)).  Applying the
    renorm of plane k+1 instead, or none, fails."""
    z, r, o, planes = _impulse_case(j)
    p = 0.4
    res = H.hankel_long(z, r, o, planes, 1.0, np.array([0.0]), "Ar", p, "Henke", True, "Henke", False, 1.0,
                        near_field_factor=False, cumulative=True)
    cum = res["cumulative_field"][:, 0, :]  # (nz - 1, No)
    c = np.array([1.0, 0.5 - 0.25j])
    dz = np.diff(z)
    a = o * p * _plasma_coef(o, H.getf2("Ar_Henke", H.omega_si_to_ev(o))) / H.C_LIGHT
    want = np.zeros_like(cum)
    for k in range(len(z) - 1):
        wk = (0.5 * dz[j - 1] if 0 < j <= k + 1 else 0.0) + (0.5 * dz[j] if j <= k else 0.0)
        want[k] = wk * p * c * np.exp(a * (z[j] - z[k]))
    assert np.max(np.abs(cum - want)) <= 1e-6 * np.max(np.abs(want))
    wj = 0.5 * ((dz[j - 1] if j > 0 else 0.0) + (dz[j] if j < len(dz) else 0.0))
    ff_want = wj * p * c * np.exp(a * (z[j] - z[-1]))
    assert np.max(np.abs(res["FF_integrated"][0] - ff_want)) <= 1e-6 * np.max(np.abs(ff_want))
    # the absorption over every plane spacing is large enough to tell the orderings apart
    assert np.min(np.exp(a[:, None] * dz[None, :])) > 1.02


def test_tabulated_pressure_needs_ascending_grids():
    """The product's pre-factor walks the plane grid and the pressure table in one forward
    pass: descending or repeated grids are refused before anything reaches the device."""
    from mma_amd import hankel as P
    prof = {"zgrid": RAMP_ZP, "value": _ramp_p(RAMP_ZP)}
    r = np.zeros(2)
    with pytest.raises(ValueError):
        P.get_propagation_pre_factor_function(RAMP_Z[::-1], r, RAMP_O, "Ar", prof)
    with pytest.raises(ValueError):
        P.get_propagation_pre_factor_function(RAMP_Z, r, RAMP_O, "Ar", {"zgrid": RAMP_ZP[::-1],
                                                                         "value": _ramp_p(RAMP_ZP)})
    zp = RAMP_ZP.copy()
    zp[5] = zp[4]
    with pytest.raises(ValueError):
        P.get_propagation_pre_factor_function(RAMP_Z, r, RAMP_O, "Ar", {"zgrid": zp, "value": _ramp_p(RAMP_ZP)})
    with pytest.raises(ValueError):
        P.get_propagation_pre_factor_function(RAMP_Z, r, RAMP_O, "Ar", {"zgrid": RAMP_ZP, "rgrid": np.array([1.0, 0.5]),
                                                                         "value": np.ones((len(RAMP_ZP), 2))})


def test_only_trapezoidal_integrators_are_accepted():
    """The product accelerates the trapezoid over the given radii only: numpy's and scipy's
    trapezoid and any callable that integrates a probe like the trapezoid (the reference's own
    wrapper) pass; a callable with a trapezoid's name but other semantics is refused."""
    from scipy import integrate
    from mma_amd import hankel as P

    def trapezoidal_integrator(y, x):  # the reference's wrapper (Hankel_transform.py:22-27)
        return integrate.trapezoid(y, x=x)

    def trapz(y, x):  # same name as numpy's, extra weighting
        return integrate.trapezoid(y * x, x=x)

    for f in (P.trapezoidal_integrator, np.trapezoid, integrate.trapezoid, trapezoidal_integrator):
        assert P._is_trapezoid(f)
    for f in (trapz, integrate.simpson, lambda y, x: 0.0):
        assert not P._is_trapezoid(f)
