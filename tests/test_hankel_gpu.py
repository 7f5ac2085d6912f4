"""GPU parity of the Hankel far-field stage (libmmahankel.so through mma_amd.hankel) against
the FP64 oracle (oracle/hankel.py) and closed forms."""
import math

import numpy as np
import pytest
import torch
from scipy import special

from conftest import rel_err
from oracle import hankel as OH

pytestmark = pytest.mark.gpu
OMEGA0 = 2.0 * math.pi * OH.C_LIGHT / 800e-9


def _l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_bessel_j0_fp64():
    """The device J0 (Chebyshev below 25, Hankel asymptotics above) against scipy over the
    arguments of the far-field integral (0 .. 700)."""
    from mma_amd import hankel as H
    from ctypes import byref, c_double
    nq, h = 4096, 700.0 / 4087
    rho = torch.tensor([1.0, 0.37], dtype=torch.float64, device="cuda")
    out = torch.empty((2, nq), dtype=torch.float64, device="cuda")
    H._check(H._L().hk_grid_bessel(2, H._ptr(rho), nq, h, H.HK_F64, H._ptr(out), nq, H._stream()))
    q = (np.arange(nq) - 4) * h
    for i, r in enumerate((1.0, 0.37)):
        ref = special.jv(0, np.abs(r * q))
        err = np.max(np.abs(out[i].cpu().numpy() - ref))
        print("J0 max abs error %.2e" % err, flush=True)
        assert err < 3e-15


@pytest.mark.parametrize("case", ["scalar", "z", "z_long", "zr", "vacuum"])
def test_prefactor_tables(case):
    """Pre-factor tables against the oracle for every pressure branch; "z" runs the walk of the
    pressure table with one thread per (frequency, plane), "z_long" (5000 table points, beyond
    the length that form takes) with one thread per frequency."""
    from mma_amd import hankel as H
    z = np.linspace(0.0, 5e-3, 17)
    r = np.linspace(0.0, 1e-4, 9)
    o = OMEGA0 * np.linspace(21, 61, 12)
    kw = dict(absorption_tables="Henke", dispersion_tables="NIST")
    if case == "scalar":
        args = ("Ar", 0.6)
    elif case in ("z", "z_long"):
        zp = np.linspace(-1e-3, 6e-3, 57 if case == "z" else 5000)
        args = ("Ar", {"zgrid": zp, "value": 0.3 + 0.7 * np.exp(-((zp - 2e-3) / 1.5e-3) ** 2)})
    elif case == "zr":
        zp = np.linspace(0.0, 5e-3, 23)
        rp = np.linspace(0.0, 1.2e-4, 7)
        args = ("Ar", {"zgrid": zp, "rgrid": rp, "value": 0.5 + np.outer(np.sin(zp * 600), np.cos(rp * 1e4)) * 0.3})
    else:
        args = ("vacuum", 1.0)
        kw = dict(include_absorption=False, include_dispersion=False)
    pre_f, ren_f = H.get_propagation_pre_factor_function(z, r, o, *args, effective_IR_refrective_index=1.0002, **kw)
    pre_o, ren_o = OH.prefactor_tables(z, r, o, *args, n_ir=1.0002, **{k.replace("effective_IR_refrective_index",
                                                                                   "n_ir"): v for k, v in kw.items()})
    for kz in (0, 7, 16):
        got = pre_f(kz)
        ref = pre_o[kz].T if case == "zr" else np.outer(np.ones(len(r)), pre_o[kz])
        e = rel_err(got, ref)
        print(case, kz, "pre %.2e" % e, flush=True)
        assert e < 1e-12
        if case != "zr":
            assert rel_err(ren_f(kz), ren_o[kz]) < 1e-12
    if case == "zr":
        assert ren_f is None


@pytest.mark.parametrize("case", ["scalar", "z", "zr"])
def test_deferred_prefactor_equals_immediate(case):
    """hk_prefactor_enqueue (tables queued, the table-error flag read later) gives the bits of
    hk_prefactor; a table error surfaces in check(), and Hankel_long (which defers) raises it
    before returning anything."""
    from mma_amd import hankel as H
    z = np.linspace(0.0, 5e-3, 17)
    r = np.linspace(0.0, 1e-4, 9)
    o = OMEGA0 * np.linspace(21, 61, 12)
    if case == "scalar":
        pressure = 0.6
    elif case == "z":
        zp = np.linspace(-1e-3, 6e-3, 57)
        pressure = {"zgrid": zp, "value": 0.3 + 0.7 * np.exp(-((zp - 2e-3) / 1.5e-3) ** 2)}
    else:
        zp = np.linspace(0.0, 5e-3, 23)
        rp = np.linspace(0.0, 1.2e-4, 7)
        pressure = {"zgrid": zp, "rgrid": rp, "value": 0.5 + np.outer(np.sin(zp * 600), np.cos(rp * 1e4)) * 0.3}
    args = (z, r, o, "Ar", pressure, "Henke", True, "NIST", True, 1.0002)
    now = H._PreFactor(*args)
    later = H._PreFactor(*args, defer=True)
    later.check()
    assert torch.equal(now.pre, later.pre)
    assert (now.renorm is None and later.renorm is None) or torch.equal(now.renorm, later.renorm)
    # 10.8 eV: Henke lists no f1 there
    bad = (z, r, OMEGA0 * np.array([7.0, 9.0]), "Ar", pressure, "Henke", True, "Henke", True, 1.0)
    with pytest.raises(ValueError):
        H._PreFactor(*bad)
    queued = H._PreFactor(*bad, defer=True)
    with pytest.raises(ValueError, match="f1-scattering coefficient not listed"):
        queued.check()
    if case == "scalar":
        S = np.ones((len(z), 2, len(r)), dtype=np.complex128)
        tgt = H.FSources_provider(np.append(z, z[-1] + 1e-4), r, OMEGA0 * np.array([7.0, 9.0]), FSource=S)
        with pytest.raises(ValueError, match="f1-scattering coefficient not listed"):
            H.Hankel_long(tgt, 1.0, np.linspace(0, 1e-3, 5), "Ar", pressure, "Henke", True, "Henke", True, 1.0,
                          contraction="fp32-mfma")


def test_prefactor_z_modulated_closed_form():
    """The device z-modulated pre-factor on the piecewise-linear pressure table with planes
    before, inside (either side of the kink) and past the table, against its closed form (the
    same case that pins the oracle's z branch)."""
    from mma_amd import hankel as H
    from test_hankel_oracle import RAMP_NIR, RAMP_O, RAMP_Z, RAMP_ZP, _ramp_closed_form, _ramp_p
    prof = {"zgrid": RAMP_ZP, "value": _ramp_p(RAMP_ZP)}
    r = np.zeros(2)
    pre_f, ren_f = H.get_propagation_pre_factor_function(RAMP_Z, r, RAMP_O, "Ar", prof, absorption_tables="Henke",
                                                         dispersion_tables="NIST",
                                                         effective_IR_refrective_index=RAMP_NIR)
    want_pre, want_ren = _ramp_closed_form()
    got_pre = np.stack([pre_f(k)[0] for k in range(len(RAMP_Z))])
    got_ren = np.stack([ren_f(k) for k in range(len(RAMP_Z))])
    e_pre = np.max(np.abs(got_pre - want_pre) / np.max(np.abs(want_pre), axis=0))
    e_ren = np.max(np.abs(got_ren - want_ren) / want_ren)
    print("device z-modulated pre-factor vs closed form: %.2e %.2e" % (e_pre, e_ren), flush=True)
    assert e_pre <= 1e-6 and e_ren <= 1e-6


def test_prefactor_errors():
    from mma_amd import hankel as H
    z = np.linspace(0, 1e-3, 3)
    r = np.linspace(0, 1e-4, 3)
    with pytest.raises(ValueError):  # 10.8 eV: Henke lists no f1 there
        H.get_propagation_pre_factor_function(z, r, OMEGA0 * np.array([7.0, 9.0]), "Ar", 1.0)
    with pytest.raises(KeyError):  # the reference fails to find 'vacuum_Henke' tables too
        H.get_propagation_pre_factor_function(z, r, OMEGA0 * np.array([21.0]), "vacuum", 1.0)
    with pytest.raises(NotImplementedError):
        H.get_propagation_pre_factor_function(z, r, OMEGA0 * np.array([21.0]), "Ar",
                                              {"rgrid": r, "value": np.ones(3)})


@pytest.mark.parametrize("uniform,near", [(True, True), (False, True), (True, False)])
def test_single_plane_transform(uniform, near):
    from mma_amd import hankel as H
    rng = np.random.default_rng(3)
    nr = 200
    r = np.linspace(0, 1.5e-4, nr) if uniform else np.sort(np.concatenate([[0.0], rng.uniform(0, 1.5e-4, nr - 1)]))
    o = OMEGA0 * np.linspace(11, 61, 9)
    F = (rng.standard_normal((9, nr)) + 1j * rng.standard_normal((9, nr))) * np.exp(-(r / 6e-5) ** 2)
    rho = np.linspace(0, 7e-3, 301)
    pre = np.exp(1j * rng.uniform(0, 6, (nr, 9)))
    got = H.HankelTransform(o, r, F, 0.98, rho, near_field_factor=near, pre_factor=pre)
    ref = OH.hankel_plane(o, r, F, 0.98, rho, near_field_factor=near, pre=pre)
    e = rel_err(got, ref)
    print("single plane uniform=%s near=%s: %.2e" % (uniform, near, e), flush=True)
    assert got.shape == (9, 301) and e < 1e-12


def test_gaussian_pair_closed_form():
    from mma_amd import hankel as H
    r = np.linspace(0.0, 2e-4, 8001)
    o = OMEGA0 * np.array([11.0, 31.0, 61.0])
    a = 1.0 / (3e-5) ** 2
    F = np.exp(-a * r ** 2)[None, :] * np.ones((3, 1))
    rho = np.linspace(0.0, 7e-3, 513)
    got = H.HankelTransform(o, r, F, 1.0, rho, near_field_factor=True)
    for i, om in enumerate(o):
        k = om / OH.C_LIGHT
        aa = a + 1j * k / 2.0
        ref = np.exp(-(k * rho) ** 2 / (4 * aa)) / (2 * aa)
        assert np.max(np.abs(got[i] - ref)) <= 2e-7 * np.max(np.abs(ref))


def _long_case(nz=9, no=24, nr=160, c64=False, seed=5):
    rng = np.random.default_rng(seed)
    zall = np.linspace(0.0, 5e-3, nz + 1)  # the provider drops the last plane
    r = np.linspace(0.0, 2e-4, nr)
    o = OMEGA0 * np.linspace(21, 61, no)
    w = 4e-5 * (1 + 0.3 * np.linspace(0, 1, nz + 1))[:, None, None]
    S = (np.exp(-(r[None, None, :] / w) ** 2) * np.exp(1j * rng.uniform(0, 6.28, (nz + 1, no, 1)))
         * (1.0 + 0.05 * rng.standard_normal((nz + 1, no, nr))))
    if c64:
        S = S.astype(np.complex64)
    zp = np.linspace(-5e-4, 5.5e-3, 41)
    pressure = {"zgrid": zp, "value": 0.2 + np.exp(-((zp - 2.5e-3) / 2e-3) ** 2)}
    rho = np.linspace(0, 7e-3, 96)
    return zall, r, o, S, pressure, rho


def _oracle_long(zall, r, o, S, pressure, rho, cumulative=False):
    zz, rr, oo, planes = OH.provider_planes(zall, r, o, S.astype(np.complex128))
    return OH.hankel_long(zz, rr, oo, planes, 1.0, rho, "Ar", pressure, "Henke", True, "NIST", True, 1.0001,
                          cumulative=cumulative)


def test_hankel_long_fp64_matches_oracle():
    from mma_amd import hankel as H
    zall, r, o, S, pressure, rho = _long_case()
    tgt = H.FSources_provider(zall, r, o, FSource=S)
    res = H.Hankel_long(tgt, 1.0, rho, "Ar", pressure, "Henke", True, "NIST", True, 1.0001,
                        store_cumulative_result=True, store_non_normalised_cumulative_result=True)
    ref = _oracle_long(zall, r, o, S, pressure, rho, cumulative=True)
    for name in ("FF_integrated", "entry_plane_transform", "exit_plane_transform", "cumulative_field"):
        e = rel_err(getattr(res, name), ref[name])
        print(name, "%.2e" % e, flush=True)
        assert e < 1e-12, name
    assert res.FF_integrated.shape == (len(rho), len(o))
    assert res.cumulative_field.shape == (len(zall) - 2, len(rho), len(o))
    # the non-cumulative path (weighted plane sums in one launch) gives the same field
    tgt = H.FSources_provider(zall, r, o, FSource=S)
    res2 = H.Hankel_long(tgt, 1.0, rho, "Ar", pressure, "Henke", True, "NIST", True, 1.0001)
    assert rel_err(res2.FF_integrated, ref["FF_integrated"]) < 1e-12


@pytest.mark.parametrize("contraction,tol", [("fp64-grid", 1e-9), ("fp32-mfma", 1e-6), ("bf16x3-mfma", 1e-6),
                                             ("bf16x6-mfma", 1e-6), ("fp16x3-mfma", 1e-6)])
def test_hankel_long_gridded(contraction, tol):
    from mma_amd import hankel as H
    zall, r, o, S, pressure, rho = _long_case()
    ref = _oracle_long(zall, r, o, S, pressure, rho)
    tgt = H.FSources_provider(zall, r, o, FSource=S)
    res = H.Hankel_long(tgt, 1.0, rho, "Ar", pressure, "Henke", True, "NIST", True, 1.0001, contraction=contraction)
    for name in ("FF_integrated", "entry_plane_transform", "exit_plane_transform"):
        e, e2 = rel_err(getattr(res, name), ref[name]), _l2(getattr(res, name), ref[name])
        print(contraction, name, "max %.2e  l2 %.2e" % (e, e2), flush=True)
        assert e < tol, name


def test_float_spreads_at_a_large_near_field_phase():
    """The two operands whose spread runs in float (two bf16 parts, two scaled fp16 parts) with
    the screen at 5 cm, where the near-field phase c r^2 reaches ~190 rad: the phases anchor on
    an FP64 sincos (a float argument would be off by ulp(190) = 1.5e-5 rad up to the next
    anchor), so the errors against the FP64 direct contraction stay those of the operands:
    1e-6 for the fp16 parts (22 bits; measured 1.3e-7), and for the bf16 parts the 2^-17
    relative rounding of an operand element, 7.6e-6 (measured up to 3.6e-6 on one plane)."""
    from mma_amd import hankel as H
    zall, r, o, S, pressure, rho = _long_case(nz=24, no=16, nr=200)
    distance = 0.05
    rho = rho * distance
    ref = None
    for contraction, tol in (("fp64", None), ("fp16x3-mfma", 1e-6), ("bf16x3-mfma", 7.6e-6)):
        tgt = H.FSources_provider(zall, r, o, FSource=S)
        res = H.Hankel_long(tgt, distance, rho, "Ar", pressure, "Henke", True, "NIST", True, 1.0001,
                            contraction=contraction)
        if ref is None:
            ref = res
            continue
        for name in ("FF_integrated", "entry_plane_transform", "exit_plane_transform"):
            e = rel_err(getattr(res, name), getattr(ref, name))
            print("5 cm", contraction, name, "max %.2e" % e, flush=True)
            assert e < tol, (contraction, name, e)


def test_hankel_long_gridded_nonuniform_r_and_scalar_pressure():
    """The gridded contraction on a non-uniform radial grid (trapezoid weights of the given
    r, bins of q = K r / D at arbitrary r) with the scalar-pressure pre-factor, and with an odd
    plane count, against the FP64 oracle."""
    from mma_amd import hankel as H
    rng = np.random.default_rng(17)
    zall, r, o, S, pressure, rho = _long_case(nz=7, no=20, nr=150, seed=9)
    r = np.sort(np.concatenate([[0.0], rng.uniform(0, 2e-4, len(r) - 1)]))
    ref = _oracle_long(zall, r, o, S, 0.7, rho)
    tgt = H.FSources_provider(zall, r, o, FSource=S)
    for contraction, tol in (("fp64-grid", 1e-9), ("fp32-mfma", 1e-6)):
        res = H.Hankel_long(tgt, 1.0, rho, "Ar", 0.7, "Henke", True, "NIST", True, 1.0001, contraction=contraction)
        for name in ("FF_integrated", "entry_plane_transform", "exit_plane_transform"):
            e = rel_err(getattr(res, name), ref[name])
            print(contraction, name, "max %.2e" % e, flush=True)
            assert e < tol, (contraction, name)
        tgt = H.FSources_provider(zall, r, o, FSource=S)  # the provider is a single pass


def test_spread_forms_agree_with_the_direct_contraction():
    """Both forms of the spread against the direct FP64 contraction, on both outputs of a
    weighted plane sum (all planes; one plane): the register-window form (a thread per radius
    walking its planes), on uniform and non-uniform radii, and the per-item form the launcher
    falls back to when the table of per-plane factors does not fit the LDS beside the bins
    (3000 planes), each with and without the near-field phase."""
    from mma_amd import hankel as H
    rng = np.random.default_rng(5)
    no, nr = 6, 70
    o = OMEGA0 * np.linspace(21, 61, no)
    rho = np.linspace(0, 7e-3, 64)
    for nz, uniform in ((40, True), (40, False), (3000, True)):
        z = np.linspace(0.0, 5e-3, nz)
        r = np.linspace(0, 2e-4, nr) if uniform else np.sort(np.concatenate([[0.0], rng.uniform(0, 2e-4, nr - 1)]))
        amp = np.exp(-(r / 6e-5) ** 2)[None, None, :]
        src = torch.tensor(amp * np.exp(1j * rng.uniform(0, 6.28, (nz, no, nr))), device="cuda")
        dz = np.diff(z)
        w = np.zeros(nz)
        w[:-1] += 0.5 * dz
        w[1:] += 0.5 * dz
        W = np.stack([w, np.eye(nz)[nz // 2]])
        for near in (True, False):
            ref = H.direct(src, z, o, r, rho, 1.0, weights=W, near_field=near).cpu().numpy()
            got = H.gridded(src, z, o, r, rho, 1.0, W, near_field=near, kind="fp64-grid").cpu().numpy()
            for q in range(2):
                e = rel_err(got[q], ref[q])
                print("spread nz %d uniform %s near %s output %d: %.2e" % (nz, uniform, near, q, e), flush=True)
                assert e < 1e-9, (nz, uniform, near, q)


def test_gemm_kernels_match_fp64():
    """hk_gemm (FP32 and bf16x3 matrix-core paths) against an FP64 product of the same FP32
    operands, with ragged m and n (tile edges)."""
    from mma_amd import hankel as H
    rng = np.random.default_rng(7)
    for m, n, k in ((200, 330, 256), (1024, 600, 5120)):  # the far-field workload contracts K = 5120
        A = torch.tensor(rng.standard_normal((m, k)), dtype=torch.float32, device="cuda")
        B = torch.tensor(rng.standard_normal((n, k)), dtype=torch.float32, device="cuda")
        ref = (A.double() @ B.double().T).cpu().numpy()
        for kind, tol in ((H.HK_GEMM_F32, 1e-6), (H.HK_GEMM_BF16X3, 2e-5), (H.HK_GEMM_BF16X6, 1e-6)):
            out = torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda")
            H._check(H._L().hk_gemm(kind, m, n, k, H._ptr(A), k, H._ptr(B), k, H._ptr(out), n, H._stream()))
            e = rel_err(out.cpu().numpy(), ref)
            print("gemm %dx%dx%d kind %d: %.2e" % (m, n, k, kind, e), flush=True)
            assert e < tol


def _bf16_parts(x, n):
    """x (float64) as n bf16 parts, each the round-to-nearest-even bf16 of the float32 of what
    the earlier parts leave (the kernels' split), as 16-bit words."""
    parts, r = [], np.asarray(x, dtype=np.float64).copy()
    for _ in range(n):
        b = r.astype(np.float32).view(np.uint32).astype(np.uint64)
        w = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)
        parts.append(w)
        r = r - (w.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return parts


def test_split_gemm_matches_fp64():
    """hk_gemm_split on operands split on the host: the products of the kept part pairs (two
    parts: p0 q0 + p0 q1 + p1 q0; three: the six of order <= 2^-16) against their FP64 sum,
    with ragged m and n (tile edges), the far-field workload's K = 5120, and K = 64 and 192: one
    and three groups of 64 K terms, whose FP32 sums alternate between two register sets (an odd
    count ends on the first set)."""
    from mma_amd import hankel as H
    rng = np.random.default_rng(11)
    pairs = {2: [(0, 0), (0, 1), (1, 0)], 3: [(0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0)]}
    for m, n, k in ((200, 330, 256), (130, 140, 64), (129, 257, 192), (1024, 600, 5120)):
        A, B = rng.standard_normal((m, k)), rng.standard_normal((n, k))
        for parts in (2, 3):
            ap, bp = _bf16_parts(A, parts), _bf16_parts(B, parts)
            val = lambda w: (w.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
            ref = sum(val(ap[i]) @ val(bp[j]).T for i, j in pairs[parts])
            Ad = torch.tensor(np.concatenate(ap, axis=1).view(np.int16), device="cuda")
            Bd = torch.tensor(np.concatenate(bp, axis=1).view(np.int16), device="cuda")
            out = torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda")
            H._check(H._L().hk_gemm_split(parts, m, n, k, H._ptr(Ad), parts * k, H._ptr(Bd), parts * k, H._ptr(out), n,
                                          H._stream()))
            e = rel_err(out.cpu().numpy(), ref)
            print("split gemm %dx%dx%d parts %d: %.2e" % (m, n, k, parts, e), flush=True)
            assert e < 1e-6


def test_f16x2_gemm_matches_fp64():
    """hk_gemm_f16x2 on operands scaled and split on the host (each row by a power of two that puts
    its largest element in [2^13, 2^14), then v s = h0 + 2^-11 h1 in fp16, the scale after the
    parts): h0 h0' + 2^-11 (h0 h1' + h1 h0') over the two scales, against its FP64 sum, with
    ragged m and n, rows of very different magnitude and a zero row."""
    from mma_amd import hankel as H
    rng = np.random.default_rng(12)

    def split(X):
        mx = np.max(np.abs(X), axis=1)
        e = np.where(mx > 0, np.floor(np.log2(np.where(mx > 0, mx, 1.0))), 0.0)
        s = np.where(mx > 0, 2.0 ** (13 - e), 1.0)
        x = X * s[:, None]
        h0 = x.astype(np.float32).astype(np.float16)
        h1 = ((x - h0.astype(np.float64)) * 2048.0).astype(np.float32).astype(np.float16)
        k = X.shape[1]
        row = np.zeros((X.shape[0], 2 * k + 8), dtype=np.uint16)
        row[:, :k] = h0.view(np.uint16)
        row[:, k:2 * k] = h1.view(np.uint16)
        row[:, 2 * k:2 * k + 4] = s.astype(np.float64).view(np.uint16).reshape(-1, 4)
        return row, h0.astype(np.float64), h1.astype(np.float64), s

    for m, n, k in ((200, 330, 256), (1024, 600, 5120)):
        A = rng.standard_normal((m, k)) * 10.0 ** rng.uniform(-6, 6, (m, 1))
        B = rng.standard_normal((n, k)) * 10.0 ** rng.uniform(-6, 6, (n, 1))
        B[3] = 0.0
        ra, a0, a1, sa = split(A)
        rb, b0, b1, sb = split(B)
        ref = (a0 @ b0.T + (a0 @ b1.T + a1 @ b0.T) / 2048.0) / np.outer(sa, sb)
        Ad = torch.tensor(ra.view(np.int16), device="cuda")
        Bd = torch.tensor(rb.view(np.int16), device="cuda")
        out = torch.full((m, n), float("nan"), dtype=torch.float64, device="cuda")
        H._check(H._L().hk_gemm_f16x2(m, n, k, H._ptr(Ad), 2 * k + 8, H._ptr(Bd), 2 * k + 8, H._ptr(out), n,
                                      H._stream()))
        got = out.cpu().numpy()
        assert np.all(got[:, 3] == 0.0)
        # each column against its own scale (the columns' magnitudes span 12 decades)
        e = max(rel_err(got[:, j], ref[:, j]) for j in range(n) if j != 3)
        print("f16x2 gemm %dx%dx%d: %.2e" % (m, n, k, e), flush=True)
        assert e < 1e-6


def test_full_size_far_field_error_bound():
    """The far-field workload at full size (256 planes x 2048 frequencies x 512 radii -> 1024
    screen radii at 1 m, Ar cell with a z-modulated pressure, complex128 and complex64 sources)
    through Hankel_long: every gridded contraction against the FP64 direct one ('fp64', pinned to
    the oracle on the small cases above).  FP32 on the matrix cores and the six-product bf16 split
    stay within 1e-6 (max |d| / max |ref|) on the integrated field and the entry and exit planes.
    The three-product bf16 split keeps 16 bits of each operand (its dropped lo*lo term is 2^-16
    of a product): measured 1.13e-6 max and 5.8e-7 Frobenius on the integrated field, whose sum
    over the planes cancels more than a single plane's; it is held to 2e-6 max and 1e-6
    Frobenius."""
    from mma_amd import hankel as H, workloads as W
    zall, r, o, rho, src, distance = W.hankel_workload()
    pressure = W.hankel_pressure()

    def run(contraction, sources):
        tgt = H.FSources_provider(zall, r, o, FSource=sources)
        return H.Hankel_long(tgt, distance, rho, "Ar", pressure, "Henke", True, "NIST", True, 1.0,
                             store_entry_and_exit_plane_transform=True, contraction=contraction)

    ref = run("fp64", src)
    names = ("FF_integrated", "entry_plane_transform", "exit_plane_transform")
    # (contraction, sources, bound on max |d| / max |ref|, bound on the Frobenius ratio)
    cases = [("fp64-grid", src, 1e-9, 1e-9), ("fp32-mfma", src, 1e-6, 1e-6), ("bf16x6-mfma", src, 1e-6, 1e-6),
             ("fp16x3-mfma", src, 1e-6, 1e-6), ("bf16x3-mfma", src, 2e-6, 1e-6)]
    src64 = src.to(torch.complex64)
    cases += [("fp32-mfma", src64, 1e-6, 1e-6), ("bf16x6-mfma", src64, 1e-6, 1e-6), ("fp16x3-mfma", src64, 1e-6, 1e-6),
              ("bf16x3-mfma", src64, 2e-6, 1e-6)]
    for contraction, sources, tol, tol_f in cases:
        res = run(contraction, sources)
        for name in names:
            got, want = getattr(res, name), getattr(ref, name)
            e, e2 = rel_err(got, want), _l2(got, want)
            print("full size %s %s %s: max %.2e frobenius %.2e" % (contraction, sources.dtype, name, e, e2),
                  flush=True)
            assert e <= tol and e2 <= tol_f, (contraction, sources.dtype, name, e, e2)
        del res


def test_complex64_sources_and_provider_selection():
    from mma_amd import hankel as H
    zall, r, o, S, pressure, rho = _long_case(nz=6, no=12, nr=96, c64=True)
    tgt = H.FSources_provider(zall, r, o, FSource=S, ko_min=2, ko_step=3, kr_step=2, kz_step=2)
    res = H.Hankel_long(tgt, 1.0, rho, "Ar", pressure, "Henke", True, "NIST", True)
    zz, rr, oo, planes = OH.provider_planes(zall, r, o, S.astype(np.complex128), ko_min=2, ko_step=3, kr_step=2,
                                            kz_step=2)
    ref = OH.hankel_long(zz, rr, oo, planes, 1.0, rho, "Ar", pressure, "Henke", True, "NIST", True, 1.0)
    assert rel_err(res.FF_integrated, ref["FF_integrated"]) < 1e-12
