"""Gradient of dipole losses with respect to the driving field.

CPU: the oracle gradient (oracle/adjoint.py, banded LAPACK solves on the reference's own CN
matrices) is pinned by central finite differences (h = 1e-6) of the C oracle's dipole, for a
random weighting of the dipole samples and for the windowed harmonic yield; the torch gate and
harmonic yield agree with their numpy restatements.
GPU: TDSEDipole (forward kernel + adjoint kernels through the C ABI) against the oracle
gradient, against finite differences of its own forward, and independent of the checkpoint
interval."""
import math

import numpy as np
import pytest
import torch

from conftest import rel_err

A_SC = 1.1893
DT = 0.25


def _field(n, amp=0.06, w=0.228, phase=0.3, seed=None):
    t = np.arange(n + 2) * DT
    E = amp * np.sin(w * t + phase) * np.sin(np.pi * t / t[-1]) ** 2
    if seed is not None:
        E = E * (1.0 + 0.1 * np.random.default_rng(seed).standard_normal(E.shape))
    return E


@pytest.fixture(scope="module")
def small_grid():
    from oracle import ctdse as O
    num_r, dx = 600, 0.2
    x, psi0, e, it, st = O.ground_state(num_r, dx, A_SC, -0.5792, 1e-25)
    assert st == 0
    return num_r, dx, x, psi0


def test_oracle_gradient_matches_finite_differences(small_grid):
    from oracle import adjoint as OA
    num_r, dx, x, psi0 = small_grid
    n = 160
    F, kg = OA.gate(_field(n))
    rng = np.random.default_rng(1)
    gobs = rng.standard_normal(n)
    g = OA.gradient(num_r, dx, A_SC, 2.0, DT, psi0, x, F, n, gobs)
    assert g[-1] == 0.0
    h = 1e-6
    for _ in range(3):
        v = rng.standard_normal(n + 2)
        v[:kg + 1] = 0.0
        lp = gobs @ OA.dipole(num_r, dx, A_SC, 2.0, DT, psi0, x, F + h * v, n)
        lm = gobs @ OA.dipole(num_r, dx, A_SC, 2.0, DT, psi0, x, F - h * v, n)
        fd = (lp - lm) / (2 * h)
        ad = g @ v[:n]
        assert abs(fd - ad) <= 1e-7 * abs(ad), (fd, ad)


def test_oracle_harmonic_yield_gradient(small_grid):
    from oracle import adjoint as OA
    num_r, dx, x, psi0 = small_grid
    n = 160
    F, kg = OA.gate(_field(n))
    win = np.zeros(n // 2 + 1)
    win[5:30] = 1.0
    s = OA.dipole(num_r, dx, A_SC, 2.0, DT, psi0, x, F, n)
    gs = OA.harmonic_yield_grad(s, DT, win)
    # dL/ds itself against finite differences of the yield
    e = np.zeros(n)
    for j in (3, 40, 111):
        e[:] = 0.0
        e[j] = 1e-4
        fd = (OA.harmonic_yield(s + e, DT, win) - OA.harmonic_yield(s - e, DT, win)) / 2e-4
        assert abs(fd - gs[j]) <= 1e-7 * np.max(np.abs(gs))
    g = OA.gradient(num_r, dx, A_SC, 2.0, DT, psi0, x, F, n, gs)
    v = np.random.default_rng(2).standard_normal(n + 2)
    v[:kg + 1] = 0.0
    h = 1e-6
    yp = OA.harmonic_yield(OA.dipole(num_r, dx, A_SC, 2.0, DT, psi0, x, F + h * v, n), DT, win)
    ym = OA.harmonic_yield(OA.dipole(num_r, dx, A_SC, 2.0, DT, psi0, x, F - h * v, n), DT, win)
    fd = (yp - ym) / (2 * h)
    assert abs(fd - g @ v[:n]) <= 1e-6 * abs(fd)


def test_torch_gate_and_yield_match_numpy():
    from mma_amd import adjoint as AD
    from oracle import adjoint as OA
    n = 120
    E = _field(n, phase=0.0)
    E[:7] = 1e-3  # no sign change until the pulse's first crossing
    Ft, mask = AD.gate(torch.tensor(E)[None, :])
    Fn, kg = OA.gate(E)
    assert np.array_equal(Ft[0].numpy(), Fn)
    assert mask[0, :kg + 1].sum() == 0 and mask[0, kg + 1:].min() == 1
    s = np.random.default_rng(3).standard_normal(n)
    win = AD.harmonic_window(n, DT, n * DT, 0.057, 3, 11)
    y = AD.harmonic_yield(torch.tensor(s)[None, :], DT, win)
    assert abs(float(y[0]) - OA.harmonic_yield(s, DT, win.numpy().astype(float))) <= 1e-12 * abs(float(y[0]))
    st = torch.tensor(s, requires_grad=True)
    AD.harmonic_yield(st, DT, win).backward()
    assert rel_err(st.grad.numpy(), OA.harmonic_yield_grad(s, DT, win.numpy().astype(float))) < 1e-12


# --------------------------------------------------------------------------- GPU

GPU_NUM_R, GPU_DX, GPU_N = 2000, 0.2, 400


@pytest.fixture(scope="module")
def gpu_case():
    from mma_amd import tdse
    geo = tdse.Geometry(GPU_NUM_R, GPU_DX, DT, A_SC, -0.5792, 1e-25, 2.0)
    x, psi0 = geo.ground_state()
    fields = np.stack([_field(GPU_N, amp=0.05, seed=1), _field(GPU_N, amp=0.07, phase=1.1, seed=2),
                       _field(GPU_N, amp=0.04, w=0.3, seed=3)])
    return geo, x, psi0, fields


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle(gpu_case):
    from mma_amd import adjoint as AD
    from oracle import adjoint as OA
    geo, x, psi0, fields = gpu_case
    print("adjoint vs oracle: num_r %d dx %g N %d inst %d" % (GPU_NUM_R, GPU_DX, GPU_N, len(fields)), flush=True)
    E = torch.tensor(fields, device="cuda", requires_grad=True)
    s = AD.TDSEDipole.apply(E, geo, GPU_N)
    gobs = np.random.default_rng(5).standard_normal((len(fields), GPU_N))
    (s * torch.tensor(gobs, device="cuda")).sum().backward()
    got = E.grad.cpu().numpy()
    s_np = s.detach().cpu().numpy()
    for i in range(len(fields)):
        F, kg = OA.gate(fields[i])
        ref_s = OA.dipole(GPU_NUM_R, GPU_DX, A_SC, 2.0, DT, psi0, x, F, GPU_N)
        ref_g = OA.gradient(GPU_NUM_R, GPU_DX, A_SC, 2.0, DT, psi0, x, F, GPU_N, gobs[i])
        ref_g[:kg + 1] = 0.0  # dL/dE is zero where the gate holds the field at zero
        es, eg = rel_err(s_np[i], ref_s), rel_err(got[i, :GPU_N], ref_g)
        print(i, "dipole %.2e grad %.2e" % (es, eg), flush=True)
        assert es < 1e-8 and eg < 1e-8
        assert np.all(got[i, :kg + 1] == 0.0) and np.all(got[i, GPU_N - 1:] == 0.0)


@pytest.mark.gpu
def test_gpu_gradient_odd_grid_matches_oracle():
    """The adjoint on a grid whose point count fills neither the last 8-point chunk nor the
    last 64-chunk tile (num_r = 2999, Nx = 3000), against the oracle gradient."""
    from mma_amd import adjoint as AD, tdse, workloads as W
    from oracle import adjoint as OA
    num_r, dx, n = 2999, 0.1, 300
    geo = tdse.Geometry(num_r, dx, DT, A_SC, -0.5792, W.UNEVEN_CV, 2.0)
    x, psi0 = geo.ground_state()
    fields = np.stack([_field(n, amp=0.06, seed=7), _field(n, amp=0.05, phase=0.4, seed=8)])
    E = torch.tensor(fields, device="cuda", requires_grad=True)
    s = AD.TDSEDipole.apply(E, geo, n)
    gobs = np.random.default_rng(9).standard_normal((len(fields), n))
    (s * torch.tensor(gobs, device="cuda")).sum().backward()
    got = E.grad.cpu().numpy()
    s_np = s.detach().cpu().numpy()
    for i in range(len(fields)):
        F, kg = OA.gate(fields[i])
        ref_s = OA.dipole(num_r, dx, A_SC, 2.0, DT, psi0, x, F, n)
        ref_g = OA.gradient(num_r, dx, A_SC, 2.0, DT, psi0, x, F, n, gobs[i])
        ref_g[:kg + 1] = 0.0
        es, eg = rel_err(s_np[i], ref_s), rel_err(got[i, :n], ref_g)
        print(i, "odd grid: dipole %.2e grad %.2e" % (es, eg), flush=True)
        assert es < 1e-8 and eg < 1e-8


@pytest.mark.gpu
def test_gpu_gradient_finite_differences(gpu_case):
    from mma_amd import adjoint as AD
    geo, x, psi0, fields = gpu_case
    E0 = torch.tensor(fields[:2], device="cuda")
    win = AD.harmonic_window(GPU_N, DT, GPU_N * DT, 0.057, 3, 15).cuda()

    def loss(E):
        return AD.harmonic_yield(AD.TDSEDipole.apply(E, geo, GPU_N), DT, win).sum()

    E = E0.clone().requires_grad_(True)
    loss(E).backward()
    g = E.grad.clone()
    _, mask = AD.gate(E0)
    rng = np.random.default_rng(7)
    h = 1e-6
    for trial in range(3):
        v = torch.tensor(rng.standard_normal(E0.shape), device="cuda") * mask
        with torch.no_grad():
            fd = (loss(E0 + h * v) - loss(E0 - h * v)).item() / (2 * h)
        ad = float((g * v).sum())
        print(trial, "fd %.12e adjoint %.12e" % (fd, ad), flush=True)
        assert abs(fd - ad) <= 1e-6 * abs(fd)


@pytest.mark.gpu
def test_gpu_gradient_independent_of_checkpoint_interval(gpu_case):
    """Stored states: the same arithmetic for every checkpoint interval K (bitwise).  Rebuilt
    states (inverse steps from the nearest later checkpoint, up to K = N steps back): equal to
    the stored-state gradient to rounding."""
    from mma_amd import adjoint as AD
    geo, x, psi0, fields = gpu_case
    gobs = torch.tensor(np.random.default_rng(9).standard_normal((len(fields), GPU_N)), device="cuda")

    def grad_of(K, mode):
        E = torch.tensor(fields, device="cuda", requires_grad=True)
        (AD.TDSEDipole.apply(E, geo, GPU_N, K, mode) * gobs).sum().backward()
        return E.grad.cpu().numpy()

    Ks = (1, 7, 20, GPU_N - 1, GPU_N, GPU_N + 3)  # K > N: no checkpoint at all
    stored = [grad_of(K, "stored") for K in Ks]
    for g in stored[1:]:
        assert np.array_equal(g, stored[0])
    for K in Ks:
        g = grad_of(K, "reconstruct")
        e = rel_err(g, stored[0])
        print("K %d rebuilt vs stored %.2e" % (K, e), flush=True)
        assert e < 1e-11


# ------------------------------------------------- gradient workload (Nx = 8001, dt = 0.1)

def _workload_geometry():
    from mma_amd import tdse, workloads as W
    g = W.ADJOINT_GRID
    return tdse.Geometry(g["num_r"], g["dx"], g["dt"], g["a"], g["e_guess"], g["cv"], g["x_int"])


@pytest.mark.gpu
def test_gpu_gradient_workload_vs_oracle():
    """Two members of the gradient workload on its grid, on a 4000-sample window, against
    the oracle gradient (harmonic-yield loss over harmonics 21-31)."""
    from mma_amd import adjoint as AD, tdse, workloads as W
    from oracle import adjoint as OA
    g = W.ADJOINT_GRID
    geo = _workload_geometry()
    x, psi0 = geo.ground_state()
    t, f = W.adjoint_batch(n_inst=8, n_t=4000, n_cycles=3)
    f = tdse.pad_fields(f[[0, 5]])
    n = f.shape[1] - 1
    print("workload vs oracle: Nx %d dt %g steps %d" % (g["num_r"] + 1, g["dt"], n), flush=True)
    win = AD.harmonic_window(n, g["dt"], n * g["dt"], W.omega_au(800e-9), 21, 31)
    E = torch.tensor(f, device="cuda", requires_grad=True)
    s = AD.TDSEDipole.apply(E, geo, n)
    AD.harmonic_yield(s, g["dt"], win.cuda()).sum().backward()
    got = E.grad.cpu().numpy()
    for i in range(2):
        F, kg = OA.gate(f[i])
        ref_s = OA.dipole(g["num_r"], g["dx"], g["a"], g["x_int"], g["dt"], psi0, x, F, n)
        gs = OA.harmonic_yield_grad(ref_s, g["dt"], win.numpy().astype(float))
        ref_g = OA.gradient(g["num_r"], g["dx"], g["a"], g["x_int"], g["dt"], psi0, x, F, n, gs)
        ref_g[:kg + 1] = 0.0
        es = rel_err(s[i].detach().cpu().numpy(), ref_s)
        eg = rel_err(got[i, :n], ref_g)
        print(i, "dipole %.2e grad %.2e" % (es, eg), flush=True)
        assert es < 1e-8 and eg < 1e-8


@pytest.mark.gpu
def test_gpu_gradient_workload_finite_differences_64_directions():
    """Every 16th member of the 1024-instance gradient workload (40000 field samples, Nx = 8001,
    dt = 0.1), one seeded random unit direction each (zero up to the gating point): central
    differences (h = 1e-6) of the harmonic yield over harmonics 21-31 against the adjoint
    directional derivative.  The error is measured against |grad| (|direction| = 1), so a
    direction nearly orthogonal to the gradient is not measured against its own tiny derivative.
    Members 0 and 20 are also compared with the oracle gradient over the full length."""
    from mma_amd import adjoint as AD, tdse, workloads as W
    from oracle import adjoint as OA
    g = W.ADJOINT_GRID
    geo = _workload_geometry()
    t, f = W.adjoint_batch(n_inst=1024)
    f = tdse.pad_fields(f[::16])
    n = f.shape[1] - 1
    win = AD.harmonic_window(n, g["dt"], n * g["dt"], W.omega_au(800e-9), 21, 31).cuda()
    print("64 directions: Nx %d steps %d" % (g["num_r"] + 1, n), flush=True)

    def losses(E):
        return AD.harmonic_yield(AD.TDSEDipole.apply(E, geo, n), g["dt"], win)

    E0 = torch.tensor(f, device="cuda")
    E = E0.clone().requires_grad_(True)
    losses(E).sum().backward()
    grad = E.grad.clone()
    print("gradient done", flush=True)
    _, mask = AD.gate(E0)
    V = torch.tensor(np.random.default_rng(11).standard_normal(f.shape), device="cuda") * mask
    V = V / V.norm(dim=1, keepdim=True)
    h = 1e-6
    with torch.no_grad():
        fd = ((losses(E0 + h * V) - losses(E0 - h * V)) / (2 * h)).cpu().numpy()
    ad = (grad * V).sum(dim=1).cpu().numpy()
    err = np.abs(fd - ad) / grad.norm(dim=1).cpu().numpy()
    print("max |fd - ad| / |grad| = %.3e, median %.3e" % (err.max(), np.median(err)), flush=True)
    assert np.all(err <= 1e-6)
    x, psi0 = geo.ground_state()
    got = grad.cpu().numpy()
    for i in (0, 20):
        F, kg = OA.gate(f[i])
        ref_s = OA.dipole(g["num_r"], g["dx"], g["a"], g["x_int"], g["dt"], psi0, x, F, n)
        gs = OA.harmonic_yield_grad(ref_s, g["dt"], win.cpu().numpy().astype(float))
        ref_g = OA.gradient(g["num_r"], g["dx"], g["a"], g["x_int"], g["dt"], psi0, x, F, n, gs)
        ref_g[:kg + 1] = 0.0
        e = rel_err(got[i, :n], ref_g)
        print("member", i, "full-length gradient vs oracle %.2e" % e, flush=True)
        assert e < 1e-8
