"""Gradients of TDSE dipole losses with respect to the driving field (no counterpart in the
reference, which has no gradient code).

``TDSEDipole.apply(E, geometry, n_steps)`` returns the dipole acceleration (sourceterm,
prop.c:328-333) of a batch of instances as a differentiable torch function of the sampled
field E.  The forward stores the wavefunction every K ~ sqrt(n_steps) steps.  The backward
pass is a hand-written adjoint Crank-Nicolson kernel, in one of two modes:
  * "reconstruct" (default, tdse_reverse.hip): one sweep rebuilds psi_k by inverse steps,
    re-anchored at the checkpoints, next to the adjoint state; no states are stored;
  * "stored" (tdse_adjoint.hip): each block of K states is recomputed from its checkpoint
    into memory and the adjoint sweeps through it.
The field gating of prop.c:125-134 (zero until the first sign change) is held fixed: gated
samples get zero gradient.

``harmonic_yield`` is the windowed harmonic yield sum_{w in window} |F[s](w)|^2 in the
calcFFTW3 convention (1DTDSE/sources/tools_fftw3.c:103-131), differentiable through torch.fft.
"""
import math
from ctypes import POINTER, c_double, c_int, c_longlong, c_void_p

import torch

from . import tdse


def _bind():
    L = tdse._lib()
    if not getattr(L, "_adj_bound", False):
        L.ctdse_state_length.restype = c_longlong
        L.ctdse_state_length.argtypes = [c_void_p]
        L.ctdse_propagate_states.restype = c_int
        L.ctdse_propagate_states.argtypes = [c_int, POINTER(c_void_p), POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                             c_void_p, POINTER(c_longlong), c_void_p, POINTER(c_longlong), c_void_p,
                                             POINTER(c_longlong), c_int, c_void_p, POINTER(c_longlong), c_void_p]
        L.ctdse_adjoint_reverse_batch.restype = c_int
        L.ctdse_adjoint_reverse_batch.argtypes = [c_int, POINTER(c_void_p), POINTER(c_int), c_void_p,
                                                  POINTER(c_longlong), c_void_p, POINTER(c_longlong), c_void_p,
                                                  POINTER(c_longlong), c_void_p, POINTER(c_longlong), c_int, c_int,
                                                  c_void_p, POINTER(c_longlong), c_void_p]
        L.ctdse_adjoint_batch.restype = c_int
        L.ctdse_adjoint_batch.argtypes = [c_int, POINTER(c_void_p), POINTER(c_int), POINTER(c_int), c_void_p,
                                          POINTER(c_longlong), c_void_p, POINTER(c_longlong), c_void_p,
                                          POINTER(c_longlong), c_void_p, POINTER(c_longlong), c_void_p,
                                          POINTER(c_longlong), c_void_p]
        L._adj_bound = True
    return L


def state_length(geometry):
    return int(_bind().ctdse_state_length(geometry.handle))


def gate(E):
    """The gated field F of prop.c:125-134: F_k = 0 for k <= k_g, where k_g is the first k with
    E_k E_{k+1} <= 0, else E_k.  Returns (F, mask) with mask = 1 where F follows E."""
    prod = E[:, :-1] * E[:, 1:]
    hit = prod <= 0
    n = E.shape[1]
    idx = torch.arange(n - 1, device=E.device).expand_as(hit)
    kg = torch.where(hit, idx, torch.full_like(idx, n)).min(dim=1).values  # first crossing
    mask = (torch.arange(n, device=E.device)[None, :] > kg[:, None]).to(E.dtype)
    return E * mask, mask


def _ptrs(lst):
    return (c_void_p * len(lst))(*[g.handle.value for g in lst])


def _i(vals):
    return (c_int * len(vals))(*[int(v) for v in vals])


def _ll(vals):
    return (c_longlong * len(vals))(*[int(v) for v in vals])


def _stream(t):
    return c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def run_states(geometry, F, n_steps, k0, n_run, init, save, save_every, out=None, init_slot=0):
    """Partial forward runs: rows of F (n_inst, L) are gated fields; init (n_inst, m, state, 2)
    or None, the start state of instance i being init[i, init_slot]; save (n_inst, m', state, 2)
    zero-initialised; states are complex128 viewed as float64 pairs in the tiled layout."""
    L = _bind()
    n_inst, Lf = F.shape
    S = state_length(geometry)
    geoms = [geometry] * n_inst
    init_p = c_void_p(init.data_ptr()) if init is not None else c_void_p()
    per_init = init.shape[1] * 2 * S if init is not None else 0
    init_off = _ll([i * per_init + init_slot * 2 * S for i in range(n_inst)])
    per_save = save.shape[1] * 2 * S if save is not None else 0
    save_p = c_void_p(save.data_ptr()) if save is not None else c_void_p()
    out_p = c_void_p(out.data_ptr()) if out is not None else c_void_p()
    out_off = _ll([i * 5 * n_steps for i in range(n_inst)])
    code = L.ctdse_propagate_states(n_inst, _ptrs(geoms), _i([n_steps] * n_inst), _i([k0] * n_inst),
                                    _i([n_run] * n_inst), c_void_p(F.data_ptr()), _ll([i * Lf for i in range(n_inst)]),
                                    init_p, init_off, save_p, _ll([i * per_save for i in range(n_inst)]),
                                    int(save_every), out_p, out_off, _stream(F))
    tdse._raise(code)


def run_adjoint(geometry, F, gobs, states, kappa, grad, k_hi, k_lo):
    L = _bind()
    n_inst, Lf = F.shape
    S = state_length(geometry)
    code = L.ctdse_adjoint_batch(n_inst, _ptrs([geometry] * n_inst), _i([k_hi] * n_inst), _i([k_lo] * n_inst),
                                 c_void_p(F.data_ptr()), _ll([i * Lf for i in range(n_inst)]),
                                 c_void_p(gobs.data_ptr()), _ll([i * gobs.shape[1] for i in range(n_inst)]),
                                 c_void_p(states.data_ptr()), _ll([i * states.shape[1] * 2 * S for i in range(n_inst)]),
                                 c_void_p(kappa.data_ptr()), _ll([i * 2 * S for i in range(n_inst)]),
                                 c_void_p(grad.data_ptr()), _ll([i * grad.shape[1] for i in range(n_inst)]),
                                 _stream(F))
    tdse._raise(code)


def run_adjoint_reverse(geometry, F, gobs, psi_hi, ckpt, every, grad, k_hi):
    """The whole backward sweep k = k_hi .. 1 in one call: psi_hi (n_inst, S, 2), rows contiguous
    (a slice of the checkpoints will do), is psi_{k_hi}; ckpt (n_inst, m, S, 2) holds
    psi_{(j+1) every} in slot j, or None; grad (n_inst, >= k_hi)."""
    L = _bind()
    n_inst, Lf = F.shape
    S = state_length(geometry)
    n_ckpt = ckpt.shape[1] if ckpt is not None else 0
    ckpt_p = c_void_p(ckpt.data_ptr()) if ckpt is not None else c_void_p()
    code = L.ctdse_adjoint_reverse_batch(n_inst, _ptrs([geometry] * n_inst), _i([k_hi] * n_inst),
                                         c_void_p(F.data_ptr()), _ll([i * Lf for i in range(n_inst)]),
                                         c_void_p(gobs.data_ptr()), _ll([i * gobs.shape[1] for i in range(n_inst)]),
                                         c_void_p(psi_hi.data_ptr()), _ll([i * psi_hi.stride(0) for i in range(n_inst)]),
                                         ckpt_p, _ll([i * n_ckpt * 2 * S for i in range(n_inst)]), int(every),
                                         int(n_ckpt), c_void_p(grad.data_ptr()),
                                         _ll([i * grad.shape[1] for i in range(n_inst)]), _stream(F))
    tdse._raise(code)


class TDSEDipole(torch.autograd.Function):
    """sourceterm(E): (n_inst, L) float64 CUDA field samples -> (n_inst, n_steps) dipole
    acceleration, n_steps = outputs.Nt of the reference (L >= n_steps + 1)."""

    @staticmethod
    def forward(ctx, E, geometry, n_steps, checkpoint_every=None, mode="reconstruct"):
        if not (E.is_cuda and E.dtype == torch.float64):
            raise TypeError("E must be a float64 CUDA tensor")
        E = E.contiguous()
        n_inst, Lf = E.shape
        if Lf < n_steps + 1:
            raise ValueError("E needs n_steps + 1 samples per row")
        if mode not in ("reconstruct", "stored"):
            raise ValueError("mode must be 'reconstruct' or 'stored'")
        K = int(checkpoint_every or max(1, math.ceil(math.sqrt(n_steps))))
        F, mask = gate(E)
        S = state_length(geometry)
        n_ckpt = n_steps // K  # the forward saves psi_K, psi_2K, ... (slot m holds psi_{(m+1)K})
        ckpt = torch.zeros((n_inst, max(n_ckpt, 1), S, 2), dtype=torch.float64, device=E.device)
        out = torch.empty((n_inst, 5, n_steps), dtype=torch.float64, device=E.device)
        run_states(geometry, F, n_steps, 0, 0, None, ckpt if n_ckpt > 0 else None, K, out)
        ctx.save_for_backward(F, mask, ckpt)
        ctx.geometry, ctx.n_steps, ctx.K, ctx.n_ckpt, ctx.mode = geometry, n_steps, K, n_ckpt, mode
        return out[:, 0, :].clone()

    @staticmethod
    def backward(ctx, gs):
        F, mask, ckpt = ctx.saved_tensors
        geometry, N, K = ctx.geometry, ctx.n_steps, ctx.K
        n_inst = F.shape[0]
        S = state_length(geometry)
        gobs = gs.contiguous().to(torch.float64)
        grad = torch.zeros((n_inst, N), dtype=torch.float64, device=F.device)
        dE = torch.zeros_like(F)
        if N < 2:
            return dE, None, None, None, None
        if ctx.mode == "reconstruct":
            # psi_{N-1}: the last checkpoint at or below N-1, run forward to N-1 if needed
            j = (N - 1) // K
            if j >= 1 and j * K == N - 1:
                psi_hi = ckpt[:, j - 1]
            else:
                buf = torch.zeros((n_inst, 1, S, 2), dtype=torch.float64, device=F.device)
                n_run = N - 1 - j * K
                run_states(geometry, F, N, j * K, n_run, None if j == 0 else ckpt, buf, n_run, init_slot=j - 1)
                psi_hi = buf[:, 0]
            run_adjoint_reverse(geometry, F, gobs, psi_hi, ckpt if ctx.n_ckpt > 0 else None, K, grad, N - 1)
            dE[:, :N] = grad * mask[:, :N]
            return dE, None, None, None, None
        kappa = torch.zeros((n_inst, S, 2), dtype=torch.float64, device=F.device)
        block = torch.zeros((n_inst, K, S, 2), dtype=torch.float64, device=F.device)
        n_blocks = (N - 2) // K + 1  # blocks cover k = 1 .. N-1
        for m in range(n_blocks - 1, -1, -1):
            k_lo = m * K
            k_hi = min((m + 1) * K, N - 1)
            if k_hi <= k_lo:
                continue
            run_states(geometry, F, N, k_lo, k_hi - k_lo, None if m == 0 else ckpt, block, 1, init_slot=m - 1)
            run_adjoint(geometry, F, gobs, block, kappa, grad, k_hi, k_lo)
        dE[:, :N] = grad * mask[:, :N]
        return dE, None, None, None, None


def harmonic_window(n, dt, tmax, omega0, h_lo, h_hi):
    """Bins of the calcFFTW3 frequency grid (omega_k = k 2pi/tmax) inside [h_lo, h_hi] omega0."""
    k = torch.arange(n // 2 + 1, dtype=torch.float64)
    w = k * 2.0 * math.pi / tmax
    return (w >= h_lo * omega0) & (w <= h_hi * omega0)


def harmonic_yield(s, dt, window):
    """sum over the window of |F_k|^2 with F = conj(rfft(s)) dt / sqrt(2 pi) (tools_fftw3.c:128-131)."""
    X = torch.fft.rfft(s, dim=-1)
    c1 = dt / math.sqrt(2.0 * math.pi)
    P = (X.real ** 2 + X.imag ** 2) * (c1 * c1)
    return (P * window.to(P.device, P.dtype)).sum(dim=-1)


__all__ = ["TDSEDipole", "gate", "harmonic_yield", "harmonic_window", "run_states", "run_adjoint",
           "run_adjoint_reverse", "state_length"]
