"""CPU FP64 reference of the TDSE gradient (test infrastructure only).

The forward states come from the C oracle (oracle/ctdse_oracle.c, a restatement of
1DTDSE/sources/prop.c); the backward sweep is a plain numpy restatement of the discrete
adjoint of prop.c's step psi_{k+1} = D(F_k) B^-1 A psi_k, written with the reference's own
(unscaled) matrix entries and a banded LAPACK solve per step:
    lambda_k    = B^T kappa_k - (dL/ds_k) V'(x) psi_k          (A^H = B^T since A = conj(B))
    dL/dF_{k-1} = dL/ds_k - 2 dt Im sum_j x_j conj(lambda_k,j) psi_k,j
    kappa_{k-1} = B^-H (exp(-i F_{k-1} dt x) lambda_k)
It shares no code with the GPU kernel (which uses the Thomas factors and segmented scans).
"""
import numpy as np
from scipy.linalg import solve_banded

from . import ctdse as O


def cn_matrix(num_r, dx, a, dt, x):
    """B = M2 + i dt/2 H as prop.c:138-159 builds it: (sub, diag, sup) with sub[j] = B[j+1, j]."""
    pot = np.array([O.lib().orc_potential(float(v), a) for v in x])
    pot_next = np.append(pot[1:], O.lib().orc_potential(float(x[-1] + dx), a))
    h = 0.5 * dt
    sub = 1 / 12. + 1j * (h * (-0.5 / (dx * dx)) + h / 12. * pot)        # dinfnew1[j]
    dia = 10 / 12. + 1j * (h * (1. / (dx * dx)) + h * 10 / 12. * pot)    # dnew1[j]
    sup = 1 / 12. + 1j * (h * (-0.5 / (dx * dx)) + h / 12. * pot_next)   # dsupnew1[j]
    return sub, dia, sup


def forward_states(num_r, dx, a, x_int, dt, psi0, x, F, n_obs):
    """psi_0 .. psi_{n_obs-1} and the n_obs dipole samples s_0 .. s_{n_obs-1} under the gated
    field F (prop.c runs n_obs - 1 steps for n_obs output samples)."""
    n_run = n_obs - 1
    res = O.propagate(num_r, dx, a, x_int, dt, psi0, x, F, n_run, tprint=dt, n_snap=n_run)
    return np.vstack([np.asarray(psi0)[None, :], res["snapshots"]]), res["sourceterm"][:n_obs]


def dipole(num_r, dx, a, x_int, dt, psi0, x, F, n_obs):
    res = O.propagate(num_r, dx, a, x_int, dt, psi0, x, F, n_obs - 1)
    return res["sourceterm"][:n_obs]


def gradient(num_r, dx, a, x_int, dt, psi0, x, F, n_obs, gobs):
    """dL/dF_k, k = 0 .. n_obs-1, of a loss with dL/ds_k = gobs[k] (s = sourceterm, n_obs
    samples); F is the gated field, so entry n_obs-1 (which reaches no observed sample) is 0."""
    states, _ = forward_states(num_r, dx, a, x_int, dt, psi0, x, F, n_obs)
    sub, dia, sup = cn_matrix(num_r, dx, a, dt, x)
    n = num_r + 1
    dpot = np.array([O.lib().orc_gradpot(float(v), a) for v in x])
    # banded storage of B^H (B[j, j-1] = sub[j-1], B[j, j+1] = sup[j]):
    # B^H[j-1, j] = conj(sub[j-1]), B^H[j+1, j] = conj(sup[j])
    ab = np.zeros((3, n), dtype=complex)
    ab[0, 1:] = np.conj(sub[:-1])
    ab[1, :] = np.conj(dia)
    ab[2, :-1] = np.conj(sup[:-1])
    kappa = np.zeros(n, dtype=complex)
    grad = np.zeros(n_obs)
    for k in range(n_obs - 1, 0, -1):
        # B^T kappa: (B^T v)_j = B[j-1, j] v_{j-1} + B[j, j] v_j + B[j+1, j] v_{j+1}
        lam = dia * kappa
        lam[1:] += sup[:-1] * kappa[:-1]
        lam[:-1] += sub[:-1] * kappa[1:]
        psik = states[k]
        lam = lam - gobs[k] * dpot * psik
        grad[k - 1] = gobs[k] - 2.0 * dt * np.sum(x * np.imag(np.conj(lam) * psik))
        mu = np.exp(-1j * F[k - 1] * dt * x) * lam
        kappa = solve_banded((1, 1), ab, mu)
    return grad


def gate(E):
    """prop.c:125-134: zero up to and including the first k with E_k E_{k+1} <= 0."""
    E = np.asarray(E, dtype=float)
    hit = np.nonzero(E[:-1] * E[1:] <= 0)[0]
    kg = hit[0] if len(hit) else len(E)
    F = E.copy()
    F[:kg + 1] = 0.0
    return F, kg


def harmonic_yield(s, dt, window):
    X = np.fft.rfft(s)
    c1 = dt / np.sqrt(2 * np.pi)
    return float(np.sum(window * np.abs(X) ** 2) * c1 * c1)


def harmonic_yield_grad(s, dt, window):
    """dL/ds for L = c1^2 sum_{k in window} |X_k|^2, X = rfft(s)."""
    n = len(s)
    X = np.fft.rfft(s)
    c1 = dt / np.sqrt(2 * np.pi)
    k = np.nonzero(window)[0]
    t = np.arange(n)
    # d|X_k|^2/ds_t = 2 Re(conj(X_k) exp(-2 pi i k t / n))
    ph = np.exp(-2j * np.pi * np.outer(k, t) / n)
    return c1 * c1 * 2.0 * np.real(np.conj(X[k])[:, None] * ph).sum(axis=0)
