"""Synthetic driving fields and instance batches for the TDSE path.

Units: atomic units inside the TDSE (the constants of 1DTDSE/sources/constants.c:26-43);
SI for laser/beam parameters.  Nothing here is measured: these generators produce
the field samples the TDSE consumes, in place of CUPRAD's output_field cube.
"""
import math

import numpy as np

# constants.c:26-43 (C set used by the TDSE)
HBAR = 1.054571800e-34
ALPHA_FINE = 1.0 / 137.035999139
C_LIGHT = 299792458.0
ELCHARGE = 1.602176565e-19
ELMASS = 9.10938356e-31
MU0 = 4.0 * math.pi * 1e-7
EPS0 = 1.0 / (MU0 * C_LIGHT * C_LIGHT)
R_BOHR = 4.0 * math.pi * EPS0 * HBAR * HBAR / (ELMASS * ELCHARGE * ELCHARGE)
TIMEAU = (ELMASS * R_BOHR * R_BOHR) / HBAR
EFIELDAU = HBAR * HBAR / (ELMASS * R_BOHR ** 3 * ELCHARGE)

# Soft-Coulomb presets, constants.c:55-116
SOFT_COULOMB_A = {"He": 0.6950, "Ne": 0.8161, "Ar": 1.1893, "Kr": 1.3676, "Xe": 1.6171}
SOFT_COULOMB_E = {"He": -0.9036, "Ne": -0.7924, "Ar": -0.5792, "Kr": -0.5145, "Xe": -0.4458}


def omega_au(wavelength_m):
    """Angular frequency in a.u. of light of the given vacuum wavelength."""
    return 2.0 * math.pi * C_LIGHT / wavelength_m * TIMEAU


def field_au(intensity_w_cm2):
    """Peak field in a.u. for a cycle-averaged intensity I = eps0 c E^2 / 2."""
    e_si = math.sqrt(2.0 * intensity_w_cm2 * 1e4 / (C_LIGHT * EPS0))
    return e_si / EFIELDAU


def sin2_pulse(t, e0, omega, n_cycles, phase=0.0):
    """E(t) = E0 sin^2(pi t/T) sin(omega t + phase) on [0, T], zero after; T = n_cycles 2pi/omega."""
    T = n_cycles * 2.0 * math.pi / omega
    t = np.asarray(t, dtype=np.float64)
    env = np.where(t <= T, np.sin(math.pi * t / T) ** 2, 0.0)
    return e0 * env * np.sin(omega * t + phase)


def single_instance(n_cycles=10, wavelength=800e-9, intensity=1.5e14, dt=0.25):
    """Single-instance case: 800 nm, sin^2 envelope, sampled at the TDSE step (k1 = 1)."""
    w = omega_au(wavelength)
    T = n_cycles * 2.0 * math.pi / w
    n_e = int(math.ceil(T / dt)) + 1
    t = np.arange(n_e) * dt
    return t, sin2_pulse(t, field_au(intensity), w, n_cycles)


# grid and atom shared by the TDSE workloads (dx = 0.1, x in [-800, 800], Ar)
GRID = dict(num_r=16000, dx=0.1, a=SOFT_COULOMB_A["Ar"], e_guess=SOFT_COULOMB_E["Ar"], cv=1e-25, x_int=2.0)


def gaussian_beam_batch(n_r=64, n_z=64, w0=100e-6, wavelength=800e-9, intensity=1.5e14, z_span=5e-3,
                        n_cycles=10, dt=0.25):
    """Uniform batch: fields E(r, z, t) of a focused Gaussian beam, following the
    Gaussian-beam field of shared_python/mynumerics.py:369-379 (amplitude w0/w(z),
    radial exp(-r^2/w^2), Gouy phase atan(z/zR), curvature phase k r^2 z/(2(zR^2+z^2)))
    with a sin^2 envelope.  r in linspace(0, 2 w0, n_r), z in linspace(-z_span/2, z_span/2, n_z).
    Returns t (n_e,), fields (n_z*n_r, n_e) ordered instance = kz*n_r + kr (MPI-TDSE.c:148)."""
    w = omega_au(wavelength)
    e0 = field_au(intensity)
    T = n_cycles * 2.0 * math.pi / w
    n_e = int(math.ceil(T / dt)) + 1
    t = np.arange(n_e) * dt
    k0 = 2.0 * math.pi / wavelength
    z_r = math.pi * w0 ** 2 / wavelength
    r = np.linspace(0.0, 2.0 * w0, n_r)
    z = np.linspace(-z_span / 2, z_span / 2, n_z)
    wz = w0 * np.sqrt(1.0 + (z / z_r) ** 2)
    gouy = np.arctan(z / z_r)
    curv = 0.5 * k0 * (r[None, :] ** 2) * (z / (z_r ** 2 + z ** 2))[:, None]
    amp = (w0 / wz)[:, None] * np.exp(-(r[None, :] / wz[:, None]) ** 2)
    phase = gouy[:, None] - curv
    env = np.where(t <= T, np.sin(math.pi * t / T) ** 2, 0.0)
    fields = (e0 * amp.reshape(-1, 1)) * env[None, :] * np.sin(w * t[None, :] + phase.reshape(-1, 1))
    return t, fields


def _uneven_draw(n_inst, seed, dt, nr_range, cycles, intensity, wavelength):
    rng = np.random.Generator(np.random.PCG64(seed))
    half = rng.integers(nr_range[0], nr_range[1] + 1, size=n_inst)
    cyc = rng.integers(cycles[0], cycles[1] + 1, size=n_inst)
    inten = np.exp(rng.uniform(math.log(intensity[0]), math.log(intensity[1]), size=n_inst))
    lam = rng.uniform(wavelength[0], wavelength[1], size=n_inst)
    n_e = np.array([int(math.ceil(cyc[i] * 2.0 * math.pi / omega_au(lam[i]) / dt)) + 1 for i in range(n_inst)])
    return 2 * half, cyc, inten, lam, n_e


UNEVEN = dict(n_inst=8192, seed=2024, dt=0.25, nr_range=(2000, 16000), cycles=(6, 20), intensity=(5e13, 3e14),
              wavelength=(800e-9, 1800e-9))
# Ground-state stopping tolerance of the uneven workload.  With the reference's CV = 1e-25 the
# inverse iteration ends in a 2-cycle instead of an exact fixed point on 10 of the draw's 6213
# distinct grids (e.g. num_r = 10198: E alternates with |dE| = 5.2e-15), where the reference
# exits after 1000 iterations (tridiag.c:139-142); 1e-15 still leaves 6 such grids.  At 1e-12
# every grid stops within 4 iterations, one iteration past |dE| ~ 1e-12, where the energy is
# at its rounding floor.
UNEVEN_CV = 1e-12


def uneven_batch(n_inst=8192, seed=2024, dt=0.25, nr_range=(2000, 16000), cycles=(6, 20),
                 intensity=(5e13, 3e14), wavelength=(800e-9, 1800e-9), indices=None):
    """Uneven batch from numpy.random.Generator(PCG64(seed)): num_r = 2*U{nr_range},
    cycles U{cycles}, intensity log-uniform, wavelength uniform; fields sampled at dt.
    Returns a list of dicts with num_r, n_cycles, intensity, wavelength, t, field (for the
    members listed in `indices` only, if given; the draw is the same either way)."""
    num_r, cyc, inten, lam, n_e = _uneven_draw(n_inst, seed, dt, nr_range, cycles, intensity, wavelength)
    out = []
    for i in (range(n_inst) if indices is None else indices):
        t = np.arange(n_e[i]) * dt
        out.append(dict(num_r=int(num_r[i]), n_cycles=int(cyc[i]), intensity=float(inten[i]),
                        wavelength=float(lam[i]), t=t,
                        field=sin2_pulse(t, field_au(inten[i]), omega_au(lam[i]), int(cyc[i]))))
    return out


def uneven_work(n_inst=8192, seed=2024, dt=0.25, nr_range=(2000, 16000), cycles=(6, 20), intensity=(5e13, 3e14),
                wavelength=(800e-9, 1800e-9)):
    """(num_r, n_e) arrays of every member of uneven_batch, without building the fields."""
    num_r, cyc, inten, lam, n_e = _uneven_draw(n_inst, seed, dt, nr_range, cycles, intensity, wavelength)
    return num_r, n_e


def stack_fields(fields, extra=3):
    """Ragged field samples -> one (n, max_len + extra) array with the reference's zero tail
    (FourInterp's calloc) and each member's step count (len + 2, the reference's outputs.Nt)."""
    width = max(len(f) for f in fields) + extra
    out = np.zeros((len(fields), width))
    for i, f in enumerate(fields):
        out[i, :len(f)] = f
    return out, [len(f) + 2 for f in fields]


# grid of the gradient workload: x in [-400, 400] a.u., dx = 0.1 (Nx = 8001), dt = 0.1 a.u.
ADJOINT_GRID = dict(num_r=8000, dx=0.1, dt=0.1, a=SOFT_COULOMB_A["Ar"], e_guess=SOFT_COULOMB_E["Ar"], cv=1e-25,
                    x_int=2.0)


def adjoint_batch(n_inst=1024, n_t=40000, dt=0.1, w0=100e-6, wavelength=800e-9, intensity=1.5e14,
                  n_cycles=30, r_max_w0=1.5):
    """Gradient workload: n_inst fields of n_t samples (dt apart) across the focal plane of a
    Gaussian beam (r in linspace(0, r_max_w0 w0, n_inst), amplitude exp(-r^2/w0^2)), each a
    sin^2 pulse of n_cycles at the given wavelength.  Returns t (n_t,), fields (n_inst, n_t)."""
    w = omega_au(wavelength)
    T = n_cycles * 2.0 * math.pi / w
    if T > (n_t - 1) * dt:
        raise ValueError("pulse longer than the sampled window")
    t = np.arange(n_t) * dt
    r = np.linspace(0.0, r_max_w0 * w0, n_inst)
    amp = field_au(intensity) * np.exp(-(r / w0) ** 2)
    env = np.where(t <= T, np.sin(math.pi * t / T) ** 2, 0.0)
    return t, amp[:, None] * (env * np.sin(w * t))[None, :]


# far-field workload: harmonics 11-61 of 800 nm, 5 mm of gas, screen at 1 m
OMEGA0_SI = 2.0 * math.pi * C_LIGHT / 800e-9


def hankel_workload(nz=256, no=2048, nr=512, nrho=1024, r_max=200e-6, rho_max=7e-3, length=5e-3, distance=1.0,
                    w0=100e-6, h_lo=11, h_hi=61, seed=7, complex64=False, device="cuda"):
    """Near-field XUV sources of a focused Gaussian beam (SI units), generated on `device`.

    The driver (w0, 800 nm, focus at the cell entrance) has intensity
    I = (w0/w)^2 exp(-2 r^2/w^2) and phase phi = atan(z/zR) - k r^2 z / (2 (zR^2 + z^2)); the
    harmonic at omega gets amplitude S(omega) (I/I0)^2 and phase (omega/omega0) phi - alpha I/I0
    plus a seeded random phase per harmonic order, S(omega) a comb of Gaussians at the odd orders
    h_lo..h_hi.  Returns (zgrid with nz+1 points, rgrid, ogrid, rho, sources (nz, no, nr)
    complex, distance): the provider keeps zgrid[0:-1], i.e. nz planes."""
    import torch
    gen = torch.Generator(device="cpu").manual_seed(seed)
    zgrid = np.linspace(0.0, length, nz + 1)
    rgrid = np.linspace(0.0, r_max, nr)
    ogrid = OMEGA0_SI * np.linspace(h_lo, h_hi, no)
    rho = np.linspace(0.0, rho_max, nrho)
    orders = np.arange(h_lo, h_hi + 1, 2)
    ph_q = torch.rand(len(orders), generator=gen, dtype=torch.float64) * 2 * math.pi
    alpha = 1.0 + torch.rand(len(orders), generator=gen, dtype=torch.float64)
    k0 = OMEGA0_SI / C_LIGHT
    z_r = math.pi * w0 ** 2 / 800e-9
    dt = torch.float64
    z = torch.tensor(zgrid[:nz], dtype=dt, device=device)[:, None, None]
    r = torch.tensor(rgrid, dtype=dt, device=device)[None, None, :]
    om = torch.tensor(ogrid, dtype=dt, device=device)[None, :, None] / OMEGA0_SI  # harmonic order (real)
    w = w0 * torch.sqrt(1.0 + (z / z_r) ** 2)
    inten = (w0 / w) ** 2 * torch.exp(-2.0 * r ** 2 / w ** 2)
    phi = torch.atan(z / z_r) - k0 * r ** 2 * z / (2.0 * (z_r ** 2 + z ** 2))
    q = torch.tensor(orders, dtype=dt, device=device)
    sig = 0.15
    dist = (om[..., None] - q) ** 2  # (1, no, 1, nq)
    comb_w = torch.exp(-dist / (2 * sig ** 2))
    env = comb_w.sum(-1)
    qphase = (comb_w * ph_q.to(device)).sum(-1) / torch.clamp(env, min=1e-300)
    qalpha = (comb_w * alpha.to(device)).sum(-1) / torch.clamp(env, min=1e-300)
    phase = om * phi - qalpha * inten + qphase
    src = (env * inten ** 2) * torch.exp(1j * phase)
    src = src.to(torch.complex64 if complex64 else torch.complex128).contiguous()
    return zgrid, rgrid, ogrid, rho, src, distance


def hankel_pressure(length=5e-3, n=101, p0=1.0):
    """z-modulated gas density of a cell with smooth edges: the pressure dict of the reference
    (keys zgrid, value), p0 [bar] in the middle."""
    zp = np.linspace(-0.1 * length, 1.1 * length, n)
    edge = 0.08 * length
    val = p0 / (1.0 + np.exp(-(zp - 0.05 * length) / edge * 4)) / (1.0 + np.exp((zp - 0.95 * length) / edge * 4))
    return {"zgrid": zp, "value": val}
