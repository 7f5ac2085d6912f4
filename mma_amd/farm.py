"""The CTDSE task farm on one GPU: the MPI driver 1DTDSE/sources/MPI-TDSE.c:148-329 and the
merge of its per-rank files (1DTDSE/python/merge.py:71-108) as one persistent launch over all
(z, r) points of a CUPRAD field.

The reference reads the field E(z, t, r) of CUPRAD (float32, C order (z, t, r),
MPI-TDSE.c:99-101,185), coarsens it to every kz_step-th plane below Nz_max and every kr_step-th
radius below Nr_max (MPI-TDSE.c:112-123), hands instance Nsim = kz * Nr_coarse + kr to one
call1DTDSE each (MPI-TDSE.c:176-186, 206-211) with one ground state for all (MPI-TDSE.c:145),
and merge.py writes every output to [kz, kr, ...] of one array under CTDSE/outputs.  Here the
instances run as one batch on the device (field refinement for k1 > 1 and the spectra with
torch.fft on the device), and the result is a mapping with the merged file's paths and layouts
('CTDSE/outputs/SourceTerm' (Nz, Nr, Nt), 'CTDSE/outputs/FSourceTerm' (Nz, Nr, Nomega, 2), ...),
which the Hankel stage's FSources_provider(data_source='dynamic', h5_handle=...) indexes as it
would the HDF5 file.
"""
import math

import numpy as np
import torch

from . import tdse

# C unit set of the TDSE code (1DTDSE/sources/constants.c:26-43)
_PI = math.pi
_HBAR = 1.054571800e-34
_C = 299792458.0
_ELCHARGE = 1.602176565e-19
_ELMASS = 9.10938356e-31
_MU0 = 4.0 * _PI * 1e-7
_EPS0 = 1.0 / (_MU0 * _C * _C)
R_BOHR = 4.0 * _PI * _EPS0 * _HBAR * _HBAR / (_ELMASS * _ELCHARGE * _ELCHARGE)
TIMEAU = (_ELMASS * R_BOHR * R_BOHR) / _HBAR
EFIELDAU = _HBAR * _HBAR / (_ELMASS * R_BOHR * R_BOHR * R_BOHR * _ELCHARGE)

OUTPUTS = "CTDSE/outputs"  # MMA_administration.paths['CTDSE_outputs']


def coarsen_grid(grid, k_step, n_max):
    """Every k_step-th point of the first n_max (tools_algorithmic.c:54-65): n_max // k_step points."""
    n = int(n_max) // int(k_step)
    return np.asarray(grid, dtype=np.float64)[0:n * int(k_step):int(k_step)].copy()


def instance_index(n_sim, n_r):
    """(kz, kr) of simulation n_sim on a grid of n_r coarse radii (MPI-TDSE.c:178-180)."""
    kr = n_sim % n_r
    return (n_sim - kr) // n_r, kr


def coarse_fields(output_field, kz_step, nz_max, kr_step, nr_max):
    """Instance fields in atomic units, in simulation order: row kz * Nr_c + kr holds
    output_field[kz * kz_step, :, kr * kr_step] / EFIELDau (MPI-TDSE.c:181-186, 206-208), with
    Nz_c = nz_max // kz_step and Nr_c = nr_max // kr_step (MPI-TDSE.c:118-119).  ValueError where
    the reference stops: nr_max beyond the field's radii (MPI-TDSE.c:193-201), nz_max beyond its
    planes (the hyperslab read fails)."""
    f = np.asarray(output_field)
    if f.ndim != 3:
        raise ValueError("output_field must be (z, t, r)")
    nz, nt, nr = f.shape
    if nr_max > nr:
        raise ValueError("'Nr_max' must be smaller than 'numerics_number_of_points_in_r in the fields array input!")
    if nz_max > nz:
        raise ValueError("Nz_max exceeds the number of planes of the field")
    nz_c, nr_c = int(nz_max) // int(kz_step), int(nr_max) // int(kr_step)
    if nz_c < 1 or nr_c < 1:
        raise ValueError("no instance: Nz_max // kz_step and Nr_max // kr_step must be positive")
    sub = f[0:nz_c * kz_step:kz_step, :, 0:nr_c * kr_step:kr_step].astype(np.float64)  # (Nz_c, Nt, Nr_c)
    return np.ascontiguousarray(np.transpose(sub, (0, 2, 1)).reshape(nz_c * nr_c, nt)) / EFIELDAU, nz_c, nr_c


def four_interp(fields, k1):
    """FourInterp (tools_fftw3.c:24-68) of every row on the device: spectrum of N samples / N
    into floor(k1 N / 2) + 1 zeroed bins (the Nyquist bin copied unhalved), unnormalised inverse
    transform of length k1 N.  k1 = 1 returns the rows unchanged (the reference's round trip
    changes them by FFT rounding only)."""
    if k1 == 1:
        return fields
    n = fields.shape[1]
    n2 = k1 * n
    spec = torch.fft.rfft(fields, dim=1) / n
    big = torch.zeros((fields.shape[0], n2 // 2 + 1), dtype=spec.dtype, device=fields.device)
    big[:, :spec.shape[1]] = spec
    return torch.fft.irfft(big, n=n2, dim=1) * n2


def spectra(signal, dt, tmax):
    """calcFFTW3 (tools_fftw3.c:84-143) of every row on the device: omega_k = k 2pi/tmax,
    F = conj(FFT) dt / sqrt(2pi) as (..., Nomega, 2) re/im, M2 = dt^2 / (2pi) |FFT|^2."""
    X = torch.fft.rfft(signal, dim=-1)
    c1 = dt / math.sqrt(2.0 * _PI)
    F = torch.stack([c1 * X.real, -c1 * X.imag], dim=-1)
    M2 = (dt * dt / (2.0 * _PI)) * (X.real * X.real + X.imag * X.imag)
    omega = torch.arange(X.shape[-1], dtype=torch.float64) * (2.0 * _PI / tmax)
    return omega, F, M2


def run(output_field, tgrid, rgrid, zgrid, kz_step, nz_max, kr_step, nr_max, num_r, dx, dt, a,
        e_guess=-0.5792, cv=1e-25, x_int=2.0, tmin=0.0, config=None, device="cuda", interp_by_nt=0, ntinterp=0):
    """All (z, r) instances of a CUPRAD field through the TDSE in one launch.

    output_field: (z, t, r) field in SI units (V/m), tgrid in s, rgrid/zgrid in m (the CUPRAD
    outputs); kz_step/nz_max/kr_step/nr_max: the coarsening of CTDSE/inputs; the rest as
    inputs_def (num_r, dx, dt, trg.a, Eguess, CV, x_int, tmin, InterpByDTorNT, Ntinterp).
    config: launch shape (default: the library's).  Returns a dict keyed like the merged file
    (OUTPUTS + '/SourceTerm', ...), torch tensors on the device for the per-instance data."""
    fields, nz_c, nr_c = coarse_fields(output_field, kz_step, nz_max, kr_step, nr_max)
    t_au = np.asarray(tgrid, dtype=np.float64) / TIMEAU  # MPI-TDSE.c:128-130
    k1, dt_ref, field_dt = _refine(t_au, dt, interp_by_nt, ntinterp)
    tmax = t_au[-1] - t_au[0]  # singleTDSE.c:58
    n_e = fields.shape[1]
    n2 = k1 * n_e
    n_out = n2 + 2  # outputs.Nt (singleTDSE.c:79,101)
    geo = tdse.Geometry(num_r, dx, dt_ref, a, e_guess, cv, x_int)
    dev = torch.tensor(fields, dtype=torch.float64, device=device)
    refined = four_interp(dev, k1)
    padded = torch.zeros((refined.shape[0], n2 + 3), dtype=torch.float64, device=device)
    padded[:, :n2] = refined  # the zero tail of FourInterp's calloc (tools_fftw3.c:51)
    out = torch.empty((refined.shape[0], 5, n_out), dtype=torch.float64, device=device)
    plan = tdse.BatchPlan(geo, padded, n_steps=[n_out] * refined.shape[0], out=out, config=config)
    plan.enqueue()
    torch.cuda.current_stream(out.device).synchronize()
    plan.close()

    def merged(row):
        return out[:, row, :].reshape(nz_c, nr_c, n_out)

    omega, F_src, M2_src = spectra(out[:, 0, :], dt_ref, tmax)
    _, F_e, M2_e = spectra(out[:, 4, :], dt_ref, tmax)
    n_om = omega.numel()
    t_out = np.cumsum(np.concatenate([[tmin], np.full(n_out - 1, dt_ref)]))  # prop.c:97,118,241
    x, psi0 = geo.ground_state()
    res = {
        OUTPUTS + "/SourceTerm": merged(0),
        OUTPUTS + "/PopTot": merged(1),
        OUTPUTS + "/PopInt": merged(2),
        OUTPUTS + "/expval_x": merged(3),
        OUTPUTS + "/Efield": merged(4),
        OUTPUTS + "/FSourceTerm": F_src.reshape(nz_c, nr_c, n_om, 2),
        OUTPUTS + "/FSourceTermM2": M2_src.reshape(nz_c, nr_c, n_om),
        OUTPUTS + "/FEfield": F_e.reshape(nz_c, nr_c, n_om, 2),
        OUTPUTS + "/FEfieldM2": M2_e.reshape(nz_c, nr_c, n_om),
        OUTPUTS + "/tgrid": t_out,
        OUTPUTS + "/omegagrid": omega.numpy(),
        OUTPUTS + "/Energy_of_the_ground_state": np.array([geo.e_init]),
        OUTPUTS + "/xgrid_micro": x,
        OUTPUTS + "/ground_state": np.stack([psi0.real, psi0.imag], axis=1),
        OUTPUTS + "/zgrid_coarse": coarsen_grid(zgrid, kz_step, nz_max),
        OUTPUTS + "/rgrid_coarse": coarsen_grid(rgrid, kr_step, nr_max),
        OUTPUTS + "/trg_a": np.array([float(a)]),
    }
    return res


def _refine(t_au, dt, interp_by_nt=0, ntinterp=0):
    """k1 rule of singleTDSE.c:52-68: k1 = Ntinterp + 1 when InterpByDTorNT is 1, else from the
    field's step and dt; returns (k1, refined step, field step)."""
    field_dt = t_au[1] - t_au[0]
    if interp_by_nt == 1:
        k1 = int(ntinterp) + 1
        if k1 < 1:
            raise ValueError("field refinement factor k1 < 1 (Ntinterp < 0)")
        return k1, field_dt / float(k1), field_dt
    k1 = int(math.floor(field_dt / dt))
    if k1 != 1:
        k1 += 1
    if k1 < 1:
        raise ValueError("field refinement factor k1 < 1 (field dt smaller than dt)")
    return k1, field_dt / float(k1), field_dt


def as_host(result):
    """The mapping with every device tensor copied to a numpy array (what an HDF5 writer or the
    Hankel provider's slicing would take)."""
    return {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else v) for k, v in result.items()}


# -- archives (HDF5, through mma_amd.h5lite) ---------------------------------------------------

CUPRAD_OUTPUTS = "CUPRAD/outputs"  # 1DTDSE/sources/h5namelist.h:13-21
CTDSE_INPUTS = "CTDSE/inputs"
GLOBAL_INPUTS = "global_inputs"
# units attached to the merged outputs (1DTDSE/python/merge.py:38-56); PopTot, which the merge
# leaves out, is written too
UNITS = {"Efield": "[a.u.]", "FEfield": "[a.u.]", "SourceTerm": "[a.u.]", "FSourceTerm": "[a.u.]",
         "FEfieldM2": "[a.u.]", "FSourceTermM2": "[a.u.]", "PopInt": "[-]", "PopTot": "[-]", "expval_x": "[a.u.]",
         "tgrid": "[a.u.]", "omegagrid": "[a.u.]", "Energy_of_the_ground_state": "[a.u.]", "xgrid_micro": "[a.u.]",
         "ground_state": "[a.u.]", "zgrid_coarse": "[a.u.]", "rgrid_coarse": "[a.u.]", "trg_a": "[a.u.]"}
# soft-Coulomb presets of the gas_preset input (1DTDSE/sources/constants.c:49-116)
GAS_PRESETS = {"He": (0.6950, -0.9036), "Ne": (0.8161, -0.7924), "Ar": (1.1893, -0.5792), "Kr": (1.3676, -0.5145),
               "Xe": (1.6171, -0.4458)}


def _scalar(f, path):
    return np.asarray(f[path]).reshape(-1)[0].item()


def read_inputs(f):
    """The CTDSE inputs of an archive (an h5lite.H5File) as ReadInputs reads them
    (1DTDSE/sources/tools_hdf5.c:65-148), with the coarsening of MPI-TDSE.c:112-116: num_r =
    2 Nx_max, dx, InterpByDTorNT, dt, Ntinterp, x_int, the soft-Coulomb a and E guess of
    global_inputs/gas_preset unless CTDSE/inputs/trg_a and Eguess are given (those take
    precedence), CV_criterion_of_GS, kz_step, kr_step, Nz_max, Nr_max.  ValueError where the
    reference stops: neither a preset nor trg_a / Eguess, or an unknown preset."""
    P = CTDSE_INPUTS + "/"
    d = {"num_r": 2 * int(_scalar(f, P + "Nx_max")), "dx": float(_scalar(f, P + "dx")),
         "interp_by_nt": int(_scalar(f, P + "InterpByDTorNT")), "dt": float(_scalar(f, P + "dt")),
         "ntinterp": int(_scalar(f, P + "Ntinterp")), "x_int": float(_scalar(f, P + "x_int")),
         "cv": float(_scalar(f, P + "CV_criterion_of_GS"))}
    a = e = None
    preset = GLOBAL_INPUTS + "/gas_preset"
    if preset in f:
        name = _scalar(f, preset)
        name = (name.decode() if isinstance(name, bytes) else str(name)).rstrip("\x00").strip()
        if name not in GAS_PRESETS:
            raise ValueError("Soft-coulomb parameter for this input not supported: %r" % name)
        a, e = GAS_PRESETS[name]
    if P + "trg_a" in f:
        a = float(_scalar(f, P + "trg_a"))
    if P + "Eguess" in f:
        e = float(_scalar(f, P + "Eguess"))
    if a is None:
        raise ValueError("The soft-Coulomb parameter in 1D-TDSE not specified.")
    if e is None:
        raise ValueError("The ground-state-energy guess in 1D-TDSE not specified.")
    d["a"], d["e_guess"] = a, e
    for k in ("kz_step", "kr_step", "Nz_max", "Nr_max"):
        d[k] = int(_scalar(f, P + k))
    return d


def run_archive(archive, out=None, config=None, device="cuda"):
    """The whole task farm on an MMA archive (MPI-TDSE.c:44-329 and merge.py): reads
    CUPRAD/outputs/{output_field (z, t, r), tgrid, rgrid, zgrid} and the CTDSE inputs, runs every
    coarsened instance in one launch, and writes CTDSE/outputs/* in the merged layout with units
    attributes to `out`; out None: into the archive itself (rewritten with everything it held;
    outputs of an earlier run are replaced).  Returns the result mapping of `run` (device
    tensors)."""
    from . import h5lite
    f = h5lite.H5File(archive)
    inp = read_inputs(f)
    C = CUPRAD_OUTPUTS + "/"
    res = run(f[C + "output_field"], f[C + "tgrid"], f[C + "rgrid"], f[C + "zgrid"], inp["kz_step"], inp["Nz_max"],
              inp["kr_step"], inp["Nr_max"], inp["num_r"], inp["dx"], inp["dt"], inp["a"], e_guess=inp["e_guess"],
              cv=inp["cv"], x_int=inp["x_int"], config=config, device=device, interp_by_nt=inp["interp_by_nt"],
              ntinterp=inp["ntinterp"])
    del f
    write_outputs(archive if out is None else out, res, keep=archive if out in (None, archive) else None)
    return res


def write_outputs(path, result, keep=None):
    """Writes the merged outputs (`run`'s mapping; device tensors are copied to the host) to the
    HDF5 file `path` with their units attributes.  keep: an HDF5 file whose datasets and
    attributes are carried over (the archive the outputs are added to); datasets of the same
    name are replaced."""
    from . import h5lite
    data, attrs = ({}, {}) if keep is None else h5lite.read_all(keep)
    for key, value in as_host(result).items():
        p = "/" + key
        data[p] = np.asarray(value)
        attrs[p] = {"units": UNITS.get(key.rsplit("/", 1)[-1], "[a.u.]").encode()}
    h5lite.write(path, data, attrs)


__all__ = ["run", "run_archive", "read_inputs", "write_outputs", "coarse_fields", "coarsen_grid", "instance_index",
           "four_interp", "spectra", "as_host", "TIMEAU", "EFIELDAU", "OUTPUTS", "UNITS", "GAS_PRESETS"]

