"""The far-field driver on one GPU: Hankel/Hankel_long_medium_parallel_cluster.py:147-370.

The reference selects the harmonic window of the TDSE spectra by bisection (mynumerics.py:35-56),
converts the frequencies to SI (:32, :207), splits the screen over processes in rho or omega
(:179-199), runs Hankel_long on every part and merges the parts (:240-305) before writing
Hankel/outputs (:307-365).  Here the whole screen is one Hankel_long call on the device (one GPU
holds the largest screen many times over), so `screen` returns what the merge produces, and
`outputs` the datasets of results_Hankel.h5 keyed by their paths.
"""
import math

import numpy as np

from . import hankel

# Python unit set (shared_python/units.py:3-20; the second TIMEau definition, :20, is the one in force)
HBAR = 1.0545718e-34
INVERSE_ALPHA_FINE = 137.035999139
ALPHA_FINE = 1.0 / INVERSE_ALPHA_FINE
C_LIGHT = 299792458.0
ELMASS = 9.10938356e-31
R_BOHR = HBAR * INVERSE_ALPHA_FINE / (C_LIGHT * ELMASS)
TIMEAU = (ELMASS * R_BOHR ** 2) / HBAR
OMEGA_AU2SI = 1.0 / TIMEAU  # ConvertPhoton(1.0, 'omegaau', 'omegaSI') (mynumerics.py:314, :324)

OUTPUTS = "Hankel/outputs"  # MMA_administration.paths['Hankel_outputs']
CTDSE_OUTPUTS = "CTDSE/outputs"


def omega_au_of_wavelength(lambda_si):
    """ConvertPhoton(lambda, 'lambdaSI', 'omegaau') (mynumerics.py:311): 2 pi hbar / (lambda m c alpha^2)."""
    return 2.0 * math.pi * HBAR / (lambda_si * ELMASS * C_LIGHT * ALPHA_FINE ** 2)


def find_interval(x, x0):
    """Index k of the interval [x_k, x_k+1) holding x0 on an ascending grid, the last interval
    closed (mynumerics.py:35-56, same bisection); LookupError outside [x_0, x_-1].  Sequences
    of x0 give an array of indices."""
    if hasattr(x0, "__len__"):
        return np.asarray([find_interval(x, v) for v in x0])
    n = len(x)
    if x0 > x[-1] or x0 < x[0]:
        raise LookupError("out of range in FindInterval")
    k1, k2, length = 1, n, n
    while length > 2:
        if x0 < x[k1 - 1 + length // 2]:
            k2 = k1 + length // 2
        else:
            k1 = k1 + length // 2
        length = k2 - k1
    if length == 2 and x0 >= x[k1]:
        return k1
    return k1 - 1


def _host(grid):
    if hasattr(grid, "detach"):  # a torch tensor, possibly on the device
        grid = grid.detach().cpu().numpy()
    return np.asarray(grid, dtype=np.float64)


def screen(ogrid_au, rgrid_macro, zgrid_macro, h5_handle, omega0_au, harmonic_range, rmax_FF, Nr_FF, distance_FF,
           h5_path=CTDSE_OUTPUTS + "/FSourceTerm", ko_step=1, kr_step=1, kr_max="end", kz_step=1, preset_gas="vacuum",
           pressure=1.0, absorption_tables="Henke", include_absorption=True, dispersion_tables="Henke",
           include_dispersion=True, effective_IR_refrective_index=1.0, near_field_factor=True,
           store_cumulative_result=False, store_non_normalised_cumulative_result=False, contraction="fp64"):
    """The far field on a screen of Nr_FF radii in [0, rmax_FF] at distance_FF for the harmonics
    harmonic_range[0] .. harmonic_range[-1] of omega0 (:147-165).  ogrid_au, rgrid_macro,
    zgrid_macro: the CTDSE outputs' omegagrid (a.u.), rgrid_coarse and zgrid_coarse; h5_handle:
    anything indexed like the merged file (an h5py file or mma_amd.farm's result).  Returns the
    Hankel_long result with the merged grids (ogrid in SI, the selected window)."""
    ogrid_au, rgrid_macro, zgrid_macro = (_host(g) for g in (ogrid_au, rgrid_macro, zgrid_macro))
    rgrid_FF = np.linspace(0.0, rmax_FF, Nr_FF)
    ko_min = find_interval(ogrid_au / omega0_au, harmonic_range[0])
    ko_max = find_interval(ogrid_au / omega0_au, harmonic_range[-1])
    target = hankel.FSources_provider(np.asarray(zgrid_macro), np.asarray(rgrid_macro), OMEGA_AU2SI * ogrid_au,
                                      h5_handle=h5_handle, h5_path=h5_path, data_source="dynamic", ko_min=ko_min,
                                      ko_max=ko_max, ko_step=ko_step, kr_max=kr_max, kr_step=kr_step, kz_step=kz_step)
    res = hankel.Hankel_long(target, distance_FF, rgrid_FF, preset_gas=preset_gas, pressure=pressure,
                             absorption_tables=absorption_tables, include_absorption=include_absorption,
                             dispersion_tables=dispersion_tables, include_dispersion=include_dispersion,
                             effective_IR_refrective_index=effective_IR_refrective_index,
                             integrator_Hankel=hankel.trapezoidal_integrator, integrator_longitudinal="trapezoidal",
                             near_field_factor=near_field_factor, store_cumulative_result=store_cumulative_result,
                             store_non_normalised_cumulative_result=store_non_normalised_cumulative_result,
                             contraction=contraction)
    res.ogrid = OMEGA_AU2SI * ogrid_au[ko_min:ko_max:ko_step]  # :245
    res.rgrid = rgrid_FF
    res.ko_min, res.ko_max = ko_min, ko_max
    return res


def outputs(res):
    """The datasets the driver writes under Hankel/outputs (:307-365): complex arrays as
    (..., 2) re/im, grids in SI; optional ones only when present."""
    def ri(a):
        a = np.asarray(a)
        return np.stack((a.real, a.imag), axis=-1)

    out = {OUTPUTS + "/FF_integrated": ri(res.FF_integrated),
           OUTPUTS + "/entry_plane_transform": ri(res.entry_plane_transform),
           OUTPUTS + "/exit_plane_transform": ri(res.exit_plane_transform),
           OUTPUTS + "/ogrid": np.asarray(res.ogrid),
           OUTPUTS + "/rgrid": np.asarray(res.rgrid)}
    for name in ("cumulative_field", "cumulative_field_no_norm"):
        if hasattr(res, name):
            out[OUTPUTS + "/" + name] = ri(getattr(res, name))
    if hasattr(res, "zgrid"):
        out[OUTPUTS + "/zgrid"] = np.asarray(res.zgrid)
    return out


def save(path, res, keep=None):
    """Writes the datasets of `outputs(res)` to the HDF5 file `path` as the driver writes
    results_Hankel.h5 (:307-365): each with its units attribute ('[arb. u.]' for the fields,
    '[SI]' for the grids, mynumerics.adddataset).  keep: an HDF5 file whose datasets and
    attributes are carried over (datasets of the same name are replaced)."""
    from . import h5lite
    data, attrs = ({}, {}) if keep is None else h5lite.read_all(keep)
    for key, value in outputs(res).items():
        p = "/" + key
        data[p] = np.asarray(value)
        attrs[p] = {"units": b"[SI]" if key.rsplit("/", 1)[-1] in ("ogrid", "rgrid", "zgrid") else b"[arb. u.]"}
    h5lite.write(path, data, attrs)


def load(path):
    """The Hankel/outputs datasets of an HDF5 file as {path: array} (complex fields as (..., 2)
    re/im, as stored)."""
    from . import h5lite
    f = h5lite.H5File(path)
    return {OUTPUTS + "/" + k: f[OUTPUTS + "/" + k] for k in f.keys(OUTPUTS)}


__all__ = ["screen", "outputs", "save", "load", "find_interval", "omega_au_of_wavelength", "OMEGA_AU2SI", "OUTPUTS"]
