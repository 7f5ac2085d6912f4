"""FP64 CPU restatement of the Hankel far-field stage (test infrastructure only: imported by
tests/ and bench.py's CPU baseline, never by the product).

Follows Hankel/Hankel_transform.py and shared_python/XUV_refractive_index.py (with the
Python unit set of shared_python/units.py and ConvertPhoton of shared_python/mynumerics.py:304-329),
vectorised with numpy; J0 is scipy.special.jv(0, .), which is what special.jn(0, .) evaluates
(Hankel_transform.py:507,512).  The XUV scattering-factor tables are the reference's
XUV_refractive_index_tables.h5, converted to mma_amd/data/xuv_tables.npz (data only).
"""
import os

import numpy as np
from scipy import special

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLES = os.path.join(ROOT, "mma_amd", "data", "xuv_tables.npz")

# shared_python/units.py:3-20 (the later TIMEau definition is the one in effect)
HBAR = 1.0545718e-34
INVERSE_ALPHA_FINE = 137.035999139
C_LIGHT = 299792458.0
ELCHARGE = 1.602176565e-19
ELMASS = 9.10938356e-31
R_BOHR = HBAR * INVERSE_ALPHA_FINE / (C_LIGHT * ELMASS)
ALPHA_FINE = 1.0 / INVERSE_ALPHA_FINE
R_ELECTRON = R_BOHR * (ALPHA_FINE ** 2)
TIMEAU = (ELMASS * R_BOHR ** 2) / HBAR
N_REF = 2.7e25  # XUV_refractive_index.py:123


def omega_si_to_ev(omega):
    """mynumerics.py:313,326: omegaSI -> a.u. -> eV."""
    w = omega * TIMEAU
    return w / (ELCHARGE / (ELMASS * ALPHA_FINE ** 2 * C_LIGHT ** 2))


def omega_si_to_lambda(omega):
    """mynumerics.py:313,323: omegaSI -> a.u. -> lambdaSI."""
    w = omega * TIMEAU
    return 2.0 * np.pi * HBAR / (w * ELMASS * C_LIGHT * ALPHA_FINE ** 2)


_tables = None


def tables():
    global _tables
    if _tables is None:
        with np.load(TABLES, allow_pickle=False) as z:
            _tables = {k: z[k] for k in z.files}
    return _tables


def gases():
    return sorted({k.split("/")[0] for k in tables()})


def _interp_checked(xt, yt, x):
    """Linear interpolation that refuses points outside the table (scipy interp1d default,
    XUV_refractive_index.py:54-55)."""
    x = np.asarray(x, dtype=np.float64)
    if np.any(x < xt[0]) or np.any(x > xt[-1]):
        raise ValueError("photon energy outside the tabulated range")
    return np.interp(x, xt, yt)


def getf1(gas, energy):
    t = tables()
    return _interp_checked(t[gas + "/Energy_f1"], t[gas + "/f1"], energy)


def getf2(gas, energy):
    t = tables()
    return _interp_checked(t[gas + "/Energy_f2"], t[gas + "/f2"], energy)


def dispersion_function(omega, pressure, gas, n_ir=1.0, n_ref=N_REF):
    """(1/v_IR) - (1/v_XUV), XUV_refractive_index.py:125-160."""
    f1 = getf1(gas, omega_si_to_ev(omega))
    if np.any(np.asarray(f1) < -9998.0):
        raise ValueError("f1-scattering coefficient not listed for this frequency.")
    lam = omega_si_to_lambda(omega)
    n_xuv = 1.0 - pressure * n_ref * R_ELECTRON * ((lam ** 2) * f1 / (2.0 * np.pi))
    v_xuv = C_LIGHT / n_xuv
    v_ir = C_LIGHT / n_ir
    return (1.0 / v_ir) - (1.0 / v_xuv)


def beta_factor_ref(omega, gas, n_ref=N_REF):
    """Imaginary part of the index at the reference density, XUV_refractive_index.py:163-196."""
    f2 = getf2(gas, omega_si_to_ev(omega))
    lam = omega_si_to_lambda(omega)
    return n_ref * R_ELECTRON * ((lam ** 2) * f2 / (2.0 * np.pi))


def _interp_edge(xt, yt, x):
    """interp1d(..., bounds_error=False, fill_value=(y[0], y[-1])) of Hankel_transform.py."""
    return np.interp(x, xt, yt, left=yt[0], right=yt[-1])


def _cumtrapz(y, x):
    """cumulative_trapezoid(y, x=x, initial=0)."""
    y = np.asarray(y, dtype=np.float64)
    out = np.zeros_like(y)
    out[1:] = np.cumsum(0.5 * (x[1:] - x[:-1]) * (y[1:] + y[:-1]))
    return out


def prefactor_tables(zgrid, rgrid, ogrid, preset_gas="vacuum", pressure=1.0, absorption_tables="Henke",
                     include_absorption=True, dispersion_tables="Henke", include_dispersion=True, n_ir=1.0):
    """The tables behind get_propagation_pre_factor_function (Hankel_transform.py:85-448):
    returns (pre, renorm) with pre[k, o] (scalar and z-modulated pressure) or pre[k, o, r]
    (z-and-r-modulated pressure) complex, and renorm[k, o] real or None."""
    zgrid = np.asarray(zgrid, dtype=np.float64)
    ogrid = np.asarray(ogrid, dtype=np.float64)
    nz, no = len(zgrid), len(ogrid)
    dgas = preset_gas + "_" + dispersion_tables
    agas = preset_gas + "_" + absorption_tables
    if isinstance(pressure, dict):
        has_z, has_r = "zgrid" in pressure, "rgrid" in pressure
        if has_z and not has_r:  # :146-239
            zp = np.asarray(pressure["zgrid"], dtype=np.float64)
            pv = np.asarray(pressure["value"], dtype=np.float64)
            pmod = _interp_edge(zp, pv, zgrid)
            pre = np.empty((nz, no), dtype=np.complex128)
            absw = np.empty((nz, no))
            cum_p = _cumtrapz(pv, zp)
            for o in range(no):
                if include_dispersion:
                    integ = _cumtrapz(dispersion_function(ogrid[o], pv, dgas, n_ir=n_ir), zp)
                    disp = 1j * _interp_edge(zp, integ, zgrid)
                else:
                    disp = 1j * (n_ir - 1.0) / C_LIGHT
                if include_absorption:
                    integ_b = beta_factor_ref(ogrid[o], agas) * cum_p
                    absf = (1.0 / C_LIGHT) * _interp_edge(zp, integ_b, zgrid)
                    absw[:, o] = ogrid[o] * absf
                    absf = absf - absf[-1]
                else:
                    absf = 0.0
                    absw[:, o] = 0.0
                pre[:, o] = pmod * np.exp(ogrid[o] * (disp + absf))
            renorm = np.exp(absw[-1, :][None, :] - absw)
            return pre, renorm
        if has_z and has_r:  # :244-399 (z and r); beta uses the dispersion tables there (:373-376)
            zp = np.asarray(pressure["zgrid"], dtype=np.float64)
            rp = np.asarray(pressure["rgrid"], dtype=np.float64)
            pv = np.asarray(pressure["value"], dtype=np.float64)
            rgrid = np.asarray(rgrid, dtype=np.float64)
            nr = len(rgrid)
            p_r = np.stack([_interp_edge(rp, pv[i], rgrid) for i in range(len(zp))])  # (Nz_table, Nr)
            pre = np.empty((nz, no, nr), dtype=np.complex128)
            for r in range(nr):
                pmod = _interp_edge(zp, p_r[:, r], zgrid)
                cum_p = _cumtrapz(p_r[:, r], zp)
                for o in range(no):
                    if include_dispersion:
                        integ = _cumtrapz(dispersion_function(ogrid[o], p_r[:, r], dgas, n_ir=n_ir), zp)
                        disp = 1j * _interp_edge(zp, integ, zgrid)
                    else:
                        disp = 1j * (n_ir - 1.0) / C_LIGHT
                    if include_absorption:
                        absf = (1.0 / C_LIGHT) * _interp_edge(zp, beta_factor_ref(ogrid[o], dgas) * cum_p, zgrid)
                    else:
                        absf = 0.0
                    pre[:, o, r] = pmod * np.exp(ogrid[o] * (disp + absf))
            return pre, None
        if has_r:
            raise NotImplementedError("radial-only pressure modulation")
        raise ValueError("Pressure modulation wrongly specified.")
    # scalar pressure, :405-448
    if include_dispersion:
        disp = 1j * dispersion_function(ogrid, pressure, dgas, n_ir=n_ir)
    else:
        disp = 1j * (n_ir - 1.0) / C_LIGHT
    if include_absorption:
        absf = (pressure / C_LIGHT) * beta_factor_ref(ogrid, agas)
    else:
        absf = 0.0
    absw = ogrid * absf
    renorm = np.exp((zgrid[-1] - zgrid)[:, None] * absw)
    pre = pressure * np.exp(ogrid[None, :] * (zgrid[:, None] * disp + (zgrid - zgrid[-1])[:, None] * absf))
    return pre.astype(np.complex128), renorm * np.ones((1, no))


def hankel_plane(ogrid, rgrid, ffield, distance, rgrid_ff, near_field_factor=True, pre=None):
    """HankelTransform (Hankel_transform.py:452-519) with the trapezoidal integrator:
    ffield[o, r] complex; pre None (factor 1) or pre[r, o].  Returns (No, Nr_FF) complex."""
    ogrid = np.asarray(ogrid, dtype=np.float64)
    rgrid = np.asarray(rgrid, dtype=np.float64)
    rff = np.asarray(rgrid_ff, dtype=np.float64)
    out = np.empty((len(ogrid), len(rff)), dtype=np.complex128)
    for o in range(len(ogrid)):
        k = ogrid[o] / C_LIGHT
        src = rgrid * ffield[o, :]
        if pre is not None:
            src = pre[:, o] * src
        if near_field_factor:
            src = np.exp(-1j * k * (rgrid ** 2) / (2.0 * distance)) * src
        J = special.jv(0, k * rgrid[None, :] * rff[:, None] / distance)  # (Nr_FF, Nr)
        y = J * src[None, :]
        out[o, :] = np.sum(0.5 * (rgrid[1:] - rgrid[:-1])[None, :] * (y[:, 1:] + y[:, :-1]), axis=1)
    return out


def provider_planes(zgrid, rgrid, ogrid, fsource, ko_min=0, ko_step=1, ko_max=None, kr_step=1, kr_max=None,
                    kz_step=1):
    """FSources_provider with data_source='static' (Hankel_transform.py:46-70): the last z
    plane is dropped (zgrid[0:-1:kz_step])."""
    ko_max = len(ogrid) if ko_max is None else ko_max
    kr_max = len(rgrid) if kr_max is None else kr_max
    z = np.asarray(zgrid)[0:-1:kz_step]
    r = np.asarray(rgrid)[0:kr_max:kr_step]
    o = np.asarray(ogrid)[ko_min:ko_max:ko_step]
    planes = [fsource[k * kz_step, ko_min:ko_max:ko_step, 0:kr_max:kr_step] for k in range(len(z))]
    return z, r, o, planes


def hankel_long(zgrid, rgrid, ogrid, planes, distance, rgrid_ff, preset_gas="vacuum", pressure=1.0,
                absorption_tables="Henke", include_absorption=True, dispersion_tables="Henke",
                include_dispersion=True, n_ir=1.0, near_field_factor=True, cumulative=False):
    """Hankel_long.__init__ (Hankel_transform.py:531-687) for given provider grids and planes.
    Returns dict with FF_integrated (Nr_FF, No), entry/exit plane transforms and, if asked,
    cumulative_field (Nz-1, Nr_FF, No) renormalised with renorm[k1] as the reference does."""
    pre, renorm = prefactor_tables(zgrid, rgrid, ogrid, preset_gas, pressure, absorption_tables,
                                   include_absorption, dispersion_tables, include_dispersion, n_ir)
    nz = len(zgrid)

    def pre_ro(k):
        if pre.ndim == 3:
            return pre[k].T  # (r, o)
        return np.outer(np.ones(len(rgrid)), pre[k])

    P = [hankel_plane(ogrid, rgrid, planes[k], distance - zgrid[k], rgrid_ff, near_field_factor, pre_ro(k)).T
         for k in range(nz)]
    ff = 0.0
    cum = []
    for k in range(nz - 1):
        ff = ff + 0.5 * (zgrid[k + 1] - zgrid[k]) * (P[k] + P[k + 1])
        if cumulative:
            cum.append(np.outer(np.ones(len(rgrid_ff)), renorm[k]) * ff)
    res = {"FF_integrated": ff, "entry_plane_transform": P[0], "exit_plane_transform": P[-1], "planes": P}
    if cumulative:
        res["cumulative_field"] = np.array(cum)
    return res
