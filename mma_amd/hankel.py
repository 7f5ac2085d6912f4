"""Hankel far-field transform on the GPU with the interface of Hankel/Hankel_transform.py.

Same names, arguments, return values and errors as the reference module:
``trapezoidal_integrator``, ``FSources_provider``, ``get_propagation_pre_factor_function``,
``HankelTransform``, ``Hankel_long``, ``Signal_cum_integrator``.  The work runs in
libmmahankel.so (include/hankel.h): the pre-factor tables, the FP64 contraction over r and
the source planes, and the gridded contraction on the matrix cores.

``Hankel_long`` takes one keyword the reference does not have, ``contraction``:
  'fp64'         (default) J0 in FP64 for every (plane, omega, rho, r) element;
  'fp64-grid'    the gridded form (Lagrange interpolation of J0 in q = K r / D from a uniform
                 grid, FP64 spreading, FP64 GEMM);
  'fp32-mfma'    the gridded form with the GEMM in FP32 on the matrix cores;
  'bf16x3-mfma'  the gridded form with the GEMM as three bf16 products (hi*hi + hi*lo + lo*hi)
                 on the matrix cores;
  'bf16x6-mfma'  the same with a three-way bf16 split and six products;
  'fp16x3-mfma'  each operand row scaled by a power of two and split into two fp16 parts
                 (22 bits), three fp16 products on the matrix cores.
The matrix-core GEMMs accumulate 32- or 64-deep stages in FP32 and the stages in FP64.
Only the trapezoidal radial integrator is accelerated; any other integrator raises
NotImplementedError (the reference accepts an arbitrary callable).
"""
import ctypes
import math
import os
from ctypes import POINTER, Structure, byref, c_double, c_int, c_longlong, c_void_p

import numpy as np
import torch

from . import HERE, LIB_DIR, NativeLibraryMissing

LIB_PATH = os.path.join(LIB_DIR, "libmmahankel.so")
TABLES = os.path.join(HERE, "data", "xuv_tables.npz")
C_LIGHT = 299792458.0  # shared_python/units.py:3

HK_OK, HK_ERR_ARGUMENT, HK_ERR_UNSUPPORTED, HK_ERR_DEVICE, HK_ERR_TABLE = 0, 3, 4, 5, 7
HK_F64, HK_F32, HK_BF16X2, HK_BF16X3, HK_F16X2 = 0, 1, 2, 3, 4
HK_GEMM_F32, HK_GEMM_BF16X3, HK_GEMM_BF16X6 = 1, 2, 3
GRID_TOL = 1e-10  # bound on the J0 interpolation error of the gridded contraction
CONTRACTIONS = ("fp64", "fp64-grid", "fp32-mfma", "bf16x3-mfma", "bf16x6-mfma", "fp16x3-mfma")


class hk_medium(Structure):
    _fields_ = [("mode", c_int), ("pressure", c_double), ("np_z", c_int), ("np_r", c_int), ("p_z", c_void_p),
                ("p_r", c_void_p), ("p_value", c_void_p), ("include_dispersion", c_int),
                ("include_absorption", c_int), ("n_ir", c_double), ("n_f1", c_int), ("n_f2", c_int),
                ("e_f1", c_void_p), ("f1", c_void_p), ("e_f2", c_void_p), ("f2", c_void_p)]


class hk_planes(Structure):
    _fields_ = [("nz", c_int), ("no", c_int), ("nr", c_int), ("nrho", c_int), ("src", c_void_p), ("src_c64", c_int),
                ("plane_stride", c_longlong), ("zgrid", c_void_p), ("ogrid", c_void_p), ("rgrid", c_void_p),
                ("tw", c_void_p), ("rho", c_void_p), ("distance", c_double), ("near_field", c_int),
                ("pre", c_void_p), ("pre_mode", c_int), ("host_zgrid", c_void_p), ("host_ogrid", c_void_p),
                ("host_rgrid", c_void_p)]


_lib = None


def _L():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing("%s is missing: build it with `python -m mma_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.hk_last_error_message.restype = ctypes.c_char_p
        L.hk_prefactor.argtypes = [POINTER(hk_medium), c_int, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p,
                                   c_void_p, c_void_p]
        L.hk_prefactor_enqueue.argtypes = [POINTER(hk_medium), c_int, c_void_p, c_int, c_void_p, c_int, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
        L.hk_prefactor_status.argtypes = [c_int]
        L.hk_direct.argtypes = [POINTER(hk_planes), c_int, c_void_p, c_int, c_void_p, c_void_p]
        L.hk_grid_size.argtypes = [c_double, c_double, c_double, POINTER(c_int), POINTER(c_double)]
        L.hk_grid_spread.argtypes = [POINTER(hk_planes), c_int, c_void_p, c_int, c_double, c_int, c_void_p,
                                     c_longlong, c_longlong, c_void_p]
        L.hk_grid_bessel.argtypes = [c_int, c_void_p, c_int, c_double, c_int, c_void_p, c_longlong, c_void_p]
        L.hk_gemm.argtypes = [c_int, c_int, c_int, c_int, c_void_p, c_longlong, c_void_p, c_longlong, c_void_p,
                              c_longlong, c_void_p]
        L.hk_gemm_split.argtypes = [c_int, c_int, c_int, c_int, c_void_p, c_longlong, c_void_p, c_longlong, c_void_p,
                                    c_longlong, c_void_p]
        L.hk_gemm_f16x2.argtypes = [c_int, c_int, c_int, c_void_p, c_longlong, c_void_p, c_longlong, c_void_p,
                                    c_longlong, c_void_p]
        _lib = L
    return _lib


def _check(code):
    if code == HK_OK:
        return
    msg = _L().hk_last_error_message().decode()
    if code == HK_ERR_TABLE or code == HK_ERR_ARGUMENT:
        raise ValueError(msg)
    if code == HK_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg)


_DEV_SMALL = {}  # device copies of small host arrays (grids, weights), keyed by their bytes
_DEV_SMALL_MAX = 64


def _dev(a, dtype=torch.float64):
    """A device copy of `a`.  Small host arrays (grids, plane weights) are cached by content,
    so repeated calls with the same grids upload nothing; the cached tensors are only read."""
    if isinstance(a, torch.Tensor):
        return a.to(device="cuda", dtype=dtype).contiguous()
    h = np.ascontiguousarray(np.asarray(a, dtype=torch.empty((), dtype=dtype).numpy().dtype))
    if h.size > 65536:
        return torch.as_tensor(h, device="cuda")
    key = (h.dtype.str, h.shape, h.tobytes(), torch.cuda.current_device())
    t = _DEV_SMALL.get(key)
    if t is None:
        if len(_DEV_SMALL) >= _DEV_SMALL_MAX:
            _DEV_SMALL.pop(next(iter(_DEV_SMALL)))
        t = _DEV_SMALL[key] = torch.as_tensor(h, device="cuda")
    return t


def _host(a):
    """A contiguous float64 host array for the ABI's optional host copies (None for tensors on
    the device, which the library then copies back itself)."""
    if isinstance(a, torch.Tensor):
        return None if a.is_cuda else np.ascontiguousarray(a.numpy(), dtype=np.float64)
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _ptr(t):
    return c_void_p(t.data_ptr()) if t is not None else c_void_p()


def _stream():
    return c_void_p(torch.cuda.current_stream().cuda_stream)


_xuv = None


def _xuv_table(name):
    global _xuv
    if _xuv is None:
        with np.load(TABLES, allow_pickle=False) as z:
            _xuv = {k: z[k] for k in z.files}
    return _xuv


def gases():
    """The '<gas>_<tables>' keys of XUV_refractive_index.gases."""
    return sorted({k.split("/")[0] for k in _xuv_table(None)})


def trapezoidal_integrator(y, x):
    """Hankel_transform.py:22-27: trapezoid of y over the abscissae x."""
    y = np.asarray(y)
    x = np.asarray(x)
    return np.sum(0.5 * (x[1:] - x[:-1]) * (y[1:] + y[:-1]))


def _known_trapezoids():
    known = [trapezoidal_integrator]
    for mod, name in ((np, "trapezoid"), (np, "trapz")):
        if hasattr(mod, name):
            known.append(getattr(mod, name))
    try:
        from scipy import integrate
        known.append(integrate.trapezoid)
    except ImportError:
        pass
    return known


_TRAPEZOID_PROBED = {}


def _is_trapezoid(integrator):
    """True for this module's trapezoidal_integrator, numpy's and scipy's trapezoid, and any other
    callable integrator(y, x) that returns the trapezoid of y over the given x on a probe: complex
    values on a non-uniform grid (e.g. Hankel_transform.trapezoidal_integrator).  The verdict is
    kept per callable."""
    if any(integrator is k for k in _known_trapezoids()):
        return True
    key = id(integrator)
    if key not in _TRAPEZOID_PROBED:
        x = np.array([0.0, 0.3, 0.35, 1.1, 1.7, 2.0])
        y = np.array([1.0, -2.0 + 1j, 0.5 - 3j, 4.0, -1.5j, 2.5 + 0.25j])
        want = np.sum(0.5 * (x[1:] - x[:-1]) * (y[1:] + y[:-1]))
        try:
            got = complex(integrator(y, x))
            ok = abs(got - want) <= 1e-12 * abs(want)
        except Exception:
            ok = False
        _TRAPEZOID_PROBED[key] = (integrator, ok)  # the callable is held so its id stays unique
    return _TRAPEZOID_PROBED[key][1]


def _trapezoid_weights(x):
    """w with sum_j w_j y_j = sum_j (x_{j+1} - x_j)(y_j + y_{j+1}) / 2."""
    x = np.asarray(x, dtype=np.float64)
    w = np.zeros_like(x)
    d = 0.5 * (x[1:] - x[:-1])
    w[:-1] += d
    w[1:] += d
    return w


class FSources_provider:
    """Hankel_transform.py:29-80.  zgrid[0:-1:kz_step] drops the last plane, as the reference
    does.  FSource may also be a torch tensor (on the GPU it is used in place)."""

    def __init__(self, zgrid, rgrid, ogrid, data_source='static', FSource=None, h5_handle=None, h5_path=None,
                 ko_min=0, ko_step=1, ko_max='end', kr_step=1, kr_max='end', kz_step=1):
        if ko_max == 'end':
            ko_max = len(ogrid)
        if kr_max == 'end':
            kr_max = len(rgrid)
        self.zgrid = zgrid[0:-1:kz_step]
        self.rgrid = rgrid[0:kr_max:kr_step]
        self.ogrid = ogrid[ko_min:ko_max:ko_step]
        self._static = None
        if data_source == 'static':
            nz = len(self.zgrid)
            self._static = (FSource, kz_step, (ko_min, ko_max, ko_step), (kr_max, kr_step), nz)

            def planes():
                for k1 in range(nz):
                    yield FSource[k1 * kz_step, ko_min:ko_max:ko_step, 0:kr_max:kr_step]
            self.Fsource_plane = planes()
        elif data_source == 'dynamic':
            def planes():
                for k1 in range(len(self.zgrid)):
                    re = h5_handle[h5_path][k1 * kz_step, 0:kr_max:kr_step, ko_min:ko_max:ko_step, 0]
                    im = h5_handle[h5_path][k1 * kz_step, 0:kr_max:kr_step, ko_min:ko_max:ko_step, 1]
                    if isinstance(re, torch.Tensor):  # a mapping of device tensors (mma_amd.farm.run)
                        yield torch.complex(re.to(torch.float64), im.to(torch.float64)).squeeze().T
                    else:
                        yield np.squeeze(re + 1j * im).T
            self.Fsource_plane = planes()
        else:
            raise ValueError('Wrongly specified input of the class.')

    def device_planes(self):
        """All planes as one contiguous (Nz, No, Nr) CUDA tensor: complex64 sources stay
        complex64, anything else becomes complex128."""
        def cast(t):
            return t if t.dtype in (torch.complex64, torch.complex128) else t.to(torch.complex128)
        if self._static is not None:
            F, kz_step, (o0, o1, os_), (r1, rs), nz = self._static
            if isinstance(F, torch.Tensor):
                return cast(F[0:nz * kz_step:kz_step, o0:o1:os_, 0:r1:rs].to(device="cuda")).contiguous()
            sel = np.asarray(F)[0:nz * kz_step:kz_step, o0:o1:os_, 0:r1:rs]
            return cast(torch.as_tensor(np.ascontiguousarray(sel), device="cuda"))
        def dev(p):
            if isinstance(p, torch.Tensor):
                return p.to(device="cuda")
            return torch.as_tensor(np.ascontiguousarray(p), device="cuda")
        return torch.stack([cast(dev(p)) for p in self.Fsource_plane]).contiguous()


_DEV_TABLES = {}


def _dev_table(t, key):
    """Device copy of one XUV table column, uploaded once per process and device."""
    dev = torch.cuda.current_device()
    if (dev, key) not in _DEV_TABLES:
        _DEV_TABLES[(dev, key)] = _dev(t[key])
    return _DEV_TABLES[(dev, key)]


def _host1d(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy().astype(np.float64).ravel()
    return np.asarray(x, dtype=np.float64).ravel()


class _PreFactor:
    """The tables behind get_propagation_pre_factor_function: pre (device, complex
    (Nz, No) or (Nz, No, Nr)), renorm (device (Nz, No) or None).  A table error (a photon energy
    outside the scattering-factor tables, or no f1 listed) raises here, or, with defer=True, in
    check(): the tables are then only queued, so work queued after them does not wait for the
    host."""

    def __init__(self, zgrid, rgrid, ogrid, preset_gas, pressure, absorption_tables, include_absorption,
                 dispersion_tables, include_dispersion, n_ir, defer=False):
        if isinstance(pressure, dict) and 'zgrid' in pressure:
            # the tables are walked in one forward pass: ascending plane grid, strictly ascending
            # pressure grids (the reference's interp1d and cumulative_trapezoid assume the same)
            if np.any(np.diff(_host1d(zgrid)) < 0):
                raise ValueError("zgrid must be ascending for a tabulated pressure")
            if np.any(np.diff(_host1d(pressure['zgrid'])) <= 0):
                raise ValueError("pressure['zgrid'] must be strictly ascending")
            if 'rgrid' in pressure and np.any(np.diff(_host1d(pressure['rgrid'])) <= 0):
                raise ValueError("pressure['rgrid'] must be strictly ascending")
        zg, rg, og = _dev(zgrid), _dev(rgrid), _dev(ogrid)
        nz, nr, no = len(zg), len(rg), len(og)
        med = hk_medium()
        med.include_dispersion = int(bool(include_dispersion))
        med.include_absorption = int(bool(include_absorption))
        med.n_ir = float(n_ir)
        keep = []
        if isinstance(pressure, dict):
            has_z, has_r = 'zgrid' in pressure, 'rgrid' in pressure
            if has_z and not has_r:
                med.mode = 1
            elif has_z and has_r:
                med.mode = 2
            elif has_r:
                raise NotImplementedError("radial-only pressure modulation (Hankel_transform.py:244-296 is not "
                                          "functional in the reference)")
            else:
                raise ValueError('Pressure modulation wrongly specified.')
            pz = _dev(pressure['zgrid'])
            pv = _dev(np.asarray(pressure['value'], dtype=np.float64))
            med.np_z, med.p_z, med.p_value = len(pz), _ptr(pz), _ptr(pv)
            keep += [pz, pv]
            if med.mode == 2:
                pr = _dev(pressure['rgrid'])
                med.np_r, med.p_r = len(pr), _ptr(pr)
                keep.append(pr)
        else:
            med.mode = 0
            med.pressure = float(pressure)
        # the z-r branch reads beta from the dispersion tables (Hankel_transform.py:373-376)
        dgas = preset_gas + '_' + dispersion_tables
        agas = preset_gas + '_' + (dispersion_tables if med.mode == 2 else absorption_tables)
        for flag, gas, ek, fk, ns in ((include_dispersion, dgas, "Energy_f1", "f1", "1"),
                                      (include_absorption, agas, "Energy_f2", "f2", "2")):
            if not flag:
                continue
            t = _xuv_table(None)
            if gas + "/" + ek not in t:
                raise KeyError(gas)
            e, f = _dev_table(t, gas + "/" + ek), _dev_table(t, gas + "/" + fk)
            keep += [e, f]
            setattr(med, "n_f" + ns, len(e))
            setattr(med, "e_f" + ns, _ptr(e))
            setattr(med, "f" + ns, _ptr(f))
        shape = (nz, no, nr) if med.mode == 2 else (nz, no)
        self.pre = torch.empty(shape, dtype=torch.complex128, device="cuda")
        self.renorm = None if med.mode == 2 else torch.empty((nz, no), dtype=torch.float64, device="cuda")
        self._flag = None
        if defer:
            self._flag = torch.empty(1, dtype=torch.int32, device="cuda")
            self._exit_scratch = torch.empty(no, dtype=torch.float64, device="cuda")
            self._keep = keep + [zg, og, rg]  # what the queued kernels read
            _check(_L().hk_prefactor_enqueue(byref(med), nz, _ptr(zg), no, _ptr(og), nr, _ptr(rg), _ptr(self.pre),
                                             _ptr(self.renorm), _ptr(self._flag), _ptr(self._exit_scratch),
                                             _stream()))
        else:
            _check(_L().hk_prefactor(byref(med), nz, _ptr(zg), no, _ptr(og), nr, _ptr(rg), _ptr(self.pre),
                                     _ptr(self.renorm), _stream()))
        self.mode = med.mode
        self.nr = nr
        self._host = None

    def check(self):
        """Raises the deferred table error, if any (waits for the tables)."""
        if self._flag is not None:
            _check(_L().hk_prefactor_status(int(self._flag.item())))

    def host(self):
        if self._host is None:
            self._host = (self.pre.cpu().numpy(), None if self.renorm is None else self.renorm.cpu().numpy())
        return self._host


def get_propagation_pre_factor_function(zgrid, rgrid, ogrid, preset_gas='vacuum', pressure=1.,
                                        absorption_tables='Henke', include_absorption=True,
                                        dispersion_tables='Henke', include_dispersion=True,
                                        effective_IR_refrective_index=1.):
    """Hankel_transform.py:85-448: returns (pre_factor, exp_renorm), functions of the plane
    index; pre_factor(kz) is (Nr, No) complex, exp_renorm(kz) is (No,) real (None for the
    z-and-r pressure table, as in the reference)."""
    tab = _PreFactor(zgrid, rgrid, ogrid, preset_gas, pressure, absorption_tables, include_absorption,
                     dispersion_tables, include_dispersion, effective_IR_refrective_index)
    nr = len(rgrid)

    def pre_factor(kz):
        pre, _ = tab.host()
        if tab.mode == 2:
            return pre[kz].T  # (Nr, No)
        return np.outer(np.ones(nr), pre[kz])

    def exp_renorm(kz):
        return tab.host()[1][kz]

    pre_factor.tables = tab
    return pre_factor, (None if tab.mode == 2 else exp_renorm)


def _planes_struct(src, zgrid, ogrid, rgrid, rho, distance, near_field, pre, pre_mode, keep):
    zg, og, rg, rh = _dev(zgrid), _dev(ogrid), _dev(rgrid), _dev(rho)
    tw = _dev(_trapezoid_weights(np.asarray(rgrid, dtype=np.float64)))
    keep += [zg, og, rg, rh, tw, src]
    p = hk_planes()
    p.nz, p.no, p.nr, p.nrho = src.shape[0], src.shape[1], src.shape[2], len(rh)
    p.src = _ptr(src)
    p.src_c64 = int(src.dtype == torch.complex64)
    p.plane_stride = src.shape[1] * src.shape[2]
    p.zgrid, p.ogrid, p.rgrid, p.tw, p.rho = _ptr(zg), _ptr(og), _ptr(rg), _ptr(tw), _ptr(rh)
    p.distance = float(distance)
    p.near_field = int(bool(near_field))
    p.pre = _ptr(pre)
    p.pre_mode = pre_mode
    hz, ho, hr = _host(zgrid), _host(ogrid), _host(rgrid)
    keep += [hz, ho, hr]
    p.host_zgrid, p.host_ogrid, p.host_rgrid = (c_void_p(x.ctypes.data) if x is not None else c_void_p()
                                                for x in (hz, ho, hr))
    return p


def direct(src, zgrid, ogrid, rgrid, rho, distance, weights=None, near_field=True, pre=None, per_plane=False):
    """FP64 contraction of the device planes src (Nz, No, Nr): returns (n_out, Nrho, No) complex
    sums over planes with the given (n_out, Nz) weights, or (Nz, Nrho, No) per plane."""
    keep = []
    pre_mode = 0 if pre is None else (1 if pre.dim() == 2 else 2)
    p = _planes_struct(src, zgrid, ogrid, rgrid, rho, distance, near_field, pre, pre_mode, keep)
    nz = src.shape[0]
    if per_plane:
        out = torch.empty((nz, p.nrho, p.no), dtype=torch.complex128, device="cuda")
        _check(_L().hk_direct(byref(p), 0, c_void_p(), 1, _ptr(out), _stream()))
        return out
    w = _dev(np.atleast_2d(weights))
    out = torch.empty((w.shape[0], p.nrho, p.no), dtype=torch.complex128, device="cuda")
    _check(_L().hk_direct(byref(p), w.shape[0], _ptr(w), 0, _ptr(out), _stream()))
    return out


def grid_size_of(zgrid, ogrid, rgrid, rho, distance, tol=GRID_TOL):
    """(nq, h) of the q grid the gridded contraction uses for these grids."""
    og = np.asarray(ogrid, dtype=np.float64)
    zg = np.asarray(zgrid, dtype=np.float64)
    q_max = float(np.max(og) / C_LIGHT / np.min(distance - zg) * np.max(np.abs(rgrid)))
    rho_max = float(np.max(np.abs(rho)))
    nq, h = c_int(), c_double()
    _check(_L().hk_grid_size(q_max if q_max > 0 else 1.0, rho_max if rho_max > 0 else 1.0, tol, byref(nq), byref(h)))
    return nq.value, h.value


def gridded(src, zgrid, ogrid, rgrid, rho, distance, weights, near_field=True, pre=None, kind="fp32-mfma",
            tol=GRID_TOL):
    """Gridded contraction: (n_out, Nrho, No) complex sums over planes with (n_out, Nz) weights.
    The B operand holds the spread sources of every output (2 No rows each)."""
    if pre is not None and pre.dim() != 2:
        raise NotImplementedError("the gridded contraction needs an r-independent pre-factor")
    keep = []
    p = _planes_struct(src, zgrid, ogrid, rgrid, rho, distance, near_field, pre, 0 if pre is None else 1, keep)
    L = _L()
    nq, h = grid_size_of(zgrid, ogrid, rgrid, rho, distance, tol)
    W = np.atleast_2d(np.asarray(weights, dtype=np.float64))
    n_out = W.shape[0]
    # operands: FP64 (library GEMM), FP32 (FP32 MFMA), pre-split bf16 parts (bf16 MFMA) or scaled
    # fp16 parts with their row scale (fp16 MFMA)
    otype = {"fp64-grid": HK_F64, "fp32-mfma": HK_F32, "bf16x3-mfma": HK_BF16X2, "bf16x6-mfma": HK_BF16X3,
             "fp16x3-mfma": HK_F16X2}[kind]
    parts = {HK_BF16X2: 2, HK_BF16X3: 3, HK_F16X2: 2}.get(otype, 1)
    ld = parts * nq + (8 if otype == HK_F16X2 else 0)  # row: the parts (and the scale, 4 words padded to 8)
    dt = {HK_F64: torch.float64, HK_F32: torch.float32}.get(otype, torch.int16)
    Bm = torch.empty((n_out * 2 * p.no, ld), dtype=dt, device="cuda")
    wq = _dev(np.ascontiguousarray(W))
    _check(L.hk_grid_spread(byref(p), n_out, _ptr(wq), nq, h, otype, _ptr(Bm), ld, 0, _stream()))
    rh = _dev(rho)
    Am = torch.empty((p.nrho, ld), dtype=dt, device="cuda")
    _check(L.hk_grid_bessel(p.nrho, _ptr(rh), nq, h, otype, _ptr(Am), ld, _stream()))
    if kind == "fp64-grid":
        out = Am @ Bm.T  # plain FP64 library GEMM
    else:
        out = torch.empty((p.nrho, Bm.shape[0]), dtype=torch.float64, device="cuda")
        if otype == HK_F16X2:
            _check(L.hk_gemm_f16x2(p.nrho, Bm.shape[0], nq, _ptr(Am), ld, _ptr(Bm), ld, _ptr(out), Bm.shape[0],
                                   _stream()))
        elif parts > 1:
            _check(L.hk_gemm_split(parts, p.nrho, Bm.shape[0], nq, _ptr(Am), ld, _ptr(Bm), ld, _ptr(out),
                                   Bm.shape[0], _stream()))
        else:
            _check(L.hk_gemm(HK_GEMM_F32, p.nrho, Bm.shape[0], nq, _ptr(Am), nq, _ptr(Bm), nq, _ptr(out), Bm.shape[0],
                             _stream()))
    # a strided view (n_out, Nrho, No) of the GEMM's rows: no copy into output order
    return torch.view_as_complex(out.view(p.nrho, n_out, p.no, 2).permute(1, 0, 2, 3))


def HankelTransform(ogrid, rgrid, FField, distance, rgrid_FF, integrator=trapezoidal_integrator,
                    near_field_factor=True, pre_factor=1.):
    """Hankel_transform.py:452-519 for one plane: FField[omega, r]; returns (No, Nr_FF)
    complex128.  A 2-D pre_factor[r, omega] multiplies the integrand; a scalar is ignored, as
    in the reference (:488,498-499)."""
    if not _is_trapezoid(integrator):
        raise NotImplementedError("only the trapezoidal radial integrator is accelerated")
    src = torch.as_tensor(np.ascontiguousarray(FField), device="cuda").to(torch.complex128)[None]
    pre = None
    if len(np.shape(pre_factor)) == 2:
        pre = torch.as_tensor(np.ascontiguousarray(np.asarray(pre_factor).T), device="cuda").to(torch.complex128)[None]
    out = direct(src, np.array([0.0]), ogrid, rgrid, rgrid_FF, distance, weights=np.ones((1, 1)),
                 near_field=near_field_factor, pre=pre)
    return out[0].T.cpu().numpy()


class Hankel_long:
    """Hankel_transform.py:522-687: the far field integrated over the source planes, with the
    same attributes (FF_integrated (Nr_FF, No), entry/exit_plane_transform, cumulative_field,
    cumulative_field_no_norm, rgrid, ogrid, zgrid)."""

    def __init__(self, target, distance, rgrid_FF, preset_gas='vacuum', pressure=1., absorption_tables='Henke',
                 include_absorption=True, dispersion_tables='Henke', include_dispersion=True,
                 effective_IR_refrective_index=1., integrator_Hankel=trapezoidal_integrator,
                 integrator_longitudinal='trapezoidal', near_field_factor=True, store_cumulative_result=False,
                 store_non_normalised_cumulative_result=False, store_entry_and_exit_plane_transform=True,
                 contraction='fp64'):
        if not (integrator_longitudinal == 'trapezoidal'):
            raise NotImplementedError('only logitudinal trapezoidal integrator implemented so far')
        if not _is_trapezoid(integrator_Hankel):
            raise NotImplementedError("only the trapezoidal radial integrator is accelerated")
        if contraction not in CONTRACTIONS:
            raise ValueError("contraction must be one of " + ", ".join(CONTRACTIONS))
        self.include_dispersion = include_dispersion
        if include_dispersion:
            self.dispersion_tables = dispersion_tables
        self.include_absorption = include_absorption
        if include_absorption:
            self.absorption_tables = absorption_tables
        self.effective_IR_refrective_index = effective_IR_refrective_index
        self.integrator_Hankel = integrator_Hankel
        self.integrator_longitudinal = integrator_longitudinal
        self.near_field_factor = near_field_factor
        self.contraction = contraction
        self.rgrid = rgrid_FF
        self.ogrid = np.copy(target.ogrid)
        if store_cumulative_result or store_non_normalised_cumulative_result:
            self.zgrid = np.copy(target.zgrid)
        gas_ok = (preset_gas + '_' + absorption_tables in gases() and preset_gas + '_' + dispersion_tables in gases())
        if not (gas_ok or preset_gas == 'vacuum'):
            raise ValueError('Wrongly specified preset gas (or tables).')
        zgrid = np.asarray(target.zgrid, dtype=np.float64)
        nz = len(zgrid)
        if nz < 2:
            raise ValueError("Hankel_long needs at least two source planes")
        tab = _PreFactor(zgrid, target.rgrid, target.ogrid, preset_gas, pressure, absorption_tables,
                         include_absorption, dispersion_tables, include_dispersion, effective_IR_refrective_index,
                         defer=True)
        src = target.device_planes()
        cumulative = store_cumulative_result or store_non_normalised_cumulative_result
        if store_cumulative_result and isinstance(pressure, dict) and 'rgrid' in pressure:
            raise NotImplementedError('Renormalisation of the signal is not implemented for radially modulated density.')
        dz = zgrid[1:] - zgrid[:-1]
        w_ff = np.zeros(nz)
        w_ff[:-1] += 0.5 * dz
        w_ff[1:] += 0.5 * dz
        args = (src, zgrid, target.ogrid, target.rgrid, rgrid_FF, distance)
        if cumulative:
            eye = np.eye(nz)
            if contraction == 'fp64':
                planes = direct(*args, near_field=near_field_factor, pre=tab.pre, per_plane=True)
            else:
                planes = gridded(*args, eye, near_field=near_field_factor, pre=tab.pre, kind=contraction)
            terms = 0.5 * torch.as_tensor(dz, device="cuda")[:, None, None] * (planes[:-1] + planes[1:])
            cum = torch.cumsum(terms, dim=0)
            ff, entry, exit_ = cum[-1], planes[0], planes[-1]
            tab.check()  # before any result reaches the host
            if store_cumulative_result:
                self.cumulative_field = (tab.renorm[:nz - 1][:, None, :] * cum).cpu().numpy()
            if store_non_normalised_cumulative_result:
                self.cumulative_field_no_norm = cum.cpu().numpy()
        else:
            W = np.stack([w_ff, np.eye(nz)[0], np.eye(nz)[-1]])
            if contraction == 'fp64':
                res = direct(*args, weights=W, near_field=near_field_factor, pre=tab.pre)
            else:
                res = gridded(*args, W, near_field=near_field_factor, pre=tab.pre, kind=contraction)
            ff, entry, exit_ = res[0], res[1], res[2]
            tab.check()  # before any result reaches the host
        self.FF_integrated = ff.cpu().numpy()
        if store_entry_and_exit_plane_transform:
            self.entry_plane_transform = entry.cpu().numpy()
            self.exit_plane_transform = exit_.cpu().numpy()


def Signal_cum_integrator(ogrid, zgrid, FSourceTerm, integrator=None):
    """Hankel_transform.py:691-698: the cumulative trapezoid over z of FSourceTerm[omega, z],
    signal[:, 0] = 0.  (The reference computes it and returns nothing; it is returned here.)"""
    F = np.asarray(FSourceTerm)
    z = np.asarray(zgrid, dtype=np.float64)
    signal = np.zeros((len(ogrid), len(z)), dtype=np.cdouble)
    if integrator is None:
        signal[:, 1:] = np.cumsum(0.5 * (z[1:] - z[:-1])[None, :] * (F[:, 1:] + F[:, :-1]), axis=1)
    else:
        for k1 in range(len(ogrid)):
            signal[k1, 1:] = integrator(F[k1, :], z)
    return signal


__all__ = ["trapezoidal_integrator", "FSources_provider", "get_propagation_pre_factor_function", "HankelTransform",
           "Hankel_long", "Signal_cum_integrator", "direct", "gridded", "gases", "CONTRACTIONS"]
