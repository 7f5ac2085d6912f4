"""Batched TDSE on the GPU, the replacement of the MPI task farm's loop over (r, z)
(1DTDSE/sources/MPI-TDSE.c:148-329): many independent instances, each with its own
field, in one persistent kernel launch.

Device memory and streams come from PyTorch; the compute is the HIP kernel behind the
C ABI ``ctdse_propagate_batch`` of libsingleTDSE.so.  Results follow the reference's
call1DTDSE conventions for k1 = 1 (the field is sampled at the TDSE step): an instance
with N_E field samples runs N_E + 2 steps and exports N_E + 2 samples of
sourceterm, PopTot, PopInt, expval and the gated Efield.
"""
import ctypes
from ctypes import POINTER, byref, c_double, c_int, c_longlong, c_void_p

import numpy as np
import torch

from . import load_library
from .ctdse.dll import CTDSEError, STATUS_NAMES

OUTPUT_ROWS = ("sourceterm", "PopTot", "PopInt", "expval", "Efield")


def _lib():
    L = load_library()
    if not getattr(L, "_tdse_bound", False):
        L.ctdse_geometry_create.restype = c_int
        L.ctdse_geometry_create.argtypes = [c_int, c_double, c_double, c_double, c_double, c_double, c_double,
                                            POINTER(c_double), POINTER(c_void_p)]
        L.ctdse_geometry_info.restype = c_int
        L.ctdse_geometry_info.argtypes = [c_void_p, POINTER(c_double), POINTER(c_int), POINTER(c_int)]
        L.ctdse_geometry_ground_state.restype = c_int
        L.ctdse_geometry_ground_state.argtypes = [c_void_p, POINTER(c_double), POINTER(c_double)]
        L.ctdse_geometry_destroy.restype = None
        L.ctdse_geometry_destroy.argtypes = [c_void_p]
        L.ctdse_propagate_batch.restype = c_int
        L.ctdse_propagate_batch.argtypes = [c_int, POINTER(c_void_p), POINTER(c_int), c_void_p, POINTER(c_longlong),
                                            c_void_p, POINTER(c_longlong), c_void_p]
        L.ctdse_last_error_message.restype = ctypes.c_char_p
        L._tdse_bound = True
    return L


def _raise(code):
    if code != 0:
        raise CTDSEError(code, _lib().ctdse_last_error_message().decode())


class Geometry:
    """Grid, ground state and factor tables of one (num_r, dx, dt, a, x_int), on the device.

    psi0=None computes the ground state by inverse iteration from (e_guess, cv) exactly as
    Initialise_grid_and_ground_state (1DTDSE/sources/tools.c:31-51)."""

    def __init__(self, num_r, dx, dt, a, e_guess=-0.5792, cv=1e-25, x_int=2.0, psi0=None, allow_unconverged=False):
        L = _lib()
        h = c_void_p()
        p0 = None
        if psi0 is not None:
            z = np.asarray(psi0, dtype=np.complex128)
            flat = np.empty(2 * z.size)
            flat[0::2] = z.real
            flat[1::2] = z.imag
            p0 = flat.ctypes.data_as(POINTER(c_double))
        code = L.ctdse_geometry_create(num_r, dx, dt, a, e_guess, cv, x_int, p0, byref(h))
        self._h = h
        if code != 0 and not (code == 1 and allow_unconverged and h.value):
            if h.value:
                L.ctdse_geometry_destroy(h)
                self._h = c_void_p()
            _raise(code)
        self.status = code
        self.num_r, self.dx, self.dt, self.a, self.x_int = num_r, dx, dt, a, x_int
        e = c_double()
        it = c_int()
        halo = c_int()
        _raise(L.ctdse_geometry_info(h, byref(e), byref(it), byref(halo)))
        self.e_init, self.gs_iterations, self.halo_points = e.value, it.value, halo.value

    @property
    def handle(self):
        return self._h

    def ground_state(self):
        n = self.num_r + 1
        x = np.zeros(n)
        psi = np.zeros(2 * n)
        _raise(_lib().ctdse_geometry_ground_state(self._h, x.ctypes.data_as(POINTER(c_double)),
                                                  psi.ctypes.data_as(POINTER(c_double))))
        return x, psi[0::2] + 1j * psi[1::2]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib().ctdse_geometry_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check_device(t, name):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise TypeError("%s must be a contiguous float64 CUDA tensor" % name)


def propagate_batch(geometry, fields, n_steps=None, out=None, stream=None):
    """Propagate a uniform-length batch.

    geometry: one Geometry for all rows, or a list with one per row.
    fields:   (n_inst, L) float64 CUDA tensor of in_field samples; row i holds at least
              n_steps[i] + 1 samples (the reference's zero padding included).
    n_steps:  steps per instance (the reference's outputs.Nt); default L - 1.
    Returns a (n_inst, 5, max_steps) float64 tensor, rows in OUTPUT_ROWS order; instance i
    uses columns [0, n_steps[i]).
    """
    _check_device(fields, "fields")
    n_inst, L_f = fields.shape
    if n_steps is None:
        n_steps = [L_f - 1] * n_inst
    n_steps = [int(s) for s in n_steps]
    if len(n_steps) != n_inst or max(n_steps) + 1 > L_f or min(n_steps) < 1:
        raise ValueError("each row needs n_steps + 1 field samples")
    geoms = geometry if isinstance(geometry, (list, tuple)) else [geometry] * n_inst
    if len(geoms) != n_inst:
        raise ValueError("one geometry per instance")
    width = max(n_steps)
    if out is None:
        out = torch.empty((n_inst, 5, width), dtype=torch.float64, device=fields.device)
    _check_device(out, "out")
    field_off = (c_longlong * n_inst)(*[i * L_f for i in range(n_inst)])
    out_off = (c_longlong * n_inst)(*[i * 5 * width for i in range(n_inst)])
    return _launch(geoms, n_steps, fields, field_off, out, out_off, stream, lambda: out)


def propagate_ragged(geometries, fields, n_steps=None, stream=None):
    """Propagate instances of different lengths: fields is a list of 1-D float64 CUDA
    tensors (each with n_steps[i] + 1 samples).  Returns a list of (5, n_steps[i]) tensors."""
    n_inst = len(fields)
    if n_steps is None:
        n_steps = [int(f.numel()) - 1 for f in fields]
    lens = [int(f.numel()) for f in fields]
    for f in fields:
        _check_device(f, "fields[i]")
    flat = torch.cat(fields) if n_inst else torch.zeros(0, dtype=torch.float64, device="cuda")
    f_off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64) if n_inst else np.zeros(0, np.int64)
    o_len = [5 * s for s in n_steps]
    o_off = np.concatenate([[0], np.cumsum(o_len)[:-1]]).astype(np.int64) if n_inst else np.zeros(0, np.int64)
    out = torch.empty(int(sum(o_len)), dtype=torch.float64, device=flat.device)
    field_off = (c_longlong * n_inst)(*f_off.tolist())
    out_off = (c_longlong * n_inst)(*o_off.tolist())

    def split():
        return [out[o_off[i]:o_off[i] + o_len[i]].view(5, n_steps[i]) for i in range(n_inst)]

    return _launch(list(geometries), [int(s) for s in n_steps], flat, field_off, out, out_off, stream, split)


def _launch(geoms, n_steps, fields, field_off, out, out_off, stream, result):
    n_inst = len(n_steps)
    handles = (c_void_p * n_inst)(*[g.handle.value for g in geoms])
    steps = (c_int * n_inst)(*n_steps)
    if stream is None:
        stream = torch.cuda.current_stream(fields.device)
    code = _lib().ctdse_propagate_batch(n_inst, handles, steps, c_void_p(fields.data_ptr()), field_off,
                                        c_void_p(out.data_ptr()), out_off, c_void_p(stream.cuda_stream))
    _raise(code)
    return result()


def pad_fields(fields_np, extra=3):
    """Append the reference's zero tail (FourInterp's calloc) to (n_inst, N_E) field samples:
    the propagation reads in_field up to index N_E + 2 (prop.c:125-133)."""
    f = np.asarray(fields_np, dtype=np.float64)
    if f.ndim == 1:
        f = f[None, :]
    return np.concatenate([f, np.zeros((f.shape[0], extra))], axis=1)


__all__ = ["Geometry", "propagate_batch", "propagate_ragged", "pad_fields", "OUTPUT_ROWS", "STATUS_NAMES"]
