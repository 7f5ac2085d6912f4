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


class LaunchConfig(ctypes.Structure):
    """Launch shape of the persistent propagation kernel (ctdse_launch_config): waves per
    workgroup (a segment of the parallel solve is 64 * waves_per_group chunks of 8 points),
    resident workgroups per compute unit, and time steps per pass over the state (1 or 2)."""
    _fields_ = [("waves_per_group", c_int), ("groups_per_cu", c_int), ("steps_per_pass", c_int)]

    def __init__(self, waves_per_group=4, groups_per_cu=3, steps_per_pass=1):
        super().__init__(int(waves_per_group), int(groups_per_cu), int(steps_per_pass))

    def as_tuple(self):
        return (self.waves_per_group, self.groups_per_cu, self.steps_per_pass)

    def __repr__(self):
        return "LaunchConfig(waves_per_group=%d, groups_per_cu=%d, steps_per_pass=%d)" % self.as_tuple()


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
        L.ctdse_propagate_one.restype = c_int
        L.ctdse_propagate_one.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p]
        L.ctdse_propagate_one_noalloc.restype = c_int
        L.ctdse_propagate_one_noalloc.argtypes = [c_void_p, c_int, c_void_p, c_void_p, c_void_p]
        L.ctdse_geometry_create_batch.restype = c_int
        L.ctdse_geometry_create_batch.argtypes = [c_int, POINTER(c_int)] + [POINTER(c_double)] * 6 + [
            POINTER(c_void_p), POINTER(c_int)]
        L.ctdse_launch_config_default.restype = c_int
        L.ctdse_launch_config_default.argtypes = [c_int, POINTER(c_void_p), POINTER(LaunchConfig)]
        L.ctdse_batch_plan_create.restype = c_int
        L.ctdse_batch_plan_create.argtypes = [c_int, POINTER(c_void_p), POINTER(c_int), c_void_p, POINTER(c_longlong),
                                              c_void_p, POINTER(c_longlong), c_longlong, POINTER(LaunchConfig),
                                              POINTER(c_void_p)]
        L.ctdse_batch_plan_enqueue.restype = c_int
        L.ctdse_batch_plan_enqueue.argtypes = [c_void_p, c_void_p]
        L.ctdse_batch_plan_graph_launch.restype = c_int
        L.ctdse_batch_plan_graph_launch.argtypes = [c_void_p, c_void_p]
        L.ctdse_batch_plan_info.restype = c_int
        L.ctdse_batch_plan_info.argtypes = [c_void_p, POINTER(LaunchConfig), POINTER(c_int)]
        L.ctdse_batch_plan_destroy.restype = None
        L.ctdse_batch_plan_destroy.argtypes = [c_void_p]
        L.ctdse_state_length.restype = c_longlong
        L.ctdse_state_length.argtypes = [c_void_p]
        L.ctdse_propagate_states.restype = c_int
        L.ctdse_propagate_states.argtypes = [c_int, POINTER(c_void_p), POINTER(c_int), POINTER(c_int), POINTER(c_int),
                                             c_void_p, POINTER(c_longlong), c_void_p, POINTER(c_longlong), c_void_p,
                                             POINTER(c_longlong), c_int, c_void_p, POINTER(c_longlong), c_void_p]
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
        self._info()

    def _info(self):
        e = c_double()
        it = c_int()
        halo = c_int()
        _raise(_lib().ctdse_geometry_info(self._h, byref(e), byref(it), byref(halo)))
        self.e_init, self.gs_iterations, self.halo_points = e.value, it.value, halo.value

    @classmethod
    def create_many(cls, num_r, dx, dt, a, e_guess=-0.5792, cv=1e-25, x_int=2.0, allow_unconverged=False):
        """Geometries for many grids at once (ctdse_geometry_create_batch): the ground states
        run side by side on the device.  num_r is a sequence; the other arguments are scalars
        or sequences of the same length.  Returns a list of Geometry."""
        n = len(num_r)

        def col(v, ctype):
            vals = list(v) if hasattr(v, "__len__") else [v] * n
            if len(vals) != n:
                raise ValueError("every per-geometry argument needs len(num_r) entries")
            return (ctype * max(n, 1))(*vals)

        L = _lib()
        hs = (c_void_p * max(n, 1))()
        st = (c_int * max(n, 1))()
        code = L.ctdse_geometry_create_batch(n, col([int(v) for v in num_r], c_int), col(dx, c_double),
                                             col(dt, c_double), col(a, c_double), col(e_guess, c_double),
                                             col(cv, c_double), col(x_int, c_double), hs, st)
        ok = [st[i] == 0 or (st[i] == 1 and allow_unconverged and hs[i]) for i in range(n)]
        if code != 0 and not all(ok):
            msg = L.ctdse_last_error_message().decode()
            for i in range(n):
                if hs[i]:
                    L.ctdse_geometry_destroy(c_void_p(hs[i]))
            raise CTDSEError(code, msg)
        out = []
        for i in range(n):
            g = cls.__new__(cls)
            g._h = c_void_p(hs[i])
            g.status = st[i]
            g.num_r = int(num_r[i])
            g.dx, g.dt, g.a, g.x_int = (float(v[i]) if hasattr(v, "__len__") else float(v) for v in (dx, dt, a, x_int))
            g._info()
            out.append(g)
        return out

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
    if out.shape != (n_inst, 5, width):
        raise ValueError("out must be (n_inst, 5, max(n_steps))")
    if min(n_steps) != width:
        # rows of width doubles, not n_steps[i]: a plan carries the row stride
        plan = BatchPlan(geoms, fields, n_steps=n_steps, out=out)
        plan.enqueue(stream)
        (stream if stream is not None else torch.cuda.current_stream(fields.device)).synchronize()
        plan.close()
        return out
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


class BatchPlan:
    """A prepared batch (ctdse_batch_plan): descriptors uploaded and the work buffer
    allocated once; enqueue() queues the persistent launch with no allocation, copy or wait
    (torch.cuda.graph can capture it) and replay() launches a HIP graph of the same work that
    the library captured on first use.  Arguments as propagate_batch; config is a
    LaunchConfig or a (waves_per_group, groups_per_cu[, steps_per_pass]) tuple (None: the
    library default).
    Runs read `fields` and write `out` in place, so refilling `fields` and replaying
    propagates new fields.  Runs of one plan must not overlap."""

    def __init__(self, geometry, fields, n_steps=None, out=None, config=None):
        _check_device(fields, "fields")
        n_inst, L_f = fields.shape
        if n_steps is None:
            n_steps = [L_f - 1] * n_inst
        n_steps = [int(v) for v in n_steps]
        if len(n_steps) != n_inst or n_inst < 1 or max(n_steps) + 1 > L_f or min(n_steps) < 1:
            raise ValueError("each row needs n_steps + 1 field samples")
        geoms = geometry if isinstance(geometry, (list, tuple)) else [geometry] * n_inst
        if len(geoms) != n_inst:
            raise ValueError("one geometry per instance")
        width = max(n_steps)
        if out is None:
            out = torch.empty((n_inst, 5, width), dtype=torch.float64, device=fields.device)
        _check_device(out, "out")
        if out.shape != (n_inst, 5, width):
            raise ValueError("out must be (n_inst, 5, max(n_steps))")
        self.fields, self.out, self.n_steps, self.geoms = fields, out, n_steps, geoms
        field_off = (c_longlong * n_inst)(*[i * L_f for i in range(n_inst)])
        out_off = (c_longlong * n_inst)(*[i * 5 * width for i in range(n_inst)])
        handles = (c_void_p * n_inst)(*[g.handle.value for g in geoms])
        cfg = None
        if config is not None:
            cfg = config if isinstance(config, LaunchConfig) else LaunchConfig(*[int(v) for v in config])
        self._p = c_void_p()
        _raise(_lib().ctdse_batch_plan_create(n_inst, handles, (c_int * n_inst)(*n_steps), c_void_p(fields.data_ptr()),
                                              field_off, c_void_p(out.data_ptr()), out_off, width,
                                              byref(cfg) if cfg is not None else None, byref(self._p)))
        info = LaunchConfig()
        wg = c_int()
        _raise(_lib().ctdse_batch_plan_info(self._p, byref(info), byref(wg)))
        self.config, self.workgroups = info, wg.value
        self.pt_steps = sum(int(g.num_r + 1) * s for g, s in zip(geoms, n_steps))

    def _stream(self, stream):
        return stream if stream is not None else torch.cuda.current_stream(self.fields.device)

    def enqueue(self, stream=None):
        _raise(_lib().ctdse_batch_plan_enqueue(self._p, c_void_p(self._stream(stream).cuda_stream)))
        return self.out

    def replay(self, stream=None):
        _raise(_lib().ctdse_batch_plan_graph_launch(self._p, c_void_p(self._stream(stream).cuda_stream)))
        return self.out

    def close(self):
        if getattr(self, "_p", None) is not None and self._p.value:
            _lib().ctdse_batch_plan_destroy(self._p)
            self._p = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_config(geometries):
    """The launch shape ctdse_propagate_batch picks for these geometries."""
    n = len(geometries)
    cfg = LaunchConfig()
    _raise(_lib().ctdse_launch_config_default(n, (c_void_p * n)(*[g.handle.value for g in geometries]), byref(cfg)))
    return cfg


def propagate_stepwise(geometry, fields, out=None, stream=None):
    """The plain form of the batched propagation, for comparison with the persistent launch:
    one kernel launch per time step over the whole batch, each instance's state held in HBM
    between launches and stepped in place (ctdse_propagate_states).  Same arguments and
    result as propagate_batch for a uniform batch (n_steps = L - 1 for every row), bit for
    bit: the same segment solve and observable order run either way."""
    from .adjoint import gate
    _check_device(fields, "fields")
    L = _lib()
    n_inst, L_f = fields.shape
    n_steps = L_f - 1
    if out is None:
        out = torch.empty((n_inst, 5, n_steps), dtype=torch.float64, device=fields.device)
    _check_device(out, "out")
    F = gate(fields)[0].contiguous()
    S = int(L.ctdse_state_length(geometry.handle))
    states = torch.zeros((n_inst, S, 2), dtype=torch.float64, device=fields.device)
    if stream is None:
        stream = torch.cuda.current_stream(fields.device)
    handles = (c_void_p * n_inst)(*([geometry.handle.value] * n_inst))
    steps = (c_int * n_inst)(*([n_steps] * n_inst))
    ones = (c_int * n_inst)(*([1] * n_inst))
    f_off = (c_longlong * n_inst)(*[i * L_f for i in range(n_inst)])
    s_off = (c_longlong * n_inst)(*[i * 2 * S for i in range(n_inst)])
    o_off = (c_longlong * n_inst)(*[i * 5 * n_steps for i in range(n_inst)])
    st = c_void_p(states.data_ptr())
    for k in range(n_steps):
        k0 = (c_int * n_inst)(*([k] * n_inst))
        init = st if k > 0 else c_void_p()  # step 0 starts from psi0 and writes the state vector
        _raise(L.ctdse_propagate_states(n_inst, handles, steps, k0, ones, c_void_p(F.data_ptr()), f_off, init,
                                        s_off, st, s_off, 1, c_void_p(out.data_ptr()), o_off,
                                        c_void_p(stream.cuda_stream)))
    return out


def propagate_one(geometry, field, n_steps=None, out=None, stream=None):
    """One instance alone (ctdse_propagate_one): the propagation call1DTDSE runs, with the
    segments of each step side by side.  field: device float64 samples (at least n_steps + 1;
    default n_steps = len(field) - 1); returns out, (5, n_steps) rows as propagate_batch's.
    Equal to propagate_batch's row to rounding, not bitwise."""
    field = field.reshape(-1)
    _check_device(field, "field")
    n_steps = field.numel() - 1 if n_steps is None else int(n_steps)
    if n_steps < 1 or field.numel() < n_steps + 1:
        raise ValueError("field needs n_steps + 1 samples")
    if out is None:
        out = torch.empty((5, n_steps), dtype=torch.float64, device=field.device)
    _check_device(out, "out")
    if out.numel() < 5 * n_steps:
        raise ValueError("out needs 5 * n_steps doubles")
    if stream is None:
        stream = torch.cuda.current_stream(field.device)
    _raise(_lib().ctdse_propagate_one(geometry.handle, n_steps, c_void_p(field.data_ptr()), c_void_p(out.data_ptr()),
                                      c_void_p(stream.cuda_stream)))
    return out


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


__all__ = ["Geometry", "BatchPlan", "LaunchConfig", "default_config", "propagate_batch", "propagate_ragged",
           "propagate_stepwise", "propagate_one",
           "pad_fields", "OUTPUT_ROWS", "STATUS_NAMES"]
