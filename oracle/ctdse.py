"""Python face of the CTDSE oracle (test infrastructure only).

* ``libctdse_oracle.so`` (oracle/ctdse_oracle.c) restates the reference C core:
  ground state, Crank-Nicolson/Numerov propagation, observables.
* The FFTW-based parts of the reference are restated with numpy's FFT:
  ``four_interp`` (1DTDSE/sources/tools_fftw3.c:24-68) and ``spectrum``
  (tools_fftw3.c:84-143).
* ``call1dtdse`` repeats the orchestration of 1DTDSE/sources/singleTDSE.c:28-117.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB_DIR: an oracle build from elsewhere (the host-sanitizer run builds one with ASan)
LIB_PATH = os.path.join(os.environ.get("ORACLE_LIB_DIR") or os.path.join(HERE, "_build"), "libctdse_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_ctdse.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

STATUS = {0: "ok", 1: "ground state not converged", 2: "zero pivot", 3: "allocation failed", 4: "bad argument"}


def build():
    """Compile the oracle library (and the reference library when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/1DTDSE/sources"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_ground_state.restype = ctypes.c_int
        L.orc_ground_state.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _ip]
        L.orc_propagate.restype = ctypes.c_int
        L.orc_propagate.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, ctypes.c_int,
                                    _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_double, _dp, ctypes.c_int]
        L.orc_propagate_batch.restype = ctypes.c_int
        L.orc_propagate_batch.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, ctypes.c_int,
                                          ctypes.c_int, _dp, ctypes.c_int]
        L.orc_findinterval.restype = None
        L.orc_findinterval.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _ip, _ip]
        L.orc_potential.restype = ctypes.c_double
        L.orc_potential.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_gradpot.restype = ctypes.c_double
        L.orc_gradpot.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def ground_state(num_r, dx, a, e_guess, cv, max_iter=1000):
    """Ground state (tools.c:31-51, tridiag.c:86-159).  Returns x, psi0 (complex), E, iterations, status."""
    x = np.zeros(num_r + 1)
    psi = np.zeros(2 * (num_r + 1))
    e = np.zeros(1)
    it = ctypes.c_int(0)
    st = lib().orc_ground_state(num_r, dx, a, e_guess, cv, max_iter, _p(x), _p(psi), _p(e), ctypes.byref(it))
    return x, psi[0::2] + 1j * psi[1::2], float(e[0]), it.value, st


def findinterval(n, v, grid):
    k1, k2 = ctypes.c_int(0), ctypes.c_int(0)
    g = np.ascontiguousarray(grid, dtype=np.float64)
    lib().orc_findinterval(n, v, _p(g), ctypes.byref(k1), ctypes.byref(k2))
    return k1.value, k2.value


def _interleave(z):
    z = np.asarray(z, dtype=np.complex128)
    out = np.empty(2 * z.size)
    out[0::2] = z.real
    out[1::2] = z.imag
    return out


def propagate(num_r, dx, a, x_int, dt, psi0, x, in_field, n_steps, tmin=0.0, tprint=0.0, n_snap=0):
    """prop.c:24-347 for n_steps steps.  in_field needs n_steps+1 samples (zero-padded here)."""
    f = np.zeros(n_steps + 2)
    m = min(len(in_field), n_steps + 2)
    f[:m] = np.asarray(in_field, dtype=np.float64)[:m]
    p0 = _interleave(psi0)
    xg = np.ascontiguousarray(x, dtype=np.float64)
    outs = [np.zeros(n_steps + 1) for _ in range(6)]
    psi = np.zeros(2 * (num_r + 1))
    snaps = np.zeros((max(n_snap, 0), 2 * (num_r + 1))) if n_snap > 0 else None
    st = lib().orc_propagate(num_r, dx, a, x_int, dt, tmin, _p(p0), _p(xg), _p(f), n_steps,
                             *[_p(o) for o in outs], _p(psi), tprint,
                             _p(snaps) if snaps is not None else None, n_snap)
    keys = ["tgrid", "Efield", "sourceterm", "PopTot", "PopInt", "expval"]
    res = dict(zip(keys, outs))
    res["psi"] = psi[0::2] + 1j * psi[1::2]
    res["status"] = st
    if snaps is not None:
        res["snapshots"] = snaps[:, 0::2] + 1j * snaps[:, 1::2]
    return res


def propagate_batch(num_r, dx, a, x_int, dt, psi0, x, fields, n_steps, n_threads=0):
    """CPU baseline: one instance per thread; returns sourceterm rows (n_inst, n_steps+1)."""
    fields = np.ascontiguousarray(fields, dtype=np.float64)
    n_inst, stride = fields.shape
    assert stride >= n_steps + 1
    out = np.zeros((n_inst, n_steps + 1))
    p0 = _interleave(psi0)
    xg = np.ascontiguousarray(x, dtype=np.float64)
    st = lib().orc_propagate_batch(n_inst, num_r, dx, a, x_int, dt, _p(p0), _p(xg), _p(fields), stride,
                                   n_steps, _p(out), n_threads)
    if st != 0:
        raise RuntimeError("oracle batch failed: " + STATUS.get(st, str(st)))
    return out


def four_interp(k, signal):
    """FourInterp (tools_fftw3.c:24-68): r2c of N samples, bins/N copied into a zeroed
    spectrum of floor(kN/2)+1 bins (Nyquist bin copied unhalved), unnormalised c2r
    of length kN into a zeroed buffer of 2*(floor(kN/2)+1) doubles."""
    signal = np.asarray(signal, dtype=np.float64)
    n = signal.size
    nc = n // 2 + 1
    n2 = k * n
    nc2 = n2 // 2 + 1
    spec = np.fft.rfft(signal)
    big = np.zeros(nc2, dtype=np.complex128)
    big[:nc] = spec / n
    out = np.zeros(2 * nc2)
    out[:n2] = np.fft.irfft(big, n=n2) * n2
    return out


def spectrum(n, dt, tmax, signal):
    """calcFFTW3 (tools_fftw3.c:84-143): omega_k = k*2pi/tmax, F = conj(FFT)*dt/sqrt(2pi),
    |F|^2-type M2 = dt^2/(2pi)|FFT|^2 over floor(N/2)+1 bins."""
    sig = np.asarray(signal, dtype=np.float64)[:n]
    nc = n // 2 + 1
    out = np.fft.rfft(sig)
    dxi = 2.0 * np.pi / tmax
    c1 = dt / np.sqrt(2.0 * np.pi)
    c2 = dt * dt / (2.0 * np.pi)
    omega = np.arange(nc, dtype=np.float64) * dxi
    F = c1 * np.conj(out)
    M2 = c2 * (out.real * out.real + out.imag * out.imag)
    return omega, F, M2


def gabor(signal, dt, n, n_freq, n_g, t_min, t_max, a):
    """GaborTransform (tools_fftw3.c:158-213): rows i = 0..n_g-1 of
    dt/sqrt(2pi) |FFT(exp(-((t_j - tau_i)/a)^2) s_j)|[0:n_freq], t_j = j dt,
    tau_i = t_min + i (t_max - t_min)/n_g."""
    s = np.asarray(signal, dtype=np.float64)[:n]
    inv_a = 1.0 / a
    dt_g = (t_max - t_min) / n_g
    t = np.arange(n) * dt
    tau = t_min + np.arange(n_g) * dt_g
    rows = np.exp(-(inv_a * (t[None, :] - tau[:, None])) ** 2) * s[None, :]
    return dt / np.sqrt(2.0 * np.pi) * np.abs(np.fft.rfft(rows, axis=1)[:, :n_freq])


def refine(tgrid_in, dt, interp_by_dt_or_nt=0, nt_interp=1):
    """k1 rule of singleTDSE.c:55-68: returns (k1, refined dt, field dt)."""
    field_dt = tgrid_in[1] - tgrid_in[0]
    if interp_by_dt_or_nt == 1:
        k1 = nt_interp + 1
    else:
        k1 = int(np.floor(field_dt / dt))
        if k1 != 1:
            k1 += 1
    return k1, field_dt / float(k1), field_dt


def call1dtdse(num_r, dx, a, x_int, dt, psi0, x, tgrid_in, field_in, tprint=0.0, writewft=0,
               interp_by_dt_or_nt=0, nt_interp=1, tmin=0.0):
    """singleTDSE.c:28-117.  k1 = 1 uses the field as given (FourInterp is the identity
    there up to FFT rounding); k1 > 1 uses four_interp."""
    tgrid_in = np.asarray(tgrid_in, dtype=np.float64)
    field_in = np.asarray(field_in, dtype=np.float64)
    n_e = field_in.size
    k1, dtr, _ = refine(tgrid_in, dt, interp_by_dt_or_nt, nt_interp)
    tmax = tgrid_in[n_e - 1] - tgrid_in[0]
    if k1 == 1:
        interp = np.zeros(n_e + 3)
        interp[:n_e] = field_in
    else:
        interp = four_interp(k1, field_in)
    n_out = k1 * n_e + 2
    n_snap = 0
    if writewft:
        spd = int(np.floor(tprint / dtr))
        n_snap = n_out // spd
    res = propagate(num_r, dx, a, x_int, dtr, psi0, x, interp, n_out, tmin=tmin, tprint=tprint, n_snap=n_snap)
    for key in ["tgrid", "Efield", "sourceterm", "PopTot", "PopInt", "expval"]:
        res[key] = res[key][:n_out]
    omega, FE, FEM2 = spectrum(n_out, dtr, tmax, res["Efield"])
    _, FS, FSM2 = spectrum(n_out, dtr, tmax, res["sourceterm"])
    res.update(omegagrid=omega, FEfield=FE, FEfieldM2=FEM2, Fsourceterm=FS, FsourcetermM2=FSM2,
               Nt=n_out, Nomega=omega.size, dt=dtr, k1=k1)
    return res


# ---------------------------------------------------------------------------
# the reference library built from /root/reference sources (oracle/_ref)
# ---------------------------------------------------------------------------
_ref = None


def ref_lib():
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
        L = ctypes.CDLL(REF_PATH)
        L.ref_ground_state.restype = ctypes.c_int
        L.ref_ground_state.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _dp, _dp, _dp]
        L.ref_call.restype = ctypes.c_int
        L.ref_call.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                               _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        _ref = L
    return _ref


def ref_ground_state(num_r, dx, a, e_guess, cv):
    x = np.zeros(num_r + 1)
    psi = np.zeros(2 * (num_r + 1))
    e = np.zeros(1)
    ref_lib().ref_ground_state(num_r, dx, a, e_guess, cv, _p(x), _p(psi), _p(e))
    return x, psi[0::2] + 1j * psi[1::2], float(e[0])


def ref_call(num_r, dx, a, x_int, dt, psi0, x, tgrid_in, field_in, writewft=0, tprint=0.0):
    tgrid_in = np.ascontiguousarray(tgrid_in, dtype=np.float64)
    field_in = np.ascontiguousarray(field_in, dtype=np.float64)
    n_e = field_in.size
    n_out = n_e + 2
    outs = [np.zeros(n_out) for _ in range(6)]
    p0 = _interleave(psi0)
    xg = np.ascontiguousarray(x, dtype=np.float64)
    n_snap = 0
    snaps = None
    if writewft:
        dtr = tgrid_in[1] - tgrid_in[0]
        n_snap = n_out // int(np.floor(tprint / dtr))
        snaps = np.zeros((n_snap, 2 * (num_r + 1)))
    got = ref_lib().ref_call(num_r, dx, a, x_int, dt, _p(p0), _p(xg), _p(tgrid_in), _p(field_in), n_e,
                             writewft, tprint, *[_p(o) for o in outs],
                             _p(snaps) if snaps is not None else None, n_snap)
    if got < 0:
        raise ValueError("reference driver supports k1 = 1 only")
    keys = ["tgrid", "Efield", "sourceterm", "PopTot", "PopInt", "expval"]
    res = dict(zip(keys, outs))
    if snaps is not None:
        res["snapshots"] = snaps[:, 0::2] + 1j * snaps[:, 1::2]
    return res
