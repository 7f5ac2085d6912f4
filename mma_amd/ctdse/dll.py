"""``TDSE_DLL``: the reference's ctypes wrapper class (1DTDSE/PythonTDSE.py:18-220) with the
same methods and argument conventions, bound to this package's libsingleTDSE.so.

Unlike the reference library, which calls exit() on fatal errors, this library records a
status; ``check_status`` turns it into a ``CTDSEError`` after each call.
"""
from ctypes import POINTER, c_char_p, c_double, c_int

import numpy as np

from .. import load_library, library_path
from .structures import (inputs_def, outputs_def, output_print_def, ctypes_arr_ptr, ctype_arr_to_numpy,
                         ctype_mtrx_to_numpy_rows)

STATUS_NAMES = {0: "ok", 1: "ground state not converged", 2: "zero pivot", 3: "bad argument", 4: "unsupported",
                5: "device error", 6: "non-finite value"}


class CTDSEError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("%s (%d): %s" % (STATUS_NAMES.get(code, "error"), code, message))
        self.code = code


def check_status(lib=None):
    lib = lib or load_library()
    lib.ctdse_last_error.restype = c_int
    lib.ctdse_last_error_message.restype = c_char_p
    code = lib.ctdse_last_error()
    if code != 0:
        raise CTDSEError(code, lib.ctdse_last_error_message().decode())


class TDSE_DLL:
    """Functions of the CTDSE library (PythonTDSE.py:18-220)."""

    def __init__(self, path_to_DLL=None, raise_on_error=True):
        self.DLL = load_library() if path_to_DLL in (None, library_path()) else __import__("ctypes").CDLL(path_to_DLL)
        self.raise_on_error = raise_on_error

    def _check(self):
        if self.raise_on_error:
            check_status(self.DLL)

    def init_GS(self, inputs):
        f = self.DLL.Initialise_grid_and_ground_state
        f.restype = None
        f.argtypes = [POINTER(inputs_def)]
        f(inputs.ptr)
        self._check()

    def call1DTDSE(self, inputs, outputs):
        f = self.DLL.call1DTDSE
        f.restype = None
        f.argtypes = [POINTER(inputs_def), POINTER(outputs_def)]
        f(inputs.ptr, outputs.ptr)
        self._check()

    def compute_PES(self, inputs, psi, E_start=-0.6, num_E=10000, epsilon=5e-4, Estep=5e-4):
        """Photoelectron spectrum (PythonTDSE.py:68-98)."""
        f = self.DLL.window_analysis
        f.restype = POINTER(c_double)
        f.argtypes = [inputs_def, POINTER(c_double), c_int, c_double, c_double, c_double]
        if isinstance(psi, np.ndarray):
            flat = np.empty(2 * len(psi))
            flat[0::2] = psi.real
            flat[1::2] = psi.imag
            psi = ctypes_arr_ptr(c_double, len(flat), flat)
        res = f(inputs, psi, c_int(num_E), c_double(epsilon), c_double(Estep), c_double(E_start))
        self._check()
        E_grid = np.linspace(E_start, E_start + (num_E - 1) * Estep, num_E)
        res_np = ctype_arr_to_numpy(res, num_E)
        self.free_arr(res)
        return E_grid, res_np

    def gabor_transform(self, signal, dt, N, omega_max, t_min, t_max, N_G, a=8.):
        """Gabor transform (PythonTDSE.py:100-135)."""
        f = self.DLL.GaborTransform
        f.restype = POINTER(POINTER(c_double))
        f.argtypes = [POINTER(c_double), c_double, c_int, c_int, c_int, c_double, c_double, c_double]
        omegas = 2 * np.pi * np.fft.fftfreq(N, dt)[0:N // 2]
        omega_range = (omegas <= omega_max)
        N_freq = len(omegas[omega_range])
        if isinstance(signal, (list, np.ndarray)):
            signal = ctypes_arr_ptr(c_double, len(signal), signal)
        res = f(signal, c_double(dt), c_int(N), c_int(N_freq), c_int(N_G), c_double(t_min), c_double(t_max),
                c_double(a))
        self._check()
        gabor_res = ctype_mtrx_to_numpy_rows(res, N_G, N_freq)
        self.free_mtrx(res, N_G)
        return np.linspace(t_min, t_max, N_G), omegas[omega_range], np.transpose(gabor_res)

    def free_mtrx(self, buffer_ptr, N_rows):
        self.DLL.free_mtrx.restype = None
        self.DLL.free_mtrx.argtypes = [POINTER(POINTER(c_double)), c_int]
        self.DLL.free_mtrx(buffer_ptr, c_int(N_rows))

    def free_arr(self, buffer_ptr):
        self.DLL.free_arr.restype = None
        self.DLL.free_arr.argtypes = [POINTER(c_double)]
        self.DLL.free_arr(buffer_ptr)

    def free_outputs(self, out_ptr):
        self.DLL.outputs_destructor.restype = None
        self.DLL.outputs_destructor.argtypes = [POINTER(outputs_def)]
        self.DLL.outputs_destructor(out_ptr)

    def free_inputs(self, in_ptr):
        self.DLL.inputs_destructor.restype = None
        self.DLL.inputs_destructor.argtypes = [POINTER(inputs_def)]
        self.DLL.inputs_destructor(in_ptr)

    def set_time_and_field(self, in_ptr, time, field, N):
        self.DLL.set_time_and_field.restype = None
        self.DLL.set_time_and_field.argtypes = [POINTER(inputs_def), POINTER(c_double), POINTER(c_double), c_int]
        time_ptr = ctypes_arr_ptr(c_double, N, time)
        field_ptr = ctypes_arr_ptr(c_double, N, field)
        self.DLL.set_time_and_field(in_ptr, time_ptr, field_ptr, N)
        self._check()

    def set_all_prints(self):
        self.DLL.Set_all_prints.restype = output_print_def
        return self.DLL.Set_all_prints()
