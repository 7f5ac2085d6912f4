"""ctypes mirrors of the CTDSE structs, laid out like 1DTDSE/sources/structures.h:16-147
(and include/ctdse.h), with the methods of the reference's Python binding
1DTDSE/PythonCTDSE/structures.py:18-650: same names, defaults and behaviour.

Array helpers copy with numpy instead of element-by-element loops
(1DTDSE/PythonCTDSE/ctypes_helper.py:14-136); the values are the same.
"""
from ctypes import (CDLL, POINTER, Structure, byref, c_char, c_double, c_int, cast)
from typing import Any

import numpy as np


def ctypes_arr_ptr(ctype_, size_, arr_):
    """ctypes array filled from a sequence (ctypes_helper.py:14-32)."""
    arr = (ctype_ * size_)()
    src = np.ascontiguousarray(np.asarray(arr_, dtype=np.float64 if ctype_ is c_double else None)[:size_])
    if ctype_ is c_double:
        np.ctypeslib.as_array(arr)[:len(src)] = src
    else:
        for i, v in enumerate(src):
            arr[i] = v
    return arr


def ctype_arr_to_numpy(c_arr, size):
    """numpy copy of a C double array (ctypes_helper.py:62-77)."""
    if size <= 0:
        return np.zeros(0)
    return np.ctypeslib.as_array(cast(c_arr, POINTER(c_double)), shape=(size,)).copy()


def ctype_cmplx_arr_to_numpy(c_arr, size):
    """complex numpy copy of an interleaved (re, im) C array (ctypes_helper.py:80-96)."""
    a = ctype_arr_to_numpy(c_arr, 2 * size)
    return a[0::2] + 1j * a[1::2]


def ctype_mtrx_to_numpy_rows(c_arr, n_rows, n_cols):
    """(n_rows, n_cols) numpy copy of a C double** (ctypes_helper.py:119-136)."""
    return np.array([ctype_arr_to_numpy(c_arr[i], n_cols) for i in range(n_rows)])


def get_wavefunction(c_arr, timesteps, wf_size):
    """complex (timesteps, wf_size) array from a C double** (ctypes_helper.py:99-117)."""
    return np.array([ctype_cmplx_arr_to_numpy(c_arr[i], wf_size) for i in range(timesteps)])


class Efield_var(Structure):
    _fields_ = [("tgrid", POINTER(c_double)), ("Field", POINTER(c_double)), ("dt", c_double), ("omega", c_double),
                ("E0", c_double), ("phi", c_double), ("ton", c_double), ("toff", c_double), ("Nt", c_int),
                ("nc", c_int)]


class trg_def(Structure):
    _fields_ = [("a", c_double)]


class analy_def(Structure):
    _fields_ = [("tprint", c_double), ("writewft", c_int)]


class output_print_def(Structure):
    _fields_ = [("Efield", c_int), ("FEfield", c_int), ("sourceterm", c_int), ("Fsourceterm", c_int),
                ("FEfieldM2", c_int), ("FsourceTermM2", c_int), ("PopTot", c_int), ("tgrid", c_int),
                ("omegagrid", c_int), ("PopInt", c_int), ("expval_x", c_int)]


class inputs_def(Structure):
    """Inputs of one TDSE run (structures.py:58-131)."""

    def __init__(self, *args: Any, **kw: Any):
        super().__init__(*args, **kw)
        self.ptr = byref(self)

    _fields_ = [("trg", trg_def), ("Efield", Efield_var), ("Eguess", c_double), ("Einit", c_double),
                ("tmin", c_double), ("Nt", c_int), ("num_t", c_int), ("dt", c_double), ("num_r", c_int),
                ("dx", c_double), ("psi0", POINTER(c_double)), ("x", POINTER(c_double)), ("gauge", c_int),
                ("x_int", c_double), ("analy", analy_def), ("InterpByDTorNT", c_int), ("Ntinterp", c_int),
                ("Print", output_print_def), ("CV", c_double), ("precision", c_char * 2)]

    def init_default_inputs(self, Eguess=-1., num_r=16000, dx=0.4, InterpByDTorNT=0, dt=0.25, trg_a=1.3677,
                            CV=1e-25, gauge=0, Ntinterp=1, writewft=0, tprint=10., x_int=2.,
                            precision=np.bytes_('d')):
        """Defaults of structures.py:176-237."""
        self.Eguess = c_double(Eguess)
        self.num_r = c_int(num_r)
        self.dx = c_double(dx)
        self.InterpByDTorNT = c_int(InterpByDTorNT)
        self.dt = c_double(dt)
        self.Ntinterp = c_int(Ntinterp)
        self.analy.writewft = c_int(writewft)
        self.analy.tprint = c_double(tprint)
        self.x_int = c_double(x_int)
        self.trg.a = c_double(trg_a)
        self.CV = c_double(CV)
        self.gauge = c_int(gauge)
        self.precision = precision

    def init_prints(self, path_to_DLL):
        DLL = CDLL(path_to_DLL)
        set_prints = DLL.Set_all_prints
        set_prints.restype = output_print_def
        self.Print = set_prints()

    def init_time_and_field(self, DLL, filename="", z_i=0, r_i=0, E=None, t=None):
        """Field and time grid from arrays (structures.py:248-299).  Reading the CUPRAD
        field from an HDF5 archive needs h5py, which is optional here."""
        if filename != "" and (E is None or t is None):
            raise NotImplementedError("reading the field from an HDF5 archive needs h5py; pass E and t")
        Nt = len(t)
        assert Nt == len(E)
        self.Efield.Nt = Nt
        DLL.set_time_and_field(self.ptr, t, E, Nt)

    def get_xgrid(self):
        return ctype_arr_to_numpy(self.x, self.num_r + 1)

    def get_GS(self):
        return ctype_cmplx_arr_to_numpy(self.psi0, self.num_r + 1)

    def get_tgrid(self):
        return ctype_arr_to_numpy(self.Efield.tgrid, self.Efield.Nt)

    def get_Efield(self):
        return ctype_arr_to_numpy(self.Efield.Field, self.Efield.Nt)

    def delete(self, DLL):
        DLL.free_inputs(self.ptr)


class outputs_def(Structure):
    """Outputs of one TDSE run (structures.py:393-452)."""

    def __init__(self, *args: Any, **kw: Any):
        super().__init__(*args, **kw)
        self.ptr = byref(self)

    _fields_ = [("tgrid", POINTER(c_double)), ("Efield", POINTER(c_double)), ("sourceterm", POINTER(c_double)),
                ("omegagrid", POINTER(c_double)), ("FEfield", POINTER(c_double)), ("Fsourceterm", POINTER(c_double)),
                ("FEfieldM2", POINTER(c_double)), ("FsourcetermM2", POINTER(c_double)), ("PopTot", POINTER(c_double)),
                ("PopInt", POINTER(c_double)), ("expval", POINTER(c_double)), ("Nt", c_int), ("Nomega", c_int),
                ("psi", POINTER(POINTER(c_double)))]

    def get_wavefunction(self, inputs, grids=True):
        """Stored wavefunctions (structures.py:547-583)."""
        if inputs.analy.writewft == 0:
            print("No wavefunction is stored!")
            return
        steps_per_dt = np.floor(inputs.analy.tprint / (self.tgrid[1] - self.tgrid[0]))
        size = int(self.Nt / steps_per_dt)
        wavefunction = get_wavefunction(self.psi, size, inputs.num_r + 1)
        if grids:
            if size == 1:
                tgrid = np.array([self.tgrid[self.Nt - 1]])
            else:
                t = self.get_tgrid()
                tgrid = np.linspace(t[0], t[-1], size)
            return tgrid, inputs.get_xgrid(), wavefunction
        return wavefunction

    def get_tgrid(self):
        return ctype_arr_to_numpy(self.tgrid, self.Nt)

    def get_Efield(self):
        return ctype_arr_to_numpy(self.Efield, self.Nt)

    def get_sourceterm(self):
        return ctype_arr_to_numpy(self.sourceterm, self.Nt)

    def get_PopTot(self):
        return ctype_arr_to_numpy(self.PopTot, self.Nt)

    def get_PopInt(self):
        return ctype_arr_to_numpy(self.PopInt, self.Nt)

    def get_Fsourceterm(self):
        return ctype_cmplx_arr_to_numpy(self.Fsourceterm, self.Nomega)

    def get_expval(self):
        return ctype_arr_to_numpy(self.expval, self.Nt)

    def get_omegagrid(self):
        return ctype_arr_to_numpy(self.omegagrid, self.Nomega)

    def get_FEfield(self):
        return ctype_cmplx_arr_to_numpy(self.FEfield, self.Nomega)

    def delete(self, DLL):
        DLL.free_outputs(self.ptr)
