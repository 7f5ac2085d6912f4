"""ctypes mirrors of the CTDSE structs, laid out like 1DTDSE/sources/structures.h:16-147
(and include/ctdse.h), with the methods of the reference's Python binding
1DTDSE/PythonCTDSE/structures.py:18-650: same names, defaults and behaviour.

Array helpers copy with numpy instead of element-by-element loops
(1DTDSE/PythonCTDSE/ctypes_helper.py:14-136); the values are the same.  HDF5 archives are read
and written with the package's own mma_amd.h5lite (h5py is not needed); arrays loaded from an
archive are allocated with the C library's malloc, so the library's destructors free them as they
free the arrays it allocates itself.
"""
import ctypes
import os
from ctypes import (CDLL, POINTER, Structure, byref, c_char, c_double, c_int, c_void_p, cast)
from typing import Any

import numpy as np

from .. import h5lite

# C unit set of PythonCTDSE/constants.py (the same as 1DTDSE/sources/constants.c:26-43)
_HBAR = 1.054571800e-34
_C = 299792458.0
_ELCHARGE = 1.602176565e-19
_ELMASS = 9.10938356e-31
_EPS0 = 1.0 / (4.0 * np.pi * 1e-7 * _C * _C)
_R_BOHR = 4.0 * np.pi * _EPS0 * _HBAR * _HBAR / (_ELMASS * _ELCHARGE * _ELCHARGE)
TIMEau = (_ELMASS * _R_BOHR * _R_BOHR) / _HBAR
EFIELDau = _HBAR * _HBAR / (_ELMASS * _R_BOHR ** 3 * _ELCHARGE)
CUPRAD_OUTPUTS = "CUPRAD/outputs"  # MMA_administration.paths['CUPRAD_outputs']

_libc = ctypes.CDLL(None)
_libc.malloc.restype = c_void_p
_libc.malloc.argtypes = [ctypes.c_size_t]


def _c_doubles(values):
    """A malloc'd C double array holding `values` (freed by the library's destructors)."""
    a = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
    p = _libc.malloc(max(a.nbytes, 8))
    if not p:
        raise MemoryError("malloc of %d bytes" % a.nbytes)
    ctypes.memmove(p, a.ctypes.data, a.nbytes)
    return cast(p, POINTER(c_double))


def _c_rows(rows):
    """A malloc'd double** of malloc'd rows (freed by free_mtrx)."""
    p = _libc.malloc(max(len(rows), 1) * ctypes.sizeof(c_void_p))
    if not p:
        raise MemoryError("malloc of the row table")
    table = cast(p, POINTER(POINTER(c_double)))
    for i, r in enumerate(rows):
        table[i] = _c_doubles(r)
    return table


def _archive(filename):
    """Datasets and attributes of an existing HDF5 file (h5py's mode 'a' appends to it), or
    empty ones."""
    if os.path.exists(filename):
        return h5lite.read_all(filename)
    return {}, {}


def _create_block(data, items):
    """h5py create_dataset calls in a try block that stops at the first name that exists (the
    reference catches that ValueError and skips the rest of the block)."""
    for path, dtype, value in items:
        if path in data:
            return
        data[path] = np.asarray(value, dtype=dtype)


def _scalar(f, path):
    return np.asarray(f[path]).reshape(-1)[0].item()


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
        """Field and time grid from arrays, or from the CUPRAD outputs of an HDF5 archive
        (structures.py:248-299): plane z_i, radius r_i of CUPRAD/outputs/output_field, in atomic
        units (tgrid / TIMEau, field / EFIELDau).  The field cube is (z, t, r), as the C driver
        reads it (MPI-TDSE.c:99-101,185), so the samples are output_field[z_i, :, r_i]; the
        reference's Python loader indexes [z_i, r_i, :], which does not match that layout.  An
        index out of range prints the reference's message and sets nothing."""
        if filename != "" and (E is None or t is None):
            f = h5lite.H5File(filename)
            cube = f[CUPRAD_OUTPUTS + "/output_field"]
            if z_i < 0 or z_i >= cube.shape[0]:
                print("Incorrect z-grid dimension selection. Select z in range (0, {})".format(cube.shape[0] - 1))
                return
            if r_i < 0 or r_i >= cube.shape[2]:
                print("Incorrect r-grid dimension selection. Select r in range (0, {})".format(cube.shape[2] - 1))
                return
            t = np.asarray(f[CUPRAD_OUTPUTS + "/tgrid"], dtype=np.float64) / TIMEau
            E = np.asarray(cube[z_i, :, r_i], dtype=np.float64) / EFIELDau
        Nt = len(t)
        assert Nt == len(E)
        self.Efield.Nt = Nt
        DLL.set_time_and_field(self.ptr, t, E, Nt)

    def save_to_hdf5(self, filename):
        """The inputs under TDSE_inputs/ of an HDF5 file (structures.py:301-355): appended to the
        file if it exists; the same names and types as the reference (float32 'f', int32 'i');
        the field and time grid only when set, the ground state, x grid and Einit only when
        Einit != 0; a block stops at the first dataset that already exists."""
        data, attrs = _archive(filename)
        P = "/TDSE_inputs/"
        _create_block(data, [(P + "trg_a", "f", self.trg.a), (P + "dx", "f", self.dx), (P + "dt", "f", self.dt),
                             (P + "Eguess", "f", self.Eguess), (P + "N_r_grid", "i", self.num_r),
                             (P + "gauge_type", "i", self.gauge), (P + "Ntinterp", "i", self.Ntinterp),
                             (P + "InterpByDTorNT", "i", self.InterpByDTorNT),
                             (P + "analy_writewft", "i", self.analy.writewft),
                             (P + "analy_tprint", "f", self.analy.tprint), (P + "CV_criterion_of_GS", "f", self.CV),
                             (P + "x_int", "f", self.x_int), (P + "num_t", "i", self.num_t)])
        if self.Efield.Nt != 0:
            _create_block(data, [(P + "Nt", "i", self.Efield.Nt), (P + "Efield", "f", self.get_Efield()),
                                 (P + "tgrid", "f", self.get_tgrid())])
        if self.Einit != 0.:
            gs = self.get_GS()
            _create_block(data, [(P + "psi0", "f", np.array([gs.real, gs.imag]).transpose()),
                                 (P + "x_grid", "f", self.get_xgrid()), (P + "Einit", "f", self.Einit)])
        h5lite.write(filename, data, attrs)

    def load_from_hdf5(self, filename):
        """The inputs from TDSE_inputs/ of an HDF5 file (structures.py:133-175): the scalars, then
        the x grid, ground state and Einit, and the field and time grid where they are stored."""
        f = h5lite.H5File(filename)
        P = "TDSE_inputs/"
        self.Eguess = c_double(_scalar(f, P + "Eguess"))
        self.num_r = c_int(int(_scalar(f, P + "N_r_grid")))
        self.dx = c_double(_scalar(f, P + "dx"))
        self.InterpByDTorNT = c_int(int(_scalar(f, P + "InterpByDTorNT")))
        self.dt = c_double(_scalar(f, P + "dt"))
        self.Ntinterp = c_int(int(_scalar(f, P + "Ntinterp")))
        self.analy.writewft = c_int(int(_scalar(f, P + "analy_writewft")))
        self.analy.tprint = c_double(_scalar(f, P + "analy_tprint"))
        self.x_int = c_double(_scalar(f, P + "x_int"))
        self.trg.a = c_double(_scalar(f, P + "trg_a"))
        self.CV = c_double(_scalar(f, P + "CV_criterion_of_GS"))
        self.gauge = c_int(int(_scalar(f, P + "gauge_type")))
        self.precision = b"d"
        try:
            x_grid = f[P + "x_grid"]
            psi0 = np.asarray(f[P + "psi0"]).reshape(-1)
            einit = _scalar(f, P + "Einit")
            self.x = _c_doubles(x_grid[:self.num_r + 1])
            self.psi0 = _c_doubles(psi0[:2 * (self.num_r + 1)])
            self.Einit = c_double(einit)
        except KeyError:
            pass
        try:
            nt = int(_scalar(f, P + "Nt"))
            field, tgrid = f[P + "Efield"], f[P + "tgrid"]
            self.Efield.Nt = c_int(nt)
            self.Efield.Field = _c_doubles(field[:nt])
            self.Efield.tgrid = _c_doubles(tgrid[:nt])
        except KeyError:
            pass

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

    def save_to_hdf5(self, filename, inputs=None):
        """The outputs under TDSE_outputs/ of an HDF5 file (structures.py:454-500): appended to
        the file if it exists; float32 series, complex spectra as (Nomega, 2); with `inputs`
        whose writewft is set, the stored wavefunctions as psi_re and psi_im."""
        data, attrs = _archive(filename)
        P = "/TDSE_outputs/"
        fe, fs = self.get_FEfield(), self.get_Fsourceterm()
        _create_block(data, [(P + "Nomega", "i", self.Nomega), (P + "Nt", "i", self.Nt),
                             (P + "tgrid", "f", self.get_tgrid()), (P + "Efield", "f", self.get_Efield()),
                             (P + "sourceterm", "f", self.get_sourceterm()),
                             (P + "omegagrid", "f", self.get_omegagrid()),
                             (P + "FEfield", "f", np.array([fe.real, fe.imag]).transpose()),
                             (P + "Fsourceterm", "f", np.array([fs.real, fs.imag]).transpose()),
                             (P + "PopTot", "f", self.get_PopTot()), (P + "PopInt", "f", self.get_PopInt()),
                             (P + "expval", "f", self.get_expval())])
        if inputs is not None and inputs.analy.writewft and self.psi:
            wf = self.get_wavefunction(inputs, grids=False)
            _create_block(data, [(P + "psi_re", "f", wf.real), (P + "psi_im", "f", wf.imag)])
        h5lite.write(filename, data, attrs)

    def load_from_hdf5(self, filename):
        """The outputs from TDSE_outputs/ of an HDF5 file (structures.py:502-546); the
        wavefunction rows where psi_re and psi_im are stored.  Prints the reference's message and
        returns when the file holds no outputs."""
        f = h5lite.H5File(filename)
        P = "TDSE_outputs/"
        try:
            nomega, nt = int(_scalar(f, P + "Nomega")), int(_scalar(f, P + "Nt"))
            arrays = {k: np.asarray(f[P + k], dtype=np.float64).reshape(-1)
                      for k in ("tgrid", "Efield", "sourceterm", "expval", "omegagrid", "FEfield", "Fsourceterm",
                                "PopTot", "PopInt")}
        except KeyError:
            print("No output data stored in the HDF5 file.")
            return
        self.Nomega = c_int(nomega)
        self.Nt = c_int(nt)
        for k in ("tgrid", "Efield", "sourceterm", "expval", "PopTot", "PopInt"):
            setattr(self, k, _c_doubles(arrays[k][:nt]))
        self.omegagrid = _c_doubles(arrays["omegagrid"][:nomega])
        self.FEfield = _c_doubles(arrays["FEfield"][:2 * nomega])
        self.Fsourceterm = _c_doubles(arrays["Fsourceterm"][:2 * nomega])
        try:
            re, im = np.asarray(f[P + "psi_re"], dtype=np.float64), np.asarray(f[P + "psi_im"], dtype=np.float64)
        except KeyError:
            return
        rows = np.empty((re.shape[0], 2 * re.shape[1]))
        rows[:, 0::2], rows[:, 1::2] = re, im
        self.psi = _c_rows(list(rows))
