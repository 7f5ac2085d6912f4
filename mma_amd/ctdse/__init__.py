"""Python face of the CTDSE library, mirroring the reference's ``1DTDSE/PythonTDSE.py``
and ``1DTDSE/PythonCTDSE`` package so that code written against them runs unchanged."""
from .structures import (Efield_var, trg_def, analy_def, output_print_def, inputs_def, outputs_def,
                         ctypes_arr_ptr, ctype_arr_to_numpy, ctype_cmplx_arr_to_numpy, get_wavefunction,
                         ctype_mtrx_to_numpy_rows)
from .dll import TDSE_DLL, CTDSEError, check_status

__all__ = ["Efield_var", "trg_def", "analy_def", "output_print_def", "inputs_def", "outputs_def", "TDSE_DLL",
           "CTDSEError", "check_status", "ctypes_arr_ptr", "ctype_arr_to_numpy", "ctype_cmplx_arr_to_numpy",
           "get_wavefunction"]
