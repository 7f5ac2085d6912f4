"""Writes tests/golden/ref_pes.npz: the photoelectron spectrum of the reference's own
window_analysis (1DTDSE/sources/tools.c:206-316, compiled from the reference sources into
oracle/_ref by oracle/Makefile) for a boosted ground state.  Run in a container that holds
/root/reference; the tests read only the fixture.

Usage: python tests/golden/make_analysis_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ctdse as O  # noqa: E402
from mma_amd import workloads as W  # noqa: E402
from mma_amd.ctdse.structures import inputs_def  # noqa: E402  (layout checked against structures.h)


def ref_window_analysis(num_r, dx, a, psi, num_E, dE, Estep, E_start):
    L = O.ref_lib()
    f = L.window_analysis
    f.restype = ctypes.POINTER(ctypes.c_double)
    f.argtypes = [inputs_def, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_double, ctypes.c_double,
                  ctypes.c_double]
    inp = inputs_def()
    inp.num_r = num_r
    inp.dx = dx
    inp.trg.a = a
    flat = np.empty(2 * len(psi))
    flat[0::2] = psi.real
    flat[1::2] = psi.imag
    res = f(inp, flat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), num_E, dE, Estep, E_start)
    out = np.array([res[i] for i in range(num_E)])
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(ctypes.cast(res, ctypes.c_void_p))
    return out


def main():
    num_r, dx, a = 1200, 0.1, W.SOFT_COULOMB_A["Ar"]
    x, psi0, e = O.ref_ground_state(num_r, dx, a, W.SOFT_COULOMB_E["Ar"], 1e-25)
    psi = psi0 * np.exp(0.7j * x)  # a boosted ground state: bound and continuum parts
    psi /= np.sqrt(np.sum(np.abs(psi) ** 2))
    cases = dict(num_E=64, dE=5e-3, Estep=0.02, E_start=-0.6)
    pes = ref_window_analysis(num_r, dx, a, psi, **cases)
    path = os.path.join(HERE, "ref_pes.npz")
    np.savez_compressed(path, num_r=num_r, dx=dx, a=a, psi_re=psi.real, psi_im=psi.imag, pes=pes,
                        **{k: np.array(v) for k, v in cases.items()})
    print("wrote", path, os.path.getsize(path), "bytes; PES range %.3e .. %.3e" % (pes.min(), pes.max()))


if __name__ == "__main__":
    main()
