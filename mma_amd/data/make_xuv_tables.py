"""Writes xuv_tables.npz: the tabulated atomic scattering factors f1, f2 (Henke and NIST) of
the noble gases that shared_python/XUV_refractive_index.py:41-57 loads from
XUV_refractive_index_tables.h5, converted with the package's own HDF5 reader.

Usage: python mma_amd/data/make_xuv_tables.py [path/to/XUV_refractive_index_tables.h5]
Arrays are named '<gas>_<tables>/<field>' with field in Energy_f1, Energy_f2, f1, f2 (eV and
electrons per atom, float64, as stored).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mma_amd import h5lite  # noqa: E402

DEFAULT = "/root/reference/shared_python/XUV_refractive_index_tables.h5"


def main(path=DEFAULT):
    f = h5lite.H5File(path)
    arrays = {}
    for gas in f.keys():
        for field in ("Energy_f1", "Energy_f2", "f1", "f2"):
            arrays[gas + "/" + field] = np.asarray(f[gas + "/" + field], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "xuv_tables.npz"), **arrays)
    print("wrote", len(arrays), "arrays")


if __name__ == "__main__":
    main(*sys.argv[1:])
