"""Writes the golden fixtures under tests/golden/ (run in a container that holds
/root/reference; the fixtures, not the reference, are what the tests read).

* ref_*.npz    outputs of the reference's own CTDSE C sources compiled by
               oracle/Makefile (target ``ref``) and driven by oracle/ref_driver.c.
* ionization_fixture.npz   the arrays of the reference's test fixture
               1DTDSE/ionization.h5 (float32 as stored), read with h5lite.

Usage: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from oracle import ctdse as O  # noqa: E402
from mma_amd import workloads as W  # noqa: E402
from mma_amd import h5lite  # noqa: E402
import layout_probe  # noqa: E402

REF_H5 = "/root/reference/1DTDSE/ionization.h5"


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print("wrote", path, os.path.getsize(path), "bytes")


def gs_case(num_r, dx, a, e_guess, cv):
    x, psi0, e = O.ref_ground_state(num_r, dx, a, e_guess, cv)
    return x, psi0, e


def run_case(name, num_r, dx, a, e_guess, cv, x_int, dt, t, field, writewft=0, tprint=0.0, snap_rows=None,
             store_psi0=True):
    x, psi0, e = gs_case(num_r, dx, a, e_guess, cv)
    r = O.ref_call(num_r, dx, a, x_int, dt, psi0, x, t, field, writewft=writewft, tprint=tprint)
    arrays = dict(num_r=num_r, dx=dx, a=a, e_guess=e_guess, cv=cv, x_int=x_int, dt=dt, t=t, field=field,
                  Einit=e, **{k: r[k] for k in ["tgrid", "Efield", "sourceterm", "PopTot", "PopInt", "expval"]})
    if store_psi0:
        arrays["psi0_re"] = psi0.real
        arrays["psi0_im"] = psi0.imag
    if writewft:
        snaps = r["snapshots"]
        rows = list(range(snaps.shape[0])) if snap_rows is None else snap_rows
        arrays["n_snap"] = snaps.shape[0]
        arrays["snap_rows"] = np.array(rows)
        arrays["snap_re"] = snaps[rows].real
        arrays["snap_im"] = snaps[rows].imag
    save(name, **arrays)


def main():
    O.build()
    g = W.GRID
    # single instance: 800 nm, 10 cycles, 1.5e14 W/cm^2, Nx = 16001, dt = 0.25
    t, f = W.single_instance()
    run_case("ref_single.npz", g["num_r"], g["dx"], g["a"], g["e_guess"], g["cv"], g["x_int"], 0.25, t, f)

    # the reference unit test's case (1DTDSE/tests.py:10-52)
    E0, w0, nc = 0.14, 0.07, 4
    T = 2 * np.pi / w0
    Tmax = nc * T
    n_t = int(Tmax / 0.25) + 1
    tt = np.linspace(0, Tmax, n_t)
    ef = E0 * np.sin(np.pi * tt / Tmax) ** 2 * np.cos(w0 * tt)
    run_case("ref_tests_py.npz", 8000, 0.4, 1.0, -1.0, 1e-15, 2.0, 0.25, tt, ef, writewft=1, tprint=10.0,
             snap_rows=[0, 1, 17, 35])

    # short cases: odd and even N_E, a field with no zero at t = 0 (gating holds until the
    # first sign change), a field that starts exactly at zero, and a non-Ar atom
    rng = np.random.Generator(np.random.PCG64(7))
    for name, n_e, phase, atom in [("ref_short_even.npz", 600, 0.0, "Ar"),
                                   ("ref_short_odd.npz", 601, 0.0, "Ar"),
                                   ("ref_short_phase.npz", 500, 0.7, "Ne"),
                                   ("ref_short_kr.npz", 431, 1.9, "Kr")]:
        tt = np.arange(n_e) * 0.25
        w = W.omega_au(800e-9)
        ef = W.field_au(2e14) * np.sin(w * tt + phase) * np.minimum(1.0, tt / 40.0)
        run_case(name, 2000, 0.2, W.SOFT_COULOMB_A[atom], W.SOFT_COULOMB_E[atom], 1e-25, 2.0, 0.25, tt, ef,
                 writewft=1, tprint=20.0, snap_rows=None if n_e < 450 else [0, 3])

    # uneven-batch members at full grid size, truncated to 300 field samples
    members = W.uneven_batch(n_inst=6, seed=11)
    arrays = {}
    for i, m in enumerate(members):
        n_e = 300
        x, psi0, e = gs_case(m["num_r"], 0.1, g["a"], g["e_guess"], g["cv"])
        r = O.ref_call(m["num_r"], 0.1, g["a"], g["x_int"], 0.25, psi0, x, m["t"][:n_e], m["field"][:n_e])
        arrays["m%d_num_r" % i] = m["num_r"]
        arrays["m%d_field" % i] = m["field"][:n_e]
        arrays["m%d_Einit" % i] = e
        for k in ["sourceterm", "PopTot", "PopInt", "expval", "Efield", "tgrid"]:
            arrays["m%d_%s" % (i, k)] = r[k]
    save("ref_uneven.npz", n=len(members), **arrays)

    # ground states of all soft-Coulomb presets at CV = 1e-25 on the 16001-point grid
    # (stored as energy, every 16th point and a SHA-256 of the full psi0 bytes).  He on a
    # 4001-point grid is NOT generated: the reference's inverse iteration falls into a
    # limit cycle there and exits after 1000 iterations (tridiag.c:139-142).
    arrays = {}
    for atom in ["He", "Ne", "Ar", "Kr", "Xe"]:
        x, psi0, e = gs_case(16000, 0.1, W.SOFT_COULOMB_A[atom], W.SOFT_COULOMB_E[atom], 1e-25)
        arrays[atom + "_Einit"] = e
        arrays[atom + "_psi0_sub"] = psi0.real[::16]
        arrays[atom + "_psi0_sha256"] = np.frombuffer(hashlib.sha256(np.ascontiguousarray(psi0.real).tobytes()).digest(), dtype=np.uint8)
    save("ref_ground_states.npz", **arrays)

    # byte layout of the reference's structs (x86-64), the ABI the drop-in must keep
    import json
    layout = layout_probe.probe("/root/reference/1DTDSE/sources", "structures.h")
    with open(os.path.join(HERE, "ref_struct_layout.json"), "w") as fh:
        json.dump(layout, fh, indent=1, sort_keys=True)

    # the reference test fixture, converted
    h = h5lite.H5File(REF_H5)
    arrays = {}
    for grp in h.keys():
        for k in h.keys(grp):
            a = h[grp + "/" + k]
            if k in ("psi_re", "psi_im"):
                a = a[[0, 1, 17, 35]]
            arrays[grp + "__" + k] = a
    save("ionization_fixture.npz", **arrays)


if __name__ == "__main__":
    main()
