"""The serial setup steps that the library runs on the host for one geometry alone: the
ground state's inverse iteration behind Initialise_grid_and_ground_state (tools.c:31-51,
tridiag.c:86-159) runs on a host core, where the serial chain is ~50x faster than in one GPU
lane, with the reference's operation order and no contraction.  Checked bit for bit against the
ground states the reference's own sources wrote (tests/golden/ref_single.npz,
ref_ground_states.npz); the library's batched GPU build of the same ground states is checked
against the same files in the GPU tests.  No device is used."""
import hashlib
import time

import numpy as np

from conftest import golden


def test_initialise_ground_state_on_host_is_bit_exact(dll):
    from mma_amd.ctdse.structures import inputs_def
    g = golden("ref_single.npz")
    inputs = inputs_def()
    inputs.init_default_inputs(Eguess=float(g["e_guess"]), num_r=int(g["num_r"]), dx=float(g["dx"]),
                               trg_a=float(g["a"]), CV=float(g["cv"]))
    t0 = time.perf_counter()
    dll.init_GS(inputs)
    print("Initialise_grid_and_ground_state, Nx %d: %.1f ms" % (int(g["num_r"]) + 1, (time.perf_counter() - t0) * 1e3))
    assert inputs.Einit == float(g["Einit"])
    np.testing.assert_array_equal(inputs.get_GS().real, g["psi0_re"])
    np.testing.assert_array_equal(inputs.get_GS().imag, g["psi0_im"])
    x = inputs.get_xgrid()
    assert x[0] == -800.0 and x[-1] == 800.0
    inputs.delete(dll)


def test_host_ground_state_presets(dll):
    from mma_amd import workloads as W
    from mma_amd.ctdse.structures import inputs_def
    g = golden("ref_ground_states.npz")
    for atom in ["He", "Ne", "Ar", "Kr", "Xe"]:
        inputs = inputs_def()
        inputs.init_default_inputs(Eguess=W.SOFT_COULOMB_E[atom], num_r=16000, dx=0.1, trg_a=W.SOFT_COULOMB_A[atom],
                                   CV=1e-25)
        dll.init_GS(inputs)
        assert inputs.Einit == float(g[atom + "_Einit"]), atom
        psi = np.ascontiguousarray(inputs.get_GS().real)
        assert hashlib.sha256(psi.tobytes()).digest() == bytes(g[atom + "_psi0_sha256"]), atom
        inputs.delete(dll)
