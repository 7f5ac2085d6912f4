"""CPU checks of the oracle: it must reproduce the compiled reference's golden vectors bit
for bit, the reference's own fixture (float32) and closed forms of the FFT-based parts."""
import numpy as np
import pytest

from conftest import golden, rel_err
from oracle import ctdse as O

SERIES = ["tgrid", "Efield", "sourceterm", "PopTot", "PopInt", "expval"]


def oracle_run(g, writewft=0, tprint=0.0):
    x, psi0, e, it, st = O.ground_state(int(g["num_r"]), float(g["dx"]), float(g["a"]), float(g["e_guess"]),
                                        float(g["cv"]))
    assert st == 0
    r = O.call1dtdse(int(g["num_r"]), float(g["dx"]), float(g["a"]), float(g["x_int"]), float(g["dt"]), psi0, x,
                     g["t"], g["field"], writewft=writewft, tprint=tprint)
    return x, psi0, e, r


@pytest.mark.parametrize("name,wft,tprint", [("ref_tests_py.npz", 1, 10.0), ("ref_short_even.npz", 1, 20.0),
                                             ("ref_short_odd.npz", 1, 20.0), ("ref_short_phase.npz", 1, 20.0),
                                             ("ref_short_kr.npz", 1, 20.0)])
def test_oracle_bit_exact_vs_reference(name, wft, tprint):
    g = golden(name)
    x, psi0, e, r = oracle_run(g, wft, tprint)
    assert e == float(g["Einit"])
    np.testing.assert_array_equal(psi0.real, g["psi0_re"])
    np.testing.assert_array_equal(psi0.imag, g["psi0_im"])
    for k in SERIES:
        np.testing.assert_array_equal(r[k], g[k], err_msg=k)
    assert r["snapshots"].shape[0] == int(g["n_snap"])
    for i, row in enumerate(g["snap_rows"]):
        np.testing.assert_array_equal(r["snapshots"][row], g["snap_re"][i] + 1j * g["snap_im"][i])


def test_oracle_bit_exact_single_instance():
    """Nx = 16001, 4416 steps (about 7 s of CPU)."""
    g = golden("ref_single.npz")
    x, psi0, e, r = oracle_run(g)
    assert e == float(g["Einit"])
    for k in SERIES:
        np.testing.assert_array_equal(r[k], g[k], err_msg=k)


def test_oracle_ground_state_presets():
    import hashlib
    from mma_amd import workloads as W
    g = golden("ref_ground_states.npz")
    for atom in ["He", "Ne", "Ar", "Kr", "Xe"]:
        x, psi0, e, it, st = O.ground_state(16000, 0.1, W.SOFT_COULOMB_A[atom], W.SOFT_COULOMB_E[atom], 1e-25)
        assert st == 0 and e == float(g[atom + "_Einit"])
        assert hashlib.sha256(np.ascontiguousarray(psi0.real).tobytes()).digest() == bytes(g[atom + "_psi0_sha256"])


def test_oracle_ground_state_limit_cycle():
    """He on a 4001-point grid: the reference exits after 1000 iterations; the oracle reports it."""
    x, psi0, e, it, st = O.ground_state(4000, 0.1, 0.695, -0.9036, 1e-25)
    assert st == 1 and it == 1000


def test_oracle_matches_reference_fixture():
    """1DTDSE/tests.py:67-88 against ionization.h5, with np.allclose as the reference does;
    the fixture is float32."""
    g = golden("ref_tests_py.npz")
    fx = golden("ionization_fixture.npz")
    x, psi0, e, r = oracle_run(g, 1, 10.0)
    assert np.allclose(fx["TDSE_inputs__x_grid"], x)
    assert np.allclose(fx["TDSE_inputs__psi0"][:, 0] + 1j * fx["TDSE_inputs__psi0"][:, 1], psi0)
    assert len(fx["TDSE_outputs__sourceterm"]) == r["Nt"]
    for k in ["Efield", "sourceterm", "expval", "PopTot", "PopInt", "tgrid"]:
        assert np.allclose(fx["TDSE_outputs__" + k], r[k]), k
    F = fx["TDSE_outputs__Fsourceterm"]
    assert np.allclose(F[:, 0] + 1j * F[:, 1], r["Fsourceterm"])
    assert np.allclose(fx["TDSE_outputs__omegagrid"], r["omegagrid"])
    assert int(fx["TDSE_outputs__Nomega"]) == r["Nomega"]
    for i, row in enumerate([0, 1, 17, 35]):
        assert np.allclose(fx["TDSE_outputs__psi_re"][i] + 1j * fx["TDSE_outputs__psi_im"][i], r["snapshots"][row])


def test_spectrum_closed_form():
    """calcFFTW3 convention on a pure cosine: one line at k0 with value conj(N/2) dt/sqrt(2 pi)."""
    n, dt, k0 = 512, 0.25, 37
    t = np.arange(n) * dt
    s = np.cos(2 * np.pi * k0 * np.arange(n) / n + 0.3)
    om, F, M2 = O.spectrum(n, dt, n * dt, s)
    assert len(om) == n // 2 + 1 and om[1] == 2 * np.pi / (n * dt)
    expect = np.zeros(n // 2 + 1, complex)
    expect[k0] = np.conj(n / 2 * np.exp(1j * 0.3)) * dt / np.sqrt(2 * np.pi)
    assert rel_err(F, expect) < 1e-13
    assert rel_err(M2, np.abs(expect) ** 2) < 1e-12


def test_four_interp_properties():
    """k = 1 is the identity up to FFT rounding; a band-limited signal is interpolated exactly
    (below Nyquist); the buffer tail is zero; the Nyquist bin is copied without halving."""
    n = 200
    s = np.sin(2 * np.pi * 7 * np.arange(n) / n) + 0.3 * np.cos(2 * np.pi * 19 * np.arange(n) / n)
    out1 = O.four_interp(1, s)
    assert rel_err(out1[:n], s) < 1e-14 and np.all(out1[n:] == 0)
    out3 = O.four_interp(3, s)
    tt = np.arange(3 * n) / 3.0
    exact = np.sin(2 * np.pi * 7 * tt / n) + 0.3 * np.cos(2 * np.pi * 19 * tt / n)
    assert rel_err(out3[:3 * n], exact) < 1e-13
    nyq = np.cos(np.pi * np.arange(n))  # pure Nyquist: reference copies the bin unhalved
    out2 = O.four_interp(2, nyq)
    assert abs(out2[0] - 2.0) < 1e-12


def test_refinement_rule():
    """k1 rule of singleTDSE.c:61-68."""
    assert O.refine(np.array([0.0, 0.25]), 0.25)[0] == 1
    assert O.refine(np.array([0.0, 0.5]), 0.25)[0] == 3
    assert O.refine(np.array([0.0, 0.3]), 0.25)[0] == 1
    assert O.refine(np.array([0.0, 1.0]), 0.25, 1, 4)[0] == 5


def test_parity_detects_phase_sign():
    """Flipping the sign of the field (equivalent to the README's exp(-iF dt x) convention) is detected."""
    g = golden("ref_short_phase.npz")
    x, psi0, e, r = oracle_run(g)
    flipped = O.call1dtdse(int(g["num_r"]), float(g["dx"]), float(g["a"]), float(g["x_int"]), float(g["dt"]), psi0,
                           x, g["t"], -g["field"])
    assert rel_err(flipped["expval"], g["expval"]) > 1e-3
