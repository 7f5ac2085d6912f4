"""CPU checks of the synthetic workloads: the uneven batch's seeded draw stays inside its
stated ranges and is reproducible, and the ragged-field stacking keeps the reference's zero
tail and step count (N_E + 2)."""
import math

import numpy as np

from mma_amd import workloads as W


def test_uneven_draw_ranges_and_reproducibility():
    num_r, n_e = W.uneven_work()
    assert len(num_r) == 8192 and np.all(num_r % 2 == 0)
    assert num_r.min() >= 4000 and num_r.max() <= 32000
    again = W.uneven_work()
    np.testing.assert_array_equal(num_r, again[0])
    np.testing.assert_array_equal(n_e, again[1])
    members = W.uneven_batch(indices=[0, 17, 8191])
    for m, i in zip(members, [0, 17, 8191]):
        assert m["num_r"] == num_r[i] and len(m["field"]) == n_e[i]
        assert 6 <= m["n_cycles"] <= 20 and 5e13 <= m["intensity"] <= 3e14 and 800e-9 <= m["wavelength"] <= 1800e-9
        T = m["n_cycles"] * 2 * math.pi / W.omega_au(m["wavelength"])
        assert (len(m["field"]) - 2) * 0.25 < T <= (len(m["field"]) - 1) * 0.25
        peak = np.max(np.abs(m["field"]))
        assert peak <= W.field_au(m["intensity"]) * (1 + 1e-12)
    work = (num_r + 1.0) * (n_e + 2)
    assert 30 < work.max() / work.min() < 80


def test_stack_fields_zero_tail():
    f, steps = W.stack_fields([np.ones(5), np.arange(1.0, 3.0)])
    assert f.shape == (2, 8) and steps == [7, 4]
    np.testing.assert_array_equal(f[1], [1, 2, 0, 0, 0, 0, 0, 0])
    np.testing.assert_array_equal(f[0, 5:], 0)
