"""ctdse_propagate_one_noalloc (one instance alone with its scratch in a fixed device array,
no device allocation) through the C ABI against ctdse_propagate_one on the same inputs: two
calls in a row, the second on a smaller grid with fewer steps (the array keeps the first call's
contents), then a call that does not fit the array, refused before anything is queued."""
import ctypes

import numpy as np
import pytest
import torch

from mma_amd import tdse, workloads as W

pytestmark = pytest.mark.gpu

NOALLOC_BYTES = 4 << 20  # CTDSE_NOALLOC_SCRATCH_BYTES


def _run(entry, geo, field, n_steps):
    out = torch.full((5, n_steps), float("nan"), dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream()
    code = entry(geo.handle, n_steps, ctypes.c_void_p(field.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                 ctypes.c_void_p(st.cuda_stream))
    tdse._raise(code)
    st.synchronize()
    return out.cpu().numpy()


def test_noalloc_equals_propagate_one_on_two_calls():
    L = tdse._lib()
    # (num_r, dx, cycles, wavelength): the second call is smaller in grid and in steps
    for num_r, dx, cycles, lam in ((16000, 0.1, 10, 800e-9), (5000, 0.1, 4, 1300e-9)):
        print("noalloc case num_r=%d cycles=%d" % (num_r, cycles), flush=True)
        geo = tdse.Geometry(num_r, dx, 0.25, W.GRID["a"], W.GRID["e_guess"], 1e-25, 2.0)
        w = W.omega_au(lam)
        n_e = int(round(cycles * 2 * np.pi / w / 0.25)) + 1
        f = W.sin2_pulse(np.arange(n_e) * 0.25, 0.06, w, cycles)
        field = torch.tensor(tdse.pad_fields(f)[0], device="cuda")
        n_steps = n_e + 2
        need = 32 * tdse._lib().ctdse_state_length(geo.handle) + 8 * (n_steps + 1)
        assert need < NOALLOC_BYTES
        ref = _run(L.ctdse_propagate_one, geo, field, n_steps)
        got = _run(L.ctdse_propagate_one_noalloc, geo, field, n_steps)
        assert np.all(np.isfinite(ref))
        np.testing.assert_array_equal(got, ref)
        # a second allocation-free run on the same inputs gives the same bits again
        np.testing.assert_array_equal(_run(L.ctdse_propagate_one_noalloc, geo, field, n_steps), ref)


def test_noalloc_refuses_runs_beyond_its_array():
    L = tdse._lib()
    geo = tdse.Geometry(2000, 0.1, 0.25, W.GRID["a"], W.GRID["e_guess"], 1e-25, 2.0)
    field = torch.zeros(16, dtype=torch.float64, device="cuda")  # never read: refused on the host
    out = torch.zeros(16, dtype=torch.float64, device="cuda")
    n_steps = NOALLOC_BYTES // 8  # the gated field alone needs more than the array
    code = L.ctdse_propagate_one_noalloc(geo.handle, n_steps, ctypes.c_void_p(field.data_ptr()),
                                         ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(0))
    assert code != 0
    msg = L.ctdse_last_error_message().decode()
    assert "scratch" in msg, msg
    with pytest.raises(tdse.CTDSEError):
        tdse._raise(L.ctdse_propagate_one_noalloc(None, 10, ctypes.c_void_p(field.data_ptr()),
                                                  ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(0)))
