"""CPU checks of the drop-in boundary: the library loads, exports every symbol that
include/ctdse.h declares, and its structs (C header and ctypes mirror) have the reference's
byte layout (tests/golden/ref_struct_layout.json, probed from 1DTDSE/sources/structures.h)."""
import ctypes
import json
import os
import re
import sys

import pytest

from conftest import ROOT, GOLDEN

sys.path.insert(0, GOLDEN)
import layout_probe  # noqa: E402


def declared_functions(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\**\s*\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "for", "return")))


def test_header_declares_reference_entry_points():
    names = declared_functions(os.path.join(ROOT, "include", "ctdse.h"))
    for n in ["Initialise_grid_and_ground_state", "call1DTDSE", "set_time_and_field", "Set_all_prints",
              "window_analysis", "GaborTransform", "free_arr", "free_mtrx", "outputs_destructor",
              "inputs_destructor", "ctdse_last_error", "ctdse_propagate_batch"]:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    from mma_amd import library_path
    if not os.path.exists(library_path()):
        from mma_amd import build
        build.build()
    lib = ctypes.CDLL(library_path())
    for n in declared_functions(os.path.join(ROOT, "include", "ctdse.h")):
        assert hasattr(lib, n), n


def test_struct_layout_matches_reference():
    ref = json.load(open(os.path.join(GOLDEN, "ref_struct_layout.json")))
    ours = layout_probe.probe(os.path.join(ROOT, "include"), "ctdse.h")
    assert ours == ref


def test_ctypes_mirror_layout_matches_reference():
    from mma_amd.ctdse import structures as S
    ref = json.load(open(os.path.join(GOLDEN, "ref_struct_layout.json")))
    for name, spec in ref.items():
        cls = getattr(S, name)
        assert ctypes.sizeof(cls) == spec["size"], name
        for f, off in spec["fields"].items():
            assert getattr(cls, f).offset == off, (name, f)


def test_hankel_library_exports_every_declared_symbol():
    from mma_amd import hankel
    if not os.path.exists(hankel.LIB_PATH):
        from mma_amd import build
        build.build()
    names = declared_functions(os.path.join(ROOT, "include", "hankel.h"))
    assert {"hk_prefactor", "hk_direct", "hk_grid_spread", "hk_grid_bessel", "hk_gemm"} <= set(names)
    lib = ctypes.CDLL(hankel.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_hankel_ctypes_structs_match_header():
    """The ctypes mirrors of hk_medium and hk_planes have the C layout (offsetof probe)."""
    from mma_amd import hankel as H
    got = layout_probe.probe(os.path.join(ROOT, "include"), "hankel.h", structs=("hk_medium", "hk_planes"))
    for name, cls in (("hk_medium", H.hk_medium), ("hk_planes", H.hk_planes)):
        assert ctypes.sizeof(cls) == got[name]["size"], name
        for f, off in got[name]["fields"].items():
            assert getattr(cls, f).offset == off, (name, f)


def test_batched_entry_points_reject_bad_arguments_before_touching_a_device():
    """Argument errors of the batched and gradient entry points come back as status 3
    (CTDSE_ERR_ARGUMENT) with a message; the checks run before any device call."""
    from mma_amd import library_path
    lib = ctypes.CDLL(library_path())
    lib.ctdse_last_error_message.restype = ctypes.c_char_p
    N = None
    assert lib.ctdse_adjoint_reverse_batch(1, N, N, N, N, N, N, N, N, N, N, 1, 0, N, N, N) == 3
    assert b"null" in lib.ctdse_last_error_message()
    assert lib.ctdse_last_error() == 3
    dummy = (ctypes.c_longlong * 1)(0)
    p = ctypes.cast(dummy, ctypes.c_void_p)
    assert lib.ctdse_adjoint_reverse_batch(1, p, p, p, p, p, p, p, p, p, p, 0, 1, p, p, N) == 3
    assert b"checkpoint spacing" in lib.ctdse_last_error_message()
    assert lib.ctdse_adjoint_batch(1, N, N, N, N, N, N, N, N, N, N, N, N, N, N) == 3
    assert lib.ctdse_propagate_states(1, N, N, N, N, N, N, N, N, N, N, 1, N, N, N) == 3
    assert lib.ctdse_propagate_batch(1, N, N, N, N, N, N, N) == 3
    assert lib.ctdse_propagate_one(N, 10, N, N, N) == 3
    assert lib.ctdse_adjoint_reverse_batch(0, N, N, N, N, N, N, N, N, N, N, 1, 0, N, N, N) in (0, 5)  # empty batch


def test_plan_and_geometry_batch_entry_points_reject_bad_arguments():
    """The plan and batched-geometry entry points check their arguments before any device call."""
    from mma_amd import library_path
    lib = ctypes.CDLL(library_path())
    lib.ctdse_last_error_message.restype = ctypes.c_char_p
    N = None
    plan = ctypes.c_void_p()
    assert lib.ctdse_batch_plan_create(1, N, N, N, N, N, N, ctypes.c_longlong(0), N, ctypes.byref(plan)) == 3
    assert plan.value is None
    assert lib.ctdse_batch_plan_create(1, N, N, N, N, N, N, ctypes.c_longlong(0), N, N) == 3
    assert lib.ctdse_batch_plan_enqueue(N, N) == 3
    assert lib.ctdse_batch_plan_graph_launch(N, N) == 3
    assert lib.ctdse_batch_plan_info(N, N, N) == 3
    lib.ctdse_batch_plan_destroy(N)
    assert lib.ctdse_geometry_create_batch(2, N, N, N, N, N, N, N, N, N) == 3
    assert lib.ctdse_geometry_create_batch(-1, N, N, N, N, N, N, N, N, N) == 3
    assert lib.ctdse_launch_config_default(0, N, N) == 3
    assert b"null" in lib.ctdse_last_error_message()


def test_autotune_candidates_and_signature():
    from mma_amd import autotune
    c = autotune.candidates(18)
    assert (4, 3, 1) in c and (8, 1, 1) in c and (2, 6, 1) in c and (4, 2, 2) in c and (4, 3, 2) not in c
    assert all(nw * 64 - 18 >= 64 and nw * per <= (12 if sp == 1 else 8) for nw, per, sp in c)
    assert autotune.candidates(100)[0][0] == 3  # two-wave segments would hold fewer than 64 own chunks
    k1, _ = autotune.signature("MI355X", 256, [16001] * 4096, [4416] * 4096)
    k2, _ = autotune.signature("MI355X", 256, [16001] * 4096, [4416] * 4096)
    k3, _ = autotune.signature("MI355X", 256, [16001] * 4095, [4416] * 4095)
    assert k1 == k2 and k1 != k3
