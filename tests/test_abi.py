"""CPU checks of the drop-in boundary: libzdx.so loads and exports every entry
point include/zdx.h declares (no compute call: no GPU here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zdx.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(zdx_[a-z_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared()
    for f in ["zdx_create", "zdx_mesh_register", "zdx_contact_pairs", "zdx_sim_load_config", "zdx_sim_step",
              "zdx_sim_checkpoint", "zdx_sim_restore"]:
        assert f in names


def test_library_exports_every_declared_symbol():
    import zdx

    if not os.path.exists(zdx.LIB_PATH):
        pytest.skip("libzdx.so not built")
    L = C.CDLL(zdx.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name
    assert sorted(zdx.EXPORTS) == declared()


def test_bad_arguments_are_rejected_without_a_device():
    import zdx

    if not os.path.exists(zdx.LIB_PATH):
        pytest.skip("libzdx.so not built")
    L = zdx.lib()
    T = C.c_int32(0)
    assert L.zdx_load_stl(b"/nonexistent.stl", None, 0, C.byref(T)) == zdx.IO
    assert L.zdx_create(None, 0) == zdx.BAD_ARG
    assert L.zdx_sim_step(None, 1, 0, None) == zdx.BAD_ARG


def test_stl_loader_matches_oracle(ref_inputs):
    import numpy as np

    import oracle as O
    import zdx

    if not os.path.exists(zdx.LIB_PATH):
        pytest.skip("libzdx.so not built")
    for f in ["low-poly-banbana.stl", "wall_box.stl", "wall_plane.stl", "convex-particle.stl"]:
        p = os.path.join(ref_inputs, "geometry", f)
        assert np.array_equal(zdx.load_stl(p), O.load_stl(p))
