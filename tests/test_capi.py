"""The C-ABI library loads, exports every symbol include/falcon_hip.h declares,
and reports errors through falcon_last_error (CPU; no compute calls)."""
import ctypes as C
import os
import re

import pytest

from falcon_amd import FalconError, Shell, configs, core

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(falcon_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(f for f in os.listdir(os.path.join(REPO, "include")) if f.endswith(".h")))
def test_every_declared_symbol_is_exported(header):
    lib = C.CDLL(core.LIB_PATH)
    names = _declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_are_reported():
    L = core.load()
    h = C.c_void_p()
    assert L.falcon_create(C.byref(h), 8, 8, 8, 3, 1.0, 2.0) == -1      # ng must be 2
    buf = C.create_string_buffer(256)
    L.falcon_last_error(buf, 256)
    assert b"ng must be 2" in buf.value
    assert L.falcon_create(C.byref(h), 8, 8, 8, 2, 2.0, 1.0) == -1      # empty radial range


def test_calls_out_of_order_raise():
    s = Shell(configs.small_mixed())
    with pytest.raises(FalconError, match="falcon_build_tables"):
        s.gc_iblocks(0)
    s.build_tables()
    with pytest.raises(FalconError, match="out of range"):
        s.gc_iblocks(s.nblocks)
    with pytest.raises(FalconError, match="set_gc_all"):
        s.get_state(0)


def test_launch_shapes_are_validated_and_kept_per_precision():
    """Host side of the launch-shape setter: defaults, per-precision storage,
    and the errors for an edge other than 8 or 16 or a negative workgroup count
    (no device call)."""
    s = Shell(configs.small_mixed())
    assert s.get_tiles(0) == (16, 0, 16, 0) and s.get_tiles(1) == (16, 0, 16, 0)
    s.set_tiles(1, 8, 16, 700, 0)
    assert s.get_tiles(1) == (8, 700, 16, 0) and s.get_tiles(0) == (16, 0, 16, 0)
    with pytest.raises(FalconError, match="tile edges must be 8 or 16"):
        s.set_tiles(0, 12, 16)
    with pytest.raises(FalconError, match="workgroup counts must be >= 0"):
        s.set_tiles(0, 16, 16, -1, 0)
    with pytest.raises(FalconError, match="precision must be"):
        s.set_tiles(2, 16, 16)
    assert s.get_tiles(0) == (16, 0, 16, 0)
