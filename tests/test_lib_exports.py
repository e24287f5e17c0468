"""CPU-side checks of the C ABI: the library loads, exports every entry
point declared in include/fg.h, and its walker record matches the oracle's
record layout.  No compute call is made (no GPU here)."""
import ctypes
import os
import re

import numpy as np

import flipgraphgpu_amd as fg
import oracle_ffi as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"^\s*(?:FG_API\s+)?(?:const\s+)?\w+\s*\*?\s*(fg_\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ["fg_create", "fg_destroy", "fg_init_naive", "fg_init_from_schemes", "fg_walk",
                     "fg_best", "fg_verify", "fg_counters"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    L = fg.load_library()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_record_layout_matches_oracle():
    assert fg.RECORD_DTYPE == O.RECORD_DTYPE
    assert fg.RECORD_DTYPE.itemsize == 64 + 4 * 3 * 256 * 8 + 16


def test_last_error_is_callable_without_gpu():
    L = fg.load_library()
    assert isinstance(L.fg_last_error(), bytes)


def test_create_rejects_bad_config_before_touching_the_device():
    L = fg.load_library()
    cfg = fg.FgConfig(fg.RING_Z2, 0, 10, 10, 0, 1, 0.0, 0.0, 1, 0, 16, 0, 0)
    ctx = ctypes.c_void_p()
    assert L.fg_create(ctypes.byref(cfg), ctypes.byref(ctx)) < 0
    assert b"walkers" in L.fg_last_error()
