"""The C-ABI library loads and exports every symbol include/janc_step.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from janc_amd import _lib
from janc_amd.mechanism import Mechanism

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "janc_step.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(jh_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_are_bound():
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_library_exports_every_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_layout_and_digest_match(mech_path):
    lib = _lib.load()
    m = Mechanism(mech_path)
    assert lib.jh_mech_const_count() == m.const_array().size
    assert lib.jh_mech_digest().decode() == m.structure_digest()
    assert lib.jh_num_species() == m.ns


def test_argument_errors_without_gpu():
    lib = _lib.load()
    out = ctypes.c_void_p()
    kinds = (ctypes.c_int * 4)(0, 0, 0, 0)
    consts = np.zeros(lib.jh_mech_const_count())
    cp = consts.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    # bad sizes / enums are rejected before any device call
    assert lib.jh_ctx_create(2, 8, 1.0, 1.0, kinds, None, 0, cp, consts.size, None,
                             ctypes.byref(out)) == _lib.JH_ERR_ARG
    assert b"nx" in lib.jh_last_error(None)
    assert lib.jh_ctx_create(8, 8, 1.0, 1.0, kinds, None, 7, cp, consts.size, None,
                             ctypes.byref(out)) == _lib.JH_ERR_ARG
    assert lib.jh_ctx_create(8, 8, 1.0, 1.0, kinds, None, 0, cp, 3, None,
                             ctypes.byref(out)) == _lib.JH_ERR_MECH
    bad = (ctypes.c_int * 4)(0, 9, 0, 0)
    assert lib.jh_ctx_create(8, 8, 1.0, 1.0, bad, None, 0, cp, consts.size, None,
                             ctypes.byref(out)) == _lib.JH_ERR_ARG
    assert lib.jh_advance(None, None, None, None, None, 1.0, 0.0, 1, None) == _lib.JH_ERR_ARG


def test_tile_variants_are_described():
    """Every stage-kernel variant has a description (no GPU call)."""
    import ctypes
    from janc_amd import _lib
    lib = _lib.load()
    n = lib.jh_tile_variants()
    assert n >= 3
    for v in range(n):
        r, t, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        assert lib.jh_tile_info(v, ctypes.byref(r), ctypes.byref(t), ctypes.byref(w)) == 0
        assert r.value in (8, 16, 32) and t.value == 32 * r.value and 1 <= w.value <= 8
    r = ctypes.c_int()
    assert lib.jh_tile_info(n, ctypes.byref(r), ctypes.byref(r), ctypes.byref(r)) == _lib.JH_ERR_ARG
