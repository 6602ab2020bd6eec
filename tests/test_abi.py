"""The C-ABI library loads and exports every entry point include/mc_abi.h
declares; host-only entry points agree with the oracle (no GPU needed)."""
import os
import re

import numpy as np

import oracle as O
from motioncor3_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "mc_abi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mc_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert _declared() == sorted(L.ABI_SYMBOLS)


def test_library_exports_every_symbol():
    lib = L.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_default_params_match_reference_defaults():
    p = L.default_params()
    o = O.default_params()
    assert (p.iter, p.kv, p.fm_ref, p.ref_quirks) == (15, 300, -1, 1)
    assert list(p.bft) == [500.0, 100.0] and list(p.group) == [1, 4]
    assert abs(p.tol - 0.1) < 1e-7 and list(p.sum_range) == [3.0, 25.0]
    assert (o.iter, list(o.bft), list(o.group)) == (p.iter, [500.0, 100.0], [1, 4])


def test_geometry_matches_oracle():
    for (W, H, px, py) in [(5760, 4092, 5, 5), (4096, 4096, 7, 5), (11520, 8184, 5, 5),
                           (1024, 1024, 0, 0), (2048, 1536, 3, 3)]:
        g = L.geometry(W, H, px, py)
        o = O.geometry(W, H, px, py)
        assert g["xcf"] == o["xcf"] and g["patch"] == o["patch"]
        assert np.allclose(g["xcf_bin"], o["xcf_bin"], rtol=1e-6)


def test_output_geometry_sum_count():
    lib = L.load()
    import ctypes as C
    ns, ow, oh = C.c_int(), C.c_int(), C.c_int()
    p = L.default_params(fm_dose=1.0, pix_size=0.83, split_sum=1)
    assert lib.mc_output_geometry(C.byref(p), 5760, 4092, 50, C.byref(ns), C.byref(ow),
                                  C.byref(oh)) == 0
    assert (ns.value, ow.value, oh.value) == (5, 5760, 4092)
    p = L.default_params(ft_bin=2.0)
    lib.mc_output_geometry(C.byref(p), 5760, 4092, 50, C.byref(ns), C.byref(ow), C.byref(oh))
    assert (ns.value, ow.value, oh.value) == (1, 2880, 2046)


def test_invalid_arguments_rejected():
    lib = L.load()
    import ctypes as C
    s = (C.c_int * 4)()
    b = (C.c_float * 2)()
    assert lib.mc_geometry(7, 8, 0, 0, s, b) == L.MC_OK - 1
    assert lib.mc_ctx_create(0, None, None) == -1
