"""The C-ABI library loads, exports every function include/nrrs_amd.h declares,
and its struct layouts match the Python binding (no device calls)."""
import ctypes as C
import os
import re

import nrrs_amd as na

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(nrrs_[a-z0-9_]+)\s*\(", text))
    return names


def test_exports_every_declared_symbol():
    lib = C.CDLL(na.LIB_PATH)
    decl = declared_functions()
    assert len(decl) >= 10
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    # the binding covers every declared entry point
    assert decl <= set(na.exported_names()), sorted(decl - set(na.exported_names()))


def test_struct_layouts_and_defaults():
    lib = na.load()  # raises on any struct-size mismatch
    cfg = na.default_config()
    assert cfg.grid.n_levels == 16 and cfg.grid.log2_table_size == 19
    assert cfg.rrs_clamp_min == C.c_float(0.05).value and cfg.rrs_clamp_max == 20.0
    assert cfg.rrs_use_ceil == 1 and cfg.rrs_no_upscale == 1 and cfg.split_cap == 8
    assert cfg.paired_tables == 1  # both nets' grids gathered in one pass by default
    assert lib.nrrs_status_string(na.ERR_OVERFLOW) == b"survivor queue overflow"


def test_invalid_arguments_rejected_without_device():
    lib = na.load()
    assert lib.nrrs_create(None, None) == na.ERR_ARG
    bad = na.default_config(grid=dict(n_features=4))
    h = C.c_void_p()
    assert lib.nrrs_grid_layout(C.byref(bad.grid), None, None, None, None, None, None) == na.ERR_ARG
    unsupported = na.default_config(precision=7)
    assert lib.nrrs_create(C.byref(unsupported), C.byref(h)) == na.ERR_UNSUPPORTED
    bad_sigma = na.default_config(train_sigma=2)
    assert lib.nrrs_create(C.byref(bad_sigma), C.byref(h)) == na.ERR_ARG
    # two-headed StatNets: a depth other than the variant's, an unknown variant, and AllSelf with
    # sigma outputs are not built
    for t, hidden in ((3, 4), (1, 4), (2, 3), (4, 2)):
        cfg = na.default_config(ll2_two_head=t, statnet=dict(n_hidden_layers=hidden, n_outputs=6))
        assert lib.nrrs_create(C.byref(cfg), C.byref(h)) == na.ERR_UNSUPPORTED
    cfg = na.default_config(ll2_two_head=3, train_sigma=1, statnet=dict(n_hidden_layers=2, n_outputs=6))
    assert lib.nrrs_create(C.byref(cfg), C.byref(h)) == na.ERR_UNSUPPORTED
    assert lib.nrrs_create(C.byref(na.default_config(rrs_aid=2)), C.byref(h)) == na.ERR_ARG
