"""CPU-side checks of the C-ABI library: it loads, exports every symbol include/lbx.h declares,
and rejects bad arguments before touching a GPU."""
import ctypes
import os
import re

import pytest

from tests.conftest import ROOT


def header_symbols():
    text = open(os.path.join(ROOT, "include", "lbx.h")).read()
    return sorted(set(re.findall(r"\b(lbx_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding():
    from lible_amd import lbx
    assert header_symbols() == sorted(lbx.SYMBOLS)


def test_library_exports_all_symbols():
    from lible_amd import lbx
    if not os.path.exists(lbx.LIB_PATH):
        pytest.skip("liblbx.so not built")
    L = ctypes.CDLL(lbx.LIB_PATH)
    for s in header_symbols():
        assert hasattr(L, s), s


def test_null_arguments_rejected():
    from lible_amd import lbx
    if not os.path.exists(lbx.LIB_PATH):
        pytest.skip("liblbx.so not built")
    L = lbx.lib()
    assert L.lbx_create(None, None, None) == -1
    assert L.lbx_dims(None, None, None, None) == -1
    assert b"null" in L.lbx_last_error()
    assert L.lbx_boys(99, 10, None, None, None) == -1
    assert L.lbx_eri4(None, None) == -1
    assert L.lbx_jk(None, None, None, None, 1e-12, 0, 1e-5, None) == -1
    assert L.lbx_jk_grad(None, None, None, 1e-12, 0, 0.0, None) == -1
    assert L.lbx_create_batch(None, 0, None, None) == -1
    assert L.lbx_ri_sums(None, None, 0, 0.0, None) == -1
    for name in ("lbx_eri4d1", "lbx_eri3d1", "lbx_eri2d1"):
        assert getattr(L, name)(None, None) == -1
    assert L.lbx_ri_jk(None, None, None, None) == -1
    assert L.lbx_one_electron(None, None, None, 0, None, None) == -1
    assert L.lbx_one_electron_grad(None, None, None, 0, None, None, None) == -1
    assert L.lbx_set_bound_window(None, 0.0, 1.0) == -1
    for name in ("lbx_eri4_batch", "lbx_eri3_batch", "lbx_eri2_batch", "lbx_eri4d1_batch",
                 "lbx_eri3d1_batch", "lbx_eri2d1_batch", "lbx_eri4soc_batch", "lbx_eri3soc_batch",
                 "lbx_eri2d2_batch", "lbx_eri4_batch_static"):
        assert getattr(L, name)(None, 1, None, None) == -1
        assert getattr(L, name)(None, -1, None, None) == -1
