"""The allocation-free block entry (lbx_eri4_batch_static) through the C-ABI against the ordinary
one (lbx_eri4_batch): the same kernel on the same task list, so the blocks are bit-identical;
two calls in a row, the second with fewer entries (the device array keeps the first call's
contents, which the second must not see), and the size limit."""
import ctypes

import numpy as np
import pytest

from lible_amd import basis as B
from tests.conftest import O3, basis_path

pytestmark = pytest.mark.gpu


def test_static_batch_matches_batch():
    import torch
    from lible_amd import lbx
    L = lbx.lib()
    nrs, xyz = O3
    m = B.molecule_from_basis_files(basis_path("def2-svp"), nrs, xyz)
    ctx = lbx.Context(m)
    l = np.asarray(m.shells.l)
    rng = np.random.default_rng(11)
    for n in (500, 37):
        q = np.ascontiguousarray(np.stack([rng.choice(np.flatnonzero(l == x), n)
                                           for x in (1, 0, 2, 1)], axis=1), dtype=np.int32)
        shape = (n, 3, 1, 5, 3)
        ref = torch.full(shape, float("nan"), dtype=torch.float64, device="cuda")
        got = torch.full(shape, float("nan"), dtype=torch.float64, device="cuda")
        assert L.lbx_eri4_batch(ctx.h, n, q.ctypes.data, ref.data_ptr()) == 0
        assert L.lbx_eri4_batch_static(ctx.h, n, q.ctypes.data, got.data_ptr()) == 0
        torch.cuda.synchronize()
        assert not torch.isnan(ref).any()
        assert torch.equal(got, ref)
    big = np.zeros(((1 << 20) + 1, 4), dtype=np.int32)
    dummy = torch.empty(1, dtype=torch.float64, device="cuda")
    assert L.lbx_eri4_batch_static(ctx.h, big.shape[0], big.ctypes.data, dummy.data_ptr()) == -1
    assert b"at most 1048576" in L.lbx_last_error()
