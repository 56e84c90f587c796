"""Edge cases of the C-ABI on the GPU: the smallest molecules, nothing to compute, and the
errors the entry points report instead of computing (checked before any kernel runs)."""
import numpy as np
import pytest

from lible_amd import basis as B
from tests.conftest import H2O, basis_path

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lbx():
    from lible_amd import lbx as L
    L.lib()
    return L


def test_single_atom_s_shells(lbx, oracle):
    """One atom with three s shells: one pair class, one class pair, same-shell quartets only
    on the diagonal; dense eri4 and J/K against the oracle."""
    import torch
    sh = B.Shells.from_lists([[(0, [3.0, 0.5], [0.4, 0.7]), (0, [1.2], [1.0]), (0, [0.2, 0.05], [0.6, 0.3])]])
    m = B.Molecule(np.zeros((1, 3)), sh)
    s = oracle.Structure(m)
    ctx = lbx.Context(m)
    e = ctx.eri4().cpu().numpy()
    ref = s.eri4()
    assert np.abs(e - ref).max() <= 1e-13 * np.abs(ref).max()
    D = B.density_from_orbitals(3, 3, 1, 0.5)
    J, K, st = ctx.jk(torch.tensor(D, device="cuda"), schwarz=0.0)
    Jr, Kr, sr = s.jk(D, thresh=0.0)
    assert st["quartets"] == sr["quartets"] == 21   # unique quartets of 6 unique pairs
    assert np.abs(J.cpu().numpy() - Jr).max() <= 1e-13 * np.abs(Jr).max()
    assert np.abs(K.cpu().numpy() - Kr).max() <= 1e-13 * np.abs(Kr).max()


def test_nothing_survives_the_screen(lbx):
    """A Schwarz threshold above every bound: no quartet, J = K = 0 exactly (also on the
    replayed graph of the second call)."""
    import torch
    m = B.synthetic_molecule(5, 3, 4.0, B.RECIPE_SMALL)
    ctx = lbx.Context(m)
    ctx.set_graphs(True)
    D = torch.tensor(B.density_from_orbitals(1, ctx.nbf, 4, 0.1), device="cuda")
    J, K = torch.full_like(D, 7.0), torch.full_like(D, 7.0)
    for _ in range(2):
        _, _, st = ctx.jk(D, schwarz=1e30, J=J, K=K)
        assert st["quartets"] == 0
        assert float(J.abs().max()) == 0.0 and float(K.abs().max()) == 0.0


def test_empty_and_invalid_block_lists(lbx):
    ctx = lbx.Context(B.synthetic_molecule(6, 2, 3.0, B.RECIPE_SMALL))
    out = ctx.blocks(np.zeros((0, 4), np.int32))
    assert out.shape[0] == 0
    nsh = len(ctx._ls[0])
    with pytest.raises(lbx.LbxError, match="out of range"):
        ctx.blocks(np.array([[0, 0, 0, 0], [nsh, 0, 0, 0]]))
    with pytest.raises(lbx.LbxError, match="out of range"):
        ctx.blocks(np.array([[0, 0, 0, nsh]]))
    ls = np.asarray(ctx._ls[0])
    s0, p0 = int(np.argmax(ls == 0)), int(np.argmax(ls == 1))
    with pytest.raises(lbx.LbxError, match="angular momenta of the first"):
        ctx.blocks(np.array([[s0, s0, s0, s0], [p0, s0, s0, s0]]))


def test_state_errors(lbx):
    """Entry points that need what the context does not hold report LBX_ERR_STATE / INVALID
    and leave the context usable."""
    import torch
    nrs, xyz = H2O
    m = B.molecule_from_basis_files(basis_path("def2-svp"), nrs, xyz)
    ctx = lbx.Context(m)
    D = torch.tensor(B.density_from_orbitals(2, ctx.nbf, 5, 0.1), device="cuda")
    with pytest.raises(lbx.LbxError, match="auxiliary"):
        ctx.ri_jk(D)
    with pytest.raises(lbx.LbxError):
        ctx.eri3()
    with pytest.raises(lbx.LbxError):
        ctx.blocks(np.array([[0, 0]]))          # eri2 blocks need an auxiliary basis
    with pytest.raises(lbx.LbxError, match="density must be"):
        ctx.jk(D[:-1, :-1])
    # still usable
    J, K, st = ctx.jk(D)
    assert st["quartets"] > 0 and float(J.abs().max()) > 0
    # a batched context holds pairs within each molecule only: no J/K, no dense tensors
    mols = [B.molecule_from_basis_files(basis_path("def2-svp"), nrs, xyz, basis_path("def2-universal-jkfit"))] * 2
    bctx = lbx.Context.batch(mols)
    with pytest.raises(lbx.LbxError):
        bctx.jk(torch.zeros((bctx.nbf, bctx.nbf), dtype=torch.float64, device="cuda"))
    with pytest.raises(lbx.LbxError):
        bctx.one_electron()
    sums, _ = bctx.ri_sums()
    assert np.allclose(sums[0].cpu().numpy(), sums[1].cpu().numpy(), rtol=1e-13)


def test_threads_share_contexts(lbx):
    """Calls on one context from several threads are serialised by the context's lock; calls on
    two contexts interleave; every result equals the single-threaded one (J/K to the rounding of
    the atomic accumulation order)."""
    import threading
    import torch
    ma = B.synthetic_molecule(41, 3, 4.0, B.RECIPE_SMALL)
    mb = B.synthetic_molecule(42, 2, 3.0, B.RECIPE_SMALL)
    ca, cb = lbx.Context(ma), lbx.Context(mb)
    D = torch.tensor(B.density_from_orbitals(6, ca.nbf, 6, 0.1), device="cuda")
    J0, K0, _ = ca.jk(D, schwarz=0.0)
    J0, K0 = J0.cpu().numpy(), K0.cpu().numpy()
    E0 = cb.eri4().cpu().numpy()
    errors, results = [], []

    def jk_worker():
        try:
            for _ in range(3):
                J, K, _ = ca.jk(D, schwarz=0.0)
                results.append(("jk", J.cpu().numpy(), K.cpu().numpy()))
        except Exception as e:   # reported below
            errors.append(e)

    def eri_worker():
        try:
            for _ in range(3):
                results.append(("eri4", cb.eri4().cpu().numpy(), None))
        except Exception as e:
            errors.append(e)

    threads = [threading.Thread(target=f) for f in (jk_worker, jk_worker, eri_worker, eri_worker)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert len(results) == 12
    for kind, x, y in results:
        if kind == "jk":
            assert np.abs(x - J0).max() <= 1e-13 * np.abs(J0).max()
            assert np.abs(y - K0).max() <= 1e-13 * np.abs(K0).max()
        else:
            assert np.array_equal(x, E0)
