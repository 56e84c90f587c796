"""Main-basis g shells on the GPU (the run-time-L kernels of the class pairs with a g shell,
next to the compiled s-f kernels) against the oracle, whose generic-L path covers any l.

Bars as in test_gpu_parity.py: FP64 element-wise within 1e-12 relative with a floor for values
that cancel to ~0 (the order-16 Hermite sums of (gg|gg) carry more rounding noise than the
order-12 ones of (ff|ff): floor 1e-13 of the largest element); gradients 1e-10 relative."""
import numpy as np
import pytest

from lible_amd import basis as B
from tests.test_gpu_parity import close

pytestmark = pytest.mark.gpu

# per atom (l, primitives, count): s to g, contracted g shells
RECIPE_G = [(0, 3, 2), (1, 2, 2), (2, 2, 1), (3, 1, 1), (4, 2, 1)]        # 29 AO per atom
RECIPE_G1 = [(0, 2, 1), (1, 1, 1), (2, 1, 1), (3, 1, 1), (4, 1, 1)]       # 25 AO per atom
RECIPE_AUX_H = [(0, 1, 2), (1, 1, 1), (2, 1, 1), (3, 1, 1), (4, 1, 1), (5, 1, 1)]


@pytest.fixture(scope="module")
def lbx():
    from lible_amd import lbx as L
    L.lib()
    return L


def test_g_eri4_dense(lbx, oracle):
    """Dense (ab|cd) over two atoms with s to g shells: every class pair up to (gg|gg)."""
    m = B.synthetic_molecule(401, 2, 3.0, RECIPE_G1)
    ctx = lbx.Context(m)
    assert ctx.nbf == 50
    close(ctx.eri4().cpu().numpy(), oracle.Structure(m).eri4(), floor=1e-13)


def test_g_jk(lbx, oracle):
    """J and K over three atoms with contracted g shells (Schwarz 1e-12 and unscreened), and the
    screened quartet count."""
    import torch
    m = B.synthetic_molecule(402, 3, 4.5, RECIPE_G)
    D = B.density_from_orbitals(9, m.shells.nbf, 20, 0.1)
    s = oracle.Structure(m)
    ctx = lbx.Context(m)
    Dt = torch.tensor(D, device="cuda")
    for thresh in (1e-12, 0.0):
        Jr, Kr, sr = s.jk(D, thresh=thresh)
        J, K, st = ctx.jk(Dt, schwarz=thresh)
        print(f"g J/K, Schwarz {thresh:g}: {st['quartets']:.0f} quartets", flush=True)
        assert st["quartets"] == sr["quartets"]
        close(J.cpu().numpy(), Jr, floor=1e-13)
        close(K.cpu().numpy(), Kr, floor=1e-13)
    # mixed precision: the g classes have no FP32 form and stay FP64, the others split
    Jm, Km, sm = ctx.jk(Dt, schwarz=1e-12, precision=lbx.LBX_MIXED, mixed_thresh=1e-5)
    Jr, Kr, _ = s.jk(D, thresh=1e-12)
    assert 0 < sm["quartets_fp32"] < sm["quartets"]
    dE = 0.5 * float((Dt * (Jm - torch.tensor(Jr, device="cuda"))).sum()) - \
        0.25 * float((Dt * (Km - torch.tensor(Kr, device="cuda"))).sum())
    assert abs(dE) <= 1e-9, dE


def test_g_eri3_eri2_sums(lbx, oracle):
    """(ab|P) with g main shells against h auxiliary shells, (P|Q), and the per-molecule sums."""
    m = B.synthetic_molecule(403, 2, 3.5, RECIPE_G1, RECIPE_AUX_H)
    s = oracle.Structure(m)
    ctx = lbx.Context(m)
    e3 = s.eri3()
    close(ctx.eri3().cpu().numpy(), e3, floor=1e-13)
    close(ctx.eri2().cpu().numpy(), s.eri2())
    got, _ = ctx.ri_sums()
    ref = s.ri_sums()
    close(got.cpu().numpy(), ref, rel=1e-12, floor=1e-12)


def test_g_jk_grad(lbx, oracle):
    """Nuclear gradient of 1/2 tr(DJ) - 1/4 tr(DK) with g shells against the oracle, and
    translational invariance."""
    import torch
    m = B.synthetic_molecule(404, 3, 4.5, RECIPE_G1)
    ctx = lbx.Context(m)
    D = B.density_from_orbitals(10, ctx.nbf, 15, 0.1)
    Dt = torch.tensor(D, device="cuda")
    g, st = ctx.jk_grad(Dt, schwarz=1e-12)
    g = g.cpu().numpy()
    gr = oracle.Structure(m).jk_grad(D, thresh=1e-12)
    close(g, gr, rel=1e-10, floor=1e-12)
    assert np.abs(g.sum(axis=0)).max() <= 1e-10 * np.abs(g).max()
    gm, sm = ctx.jk_grad(Dt, schwarz=1e-12, precision=lbx.LBX_MIXED, mixed_thresh=1e-5)
    assert sm["quartets_fp32"] > 0
    assert np.abs(gm.cpu().numpy() - gr).max() <= 1e-7


def test_g_one_electron(lbx, oracle):
    """S, T and the attraction to point charges with g shells against the oracle; the
    one-electron gradient against five-point differences of the GPU integrals."""
    import torch
    m = B.synthetic_molecule(405, 3, 4.5, RECIPE_G)
    rng = np.random.default_rng(7)
    charges = np.column_stack([m.coords, [6.0, 7.0, 8.0]])
    charges = np.vstack([charges, np.column_stack([rng.uniform(-2, 6, (3, 3)), rng.uniform(-1, 1, 3)])])
    s = oracle.Structure(m)
    ctx = lbx.Context(m)
    S, T, V = (x.cpu().numpy() for x in ctx.one_electron(charges=charges))
    close(S, s.overlap())
    close(T, s.kinetic())
    close(V, s.external_charges(charges))
    assert np.abs(np.diag(S) - 1.0).max() < 1e-12

    n = ctx.nbf
    D = torch.tensor(B.density_from_orbitals(11, n, 10, 0.1), device="cuda")
    W = torch.tensor(B.density_from_orbitals(12, n, 10, 0.1), device="cuda")
    g, gq = ctx.one_electron_grad(D, W, charges)
    g = g.cpu().numpy()
    gq = gq.cpu().numpy()

    def energy(coords, ch):
        c = lbx.Context(B.Molecule(coords, m.shells))
        S_, T_, V_ = c.one_electron(charges=ch)
        c.close()
        return float((D * (T_ + V_)).sum()) - float((W * S_).sum())

    # five-point stencil: the steep g primitives (exponents up to ~1e3) give the central
    # difference at h = 1e-4 a truncation error of ~1e-5 relative; the O(h^4) stencil's is ~1e-10
    h = 1e-4

    def five_point(f):
        e1, e2 = f(h) - f(-h), f(2 * h) - f(-2 * h)
        return (8 * e1 - e2) / (12 * h)

    def moved(arr, i, x, step):
        c = arr.copy()
        c[i, x] += step
        return c

    for ia, x in ((0, 0), (1, 2), (2, 1)):
        fd = five_point(lambda d: energy(moved(m.coords, ia, x, d), charges))
        assert abs(g[ia, x] - fd) <= 1e-7 * max(1.0, abs(fd)), (ia, x, g[ia, x], fd)
    for iq, x in ((3, 0), (5, 2)):
        fd = five_point(lambda d: energy(m.coords, moved(charges, iq, x, d)))
        assert abs(gq[iq, x] - fd) <= 1e-7 * max(1.0, abs(fd)), (iq, x, gq[iq, x], fd)


def test_g_derivatives_blocks_and_limits(lbx, oracle):
    """Dense first-derivative tensor with g shells (every class pair up to (gg|gg), whose blocks
    the derivative kernel takes in column chunks) and per-quartet blocks ((gs|ps), (gp|ss),
    (sg|ss), (gg|gg); integrals and first derivatives) against the oracle; a main h shell is
    rejected."""
    from tests.test_gpu_parity import shell_offsets

    def close_block(got, ref, scale):
        # within 1e-12 of each element plus 1e-13 of the whole tensor's scale (the floor of the
        # dense comparisons here; a block may be all but zero by symmetry)
        err = np.abs(np.asarray(got) - ref)
        assert (err <= 1e-12 * np.abs(ref) + 1e-13 * scale).all(), f"max err {err.max():.3e}"

    m = B.synthetic_molecule(401, 2, 3.0, RECIPE_G1)
    ctx = lbx.Context(m)
    s = oracle.Structure(m)
    e4, d4 = s.eri4(), s.eri4d1()
    close(ctx.eri4d1().cpu().numpy(), d4, floor=1e-13)
    l, off = shell_offsets(m.shells)
    g, sh, p = np.flatnonzero(l == 4), np.flatnonzero(l == 0), np.flatnonzero(l == 1)
    # every entry of one call has the same angular momenta per position; the third list has the
    # g shell second (the library stores the pair as (g, s) and swaps the block back)
    for q in (np.array([[g[0], sh[0], p[1], sh[1]], [g[1], sh[1], p[0], sh[0]]]),
              np.array([[g[1], p[0], sh[0], sh[1]], [g[0], p[1], sh[1], sh[1]]]),
              np.array([[sh[0], g[0], sh[1], sh[1]], [sh[1], g[1], sh[0], sh[0]]]),
              np.array([[g[0], g[1], g[1], g[0]], [g[1], g[1], g[0], g[0]]])):
        got = ctx.blocks(q).cpu().numpy()
        gd = ctx.blocks(q, deriv=True).cpu().numpy()
        for t, (a, b, c, d) in enumerate(q):
            S = [2 * int(l[x]) + 1 for x in (a, b, c, d)]
            sl = tuple(slice(off[x], off[x] + n) for x, n in zip((a, b, c, d), S))
            close_block(got[t, 0], e4[sl], np.abs(e4).max())
            close_block(gd[t], d4[(slice(None),) + sl], np.abs(d4).max())
    mh = B.synthetic_molecule(406, 1, 1.0, [(0, 1, 1), (5, 1, 1)])
    with pytest.raises(lbx.LbxError):
        lbx.Context(mh)


def _oracle_rhf(s, S, H, e_nuc, nocc):
    """Plain RHF with DIIS over the oracle's one-electron integrals and J/K."""
    sv, U = np.linalg.eigh(S)
    X = U / np.sqrt(sv)

    def dens(F):
        _, C = np.linalg.eigh(X.T @ F @ X)
        C = X @ C
        return 2 * C[:, :nocc] @ C[:, :nocc].T

    D, fs, es, E_old = dens(H), [], [], 0.0
    for _ in range(60):
        J, K, _ = s.jk(D, thresh=1e-12)
        F = H + J - 0.5 * K
        E = 0.5 * (D * (H + F)).sum() + e_nuc
        err = X.T @ (F @ D @ S - S @ D @ F) @ X
        if abs(E - E_old) < 1e-11 and np.abs(err).max() < 1e-8:
            return E
        E_old = E
        fs, es = (fs + [F])[-8:], (es + [err])[-8:]
        if len(fs) >= 2:
            n = len(fs)
            Bm = -np.ones((n + 1, n + 1))
            Bm[n, n] = 0.0
            for i in range(n):
                for j in range(n):
                    Bm[i, j] = (es[i] * es[j]).sum()
            rhs = np.zeros(n + 1)
            rhs[n] = -1.0
            c = np.linalg.solve(Bm, rhs)
            F = sum(c[i] * fs[i] for i in range(n))
        D = dens(F)
    raise AssertionError("oracle SCF did not converge")


def test_g_water_qzvp(lbx, oracle):
    """H2O in def2-QZVP (O 7s4p3d2f1g, H 4s3p2d1f; 117 functions): J/K and S, T, V against the
    oracle, and the RHF energy of the device integrals (lible_amd.scf) against an SCF on the
    oracle's integrals, 1e-9 Eh; the energy, -76.0668 Eh, lies just above the Hartree-Fock
    limit of water (about -76.0675 Eh), as a quadruple-zeta energy must."""
    import torch
    from lible_amd import scf
    from tests.conftest import H2O, basis_path
    nrs, xyz = H2O
    m = B.molecule_from_basis_files(basis_path("def2-qzvp"), nrs, xyz)
    assert m.shells.nbf == 117 and int(m.shells.l.max()) == 4
    s = oracle.Structure(m)
    ctx = lbx.Context(m)
    D = B.density_from_orbitals(3, ctx.nbf, 5, 0.1)
    Jr, Kr, sr = s.jk(D, thresh=1e-12)
    J, K, st = ctx.jk(torch.tensor(D, device="cuda"), schwarz=1e-12)
    assert st["quartets"] == sr["quartets"]
    close(J.cpu().numpy(), Jr, floor=1e-13)
    close(K.cpu().numpy(), Kr, floor=1e-13)
    nuc = np.column_stack([m.coords, np.asarray(nrs, np.float64)])
    S, T, V = (x.cpu().numpy() for x in ctx.one_electron(charges=nuc))
    So, To, Vo = s.overlap(), s.kinetic(), s.external_charges(nuc)
    close(S, So)
    close(T, To)
    close(V, Vo)
    ctx.close()
    rhf = scf.RHF(m, nrs)
    res = rhf.kernel(e_tol=1e-11, g_tol=1e-8)
    rhf.close()
    assert res.converged
    E_ref = _oracle_rhf(s, So, To + Vo, scf.nuclear_repulsion(m.coords, nrs), 5)
    print(f"H2O def2-QZVP RHF: GPU {res.energy:.10f}, oracle {E_ref:.10f} Eh", flush=True)
    assert abs(res.energy - E_ref) <= 1e-9
    assert -76.0676 < res.energy < -76.066


def test_g_ri_sums_batch(lbx, oracle):
    """Batched RI sums (one context over several molecules) with g main shells: every molecule's
    sums equal the oracle's for that molecule alone, in FP64 and with the mixed split (the g
    classes have no FP32 form and stay FP64)."""
    mols = [B.synthetic_molecule(410 + k, 2 + k, 3.5, RECIPE_G1, RECIPE_AUX_H) for k in range(3)]
    ctx = lbx.Context.batch(mols)
    got, _ = ctx.ri_sums()
    got = got.cpu().numpy()
    for k, m in enumerate(mols):
        close(got[k], oracle.Structure(m).ri_sums(), rel=1e-12, floor=1e-12)
    mixed, sm = ctx.ri_sums(precision=lbx.LBX_MIXED, mixed_thresh=1e-5)
    gm = mixed.cpu().numpy()
    assert np.abs(gm[:, [0, 2]] - got[:, [0, 2]]).max() <= 1e-8 * np.abs(got[:, [0, 2]]).max()
