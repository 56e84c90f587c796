"""The benchmark workloads at their full sizes on the GPU.

40-atom J/K (1200 functions): against oracle fixtures written once in the build container
(tests/golden/make_config_fixtures.py): screened quartet count, sums, traces, the energy, the
diagonals and seeded sampled elements of J and K, for the default kernel choice and for the
block kernel on every class pair that has one.
120-atom J/K (3600 functions): no oracle run fits; the mixed build against the FP64 build,
the energy change 1/2 tr(D dJ) - 1/4 tr(D dK) of the FP32 tier within 1e-10 Eh.
256-molecule batched RI: every molecule's eri3 / eri2 sums against the oracle fixture, graph
replays, and the mixed-precision sums.
60-atom gradient (1800 functions): translational invariance and central differences of the
energy on the three largest components.
"""
import json
import os

import numpy as np
import pytest

from lible_amd import basis as B
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

SCHWARZ = 1e-12


def fixture(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.fail(f"missing fixture {path} (tests/golden/make_config_fixtures.py)")
    with open(path) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def lbx():
    from lible_amd import lbx as L
    L.lib()
    return L


def rel_close(got, ref, rel, what):
    got, ref = np.asarray(got, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max()
    assert err <= rel * scale, f"{what}: max |diff| {err:.3e} > {rel:g} x {scale:.3e}"


def check_c2(J, K, st, D, fx):
    assert st["quartets"] == fx["quartets"]
    idx = np.asarray(fx["sample_idx"])
    # FP64 with different summation orders: elements agree to ~1e-14 of the largest; the bar is
    # 1e-12 of the largest element (a dropped quartet class moves them by 1e-3 or more)
    rel_close(np.diag(J), fx["diag_J"], 1e-12, "diag J")
    rel_close(np.diag(K), fx["diag_K"], 1e-12, "diag K")
    rel_close(J[idx[:, 0], idx[:, 1]], fx["sample_J"], 1e-12, "sampled J")
    rel_close(K[idx[:, 0], idx[:, 1]], fx["sample_K"], 1e-12, "sampled K")
    for key, val in (("sum_abs_J", np.abs(J).sum()), ("sum_abs_K", np.abs(K).sum()),
                     ("tr_DJ", float((D * J).sum())), ("tr_DK", float((D * K).sum()))):
        assert abs(val - fx[key]) <= 1e-12 * abs(fx[key]), (key, val, fx[key])


def test_c2_jk_full_size(lbx):
    """40 atoms, 1200 functions, Schwarz 1e-12: the default kernels and the block kernel."""
    import torch
    fx = fixture("c2_jk_fp64.json")
    cfg = B.CONFIGS["c2_jk_fp64"]
    m = B.config_molecule("c2_jk_fp64")
    ctx = lbx.Context(m)
    assert ctx.nbf == fx["nbf"]
    Dh = B.density_from_orbitals(cfg["seed"] + 100, ctx.nbf, cfg["nocc"], cfg["sigma"])
    D = torch.tensor(Dh, device="cuda")
    J, K, st = ctx.jk(D, schwarz=SCHWARZ)
    print(f"c2 default: {st['quartets']:.0f} quartets", flush=True)
    check_c2(J.cpu().numpy(), K.cpu().numpy(), st, Dh, fx)
    classes = [(la, lb, lc, ld) for la in range(4) for lb in range(la + 1)
               for lc in range(4) for ld in range(lc + 1)
               if la * (la + 1) // 2 + lb >= lc * (lc + 1) // 2 + ld]
    used = [c for c in classes if ctx.set_variant(c, 6)]
    assert len(used) == len(classes)
    J, K, st = ctx.jk(D, schwarz=SCHWARZ)
    print("c2 block kernel: done", flush=True)
    check_c2(J.cpu().numpy(), K.cpu().numpy(), st, Dh, fx)


C3_SPLIT = 1e-6


@pytest.mark.timeout(600)
def test_c3_jk_mixed_full_size(lbx):
    """120 atoms, 3600 functions: FP64 and mixed builds (FP32 below Q_ab Q_cd = 1e-6).  The
    FP32 tier's own error is isolated with a bound window: the quartets below the split built in
    FP64 and in FP32, whose partial J and K are small enough that FP64 rounding does not hide the
    difference; it must move the energy by at most 1e-10 Eh.  (At this size the split of 1e-5
    puts 8.7e9 quartets in FP32 and their rounding adds up to ~3e-10 Eh, so the split is 1e-6
    here.)  The whole mixed build differs from the FP64 build also by the FP64 rounding of ~1e10
    atomic additions in another order (both builds' energies are ~4e5 Eh, one ulp 6e-11), so its
    energy change is checked to 1e-9 Eh."""
    import torch
    cfg = B.CONFIGS["c3_jk_mixed"]
    m = B.config_molecule("c3_jk_mixed")
    ctx = lbx.Context(m)
    D = torch.tensor(B.density_from_orbitals(cfg["seed"] + 100, ctx.nbf, cfg["nocc"], cfg["sigma"]),
                     device="cuda")
    J, K, st = ctx.jk(D, schwarz=SCHWARZ)
    print(f"c3 FP64: {st['quartets']:.3e} quartets", flush=True)
    Jm, Km, sm = ctx.jk(D, schwarz=SCHWARZ, precision=lbx.LBX_MIXED, mixed_thresh=C3_SPLIT)
    print(f"c3 mixed: {sm['quartets_fp32'] / sm['quartets']:.3f} FP32", flush=True)
    assert sm["quartets"] == st["quartets"] and sm["quartets_fp32"] > 0.5 * sm["quartets"]
    dE = 0.5 * float((D * (Jm - J)).sum()) - 0.25 * float((D * (Km - K)).sum())
    dJ = float((Jm - J).abs().max())
    dK = float((Km - K).abs().max())
    print(f"c3 whole build: dE {dE:.3e} Eh, max dJ {dJ:.3e}, max dK {dK:.3e}", flush=True)
    assert abs(dE) <= 1e-9
    assert dJ <= 1e-9 and dK <= 1e-9
    assert float((J - J.T).abs().max()) == 0.0 and float((K - K.T).abs().max()) == 0.0
    del Jm, Km
    ctx.set_bound_window(0.0, C3_SPLIT)
    Jw, Kw, sw = ctx.jk(D, schwarz=SCHWARZ)
    Jw32, Kw32, sw32 = ctx.jk(D, schwarz=SCHWARZ, precision=lbx.LBX_MIXED, mixed_thresh=C3_SPLIT)
    ctx.set_bound_window()
    assert sw["quartets"] == sw32["quartets_fp32"] == sm["quartets_fp32"]
    dE32 = 0.5 * float((D * (Jw32 - Jw)).sum()) - 0.25 * float((D * (Kw32 - Kw)).sum())
    print(f"c3 FP32 tier alone: dE {dE32:.3e} Eh, max dJ {float((Jw32 - Jw).abs().max()):.3e}, "
          f"max dK {float((Kw32 - Kw).abs().max()):.3e}", flush=True)
    assert abs(dE32) <= 1e-10


def test_c4_ri_batch_full_size(lbx):
    """256 molecules of 3-30 atoms in one context and one captured graph."""
    import torch
    fx = fixture("c4_ri_batch.json")
    mols = B.batch_molecules()
    assert [m.natoms for m in mols] == fx["natoms"]
    ctx = lbx.Context.batch(mols)
    ctx.set_graphs(True)
    out, st = ctx.ri_sums()
    got = out.cpu().numpy()
    ref = np.asarray(fx["sums"])
    print(f"c4: {st['triplets']:.3e} triplets, {st['doublets']:.3e} doublets", flush=True)
    # sums of |x| of each molecule at 1e-12 relative; the signed sums cancel, so their bar is
    # 1e-12 of the matching sum of |x|
    for col, abscol in ((0, 0), (1, 0), (2, 2), (3, 2)):
        err = np.abs(got[:, col] - ref[:, col]) / np.abs(ref[:, abscol])
        assert err.max() <= 1e-12, (col, err.max(), int(err.argmax()))
    again = torch.empty_like(out)
    ctx.ri_sums(again)
    ctx.ri_sums(again)
    assert np.abs(again.cpu().numpy() - got).max() <= 1e-13 * np.abs(got).max()
    mixed, sm = ctx.ri_sums(precision=lbx.LBX_MIXED, mixed_thresh=1e-5)
    gm = mixed.cpu().numpy()
    assert sm["triplets_fp32"] > 0
    rel = np.abs(gm[:, [0, 2]] - ref[:, [0, 2]]) / np.abs(ref[:, [0, 2]])
    assert rel.max() <= 1e-8, rel.max()


@pytest.mark.timeout(600)
def test_c5_gradient_full_size(lbx):
    """60 atoms, 1800 functions: translational invariance and central differences (h = 1e-4
    bohr) on the three largest components, each energy difference formed from the J and K
    differences."""
    import torch
    cfg = B.CONFIGS["c5_gradient"]
    m = B.config_molecule("c5_gradient")
    ctx = lbx.Context(m)
    D = torch.tensor(B.density_from_orbitals(cfg["seed"] + 100, ctx.nbf, cfg["nocc"], cfg["sigma"]),
                     device="cuda")
    g, st = ctx.jk_grad(D, schwarz=SCHWARZ)
    g = g.cpu().numpy()
    print(f"c5: {st['quartets']:.3e} quartets, max |g| {np.abs(g).max():.3e}", flush=True)
    assert np.abs(g.sum(axis=0)).max() <= 1e-10 * np.abs(g).max()
    h = 1e-4

    def jk_at(ia, x, step):
        c = m.coords.copy()
        c[ia, x] += step
        cx = lbx.Context(B.Molecule(c, m.shells))
        J, K, _ = cx.jk(D, schwarz=SCHWARZ)
        cx.close()
        return J, K

    def diff(p, q):
        return 0.5 * float((D * (p[0] - q[0])).sum()) - 0.25 * float((D * (p[1] - q[1])).sum())

    for flat in np.argsort(-np.abs(g).ravel())[:3]:
        ia, x = divmod(int(flat), 3)
        e1 = diff(jk_at(ia, x, h), jk_at(ia, x, -h))
        e2 = diff(jk_at(ia, x, 2 * h), jk_at(ia, x, -2 * h))
        fd2 = e1 / (2 * h)                         # central difference, error O(h^2 E''')
        fd4 = (8 * e1 - e2) / (12 * h)             # five-point stencil, error O(h^4)
        print(f"c5: atom {ia} axis {x}: analytic {g[ia, x]:.10f} central {fd2:.10f} "
              f"five-point {fd4:.10f}", flush=True)
        assert abs(g[ia, x] - fd4) <= 1e-7, (ia, x, g[ia, x], fd4)
        assert abs(g[ia, x] - fd2) <= 1e-5, (ia, x, g[ia, x], fd2)
