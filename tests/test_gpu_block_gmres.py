"""Block GMRES (one block Krylov space for all columns, as Krylov.block_gmres called by
_block_gmres_solve, src/solver/multi_rhs.jl:144-148) against the oracle's LU solution and the
per-column solver.  Bars: 1e-10 relative for FP64 solutions, 1e-8 for mixed precision against FP64."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG1 = (1.2, 0.8, 0.6, 4, 0.2)


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def cfg(dbie, O):
    say("build config-1 operators (N=2560) for block GMRES")
    g = dbie.single_dielectric_box3d(*CFG1, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    ref_if = O.Interface.box(*CFG1, 4.0, 1.0)
    ref = O.CorrectedOperator(ref_if)
    Ad = ref_if.dense_DT() + ref.Cs.toarray() + np.diag(ref.t)
    return g, op, ref_if, Ad


@pytest.mark.parametrize("K,pre", [(1, True), (5, True), (16, False), (40, True), (1, 2), (5, 2), (40, 2), (5, 3),
                                   (40, 3)])
def test_block_gmres_matches_oracle(dbie, O, cfg, K, pre):
    g, op, ref_if, Ad = cfg
    rng = np.random.default_rng(60 + K)
    pts = rng.uniform(-0.25, 0.25, (K, 3))
    q = rng.uniform(-1, 1, K)
    F = dbie.rhs_point_charges(g, pts, q, 4.0)
    S_ref = O.solve_lu(Ad, ref_if.rhs_points(pts, q, 4.0))
    say(f"block GMRES K={K} precondition={pre}")
    Sb, stb = op.solve(F, rtol=1e-12, precondition=pre, krylov="block")
    Sc, stc = op.solve(F, rtol=1e-12, precondition=pre, krylov="columns")
    say(f"  iterations block {stb['niter']}, per column {stc['niter']}")
    assert stb["solved"] and stb["max_rel_residual"] < 1e-11
    for k in range(K):
        assert rel(Sb[:, k], S_ref[:, k]) < 1e-10
        assert rel(Sb[:, k], Sc[:, k]) < 1e-10


def test_block_gmres_restarts_and_mixed(dbie, cfg):
    g, op, ref_if, Ad = cfg
    rng = np.random.default_rng(70)
    K = 24
    F = dbie.rhs_point_charges(g, rng.uniform(-0.25, 0.25, (K, 3)), rng.uniform(-1, 1, K), 4.0)
    S64, _ = op.solve(F, rtol=1e-12)
    say("block GMRES restarted every 3 block iterations")
    Sr, st = op.solve(F, rtol=1e-12, restart=3, krylov="block")
    assert st["solved"]
    assert max(rel(Sr[:, k], S64[:, k]) for k in range(K)) < 1e-10
    say("mixed precision with block GMRES inner solves")
    Sm, stm = op.solve(F, rtol=1e-11, precision="fp16x3", krylov="block")
    assert stm["solved"] and stm["nrefine"] >= 1
    assert max(rel(Sm[:, k], S64[:, k]) for k in range(K)) < 1e-8
