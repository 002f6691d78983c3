"""The corrected double layer on the MI355X: laplace3d_D_fmm3d_corrected (src/kernel/laplace3d_near.jl:421-445)
with the D correction blocks of laplace3d_D_corrections (src/kernel/laplace3d_near_upsampling.jl:278-282:
source normal, the D direct term), built on the device on the operator's near-pair list, against the
oracle's D blocks and its dense D; and the corrected Dt map laplace3d_DT_fmm3d_corrected (:395-419)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module", params=[4, 5])
def case(request, dbie, O):
    nq = request.param
    say(f"config-1 box, n_quad={nq}: operator and oracle D corrections")
    g = dbie.single_dielectric_box3d(1.2, 0.8, 0.6, nq, 0.2, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    ref_if = O.Interface.box(1.2, 0.8, 0.6, nq, 0.2, 4.0, 1.0)
    nl = ref_if.neighbor_list(8, 1e-9, True)
    BD = nl.corrections(mode="D").copy()
    CD = nl.sparse()
    BT = nl.corrections(mode="DT").copy()
    CT = nl.sparse()
    return g, op, ref_if, nl, BD, CD, BT, CT


def test_d_correction_blocks_vs_oracle(case):
    g, op, ref_if, nl, BD, CD, BT, CT = case
    say("D correction blocks (device build) vs oracle")
    src, trg, blk = op.corrections("D")
    s2, t2 = nl.pairs()
    order = np.lexsort((s2, t2))  # the product stores block rows by target, then source
    assert np.array_equal(src, s2[order]) and np.array_equal(trg, t2[order])
    scale = np.abs(BD).max()
    err = np.abs(blk - BD[order]).max() / scale
    say(f"  max |device - oracle| / scale {err:.2e}")
    assert err <= 1e-13
    # the Dt blocks are still the operator's own
    _, _, bt = op.corrections("DT")
    assert np.abs(bt - BT[order]).max() <= 1e-13 * np.abs(BT).max()


def test_d_blocks_host_build_equals_device(dbie, case):
    g, op, ref_if, nl, BD, CD, BT, CT = case
    say("D correction blocks: host build == device build")
    oph = dbie.BatchedDielectricOperator(g, device_build=False)
    _, _, bh = oph.corrections("D")
    _, _, bd_ = op.corrections("D")
    assert np.abs(bd_ - bh).max() <= 1e-14 * np.abs(bh).max()


@pytest.mark.parametrize("K", [1, 5, 20])
def test_corrected_d_apply_vs_oracle(case, K):
    g, op, ref_if, nl, BD, CD, BT, CT = case
    say(f"corrected D and Dt apply, K={K}, vs oracle dense + corrections")
    X = np.random.default_rng(50 + K).standard_normal((g.n_points, K))
    Yd = op.apply_kernel(X, "D", corrected=True)
    assert rel(Yd, ref_if.dense_D() @ X + CD @ X) < 1e-12
    Yt = op.apply_kernel(X, "DT", corrected=True)
    assert rel(Yt, ref_if.dense_DT() @ X + CT @ X) < 1e-12
    # the corrected Dt map is the solve operator without its diagonal
    assert rel(Yt, op @ X - g.diag()[:, None] * X) < 1e-12


def test_corrected_d_constant_density(dbie, case):
    """D 1 = 1/2 on a closed surface (test/kernel/laplace3d.jl:49-56); the corrections resolve the
    near-singular panel integrals that the plain quadrature misses by up to 0.26."""
    g, op, ref_if, nl, BD, CD, BT, CT = case
    say("corrected D * 1 ~ 1/2")
    Dm = dbie.laplace3d_D_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    y = Dm @ np.ones(g.n_points)
    y0 = op.apply_kernel(np.ones(g.n_points), "D")
    say(f"  max |D 1 - 1/2|: plain {np.abs(y0 - 0.5).max():.2e}, corrected {np.abs(y - 0.5).max():.2e}")
    assert np.abs(y - 0.5).max() < 1e-3
    assert np.dot(np.abs(y - 0.5), g.weights) < 1e-4
    assert np.abs(y0 - 0.5).max() > 0.1
