"""HIP-graph replay of repeated operator applies: every replay equals the eager apply bit for bit,
with eager work and synchronisations between replays, and a graph whose workspace has been
reallocated is retired instead of replayed."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG1 = (1.2, 0.8, 0.6, 4, 0.2)


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def op(dbie):
    say("build config-1 operator (N=2560) for the graph checks")
    g = dbie.single_dielectric_box3d(*CFG1, 4.0, 1.0)
    o = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    yield o
    o.graphs(True)


@pytest.mark.parametrize("prec,K", [("fp64", 8), ("fp64", 40), ("fp16x3", 48), ("fp16x3", 256), ("fp32", 16)])
def test_graph_replays_equal_eager(dbie, op, prec, K):
    n = op.n
    rng = np.random.default_rng(40 + K)
    # tune the class first so that the launch shape is settled (the tuner's candidates run eagerly)
    X0 = rng.standard_normal((n, K))
    for _ in range(80):
        op.apply(X0, prec)
        if op.tuned_shape(K, prec) is not None:
            break
    c0, r0 = op.graphs(True)
    say(f"graph replays {prec} K={K}")
    for it in range(6):
        X = rng.standard_normal((n, K))
        Yg = op.apply(X, prec)        # first sighting eager, then captured, then replayed
        op.graphs(False)
        Ye = op.apply(X, prec)        # eager launches of the same apply
        op.graphs(True)
        assert np.array_equal(Yg, Ye), it
    c1, r1 = op.graphs()
    say(f"  captures {c1 - c0}, replays {r1 - r0}")
    assert c1 - c0 >= 1 and r1 - r0 >= 4


def test_graph_retired_when_workspace_moves(dbie, op):
    n = op.n
    rng = np.random.default_rng(50)
    op.graphs(True)
    X = rng.standard_normal((n, 24))
    for _ in range(3):
        op.apply(X, "fp64")
    _, r0 = op.graphs()
    # a wider apply grows the partial-sum workspace; the K = 24 graph must not replay into the old one
    op.apply(rng.standard_normal((n, 300)), "fp64")
    Yg = op.apply(X, "fp64")
    op.graphs(False)
    Ye = op.apply(X, "fp64")
    op.graphs(True)
    assert np.array_equal(Yg, Ye)
    Yg2 = op.apply(X, "fp64")
    Yg3 = op.apply(X, "fp64")
    assert np.array_equal(Yg2, Ye) and np.array_equal(Yg3, Ye)
    assert op.graphs()[1] > r0


def test_solve_with_graphs_matches_eager(dbie, op):
    """GMRES replays the captured apply every iteration; the solution equals the eager solve."""
    rng = np.random.default_rng(51)
    g_pts = rng.uniform(-0.2, 0.2, (6, 3))
    q = rng.uniform(-1, 1, 6)
    iface = dbie.single_dielectric_box3d(*CFG1, 4.0, 1.0)
    F = dbie.rhs_point_charges(iface, g_pts, q, 4.0)
    op.graphs(True)
    Sg, stg = op.solve(F, rtol=1e-12)
    op.graphs(False)
    Se, ste = op.solve(F, rtol=1e-12)
    op.graphs(True)
    assert stg["niter"] == ste["niter"]
    assert np.array_equal(Sg, Se)
