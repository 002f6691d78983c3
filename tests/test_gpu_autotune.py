"""GPU checks of the far-field launch-shape autotuner: every shape it may pick gives the same
operator (to the bar of its precision), and the online tuner settles on one shape per class.

Bars: 1e-12 relative against the oracle for FP64 applies, 1e-5 for FP32 and 2e-6 for the
split-FP16 form against the FP64 apply (per column)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CFG1 = (1.2, 0.8, 0.6, 4, 0.2)
BOUND = {"fp32": 1e-5, "fp16x3": 2e-6}


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def col_err(Y, Yref):
    return max(rel(Y[:, k], Yref[:, k]) for k in range(Y.shape[1]))


@pytest.fixture(scope="module")
def cfg(dbie, O):
    say("build config-1 operator (N=2560) for the autotuner checks")
    g = dbie.single_dielectric_box3d(*CFG1, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    ref = O.CorrectedOperator(O.Interface.box(*CFG1, 4.0, 1.0))
    yield g, op, ref
    dbie.autotune(True, clear=True)  # pinned shapes stay out of later tests


@pytest.mark.parametrize("prec,K", [("fp64", 8), ("fp64", 40), ("fp32", 40), ("fp16x3", 32), ("fp16x3", 256)])
def test_every_split_count_gives_the_same_operator(dbie, cfg, prec, K):
    g, op, ref = cfg
    X = np.random.default_rng(21).standard_normal((g.n_points, K))
    Yref = ref @ X
    variants = (-1, 0, 1) if prec == "fp16x3" and K > 128 else (-1,)
    for splits in (1, 2, 3, 4, 8):
        for var in variants:
            say(f"pinned shape {prec} K={K} splits={splits} variant={var}")
            op.set_tuned_shape(K, prec, splits, var)
            assert op.tuned_shape(K, prec) == (splits, var)
            Y = op.apply(X, prec)
            if prec == "fp64":
                assert rel(Y, Yref) < 1e-12
            else:
                assert col_err(Y, Yref) < BOUND[prec]


@pytest.mark.parametrize("prec,K", [("fp64", 24), ("fp16x3", 48), ("fp16x3", 256)])
def test_tuner_picks_and_caches_a_shape(dbie, cfg, prec, K):
    """The online tuner tries one candidate per apply (every result stays within the bar), settles
    once every candidate is timed, and then replays the chosen shape bit for bit."""
    g, op, ref = cfg
    X = np.random.default_rng(22).standard_normal((g.n_points, K))
    Yref = ref @ X
    dbie.autotune(True, clear=True)
    assert op.tuned_shape(K, prec) is None

    def check(Y):
        if prec == "fp64":
            assert rel(Y, Yref) < 1e-12
        else:
            assert col_err(Y, Yref) < BOUND[prec]

    say(f"online tuning {prec} K={K}")
    shape = None
    for n in range(60):
        check(op.apply(X, prec))
        shape = op.tuned_shape(K, prec)
        if shape is not None:
            break
    say(f"  chosen shape {shape} after {n + 1} applies")
    assert shape is not None and shape[0] >= 1
    Y1 = op.apply(X, prec)
    Y2 = op.apply(X, prec)  # settled: same shape, same launches, same bits
    assert np.array_equal(Y1, Y2)
    assert op.tuned_shape(K, prec) == shape
    check(Y1)
    # switched off: default shapes, cached choices kept
    dbie.autotune(False)
    try:
        check(op.apply(X, prec))
        assert op.tuned_shape(K, prec) == shape
    finally:
        dbie.autotune(True)


@pytest.mark.parametrize("K", [256, 300])
def test_warp_specialised_kernel_per_source_exclusion(dbie, O, K):
    """n_quad = 5 (25-node panels): the warp-specialised split-FP16 kernel leaves the near pairs
    out source by source (a lane's 8 sources may straddle two panels); both consumer layouts and
    the general kernel give the same operator to the split-FP16 bar."""
    say(f"n_quad=5 split-FP16 far field, K={K}: variants -1, 0, 1")
    g = dbie.single_dielectric_box3d(1.2, 0.8, 0.6, 5, 0.2, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    ref = O.CorrectedOperator(O.Interface.box(1.2, 0.8, 0.6, 5, 0.2, 4.0, 1.0))
    X = np.random.default_rng(23).standard_normal((g.n_points, K))
    Yref = ref @ X
    out = {}
    try:
        for splits in (1, 3):
            for var in (-1, 0, 1):
                op.set_tuned_shape(K, "fp16x3", splits, var)
                Y = op.apply(X, "fp16x3")
                err = col_err(Y, Yref)
                assert err < BOUND["fp16x3"], (splits, var, err)
                out[(splits, var)] = Y
    finally:
        dbie.autotune(True, clear=True)
    say(f"  max column error vs oracle {max(col_err(Y, Yref) for Y in out.values()):.2e}")


def test_split_fp16_applies_are_bit_reproducible(dbie):
    """Repeated split-FP16 applies at N = 64,384, K = 256 give the same bits, eager and replayed as a
    HIP graph, for the warp-specialised kernel with one (variant 1) and two (variant -1) consumer waves
    per SIMD and the general kernel (variant 0), at 1 source split (the longest stage walk: 2012 stages
    of 32 sources per workgroup) and at 6 splits.  Each stage's source coordinates are copied into LDS
    by three producer waves and read by all four, so the copies must have landed in every wave before
    the barrier that precedes their use."""
    say("bit-reproducible split-FP16 applies, N=64384, K=256")
    g = dbie.single_dielectric_box3d(2.4, 1.6, 1.2, 4, 0.05, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    X = np.random.default_rng(24).standard_normal((g.n_points, 256))
    try:
        for graphs in (False, True):
            op.graphs(graphs)
            for splits, var in ((1, 1), (1, -1), (1, 0), (6, 1), (6, -1), (6, 0)):
                op.set_tuned_shape(256, "fp16x3", splits, var)
                ref = op.apply(X, "fp16x3")
                for it in range(6):
                    assert np.array_equal(op.apply(X, "fp16x3"), ref), (graphs, splits, var, it)
    finally:
        op.graphs(True)
        dbie.autotune(True, clear=True)
