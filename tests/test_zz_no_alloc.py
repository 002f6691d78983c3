"""dbie_apply_device_noalloc: the operator apply whose workspaces come from the library's fixed
device scratch array instead of an allocation.  Through the C ABI it must give the bits of
dbie_apply_device with the same launch shape (two calls in a row, the second on a smaller input),
match the oracle, and reject a size beyond the scratch array with DBIE_BAD_ARG."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 256 << 20


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def noalloc_fn(dbie):
    from dbie import _lib as L
    fn = L.lib().dbie_apply_device_noalloc
    fn.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int]
    fn.restype = C.c_int
    return fn


def run_pair(dbie, op, X, prec, shape):
    """(existing entry point, no-alloc entry point) results for X with the launch shape pinned."""
    from dbie import _lib as L
    N, K = X.shape
    op.set_tuned_shape(K, prec, *shape)
    dX = dbie.DeviceMatrix.from_host(X)
    dY1 = dbie.DeviceMatrix(N, K)
    dY2 = dbie.DeviceMatrix(N, K)
    code = L.PRECISIONS[prec]
    L.check(L.lib().dbie_apply_device(op._h, K, dX.ptr, N, dY1.ptr, N, code))
    op.synchronize()
    L.check(noalloc_fn(dbie)(op._h, K, dX.ptr, N, dY2.ptr, N, code))
    op.synchronize()
    return dY1.to_host(), dY2.to_host()


@pytest.mark.parametrize("prec,Ks,shape", [("fp64", (24, 8), (2, -1)), ("fp16x3", (256, 40), (1, 1)),
                                           ("fp16x3", (12, 3), (1, -1)), ("fp32", (20, 3), (2, -1))])
def test_noalloc_apply_equals_apply_device(dbie, O, prec, Ks, shape):
    g = dbie.single_dielectric_box3d(1.2, 0.8, 0.6, 4, 0.2, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    ref = O.CorrectedOperator(O.Interface.box(1.2, 0.8, 0.6, 4, 0.2, 4.0, 1.0))
    rng = np.random.default_rng(41)
    bound = {"fp64": 1e-12, "fp32": 1e-5, "fp16x3": 2e-6}[prec]
    try:
        for K in Ks:  # two calls in a row, the second on a smaller input
            say(f"no-alloc apply {prec} N={g.n_points} K={K} vs dbie_apply_device and the oracle")
            X = rng.standard_normal((g.n_points, K))
            Y1, Y2 = run_pair(dbie, op, X, prec, shape)
            assert np.array_equal(Y1, Y2), (prec, K, float(np.abs(Y1 - Y2).max()))
            Yr = ref @ X
            err = (np.linalg.norm(Y2 - Yr, axis=0) / np.linalg.norm(Yr, axis=0)).max()
            assert err < bound, (prec, K, err)
    finally:
        dbie.autotune(True, clear=True)


def test_noalloc_config3_size(dbie):
    """N = 64,384, K = 256 split-FP16 (the batched solve's shape) fits the scratch array."""
    say("no-alloc apply fp16x3 N=64384 K=256, then K=16 fp64")
    g = dbie.single_dielectric_box3d(2.4, 1.6, 1.2, 4, 0.05, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    rng = np.random.default_rng(42)
    try:
        for prec, K, shape in (("fp16x3", 256, (1, 1)), ("fp64", 16, (1, -1))):
            X = rng.standard_normal((g.n_points, K))
            Y1, Y2 = run_pair(dbie, op, X, prec, shape)
            assert np.array_equal(Y1, Y2), (prec, K)
    finally:
        dbie.autotune(True, clear=True)


def test_noalloc_rejects_oversize(dbie):
    from dbie import _lib as L
    say("no-alloc apply beyond the scratch array is refused")
    g = dbie.single_dielectric_box3d(1.2, 0.8, 0.6, 4, 0.2, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    N = g.n_points
    K = SCRATCH_BYTES // (8 * N) + 8  # one split of partial sums alone exceeds the array
    dX = dbie.DeviceMatrix(N, K)
    dY = dbie.DeviceMatrix(N, K)
    st = noalloc_fn(dbie)(op._h, K, dX.ptr, N, dY.ptr, N, L.FP64)
    assert st == L.BAD_ARG
    assert "scratch" in L.last_error()
