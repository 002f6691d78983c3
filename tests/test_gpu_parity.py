"""GPU parity of libdbie against the CPU oracle (run with -m gpu on the MI355X box).

Bars: 1e-12 relative for operator applies and blocks, 1e-10 for FP64 solutions,
1e-8 for mixed-precision solutions against the FP64 solution."""
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
def cfg1(dbie, O):
    say("build config-1 operators (N=2560)")
    g = dbie.single_dielectric_box3d(*CFG1, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    ref_if = O.Interface.box(*CFG1, 4.0, 1.0)
    ref = O.CorrectedOperator(ref_if)
    return g, op, ref_if, ref


def test_corrected_apply_parity(cfg1):
    g, op, ref_if, ref = cfg1
    say("corrected apply K=1,5 vs oracle")
    rng = np.random.default_rng(0)
    for K in (1, 5):
        X = rng.standard_normal((g.n_points, K))
        assert rel(op @ X, ref @ X) < 1e-12
    x = np.linspace(0.1, 1.0, g.n_points)  # test/solver/multi_rhs.jl:47-57
    assert rel(op @ x, ref @ x) < 1e-12


def test_correction_blocks_parity(cfg1):
    g, op, ref_if, ref = cfg1
    say("correction blocks vs oracle")
    src, trg, blk = op.corrections()
    s2, t2 = ref.nl.pairs()
    B2 = ref.nl.blocks
    order = np.lexsort((s2, t2))  # product stores block rows by target, then source
    assert np.array_equal(src, s2[order]) and np.array_equal(trg, t2[order])
    scale = np.abs(B2).max()
    assert np.abs(blk - B2[order]).max() <= 1e-13 * scale


def test_plain_kernels_parity(dbie, O, cfg1):
    g, op, ref_if, ref = cfg1
    say("plain DT/S/D apply vs oracle dense")
    X = np.random.default_rng(1).standard_normal((g.n_points, 3))
    for name, M in (("DT", ref_if.dense_DT()), ("S", ref_if.dense_S()), ("D", ref_if.dense_D())):
        assert rel(op.apply_kernel(X, name), M @ X) < 1e-12, name


def test_dense_lhs_and_lu(dbie, O, cfg1):
    g, op, ref_if, ref = cfg1
    say("dense lhs + rocSOLVER LU (config 1, README quickstart)")
    A = dbie.lhs_dielectric_box3d(g)
    A_ref = ref_if.lhs_dense()
    assert np.abs(A - A_ref).max() <= 1e-13 * np.abs(A_ref).max()
    b = dbie.rhs_dielectric_box3d(g, dbie.PointSource((0.1, 0.1, 0.1), 1.0), 4.0)
    b_ref = ref_if.rhs_points([0.1, 0.1, 0.1], [1.0], 4.0)[:, 0]
    assert rel(b, b_ref) < 1e-13
    x = dbie.solve_lu(A, b)
    x_ref = O.solve_lu(A_ref, b_ref)
    assert rel(x, x_ref) < 1e-10
    assert np.linalg.norm(A_ref @ x - b_ref) < 1e-10
    assert abs(np.dot(g.weights, x) + 0.25 - 1.0) < 0.1


def test_fp64_solve_matches_oracle(dbie, O, cfg1):
    g, op, ref_if, ref = cfg1
    say("FP64 GMRES on the corrected operator vs oracle LU")
    F = dbie.rhs_point_charges(g, [[0.1, 0.1, 0.1], [-0.2, 0.05, 0.1]], [1.0, -0.5], 4.0)
    Sig, st = op.solve(F, rtol=1e-12)
    assert st["solved"] and st["max_rel_residual"] < 1e-11
    Ad = ref_if.dense_DT() + ref.Cs.toarray() + np.diag(ref.t)
    S_ref = O.solve_lu(Ad, ref_if.rhs_points([[0.1, 0.1, 0.1], [-0.2, 0.05, 0.1]], [1.0, -0.5], 4.0))
    for k in range(2):
        assert rel(Sig[:, k], S_ref[:, k]) < 1e-10


def test_block_equals_per_column(dbie, cfg1):
    # test/solver/multi_rhs.jl:59-66 and multi_rhs_vector.jl:38-48
    g, op, ref_if, ref = cfg1
    say("block solve == per-column solves")
    rng = np.random.default_rng(3)
    pts = rng.uniform(-0.25, 0.25, (4, 3))
    F = dbie.rhs_point_charges(g, pts, rng.uniform(-1, 1, 4), 4.0)
    S, _ = op.solve(F, rtol=1e-12)
    for k in range(4):
        sk, _ = op.solve(F[:, k], rtol=1e-12)
        assert rel(S[:, k], sk) < 1e-10


def test_mfma_apply_parity(cfg1):
    """K >= 16 routes the far field through the MFMA kernels (FP64: v_mfma_f64_16x16x4)."""
    g, op, ref_if, ref = cfg1
    rng = np.random.default_rng(11)
    for K in (16, 40, 256):
        say(f"MFMA FP64 apply K={K} vs oracle")
        X = rng.standard_normal((g.n_points, K))
        Y = op @ X
        Yr = ref @ X
        for k in (0, K // 2, K - 1):
            assert rel(Y[:, k], Yr[:, k]) < 1e-12
        assert rel(Y, Yr) < 1e-12
    X = rng.standard_normal((g.n_points, 48))
    Y64 = op @ X
    for prec, bound in (("fp32", 1e-5), ("bf16", 2e-2), ("bf16x3", 2e-4), ("fp16x3", 2e-6)):
        say(f"low-precision apply {prec} K=48")
        Yp = op.apply(X, prec)
        errs = [rel(Yp[:, k], Y64[:, k]) for k in range(48)]
        say(f"  max column error {max(errs):.3e}")
        for k in range(48):
            assert errs[k] < bound, (prec, k)


@pytest.mark.parametrize("K", [32, 160])
def test_fp16x3_localized_inputs(cfg1, K):
    """Inputs concentrated on edge panels (like iterative-refinement corrections): the split-FP16
    operator leaves the near panel pairs to the full FP32 near blocks, so its error stays at the
    entry level instead of being amplified by the near-pair cancellation.  K = 160 takes the
    256-column (warp-specialised) kernel."""
    g, op, ref_if, ref = cfg1
    rng = np.random.default_rng(13)
    _, _, is_edge = g.panels()
    mask = np.repeat(is_edge, g.n_quad ** 2).astype(float)
    X = rng.standard_normal((g.n_points, K)) * mask[:, None]
    Y64 = op @ X
    errs = {}
    for prec in ("fp16x3", "bf16x3"):
        Yp = op.apply(X, prec)
        errs[prec] = max(rel(Yp[:, k], Y64[:, k]) for k in range(K))
    say(f"localized inputs K={K}: max column error {errs}")
    assert errs["fp16x3"] < 1e-6


@pytest.mark.parametrize("K", [256, 300])
def test_fp16x3_wide_blocks(cfg1, K):
    """256-column blocks (one full block; a full and a partial one) against the FP64 operator."""
    g, op, ref_if, ref = cfg1
    say(f"split-FP16 apply K={K} vs FP64")
    X = np.random.default_rng(14).standard_normal((g.n_points, K))
    Y64 = op @ X
    Yp = op.apply(X, "fp16x3")
    errs = np.array([rel(Yp[:, k], Y64[:, k]) for k in range(K)])
    say(f"  max column error {errs.max():.3e}")
    assert errs.max() < 2e-6


@pytest.mark.parametrize("K", [32, 256])
def test_fp16x3_column_scaling(cfg1, K):
    """Columns spanning 40 decades keep their relative accuracy (per-column power-of-two scaling)."""
    g, op, ref_if, ref = cfg1
    rng = np.random.default_rng(12)
    X = rng.standard_normal((g.n_points, K)) * (10.0 ** np.linspace(-20, 20, K))[None, :]
    X[:, 5] = 0.0  # an all-zero column stays exactly zero
    Y64 = op @ X
    Yp = op.apply(X, "fp16x3")
    assert np.all(Yp[:, 5] == 0.0)
    for k in range(K):
        if k != 5:
            assert rel(Yp[:, k], Y64[:, k]) < 2e-6, k


@pytest.mark.parametrize("precision,K", [("fp32", 8), ("bf16", 8), ("bf16", 16), ("bf16x3", 24), ("fp32", 20),
                                               ("fp16x3", 16), ("fp16x3", 40)])
def test_mixed_precision_solve(dbie, cfg1, precision, K):
    g, op, ref_if, ref = cfg1
    say(f"mixed-precision IR ({precision}, K={K}) vs FP64 solve")
    rng = np.random.default_rng(4)
    F = dbie.rhs_point_charges(g, rng.uniform(-0.25, 0.25, (K, 3)), rng.uniform(-1, 1, K), 4.0)
    S64, _ = op.solve(F, rtol=1e-12)
    Sm, st = op.solve(F, rtol=1e-11, precision=precision)
    assert st["solved"] and st["nrefine"] >= 1
    for k in range(K):
        assert rel(Sm[:, k], S64[:, k]) < 1e-8


@pytest.mark.parametrize("precision,restart", [("fp16x3", 0), ("fp16x3", 4), ("bf16", 0), ("fp32", 3)])
def test_recycled_refinement(dbie, cfg1, precision, restart):
    """Refinements that deflate the first inner solve's Krylov space (GCRO) reach the same
    FP64-verified tolerance as restarted ones, with no more inner iterations; a zero column and
    short restart cycles (deflation re-applied at each cycle start) included."""
    g, op, ref_if, ref = cfg1
    K = 12
    say(f"recycled vs restarted refinement ({precision}, restart={restart})")
    rng = np.random.default_rng(14)
    F = dbie.rhs_point_charges(g, rng.uniform(-0.25, 0.25, (K, 3)), rng.uniform(-1, 1, K), 4.0)
    F[:, 3] = 0.0
    S64, _ = op.solve(F, rtol=1e-12)
    out = {}
    for rec in (False, True):
        Sm, st = op.solve(F, rtol=1e-11, precision=precision, inner_rtol=1e-5, restart=restart, recycle=rec)
        assert st["solved"] and st["max_rel_residual"] <= 1e-11, st
        assert np.all(Sm[:, 3] == 0.0)
        for k in range(K):
            if k != 3:
                assert rel(Sm[:, k], S64[:, k]) < 1e-8, (rec, k)
        out[rec] = st
    say(f"  inner iterations: restarted {out[False]['niter']}, recycled {out[True]['niter']}")
    assert out[True]["niter"] <= out[False]["niter"] + 1


@pytest.mark.parametrize("nq", [2, 3, 5])
def test_near_mfma_row_major_blocks(dbie, O, nq):
    """n_quad = 2, 3, 5 (4x4, 9x9, 25x25 near blocks, row-major): the FP64 MFMA near kernel (K >= 16;
    256-column groups, a partial last group) matches the oracle to 1e-12 and the one-thread-per-output
    kernel (K < 16) to rounding; the FP32-block forms stay within their precision's bar."""
    say(f"near MFMA kernel, row-major {nq * nq}x{nq * nq} blocks")
    g = dbie.single_dielectric_box3d(1.2, 0.8, 0.6, nq, 0.2, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    ref = O.CorrectedOperator(O.Interface.box(1.2, 0.8, 0.6, nq, 0.2, 4.0, 1.0))
    rng = np.random.default_rng(30 + nq)
    for K in (16, 40, 300):
        X = rng.standard_normal((g.n_points, K))
        Y = op.apply(X, "fp64")
        assert rel(Y, ref @ X) < 1e-12, K
        Yv = op.apply(X[:, :8], "fp64")  # K < 16: one thread per output
        assert rel(Y[:, :8], Yv) < 1e-14, K
        for prec, bound in (("fp32", 1e-5), ("fp16x3", 2e-6)):
            Yp = op.apply(X, prec)
            err = max(rel(Yp[:, k], Y[:, k]) for k in range(K))
            assert err < bound, (K, prec, err)


def test_rhs_volume_screened(dbie, O, cfg1):
    g, op, ref_if, ref = cfg1
    say("volume RHS with sharp screening vs oracle")
    rng = np.random.default_rng(5)
    vss, cols = [], []
    for k in range(3):
        pos = rng.uniform(-0.8, 0.8, (50 + 10 * k, 3))
        w = rng.uniform(0.1, 1.0, len(pos)); rho = rng.standard_normal(len(pos))
        vss.append(dbie.VolumeSource(pos.T, w, rho))
        cols.append(ref_if.rhs_volume(pos, w, ref_if.screen_sharp(pos, rho), 1.0))
    F = dbie.rhs_dielectric_box3d_fmm3d(g, vss, 1e-9)
    for k in range(3):
        assert rel(F[:, k], cols[k]) < 1e-12
    # unscreened single source with eps_src
    f1 = dbie.rhs_dielectric_box3d(g, vss[0], 2.0)
    assert rel(f1, ref_if.rhs_volume(vss[0].positions.T, vss[0].weights, vss[0].density, 2.0)) < 1e-12


def test_pottrg_parity(dbie, O, cfg1):
    g, op, ref_if, ref = cfg1
    say("pottrg vs oracle")
    rng = np.random.default_rng(6)
    trg = rng.uniform(-0.5, 0.5, (300, 3))
    Sg = rng.standard_normal((g.n_points, 2))
    assert rel(op.pottrg(trg, Sg), ref_if.pottrg(trg, Sg)) < 1e-12


def test_config2_sampled_rows(dbie, O):
    say("config-2 size (N=64384) corrected apply, sampled rows vs oracle")
    box = (2.4, 1.6, 1.2, 4, 0.05)
    g = dbie.single_dielectric_box3d(*box, 10.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    ref_if = O.Interface.box(*box, 10.0, 1.0)
    nl = ref_if.neighbor_list(8, 1e-9, True)
    nl.corrections()
    x = np.random.default_rng(7).standard_normal(g.n_points)
    y = op @ x
    rows = np.sort(np.random.default_rng(8).choice(g.n_points, 512, replace=False)).astype(np.int32)
    y_ref = ref_if.apply_DT_rows(rows, x) + nl.apply(x)[rows] + ref_if.diag()[rows] * x[rows]
    assert rel(y[rows], y_ref) < 1e-12
    assert op.n_blocks == 196768 + 4080


def test_device_resident_solve(dbie, cfg1):
    g, op, ref_if, ref = cfg1
    say("device-resident solve == host solve (FP64 and mixed)")
    rng = np.random.default_rng(9)
    F = dbie.rhs_point_charges(g, rng.uniform(-0.25, 0.25, (32, 3)), rng.uniform(-1, 1, 32), 4.0)
    dF = dbie.DeviceMatrix.from_host(F)
    dS = dbie.DeviceMatrix(g.n_points, 32)
    S_host, _ = op.solve(F, rtol=1e-12)
    st = op.solve_device(dF, dS, rtol=1e-12)
    assert st["solved"] and rel(dS.to_host(), S_host) < 1e-12
    st = op.solve_device(dF, dS, rtol=1e-11, precision="bf16x3", inner_rtol=1e-5)
    assert st["solved"]
    Sm = dS.to_host()
    for k in range(32):
        assert rel(Sm[:, k], S_host[:, k]) < 1e-8


@pytest.mark.parametrize("nq,max_order,edges", [(4, 8, True), (4, 16, True), (5, 16, True), (4, 8, False)])
def test_device_built_blocks_match_host(dbie, nq, max_order, edges):
    """Pair list and correction blocks built on the device equal the host-built ones (same
    decisions, same sums in the same order)."""
    say(f"device-built vs host-built correction blocks, n_quad={nq}, max_order={max_order}, edges={edges}")
    g = dbie.single_dielectric_box3d(1.0, 0.8, 0.6, nq, 0.2, 4.0, 1.0)
    opd = dbie.BatchedDielectricOperator(g, max_order=max_order, correct_edges=edges, device_build=True)
    oph = dbie.BatchedDielectricOperator(g, max_order=max_order, correct_edges=edges, device_build=False)
    sd, td, bd_ = opd.corrections()
    sh, th, bh = oph.corrections()
    assert np.array_equal(sd, sh) and np.array_equal(td, th)
    scale = np.abs(bh).max()
    say(f"  max |device - host| / scale {np.abs(bd_ - bh).max() / scale:.2e}, "
        f"bit-equal entries {np.mean(bd_ == bh):.6f}")
    assert np.abs(bd_ - bh).max() <= 1e-14 * scale
    X = np.random.default_rng(21).standard_normal((g.n_points, 32))
    assert rel(opd @ X, oph @ X) < 1e-13
    Y64 = oph @ X
    for prec in ("fp16x3", "fp32"):
        Yd, Yh = opd.apply(X, prec), oph.apply(X, prec)
        assert rel(Yd, Yh) < 1e-7 * (1 + 100 * (prec == "fp32")), prec
        assert rel(Yd, Y64) < (2e-6 if prec == "fp16x3" else 1e-5), prec


def test_operator_from_caller_nodes(dbie, O, cfg1):
    """An interface built from the panels' own quadrature (points, weights, GL rule as given) gives the
    mesher's operator bit for bit when the nodes are the mesher's; with nodes moved by 1e-9 the plain Dt
    apply follows the caller's nodes (numpy direct sum on those nodes)."""
    g, op, ref_if, ref = cfg1
    say("operator from caller-supplied panel nodes")
    corners, normals, is_edge = g.panels()
    nd = g.nodes
    x, w = dbie.gausslegendre(4)
    g2 = dbie.interface_from_panels(corners, normals, 4, 4.0, 1.0, is_edge, points=nd[0:3].T, weights=nd[6],
                                    gl_xs=x, gl_ws=w)
    op2 = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g2, 1e-9, 1e-9, 8)
    X = np.random.default_rng(60).standard_normal((g.n_points, 3))
    assert np.array_equal(op2 @ X, op @ X)
    pts = nd[0:3].T + 1e-9 * np.random.default_rng(61).standard_normal((g.n_points, 3))
    g3 = dbie.interface_from_panels(corners, normals, 4, 4.0, 1.0, is_edge, points=pts, weights=nd[6],
                                    gl_xs=x, gl_ws=w)
    op3 = dbie.BatchedDielectricOperator(g3, corrected=False)
    d = pts[:, None, :] - pts[None, :, :]
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, np.inf)
    nrm = nd[3:6].T
    DT = np.einsum("ik,ijk->ij", nrm, d) / (4 * np.pi * r ** 3) * nd[6][None, :]
    assert rel(op3.apply_kernel(X, "DT"), DT @ X) < 1e-12


@pytest.mark.parametrize("precision,krylov", [("fp64", "columns"), ("fp64", "block"), ("fp16x3", "columns"),
                                              ("fp16x3", "block")])
def test_residual_history(dbie, cfg1, precision, krylov):
    """Per-solve residual history (the Krylov stats of src/solver/multi_rhs.jl:285): one Krylov entry per
    (inner) iteration, one FP64 entry per true residual (nrefine of them for a mixed solve, the final
    check of an FP64 solve), ending at the reported residual below the tolerance."""
    g, op, ref_if, ref = cfg1
    say(f"residual history ({precision}, {krylov})")
    rng = np.random.default_rng(17)
    F = dbie.rhs_point_charges(g, rng.uniform(-0.25, 0.25, (20, 3)), rng.uniform(-1, 1, 20), 4.0)
    S, st = op.solve(F, rtol=1e-10, precision=precision, inner_rtol=1e-5, krylov=krylov)
    h = st["residual_history"]
    kinds = [k for k, _ in h]
    assert kinds.count("krylov") == st["niter"]
    assert kinds.count("fp64") == (st["nrefine"] if precision != "fp64" else 1)
    assert h[-1][0] == "fp64" and abs(h[-1][1] - st["max_rel_residual"]) <= 1e-3 * st["max_rel_residual"]
    assert h[-1][1] <= 1e-10
    if precision == "fp64":  # GMRES residuals never increase
        kr = [v for k, v in h if k == "krylov"]
        assert all(b <= a * (1 + 1e-12) for a, b in zip(kr, kr[1:]))


@pytest.mark.parametrize("drop", [1, 2, 3])
def test_split_fp16_near_super_rows_ragged(dbie, O, drop):
    """The split-FP16 near field walks super-rows of four block rows; an interface whose panel count is
    not a multiple of four (a box with its last `drop` panels removed) leaves a short last super-row.
    FP64 apply against the oracle (1e-12), split-FP16 against FP64 per column (2e-6) at widths that take
    one, two and four column blocks per wave, with an all-zero column exactly zero."""
    g = dbie.single_dielectric_box3d(*CFG1, 4.0, 1.0)
    corners, normals, is_edge = g.panels()
    P = corners.shape[0] - drop
    say(f"ragged super-rows: P={P} ({P % 4} rows in the last)")
    g2 = dbie.interface_from_panels(corners[:P], normals[:P], 4, 4.0, 1.0, is_edge[:P])
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g2, 1e-9, 1e-9, 8)
    ref = O.CorrectedOperator(O.Interface.from_panels(corners[:P], normals[:P], 4, 4.0, 1.0, is_edge[:P]))
    rng = np.random.default_rng(40 + drop)
    X = rng.standard_normal((g2.n_points, 16))
    assert rel(op @ X, ref @ X) < 1e-12
    for K in (16, 24, 100, 256):
        X = rng.standard_normal((g2.n_points, K))
        X[:, 5] = 0.0
        Y64 = op @ X
        Yp = op.apply(X, "fp16x3")
        assert np.all(Yp[:, 5] == 0.0), K
        errs = [rel(Yp[:, k], Y64[:, k]) for k in range(K) if k != 5]
        say(f"  K={K}: max column error {max(errs):.2e}")
        assert max(errs) < 2e-6, K
