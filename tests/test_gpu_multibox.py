"""GPU parity on several touching boxes (multi_dielectric_box3d, src/shape/box3d_multi.jl:358-386)
with per-panel contrasts, and box-by-box screening (src/core/sources.jl:408-515).

Bars: 1e-12 relative for applies, dense matrices and right-hand sides against the oracle, 1e-10 for
FP64 solutions, 1e-8 for mixed-precision solutions against the FP64 solution; SharpScreening
densities bit-exact, soft screening 1e-13 (erf from different libraries)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX2 = [((0.0, 0.0, 0.0), 1.0, 1.0, 1.0), ((1.0, 0.0, 0.0), 1.0, 1.0, 1.0)]
EPS2 = [2.0, 4.0]
# an L of three boxes with unequal sides and a partly shared face
BOXL = [((0.0, 0.0, 0.0), 0.8, 0.6, 0.5), ((0.7, 0.0, 0.0), 0.6, 0.4, 0.5), ((0.0, 0.55, 0.0), 0.8, 0.5, 0.5)]
EPSL = [4.0, 10.0, 2.5]


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


GEOMS = {"two_cubes": (BOX2, EPS2, 4, 0.3), "l_shape": (BOXL, EPSL, 3, 0.15)}


@pytest.fixture(scope="module", params=["two_cubes", "l_shape"])
def mb(request, dbie, O):
    boxes, epses, nq, lec = GEOMS[request.param]
    say(f"build multi-box geometry {request.param}")
    g = dbie.multi_dielectric_box3d(nq, lec, boxes, epses, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
    ref_if = O.Interface.multi_box(nq, lec, boxes, epses, 1.0)
    ref = O.CorrectedOperator(ref_if)
    say(f"  N={g.n_points} panels={g.n_panels}")
    return g, op, ref_if, ref, boxes, epses


def test_multibox_apply_parity(mb):
    g, op, ref_if, ref, _, _ = mb
    say("multi-box corrected apply K=1,6 vs oracle")
    rng = np.random.default_rng(30)
    for K in (1, 6):
        X = rng.standard_normal((g.n_points, K))
        assert rel(op @ X, ref @ X) < 1e-12


def test_multibox_dense_lu(dbie, O, mb):
    """test/shape/multi_box3d.jl:334-358 (point charge outside both boxes, eps_src = eps_out)"""
    g, op, ref_if, ref, _, _ = mb
    say("multi-box dense lhs + LU vs oracle")
    A = dbie.lhs_dielectric_box3d(g)
    A_ref = ref_if.lhs_dense()
    assert np.abs(A - A_ref).max() <= 1e-13 * np.abs(A_ref).max()
    b = dbie.rhs_dielectric_box3d(g, dbie.PointSource((3.0, 0.0, 0.0), 1.0), 1.0)
    b_ref = ref_if.rhs_points([3.0, 0.0, 0.0], [1.0], 1.0)[:, 0]
    assert rel(b, b_ref) < 1e-13
    x = dbie.solve_lu(A, b)
    assert np.all(np.isfinite(x))
    assert rel(x, O.solve_lu(A_ref, b_ref)) < 1e-10


def test_multibox_gmres_with_contrast_scaling(dbie, O, mb):
    """dielectric_diagonal_scaling (src/solver/multi_rhs.jl:213-247) matters with several contrasts:
    every form (none, diagonal scaling, the near-field Neumann term) converges to the oracle's LU
    solution."""
    g, op, ref_if, ref, _, _ = mb
    pts = [[3.0, 0.0, 0.0], [0.1, 0.1, 0.1], [-0.1, 0.05, -0.1]]
    say("multi-box FP64 GMRES (with and without diagonal scaling) vs oracle LU")
    F = dbie.rhs_point_charges(g, pts, [1.0, -0.5, 0.25], 1.0)
    Ad = ref_if.dense_DT() + ref.Cs.toarray() + np.diag(ref.t)
    S_ref = O.solve_lu(Ad, ref_if.rhs_points(pts, [1.0, -0.5, 0.25], 1.0))
    iters = {}
    for pre in (2, 1, 0):
        S, st = op.solve(F, rtol=1e-12, precondition=pre)
        assert st["solved"]
        iters[pre] = st["niter"]
        for k in range(3):
            assert rel(S[:, k], S_ref[:, k]) < 1e-10
    say(f"  iterations near-field {iters[2]}, with scaling {iters[1]}, without {iters[0]}")
    assert iters[2] < iters[1]


def test_multibox_mixed_precision(dbie, mb):
    g, op, _, _, _, _ = mb
    say("multi-box mixed-precision IR (fp16x3, K=24) vs FP64")
    rng = np.random.default_rng(31)
    F = dbie.rhs_point_charges(g, rng.uniform(-0.2, 0.2, (24, 3)), rng.uniform(-1, 1, 24), 1.0)
    S64, _ = op.solve(F, rtol=1e-12)
    for pre in (1, 2):
        Sm, st = op.solve(F, rtol=1e-11, precision="fp16x3", precondition=pre)
        assert st["solved"]
        assert max(rel(Sm[:, k], S64[:, k]) for k in range(24)) < 1e-8


def test_multibox_screened_volume_rhs(dbie, O, mb):
    """rhs_dielectric_box3d(interface, boxes, epses, eps_out, vs) (src/solver/dielectric_box3d.jl:224-232)
    and the screen_boxes form of the multi-source RHS (src/solver/multi_rhs.jl:168-211)."""
    g, op, ref_if, ref, boxes, epses = mb
    say("multi-box screened volume RHS vs oracle")
    rng = np.random.default_rng(32)
    lo = np.min([np.array(b[0]) - np.array(b[1:]) / 2 for b in boxes], axis=0) - 0.2
    hi = np.max([np.array(b[0]) + np.array(b[1:]) / 2 for b in boxes], axis=0) + 0.2
    vss, cols = [], []
    for k in range(3):
        pos = rng.uniform(lo, hi, (40 + 7 * k, 3))
        w = rng.uniform(0.1, 1.0, len(pos))
        rho = rng.standard_normal(len(pos))
        vss.append(dbie.VolumeSource(pos.T, w, rho))
        cols.append(ref_if.rhs_volume(pos, w, O.screen_boxes(pos, rho, boxes, epses, 1.0, "sharp"), 1.0))
    F = dbie.rhs_dielectric_box3d_fmm3d(g, vss, 1e-9, screen_boxes=boxes, screen_epses=epses, screen_eps_out=1.0)
    for k in range(3):
        assert rel(F[:, k], cols[k]) < 1e-12
    f0 = dbie.rhs_dielectric_box3d(g, boxes, epses, 1.0, vss[0])
    assert rel(f0, cols[0]) < 1e-12
    # soft screening through the same RHS entry point
    Fm = dbie.rhs_volume_boxes(g, vss[:1], boxes, epses, 1.0, dbie.SoftMixPermittivity(0.05))
    col = ref_if.rhs_volume(vss[0].positions.T, vss[0].weights,
                            O.screen_boxes(vss[0].positions.T, vss[0].density, boxes, epses, 1.0, "mix", 0.05), 1.0)
    assert rel(Fm[:, 0], col) < 1e-12


def test_screened_volume_source_forms(dbie, O):
    """screened_volume_source's three forms on the GPU against the oracle and the reference's
    known answers (test/solver/dielectric_box3d.jl:173-234, test/core/sources.jl:184-245)."""
    say("screened_volume_source forms vs oracle")
    eps_in, eps_out, rho0, bw = 4.0, 1.0, 2.0, 0.05
    pts = np.array([(0.0, 0.0, 0.0), (0.5, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.5, 0.5), (1.5, 0.0, 0.0)])
    vs = dbie.VolumeSource(pts.T, np.ones(5), np.full(5, rho0))
    itf = dbie.single_dielectric_box3d(1.0, 1.0, 1.0, 2, 0.2, eps_in, eps_out)
    tol = math.sqrt(np.finfo(float).eps)
    for mode, name in ((dbie.SharpScreening(), "sharp"), (dbie.SoftMixPermittivity(bw), "mix"),
                       (dbie.SoftMixInversePermittivity(bw), "inv")):
        a = dbie.screened_volume_source(itf, vs, mode).density
        b = dbie.screened_volume_source(1.0, 1.0, 1.0, vs, eps_in, eps_out, mode).density
        r = O.screen_one_region(pts, vs.density, (-0.5,) * 3, (0.5,) * 3, eps_in, eps_out, name,
                                getattr(mode, "bandwidth", 0.0), tol)
        if name == "sharp":
            assert np.array_equal(a, r) and np.array_equal(b, r)
            assert np.array_equal(a, [rho0 / eps_in] * 4 + [rho0 / eps_out])
        else:
            assert np.allclose(a, r, rtol=1e-13, atol=0) and np.allclose(b, r, rtol=1e-13, atol=0)
    # several boxes: order independence of the soft modes and the first-box rule of the sharp one
    b1, b2 = ((-0.5, 0.0, 0.0), 1.0, 1.0, 1.0), ((0.5, 0.0, 0.0), 1.0, 1.0, 1.0)
    p2 = np.array([(-0.5, 0.0, 0.0), (0.5, 0.0, 0.0), (0.0, 0.0, 0.0), (5.0, 0.0, 0.0)])
    v2 = dbie.VolumeSource(p2.T, np.ones(4), np.ones(4))
    for mode, name in ((dbie.SoftMixPermittivity(0.5), "mix"), (dbie.SoftMixInversePermittivity(0.5), "inv")):
        d12 = dbie.screened_volume_source([b1, b2], [2.0, 4.0], 1.0, v2, mode).density
        d21 = dbie.screened_volume_source([b2, b1], [4.0, 2.0], 1.0, v2, mode).density
        assert np.allclose(d12, d21, atol=1e-12, rtol=0)
        assert np.allclose(d12, O.screen_boxes(p2, np.ones(4), [b1, b2], [2.0, 4.0], 1.0, name, 0.5), rtol=1e-13)
    d = dbie.screened_volume_source([b1, b2], [2.0, 4.0], 1.0, v2, dbie.SharpScreening()).density
    assert np.array_equal(d, O.screen_boxes(p2, np.ones(4), [b1, b2], [2.0, 4.0], 1.0, "sharp"))
    with pytest.raises(TypeError):
        dbie.screened_volume_source(itf, vs)
    with pytest.raises(ValueError):
        dbie.SoftMixPermittivity(0.0)


def test_lattice_batch_solve_and_result_file(dbie, O, mb, request, tmp_path):
    """solve_dielectric_lattice_batch (src/solver/lattice_batch.jl:285-304): pair densities on a
    shared grid, screened box by box, one block solve; then the batch result file round trip."""
    g, op, ref_if, ref, boxes, epses = mb
    _, _, nq, lec = GEOMS[request.node.callspec.params["mb"]]
    say("lattice batch solve (3 pair densities) vs oracle LU")
    rng = np.random.default_rng(33)
    gidx = np.stack(np.meshgrid(np.arange(6), np.arange(4), np.arange(3), indexing="ij"), -1).reshape(-1, 3)
    lo = np.min([np.array(b[0]) - np.array(b[1:]) / 2 for b in boxes], axis=0)
    pos = (lo[:, None] + 0.05 + gidx.T * 0.17).astype(float)
    dens = rng.standard_normal((len(gidx), 3))
    b = dbie.LatticeBatch(np.array([(1, 2), (1, 3), (2, 3)]), gidx, pos, np.full(len(gidx), 2e-3), dens)
    res = dbie.solve_dielectric_lattice_batch(boxes, epses, 1.0, b, n_quad=nq, l_ec=lec, fmm_tol=1e-9,
                                              gmres_rtol=1e-12)
    assert np.array_equal(res["interface"].nodes, g.nodes)
    Ad = ref_if.dense_DT() + ref.Cs.toarray() + np.diag(ref.t)
    F_ref = np.stack([ref_if.rhs_volume(pos.T, b.weights, O.screen_boxes(pos.T, dens[:, k], boxes, epses, 1.0,
                                                                          "sharp"), 1.0) for k in range(3)], 1)
    S_ref = O.solve_lu(Ad, F_ref)
    for k in range(3):
        assert rel(res["sigma"][:, k], S_ref[:, k]) < 1e-10
    resm = dbie.solve_dielectric_lattice_batch(boxes, epses, 1.0, b, n_quad=nq, l_ec=lec, fmm_tol=1e-9,
                                               gmres_rtol=1e-11, interface=g, precision="fp16x3")
    for k in range(3):
        assert rel(resm["sigma"][:, k], S_ref[:, k]) < 1e-8
    path = str(tmp_path / "batch_0003.npz")
    dbie.save_batch_result(path, dbie.batch_result(3, b, res))
    assert dbie.is_complete_batch(path)
    br = dbie.load_batch_result(path)
    assert np.array_equal(br.sigma, res["sigma"]) and br.stats["K"] == 3 and br.stats["dof"] == g.n_points
