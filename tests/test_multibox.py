"""Several touching boxes (src/shape/box3d_multi.jl) and box-by-box screening (src/core/sources.jl:408-515).

The oracle's restatement is pinned by the reference's own known answers
(test/shape/multi_box3d.jl, test/solver/dielectric_box3d.jl:173-234, test/core/sources.jl:184-245);
the product mesher must then reproduce the oracle's panels bit for bit (CPU only: meshing is host work)."""
import math

import numpy as np
import pytest

BOX2 = [((0.0, 0.0, 0.0), 1.0, 1.0, 1.0), ((1.0, 0.0, 0.0), 1.0, 1.0, 1.0)]
BOX3L = [((0.0, 0.0, 0.0), 1.0, 1.0, 1.0), ((1.0, 0.0, 0.0), 1.0, 1.0, 1.0), ((0.0, 1.0, 0.0), 1.0, 1.0, 1.0)]
BOXBIG = [((0.0, 0.0, 0.0), 2.0, 2.0, 2.0), ((1.5, 0.0, 0.0), 1.0, 1.0, 1.0)]
# an off-centre stack with unequal sides and a partly shared face
BOXSTACK = [((0.1, -0.2, 0.3), 1.2, 0.8, 0.6), ((0.1, -0.2, 0.9), 0.6, 0.4, 0.6), ((1.0, -0.2, 0.3), 0.6, 0.8, 0.6)]
# a 2 x 2 x 1 lattice (shared faces meet at a common edge)
LATTICE = [((x, y, 0.0), 0.5, 0.5, 0.4) for x in (-0.25, 0.25) for y in (-0.25, 0.25)]


def area(r):
    a, b, _, d = r[:4]
    return math.dist(b, a) * math.dist(a, d)


# ---- oracle helpers against the reference's known answers -------------------------------------
def test_faces_at_center(O):
    """test/shape/multi_box3d.jl:7-40"""
    faces = O.box_faces_at_center((1.0, 2.0, 3.0), 2.0, 4.0, 6.0)
    assert len(faces) == 6
    for a, b, c, d, n in faces:
        fc = np.mean([a, b, c, d], axis=0)
        diff = fc - np.array([1.0, 2.0, 3.0])
        assert abs(diff @ np.array(n)) > 0
        assert np.linalg.norm(diff - np.array(n) * (diff @ np.array(n))) < 1e-14
        assert np.linalg.norm(n) == pytest.approx(1.0)
    areas = sorted(area(f) for f in faces)
    assert areas[0] == pytest.approx(8.0) and areas[1] == pytest.approx(8.0)
    assert areas[2] == pytest.approx(12.0) and areas[3] == pytest.approx(12.0)
    assert areas[4] == pytest.approx(24.0) and areas[5] == pytest.approx(24.0)


def test_rect_overlap(O):
    """test/shape/multi_box3d.jl:43-104"""
    f1 = ((0.5, -0.5, -0.5), (0.5, 0.5, -0.5), (0.5, 0.5, 0.5), (0.5, -0.5, 0.5))
    ok, reg = O.rect_overlap_3d(*f1, (1.0, 0.0, 0.0), *f1, (-1.0, 0.0, 0.0))
    assert ok and area(reg) == pytest.approx(1.0)
    f3 = ((0.0, -0.5, 0.5), (0.0, 0.5, 0.5), (1.0, 0.5, 0.5), (1.0, -0.5, 0.5))
    ok, _ = O.rect_overlap_3d(*f1, (1.0, 0.0, 0.0), *f3, (0.0, 0.0, 1.0))
    assert not ok
    fa = ((0.5, 0.5, -0.5), (-0.5, 0.5, -0.5), (-0.5, 0.5, 0.5), (0.5, 0.5, 0.5))
    fb = ((-0.5, 0.5, -0.5), (0.5, 0.5, -0.5), (0.5, 0.5, 0.5), (-0.5, 0.5, 0.5))
    ok, reg = O.rect_overlap_3d(*fa, (0.0, 1.0, 0.0), *fb, (0.0, -1.0, 0.0))
    assert ok and area(reg) == pytest.approx(1.0)
    # the overlap keeps the first rectangle's winding
    assert reg[0] == fa[0] and reg[1] == fa[1]


def test_detect_shared_faces(O):
    """test/shape/multi_box3d.jl:106-137"""
    sh = O.detect_shared_faces_3d(BOX2)
    assert len(sh) == 1
    reg, i, j, n = sh[0]
    assert i < j and area(reg) == pytest.approx(1.0)
    assert n == (-1.0, 0.0, 0.0)  # from box j towards box i
    line = BOX2 + [((2.0, 0.0, 0.0), 1.0, 1.0, 1.0)]
    assert len(O.detect_shared_faces_3d(line)) == 2
    assert len(O.detect_shared_faces_3d([BOX2[0], ((5.0, 0.0, 0.0), 1.0, 1.0, 1.0)])) == 0


def test_subtract_rects(O):
    """test/shape/multi_box3d.jl:139-222"""
    f = ((0.5, -0.5, -0.5), (0.5, 0.5, -0.5), (0.5, 0.5, 0.5), (0.5, -0.5, 0.5))
    n = (1.0, 0.0, 0.0)
    rem = O.subtract_rects_from_face_3d(*f, n, [])
    assert len(rem) == 1 and area(rem[0]) == pytest.approx(1.0)
    assert rem[0][4] == (True,) * 4 and rem[0][5] == (True,) * 4
    assert O.subtract_rects_from_face_3d(*f, n, [f]) == []
    half = [((0.5, -0.5, -0.5), (0.5, 0.5, -0.5), (0.5, 0.5, 0.0), (0.5, -0.5, 0.0))]
    rem = O.subtract_rects_from_face_3d(*f, n, half)
    assert sum(area(r) for r in rem) == pytest.approx(0.5)
    assert rem[0][4][0] and rem[0][4][2]
    # a hole in the middle of a 2 x 2 face
    f2 = ((1.0, -1.0, -1.0), (1.0, 1.0, -1.0), (1.0, 1.0, 1.0), (1.0, -1.0, 1.0))
    hole = [((1.0, -0.5, -0.5), (1.0, 0.5, -0.5), (1.0, 0.5, 0.5), (1.0, -0.5, 0.5))]
    rem = O.subtract_rects_from_face_3d(*f2, n, hole)
    assert len(rem) == 8 and sum(area(r) for r in rem) == pytest.approx(3.0)
    assert all(sum(r[4]) == 2 for r in rem)
    assert all(sum(r[5]) == 4 for r in rem)
    # two edge-attached shared regions leave a smooth interior grid point
    f3 = ((1.0, -2.0, -2.0), (1.0, 2.0, -2.0), (1.0, 2.0, 2.0), (1.0, -2.0, 2.0))
    cut = [((1.0, -2.0, -1.0), (1.0, -1.0, -1.0), (1.0, -1.0, 1.0), (1.0, -2.0, 1.0)),
           ((1.0, -1.0, -2.0), (1.0, 1.0, -2.0), (1.0, 1.0, -1.0), (1.0, -1.0, -1.0))]
    rem = O.subtract_rects_from_face_3d(*f3, n, cut)
    centre = [r for r in rem if np.allclose(np.mean(r[:4], axis=0)[1:], 0.0, atol=1e-12)]
    assert len(centre) == 1
    assert centre[0][4] == (True, False, False, True)
    assert centre[0][5] == (True, True, False, True)


def test_oracle_multi_box(O):
    """test/shape/multi_box3d.jl:224-332: single-box equivalence, eps pairs, total areas."""
    single = O.Interface.box(1.0, 1.0, 1.0, 4, 0.3, 2.0, 1.0)
    multi = O.Interface.multi_box(4, 0.3, [((0.0, 0.0, 0.0), 1.0, 1.0, 1.0)], [2.0], 1.0)
    assert multi.P == single.P and multi.N == single.N
    assert np.array_equal(multi.nodes(), single.nodes())
    for boxes, epses, total, pairs in ((BOX2, [2.0, 4.0], 11.0, {(2.0, 1.0), (4.0, 1.0), (4.0, 2.0)}),
                                       (BOX3L, [2.0, 3.0, 4.0], 16.0, None),
                                       (BOXBIG, [2.0, 4.0], 29.0, {(2.0, 1.0), (4.0, 1.0), (4.0, 2.0)})):
        f = O.Interface.multi_box(4, 0.3, boxes, epses, 1.0)
        assert f.nodes()[6].sum() == pytest.approx(total, abs=0.01)
        ei, eo = panel_eps(O, boxes, epses, 1.0, 0.3)
        assert len(ei) == f.P
        if pairs is not None:
            assert pairs <= set(zip(ei, eo))


def panel_eps(O, boxes, epses, eps_out, l_ec):
    """per-panel (eps_in, eps_out) from the oracle's region list (its panel order)"""
    ei, eo = [], []
    for (ra, rb, rc, rd, fn, e_in, e_out, ie, ic) in O.multi_box3d_face_regions(boxes, epses, eps_out):
        c, _ = O.rect_adaptive_panels(ra, rb, rc, rd, fn, ie, ic, math.sqrt(2.0), l_ec)
        ei += [e_in] * len(c)
        eo += [e_out] * len(c)
    return ei, eo


def test_oracle_multi_box_lu_solve(O):
    """test/shape/multi_box3d.jl:334-358: dense LHS and LU solve for a point charge outside."""
    f = O.Interface.multi_box(4, 0.3, BOX2, [2.0, 4.0], 1.0)
    A = f.lhs_dense()
    b = f.rhs_points([(3.0, 0.0, 0.0)], [1.0], 1.0)[:, 0]
    x = O.solve_lu(A, b)
    assert x.shape == (f.N,) and np.all(np.isfinite(x))
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)


# ---- product mesher against the oracle ---------------------------------------------------------
@pytest.mark.parametrize("case", ["box2", "l_shape", "big_small", "stack", "lattice"])
def test_multi_box_mesh_matches_oracle(dbie, O, case):
    boxes, epses, eps_out, nq, lec = {
        "box2": (BOX2, [2.0, 4.0], 1.0, 4, 0.3),
        "l_shape": (BOX3L, [2.0, 3.0, 4.0], 1.0, 4, 0.3),
        "big_small": (BOXBIG, [2.0, 4.0], 1.0, 4, 0.3),
        "stack": (BOXSTACK, [4.0, 10.0, 2.5], 1.5, 3, 0.15),
        "lattice": (LATTICE, [2.0, 3.0, 5.0, 7.0], 1.0, 4, 0.1),
    }[case]
    g = dbie.multi_dielectric_box3d(nq, lec, [dict(center=b[0], Lx=b[1], Ly=b[2], Lz=b[3]) for b in boxes], epses,
                                    eps_out)
    f = O.Interface.multi_box(nq, lec, boxes, epses, eps_out)
    c1, n1, e1 = g.panels()
    c2, n2, e2 = f.panels()
    assert g.n_panels == f.P and g.n_points == f.N
    assert np.array_equal(c1, c2) and np.array_equal(n1, n2) and np.array_equal(e1.astype(np.int32), e2)
    assert np.array_equal(g.nodes, f.nodes())
    ei, eo = panel_eps(O, boxes, epses, eps_out, lec)
    assert np.array_equal(g.eps_in, ei) and np.array_equal(g.eps_out, eo)


def test_multi_box_known_answers_and_errors(dbie):
    g = dbie.multi_dielectric_box3d(4, 0.3, BOXBIG, [2.0, 4.0], 1.0)
    assert g.weights.sum() == pytest.approx(29.0, abs=0.1)
    assert {(2.0, 1.0), (4.0, 1.0), (4.0, 2.0)} <= set(zip(g.eps_in.tolist(), g.eps_out.tolist()))
    s = dbie.single_dielectric_box3d(1.0, 1.0, 1.0, 4, 0.3, 2.0, 1.0)
    m = dbie.multi_dielectric_box3d(4, 0.3, [((0.0, 0.0, 0.0), 1.0, 1.0, 1.0)], [2.0])
    assert np.array_equal(s.nodes, m.nodes)
    with pytest.raises(ValueError):
        dbie.multi_dielectric_box3d(4, 0.3, BOX2, [2.0])
    with pytest.raises(ValueError):
        dbie.multi_dielectric_box3d(4, 0.3, [((0.0, 0.0, 0.0), -1.0, 1.0, 1.0)], [2.0])
    with pytest.raises(ValueError):
        dbie.multi_dielectric_box3d(4, 0.0, BOX2, [2.0, 4.0])


# ---- oracle screening against the reference's known answers ------------------------------------
def test_oracle_screening_one_region(O):
    """test/solver/dielectric_box3d.jl:187-229"""
    eps_in, eps_out, rho0, bw = 4.0, 1.0, 2.0, 0.05
    pts = np.array([(0.0, 0.0, 0.0), (0.5, 0.0, 0.0), (0.5, 0.5, 0.0), (0.5, 0.5, 0.5), (1.5, 0.0, 0.0)])
    rho = np.full(5, rho0)
    lo, hi = (-0.5,) * 3, (0.5,) * 3
    tol = math.sqrt(np.finfo(float).eps)
    sharp = O.screen_one_region(pts, rho, lo, hi, eps_in, eps_out, "sharp", tol=tol)
    assert np.allclose(sharp, [rho0 / eps_in] * 4 + [rho0 / eps_out], rtol=1e-15)

    def H(t):
        return (1 + math.erf(t / bw)) / 2

    def scr(p):
        return H(p[0] + 0.5) * H(0.5 - p[0]) * H(p[1] + 0.5) * H(0.5 - p[1]) * H(p[2] + 0.5) * H(0.5 - p[2])

    mix = O.screen_one_region(pts, rho, lo, hi, eps_in, eps_out, "mix", bw)
    assert np.allclose(mix, [rho0 / (eps_in * scr(p) + eps_out * (1 - scr(p))) for p in pts], atol=1e-12, rtol=0)
    assert abs(mix[0] - rho0 / eps_in) < 1e-12 and abs(mix[-1] - rho0 / eps_out) < 1e-12
    inv = O.screen_one_region(pts, rho, lo, hi, eps_in, eps_out, "inv", bw)
    assert np.allclose(inv, [rho0 * (scr(p) / eps_in + (1 - scr(p)) / eps_out) for p in pts], atol=1e-12, rtol=0)
    # the several-box form with one box reduces to the one-region form
    one = [((0.0, 0.0, 0.0), 1.0, 1.0, 1.0)]
    assert np.array_equal(O.screen_boxes(pts, rho, one, [eps_in], eps_out, "sharp"), sharp)
    assert np.allclose(O.screen_boxes(pts, rho, one, [eps_in], eps_out, "mix", bw), mix, rtol=1e-15)
    assert np.allclose(O.screen_boxes(pts, rho, one, [eps_in], eps_out, "inv", bw), inv, rtol=1e-14)


def test_oracle_screening_boxes(O):
    """test/core/sources.jl:184-245: order independence and the partition-of-unity closed form."""
    b1, b2 = ((-0.5, 0.0, 0.0), 1.0, 1.0, 1.0), ((0.5, 0.0, 0.0), 1.0, 1.0, 1.0)
    e1, e2, e0, bw = 2.0, 4.0, 1.0, 0.5
    pts = np.array([(-0.5, 0.0, 0.0), (0.5, 0.0, 0.0), (0.0, 0.0, 0.0), (5.0, 0.0, 0.0)])
    rho = np.ones(4)

    def P(p, b):
        lo = [b[0][k] - b[1 + k] / 2 for k in range(3)]
        hi = [b[0][k] + b[1 + k] / 2 for k in range(3)]
        sm = lambda d: (1 + math.erf(d / bw)) / 2  # noqa: E731
        return math.prod(sm(p[k] - lo[k]) * sm(hi[k] - p[k]) for k in range(3))

    for mode, f in (("mix", lambda p1, p2: 1 / (e1 * p1 + e2 * p2 + e0 * (1 - p1 - p2))),
                    ("inv", lambda p1, p2: p1 / e1 + p2 / e2 + (1 - p1 - p2) / e0)):
        a = O.screen_boxes(pts, rho, [b1, b2], [e1, e2], e0, mode, bw)
        b = O.screen_boxes(pts, rho, [b2, b1], [e2, e1], e0, mode, bw)
        assert np.allclose(a, b, atol=1e-12, rtol=0)
        assert np.allclose(a, [f(P(p, b1), P(p, b2)) for p in pts], atol=1e-12, rtol=0)
    # sharp: the first box holding the point wins (the shared face belongs to box 1 here)
    s = O.screen_boxes(pts, rho, [b1, b2], [e1, e2], e0, "sharp")
    assert np.array_equal(s, [1 / e1, 1 / e2, 1 / e1, 1 / e0])
