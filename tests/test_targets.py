"""Off-surface targets, Gaussian blobs and the reaction matrix (CPU: oracle pins and host parity).

The oracle's adaptive cubature and near-target corrections are pinned to closed forms (the
potential of a uniformly charged rectangle, summed over the six faces of the box), the
prolongation to exact polynomial reproduction, the Gauss-Hermite blobs to numpy's rule and to
the analytic potential of a Gaussian charge, and the reaction matrix to reciprocity.  The
product's host-side build (refinement, prolongation, corrections) must equal the oracle's bit
for bit."""
import math

import numpy as np
import pytest
from scipy.special import erf, gamma

BOX1 = (1.2, 0.8, 0.6)


def rect_F(X, Y, z):
    R = math.sqrt(X * X + Y * Y + z * z)
    rx, ry = math.sqrt(X * X + z * z), math.sqrt(Y * Y + z * z)
    t1 = X * math.asinh(Y / rx) if rx > 0 else 0.0
    t2 = Y * math.asinh(X / ry) if ry > 0 else 0.0
    t3 = z * math.atan(X * Y / (z * R)) if z != 0 else 0.0
    return t1 + t2 - t3


def rect_pot(x1, x2, y1, y2, x0, y0, z0):
    """Potential of unit surface density on [x1,x2] x [y1,y2] (plane z=0) at (x0,y0,z0), 1/(4 pi r)."""
    v = (rect_F(x2 - x0, y2 - y0, z0) - rect_F(x1 - x0, y2 - y0, z0) - rect_F(x2 - x0, y1 - y0, z0)
         + rect_F(x1 - x0, y1 - y0, z0))
    return v / (4 * math.pi)


def box_pot(L, t):
    Lx, Ly, Lz = L
    x, y, z = t
    s = 0.0
    for sgn in (1, -1):
        s += rect_pot(-Lx / 2, Lx / 2, -Ly / 2, Ly / 2, x, y, z - sgn * Lz / 2)
        s += rect_pot(-Ly / 2, Ly / 2, -Lz / 2, Lz / 2, y, z, x - sgn * Lx / 2)
        s += rect_pot(-Lx / 2, Lx / 2, -Lz / 2, Lz / 2, x, z, y - sgn * Ly / 2)
    return s


def blob_cloud(dbie, centers, sigma, n=6):
    srcs = [dbie.gaussian_blob_gh(c, sigma, n) for c in centers]
    pos = np.concatenate([s.positions.T for s in srcs])
    q = np.concatenate([s.weights * s.density for s in srcs])
    grp = np.concatenate([[0], np.cumsum([len(s.weights) for s in srcs])]).astype(np.int64)
    return srcs, pos, q, grp


# ---------------------------------------------------------------- Gauss-Hermite and blobs
@pytest.mark.parametrize("n", [1, 2, 3, 6, 11, 20])
def test_gauss_hermite_rule(dbie, n):
    x, w = dbie.gausshermite(n)
    xr, wr = np.polynomial.hermite.hermgauss(n)
    assert np.all(np.diff(x) > 0)
    assert np.max(np.abs(x - xr)) <= 1e-13 * max(1.0, np.abs(xr).max())
    assert np.max(np.abs(w - wr) / wr.max()) <= 1e-12
    for k in range(n):  # exact for degree <= 2n-1: sum w x^(2k) = Gamma(k + 1/2)
        assert abs(np.sum(w * x ** (2 * k)) - gamma(k + 0.5)) <= 1e-12 * gamma(k + 0.5)


def test_gauss_hermite_bad_args(dbie):
    with pytest.raises(ValueError):
        dbie.gausshermite(0)


def test_gaussian_blob_charge_and_potential(dbie):
    c, sigma = (0.1, -0.2, 0.05), 0.02
    vs = dbie.gaussian_blob_gh(c, sigma, 6)
    assert vs.positions.shape == (3, 216)
    assert abs(np.sum(vs.weights * vs.density) - 1.0) < 1e-14
    # 6-point Gauss-Hermite error falls off like (sigma / d)^12
    for d, tol in ((0.15, 5e-8), (0.3, 1e-10), (1.0, 1e-13)):
        t = np.array(c) + d * np.array([0.6, 0.0, 0.8])
        r = np.linalg.norm(vs.positions.T - t, axis=1)
        u = np.sum(vs.weights * vs.density / (4 * np.pi * r))
        exact = erf(d / (math.sqrt(2) * sigma)) / (4 * np.pi * d)
        assert abs(u - exact) < tol * exact, d
    # the reference's midpoint-grid GaussianVolumeSource (src/core/sources.jl:517-548)
    g = dbie.GaussianVolumeSource(c, sigma, 12, 1e-8)
    assert g.positions.shape == (3, 1728)
    assert abs(np.sum(g.weights * g.density) - 1.0) < 1e-6
    assert np.allclose(g.positions[:, 0], np.array(c) - math.sqrt(2 * sigma ** 2 * math.log(1e8)) * (1 - 1 / 12))
    # ix outer, iz inner
    assert g.positions[2, 1] > g.positions[2, 0] and g.positions[0, 1] == g.positions[0, 0]


def test_blob_lattice(dbie):
    C = dbie.blob_lattice_centers(1.2, 0.8, 0.6)
    assert C.shape == (1024, 3)
    assert np.allclose(C[0], [-0.6 + 0.0375, -0.4 + 0.05, -0.3 + 0.0375])
    assert np.allclose(C[1], [-0.6 + 0.0375, -0.4 + 0.05, -0.3 + 0.0375 + 0.075])
    assert np.all(np.abs(C) < np.array([0.6, 0.4, 0.3]))


# ---------------------------------------------------------------- oracle pins
@pytest.mark.parametrize("z0", [1e-1, 1e-2, 1e-3, -4e-3])
def test_oracle_cubature_rectangle_closed_form(O, z0):
    c = np.array([[-0.1, -0.1, 0], [0.1, -0.1, 0], [0.1, 0.1, 0], [-0.1, 0.1, 0]])
    for nq in (4, 5):
        f = O.Interface.from_panels(c[None], np.array([[0, 0, 1.0]]), nq, 4.0, 1.0)
        for t in ([0.03, -0.02, z0], [0.099, 0.0, z0], [0.2, 0.15, z0]):
            I, ok = O.pot_panel_hcub(f, 0, np.array(t), 1e-11)
            assert ok
            exact = rect_pot(-0.1, 0.1, -0.1, 0.1, *t)
            assert abs(I.sum() - exact) < 1e-10 * abs(exact) + 1e-11, (nq, t)


def test_oracle_cubature_basis_moments(O):
    """sum_j I_j x_j reproduces the integral of x G (linear functions are in the basis):
    compared with a 64 x 64 tensor Gauss-Legendre rule on 16 x 16 sub-squares (far target)."""
    c = np.array([[-0.1, -0.1, 0], [0.1, -0.1, 0], [0.1, 0.1, 0], [-0.1, 0.1, 0]])
    f = O.Interface.from_panels(c[None], np.array([[0, 0, 1.0]]), 4, 4.0, 1.0)
    t = np.array([0.05, 0.02, 0.04])
    I, ok = O.pot_panel_hcub(f, 0, t, 1e-12)
    nodes = f.nodes()
    xg, wg = np.polynomial.legendre.leggauss(64)
    edges = np.linspace(-0.1, 0.1, 17)
    ref = 0.0
    for a, b in zip(edges[:-1], edges[1:]):
        xs = (a + b) / 2 + (b - a) / 2 * xg
        for cc, dd in zip(edges[:-1], edges[1:]):
            ys = (cc + dd) / 2 + (dd - cc) / 2 * xg
            X, Y = np.meshgrid(xs, ys, indexing="ij")
            W = np.outer(wg, wg) * (b - a) / 2 * (dd - cc) / 2
            r = np.sqrt((X - t[0]) ** 2 + (Y - t[1]) ** 2 + t[2] ** 2)
            ref += np.sum(W * X / (4 * np.pi * r))
    assert abs(np.sum(I * nodes[0]) - ref) < 1e-11


def test_oracle_target_map_box_closed_form(O):
    """sigma = 1 on the whole box: the corrected potential equals the sum of six rectangle
    potentials for targets inside, outside and close to faces."""
    f = O.Interface.box(*BOX1, 4, 0.2, 4.0, 1.0)
    t = np.array([[0.1, 0.05, 0.28], [0.1, 0.05, 0.33], [0.0, 0.39, 0.0], [0.605, 0.1, -0.05], [0.0, 0.0, 0.0],
                  [-0.2, -0.37, 0.1]])
    tm = O.TargetMap(f, t, 1e-10)
    assert tm.n_pairs > 0 and tm.n_capped == 0
    u = tm.apply(np.ones(f.N))[:, 0]
    ex = np.array([box_pot(BOX1, ti) for ti in t])
    assert np.all(np.abs(u[:5] - ex[:5]) < 1e-7 * np.abs(ex[:5]))
    # the last target is near panels flagged is_edge, which build_target_neighbor_list skips
    # (include_edges_src = false, src/kernel/laplace3d_near_hcubature.jl:355): less accurate
    assert abs(u[5] - ex[5]) < 1e-5 * abs(ex[5])
    # without the refinement and corrections the near targets are far less accurate
    from oracle import lib  # direct sum over the coarse nodes for comparison
    U0 = np.zeros(len(t))
    lib().orc_pottrg(f.h, len(t), np.ascontiguousarray(t.ravel()), 1, np.ones(f.N), U0)
    assert np.abs(U0[:2] - ex[:2]).max() > 100 * np.abs(u[:2] - ex[:2]).max()


@pytest.mark.parametrize("nq", [3, 4, 5])
def test_oracle_prolongation_polynomials(O, nq):
    f = O.Interface.box(*BOX1, nq, 0.2, 4.0, 1.0)
    t = np.array([[0.1, 0.05, 0.28], [0.59, -0.3, 0.2], [-0.4, 0.39, -0.29]])
    tm = O.TargetMap(f, t, 1e-8)
    xyzw, par, spl, (pp, pc, pv), _ = tm.parts()
    assert spl.any()
    nd = f.nodes()

    def poly(x, y, z):  # total degree nq - 1
        return 1 + 0.3 * x - 0.7 * y + 0.2 * z + (nq > 2) * (x * y - 0.5 * z * z) + (nq > 3) * (x * y * z)

    s = poly(*nd[0:3])
    Ps = np.array([np.dot(pv[pp[r]:pp[r + 1]], s[pc[pp[r]:pp[r + 1]]]) for r in range(tm.N_refined)])
    assert np.max(np.abs(Ps - poly(*xyzw[0:3]))) < 1e-12


def test_oracle_refinement_invariants(O):
    f = O.Interface.box(*BOX1, 4, 0.2, 4.0, 1.0)
    t = np.array([[0.1, 0.05, 0.28], [0.59, -0.3, 0.2], [-0.4, 0.39, -0.29], [2.0, 2.0, 2.0]])
    rf = 5.0
    tm = O.TargetMap(f, t, 1e-8, rf)
    xyzw, par, spl, _, _ = tm.parts()
    assert abs(xyzw[3].sum() - f.nodes()[6].sum()) < 1e-12 * f.nodes()[6].sum()  # area preserved
    assert abs(f.nodes()[6].sum() - 2 * (1.2 * 0.8 + 1.2 * 0.6 + 0.8 * 0.6)) < 1e-12
    corners, _, _ = f.panels()
    limit = min(max(np.linalg.norm(c[0] - c[1]), np.linalg.norm(c[1] - c[2])) for c in corners)
    npp = 16
    for q in range(tm.P_refined):
        pts = xyzw[0:3, q * npp:(q + 1) * npp].T
        pc = corners[par[q]]
        lo, hi = pc.min(0) - 1e-12, pc.max(0) + 1e-12
        assert np.all((pts >= lo) & (pts <= hi))
    assert tm.N_refined > f.N


# ---------------------------------------------------------------- product host == oracle
def _host_cases(dbie):
    _, pos, _, _ = blob_cloud(dbie, [(0.2, 0.1, 0.2), (-0.45, -0.3, -0.2), (0.5, 0.3, 0.25)], 0.02)
    return [
        ((1.2, 0.8, 0.6, 4, 0.2), pos, 1e-9, 5.0),
        ((1.2, 0.8, 0.6, 5, 0.2), np.array([[0.61, 0.0, 0.0], [0.0, 0.0, 0.3001], [0.59, 0.39, 0.29]]), 1e-10, 5.0),
        ((1.0, 1.0, 1.0, 3, 0.3), np.array([[0.1, 0.2, 0.45], [3.0, 0.0, 0.0]]), 1e-8, 2.0),
        ((1.0, 1.0, 1.0, 3, 0.3), np.zeros((0, 3)), 1e-8, 5.0),
    ]


def test_host_target_map_equals_oracle(dbie, O):
    for (Lx, Ly, Lz, nq, l_ec), pos, atol, rf in _host_cases(dbie):
        itf = dbie.single_dielectric_box3d(Lx, Ly, Lz, nq, l_ec, 4.0, 1.0)
        f = O.Interface.box(Lx, Ly, Lz, nq, l_ec, 4.0, 1.0)
        tm = dbie.TargetPotentialMap(itf, pos, atol, rf, n_threads=3, device_cubature=False)
        to = O.TargetMap(f, pos, atol, rf)
        assert (tm.n_refined, tm.n_pairs, tm.nnz, tm.n_capped) == (to.N_refined, to.n_pairs, to.nnz, to.n_capped)
        xyzw, par, spl = tm.refined()
        oxyzw, opar, ospl, (pp, pc, pv), (cp, cc, cv) = to.parts()
        assert np.array_equal(xyzw, oxyzw) and np.array_equal(par, opar) and np.array_equal(spl, ospl.astype(bool))
        (Pp, Pc, Pv), (Cp, Cc, Cv) = tm.matrices()
        assert np.array_equal(Pp, pp) and np.array_equal(Pc, pc) and np.array_equal(Pv, pv)
        assert np.array_equal(Cp, cp) and np.array_equal(Cc, cc)
        assert np.max(np.abs(Cv - cv), initial=0.0) <= 1e-15 * max(np.abs(cv).max(initial=0.0), 1.0)


def test_target_map_bad_args(dbie):
    itf = dbie.single_dielectric_box3d(1.0, 1.0, 1.0, 3, 0.3, 4.0, 1.0)
    with pytest.raises(ValueError):
        dbie.TargetPotentialMap(itf, np.zeros((2, 3)), 0.0, device_cubature=False)
    with pytest.raises(ValueError):
        dbie.TargetPotentialMap(itf, np.zeros((2, 3)), 1e-8, range_factor=-1.0, device_cubature=False)


def test_oracle_reaction_reciprocity(dbie, O):
    """For blobs inside the dielectric the reaction matrix is symmetric up to discretisation
    error (reciprocity of the reaction Green's function)."""
    f = O.Interface.box(*BOX1, 4, 0.2, 4.0, 1.0)
    srcs, pos, q, grp = blob_cloud(dbie, [[0.2, 0.1, 0.05], [-0.3, -0.1, 0.0], [0.0, 0.15, -0.1],
                                          [0.45, -0.25, 0.15]], 0.03)
    F = np.stack([f.rhs_volume(s.positions.T, s.weights, f.screen_sharp(s.positions.T, s.density), 1.0)
                  for s in srcs], 1)
    S = O.solve_lu(f.lhs_dense(), F)
    tm = O.TargetMap(f, pos, 1e-10)
    R = tm.reaction(grp, q, S)
    assert np.abs(R - R.T).max() < 2e-2 * np.abs(R).max()
    # a charge inside the higher-permittivity box sees a reaction potential of its own sign
    assert np.all(np.diag(R) > 0)
