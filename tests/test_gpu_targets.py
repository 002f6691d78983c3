"""GPU parity of the off-surface target map and the reaction matrix (run with -m gpu).

Bars: 1e-12 relative for the corrected potential and the reaction matrix against the oracle on
the same densities; 1e-8 (relative to max |R|) for the reaction matrix from mixed-precision
densities against the one from FP64 densities."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BOX1 = (1.2, 0.8, 0.6)


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def blob_cloud(dbie, centers, sigma, n=6):
    srcs = [dbie.gaussian_blob_gh(c, sigma, n) for c in centers]
    pos = np.concatenate([s.positions.T for s in srcs])
    q = np.concatenate([s.weights * s.density for s in srcs])
    grp = np.concatenate([[0], np.cumsum([len(s.weights) for s in srcs])]).astype(np.int64)
    return srcs, pos, q, grp


@pytest.fixture(scope="module")
def small(dbie, O):
    say("target map on the config-1 box: 4 blobs (one crossing a face), 864 targets")
    centers = [(0.2, 0.1, 0.05), (-0.3, -0.1, 0.0), (0.0, 0.15, -0.1), (0.55, -0.25, 0.27)]
    srcs, pos, q, grp = blob_cloud(dbie, centers, 0.02)
    itf = dbie.single_dielectric_box3d(*BOX1, 4, 0.2, 4.0, 1.0)
    tm = dbie.TargetPotentialMap(itf, pos, 1e-10)
    f = O.Interface.box(*BOX1, 4, 0.2, 4.0, 1.0)
    to = O.TargetMap(f, pos, 1e-10)
    return itf, tm, f, to, srcs, pos, q, grp


def test_target_map_apply_parity(small):
    itf, tm, f, to, srcs, pos, q, grp = small
    say(f"tmap apply K=5 (N={itf.n_points}, refined {tm.n_refined}, pairs {tm.n_pairs})")
    S = np.random.default_rng(3).standard_normal((itf.n_points, 5))
    U = tm.apply(S)
    Ur = to.apply(S)
    for k in range(5):
        assert np.linalg.norm(U[:, k] - Ur[:, k]) <= 1e-12 * np.linalg.norm(Ur[:, k]), k
    u1 = tm.apply(S[:, 2])
    assert u1.shape == (tm.M,) and np.allclose(u1, U[:, 2], rtol=0, atol=1e-15 * np.abs(U).max())


def test_reaction_matrix_parity(dbie, small):
    itf, tm, f, to, srcs, pos, q, grp = small
    say("reaction matrix (host and device densities) vs oracle")
    S = np.random.default_rng(4).standard_normal((itf.n_points, 6))
    R = tm.reaction(grp, q, S)
    Rr = to.reaction(grp, q, S)
    assert R.shape == (4, 6)
    assert np.abs(R - Rr).max() <= 1e-12 * np.abs(Rr).max()
    dS = dbie.DeviceMatrix.from_host(S)
    Rd = tm.reaction(grp, q, dS)
    assert np.abs(Rd - R).max() <= 1e-15 * np.abs(R).max()
    with pytest.raises(ValueError):
        tm.reaction(grp[:-1], q, S)  # groups must cover every target


def test_reaction_matrix_mixed_vs_fp64(dbie, O):
    """Scaled-down volume-source reaction workload: 16 Gauss-Hermite blobs on a 4 x 2 x 2 lattice (plus two near faces) in
    the n_quad = 5 box, screened volume RHS, FP64 and mixed-precision block solves, R from each."""
    Lx, Ly, Lz = BOX1
    say("mini reaction workload: box 1.2x0.8x0.6, n_quad=5, l_ec=0.1, 18 blobs")
    itf = dbie.single_dielectric_box3d(Lx, Ly, Lz, 5, 0.1, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(itf, 1e-9, 1e-9, 8)
    # the lattice plus two blobs close to faces (one crossing the interface) so that near-target
    # corrections and screening take part
    C = np.vstack([dbie.blob_lattice_centers(Lx, Ly, Lz, 4, 2, 2), [[0.0, 0.0, 0.27], [0.25, 0.365, 0.0]]])
    srcs, pos, q, grp = blob_cloud(dbie, C, 0.02)
    F = dbie.rhs_dielectric_box3d_fmm3d(itf, srcs, 1e-9)
    assert F.shape == (itf.n_points, 18)
    S64, st64 = op.solve(F, rtol=1e-12)
    Sm, stm = op.solve(F, rtol=1e-11, precision="fp16x3")
    assert st64["solved"] and stm["solved"]
    R64, tm = dbie.reaction_matrix(itf, srcs, S64, lhs_tol=1e-10)
    Rm, _ = dbie.reaction_matrix(itf, srcs, Sm, tmap=tm)
    say(f"  N={itf.n_points} refined={tm.n_refined} pairs={tm.n_pairs} "
        f"max|Rm-R64|/max|R64| = {np.abs(Rm - R64).max() / np.abs(R64).max():.2e}")
    assert np.abs(Rm - R64).max() <= 1e-8 * np.abs(R64).max()
    # the FP64 reaction matrix against the oracle on the same densities
    f = O.Interface.box(Lx, Ly, Lz, 5, 0.1, 4.0, 1.0)
    to = O.TargetMap(f, pos, 1e-10)
    Rr = to.reaction(grp, q, S64)
    assert to.n_pairs > 0
    assert np.abs(R64 - Rr).max() <= 1e-12 * np.abs(Rr).max()
    # blobs wholly inside the box: reciprocity holds to discretisation accuracy
    inside = np.array([np.all(np.abs(s.positions.T) < np.array([Lx, Ly, Lz]) / 2) for s in srcs])
    Ri = R64[np.ix_(inside, inside)]
    assert inside.sum() >= 4 and np.abs(Ri - Ri.T).max() < 1e-2 * np.abs(Ri).max()


@pytest.mark.parametrize("nq,l_ec,atol", [(4, 0.2, 1e-10), (5, 0.2, 1e-9)])
def test_device_cubatures_match_host(dbie, nq, l_ec, atol):
    """Near-target cubatures integrated on the device give the host-built map bit for bit."""
    say(f"target map: device vs host cubatures (n_quad={nq}, atol={atol})")
    centers = [(0.2, 0.1, 0.05), (0.55, -0.25, 0.27), (-0.58, 0.0, 0.0)]
    srcs, pos, q, grp = blob_cloud(dbie, centers, 0.02)
    itf = dbie.single_dielectric_box3d(*BOX1, nq, l_ec, 4.0, 1.0)
    td = dbie.TargetPotentialMap(itf, pos, atol, device_cubature=True)
    th = dbie.TargetPotentialMap(itf, pos, atol, device_cubature=False)
    say(f"  pairs {td.n_pairs}, capped {td.n_capped}, nnz {td.nnz}")
    assert (td.n_refined, td.n_pairs, td.nnz, td.n_capped) == (th.n_refined, th.n_pairs, th.nnz, th.n_capped)
    (Pp, Pc, Pv), (Cp, Cc, Cv) = td.matrices()
    (pp, pc, pv), (cp, cc, cv) = th.matrices()
    assert np.array_equal(Cp, cp) and np.array_equal(Cc, cc)
    assert np.array_equal(Cv, cv), np.abs(Cv - cv).max()


@pytest.mark.parametrize("nq,l_ec", [(4, 0.2), (5, 0.2), (4, 0.1)])
def test_device_refinement_matches_oracle(dbie, O, nq, l_ec):
    """The refinement around the targets and the prolongation built on the device
    (_refine_interface_for_targets / _refined_interface_prolongation, src/kernel/laplace3d_near.jl:99-225)
    equal the oracle's TargetMap and the host build bit for bit: refined nodes and weights, parent and
    split flags in the same order, prolongation rows, columns and values; corrections and evaluation
    follow."""
    say(f"target map: device refinement + prolongation vs oracle and host (n_quad={nq}, l_ec={l_ec})")
    centers = [(0.2, 0.1, 0.05), (0.55, -0.25, 0.27), (-0.58, 0.0, 0.0), (0.0, 0.39, -0.29)]
    srcs, pos, q, grp = blob_cloud(dbie, centers, 0.02)
    itf = dbie.single_dielectric_box3d(*BOX1, nq, l_ec, 4.0, 1.0)
    td = dbie.TargetPotentialMap(itf, pos, 1e-9, device_refine=True)
    th = dbie.TargetPotentialMap(itf, pos, 1e-9, device_refine=False)
    to = O.TargetMap(O.Interface.box(*BOX1, nq, l_ec, 4.0, 1.0), pos, 1e-9)
    oxyzw, opar, ospl, (opp, opc, opv), (ocp, occ, ocv) = to.parts()
    xyzw, par, spl = td.refined()
    say(f"  refined nodes {td.n_refined} (oracle {oxyzw.shape[1]}), pairs {td.n_pairs}")
    assert np.array_equal(xyzw, oxyzw)
    assert np.array_equal(par, opar) and np.array_equal(spl, ospl.astype(bool))
    (pp, pc, pv), (cp, cc, cv) = td.matrices()
    assert np.array_equal(pp, opp.astype(np.int64)) and np.array_equal(pc, opc) and np.array_equal(pv, opv)
    (hpp, hpc, hpv), (hcp, hcc, hcv) = th.matrices()
    hxyzw, hpar, hspl = th.refined()
    assert np.array_equal(xyzw, hxyzw) and np.array_equal(pp, hpp) and np.array_equal(pc, hpc)
    assert np.array_equal(pv, hpv)
    assert np.array_equal(cp, hcp) and np.array_equal(cc, hcc) and np.array_equal(cv, hcv)
    S = np.random.default_rng(7).standard_normal((itf.n_points, 3))
    U, Uh = td.apply(S), th.apply(S)
    assert np.abs(U - Uh).max() <= 1e-13 * np.abs(Uh).max()
    R, Rh = td.reaction(grp, q, S), th.reaction(grp, q, S)
    assert np.abs(R - Rh).max() <= 1e-13 * np.abs(Rh).max()
    ms_d, ms_h = td.setup_ms(), th.setup_ms()
    say(f"  setup ms device {ms_d} host {ms_h}")
