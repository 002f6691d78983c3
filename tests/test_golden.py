"""Parity against the committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the CPU oracle).

CPU tests: the fixtures load without pickle; the product's host setup (mesher nodes, near-pair
decisions) reproduces them bit for bit; the oracle still reproduces them (a change of the oracle
shows here); and the properties the mathematics fixes hold on the stored vectors (Gauss's law
for the induced charge, reciprocity of the reaction matrix, the box's reflection symmetry).
GPU tests: the device RHS, corrections, corrected apply, dense solve and reaction matrix against
the stored vectors."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
import make_golden as MG  # noqa: E402  (geometry constants and the blob / lattice helpers)


def load(name):
    with np.load(os.path.join(GOLD, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relmax(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def test_fixtures_load_and_match_meta():
    meta = json.load(open(os.path.join(GOLD, "golden_meta.json")))
    for name, shapes in meta["fixtures"].items():
        z = load(name)
        assert sorted(z) == sorted(shapes), name
        for k, s in shapes.items():
            assert list(z[k].shape) == s, (name, k)
            assert z[k].dtype != object


# ----------------------------------------------------------------- CPU: product host setup
def test_cfg1_nodes_bit_exact(dbie):
    z = load("cfg1")
    g = dbie.single_dielectric_box3d(*MG.CFG1)
    assert np.array_equal(g.nodes, z["nodes"])


def test_tiny_neighbor_decisions(dbie):
    z = load("tiny")
    for name in ("parallel", "edge"):
        g = dbie.interface_from_panels(z[f"{name}_corners"], z[f"{name}_normals"], 4, 2.0, 1.0)
        up, ad = dbie.build_neighbor_list(g, 8, 1e-9, correct_edges=True)
        dec = z[f"{name}_decisions"]
        want_up = {(int(i), int(j)): int(p) for i, j, p in dec if p >= 0}
        want_ad = sorted((int(i), int(j)) for i, j, p in dec if p < 0)
        assert up == want_up, name
        assert sorted(ad) == want_ad, name
    # the parallel pair is upsampled (order 8 is the first to meet 1e-9), the pair sharing an edge
    # goes to adaptive cubature
    assert set(z["parallel_decisions"][:, 2]) == {8}
    assert set(z["edge_decisions"][:, 2]) == {-1}


# ----------------------------------------------------------------- CPU: oracle still reproduces
def test_oracle_reproduces_cfg1(O):
    z = load("cfg1")
    f = O.Interface.box(*MG.CFG1)
    assert np.array_equal(f.nodes(), z["nodes"])
    F = f.rhs_points(np.array([[0.1, 0.1, 0.1]]), np.array([1.0]), 4.0)[:, 0]
    assert relmax(F, z["F"]) <= 1e-15
    sigma = O.solve_lu(f.lhs_dense(), F)
    assert relmax(sigma, z["sigma"]) <= 1e-12


def test_oracle_reproduces_tiny(O):
    z = load("tiny")
    for name in ("parallel", "edge"):
        f = O.Interface.from_panels(z[f"{name}_corners"], z[f"{name}_normals"], 4, 2.0, 1.0)
        nl = f.neighbor_list(8, 1e-9, True)
        B = nl.corrections()
        assert np.array_equal(np.stack(nl.pairs()), z[f"{name}_pairs"])
        assert relmax(B, z[f"{name}_blocks"]) <= 1e-13, name


def test_oracle_reproduces_cfg2_rows(O):
    z = load("cfg2_rows")
    f = O.Interface.box(*MG.CFG2)
    assert f.N == int(z["N"][0])
    nl = f.neighbor_list(8, 1e-9, True)
    nl.corrections()
    x = np.random.default_rng(7).standard_normal(f.N)
    rows = z["rows"]
    y = f.apply_DT_rows(rows, x) + nl.apply(x)[rows] + f.diag()[rows] * x[rows]
    assert relmax(y, z["y"]) <= 1e-13


def test_oracle_reproduces_cfg5_mini(O):
    z = load("cfg5_mini")
    f = O.Interface.box(*MG.CFG5)
    cl = [MG.blob(c, 0.02) for c in z["centers"]]
    assert np.array_equal(np.concatenate([p for p, _ in cl]), z["pos"])
    F = np.stack([f.rhs_volume(p, w, f.screen_sharp(p, np.ones(len(w))), 1.0) for p, w in cl], 1)
    assert relmax(F, z["F"]) <= 1e-14
    R = O.TargetMap(f, z["pos"], 1e-10).reaction(z["grp"], z["w"], z["sigma"])
    assert relmax(R, z["R"]) <= 1e-12


# ----------------------------------------------------------------- CPU: what the mathematics fixes
def test_cfg1_gauss_law():
    # total induced charge dot(w, sigma) + 1/eps_in -> 1/eps_out = 1 (test/solver/dielectric_box3d.jl:27,30)
    z = load("cfg1")
    q = float(z["nodes"][6] @ z["sigma"] + 0.25)
    assert q == pytest.approx(float(z["induced_charge"][0]), abs=1e-14)
    assert abs(q - 1.0) < 0.1


def test_cfg5_reaction_symmetries():
    z = load("cfg5_mini")
    R = z["R"]
    # reciprocity: the induced potential of blob b at blob a equals that of a at b
    assert np.abs(R - R.T).max() <= 1e-9 * np.abs(R).max()
    # a charge inside the box (eps_in 4 > eps_out 1) induces surface charge of its own sign, so every
    # self term is positive, and so is the matrix (an energy)
    assert np.all(np.diag(R) > 0)
    assert np.linalg.eigvalsh(0.5 * (R + R.T)).min() > 0
    # the eight lattice sites are images of each other under the box's three reflections, and the
    # graded mesh is reflection symmetric, so every diagonal entry is the same
    d = np.diag(R)
    assert np.abs(d - d[0]).max() <= 1e-9 * d[0]
    # pairs related by the same reflection carry the same coupling: sites differing in one axis only
    C = z["centers"]
    for axis in range(3):
        vals = [R[a, b] for a in range(8) for b in range(8)
                if np.count_nonzero(np.abs(C[a] - C[b]) > 1e-12) == 1 and abs(C[a, axis] - C[b, axis]) > 1e-12]
        assert len(vals) == 8
        assert np.ptp(vals) <= 1e-9 * abs(vals[0])
    # every blob carries unit charge
    assert np.allclose(np.add.reduceat(z["w"], z["grp"][:-1]), 1.0, atol=1e-13)


# ----------------------------------------------------------------- GPU
@pytest.mark.gpu
def test_gpu_cfg1_rhs_and_dense_solve(dbie):
    z = load("cfg1")
    say("golden cfg1: point RHS + uncorrected solve vs fixture")
    g = dbie.single_dielectric_box3d(*MG.CFG1)
    F = dbie.rhs_dielectric_box3d(g, dbie.PointSource((0.1, 0.1, 0.1), 1.0), 4.0)
    assert relmax(F, z["F"]) <= 1e-14
    op = dbie.BatchedDielectricOperator(g, corrected=False)
    S, st = op.solve(F[:, None], rtol=1e-13)
    assert st["solved"]
    assert relmax(S[:, 0], z["sigma"]) <= 1e-11
    A = dbie.lhs_dielectric_box3d(g)
    assert relmax(dbie.solve_lu(A, F), z["sigma"]) <= 1e-12


@pytest.mark.gpu
def test_gpu_tiny_corrections(dbie):
    z = load("tiny")
    for name in ("parallel", "edge"):
        say(f"golden tiny {name}: correction blocks vs fixture")
        g = dbie.interface_from_panels(z[f"{name}_corners"], z[f"{name}_normals"], 4, 2.0, 1.0)
        op = dbie.BatchedDielectricOperator(g, up_tol=1e-9, max_order=8)
        src, trg, blk = op.corrections()
        want = {(int(s), int(t)): b for (s, t), b in zip(z[f"{name}_pairs"].T, z[f"{name}_blocks"])}
        got = {(int(s), int(t)): b for s, t, b in zip(src, trg, blk)}
        assert sorted(got) == sorted(want), name
        scale = max(np.abs(b).max() for b in want.values())
        for k in want:
            assert np.abs(got[k] - want[k]).max() <= 1e-13 * scale, (name, k)


@pytest.mark.gpu
def test_gpu_cfg2_rows(dbie):
    z = load("cfg2_rows")
    say("golden config 2: 1024 corrected-apply rows vs fixture")
    g = dbie.single_dielectric_box3d(*MG.CFG2)
    assert g.n_points == int(z["N"][0])
    op = dbie.BatchedDielectricOperator(g, up_tol=1e-9, max_order=8)
    x = np.random.default_rng(7).standard_normal(g.n_points)
    y = op @ x
    assert relmax(y[z["rows"]], z["y"]) <= 1e-12
    # the split-FP16 operator stays within its error model on the same rows
    y3 = op.apply(np.repeat(x[:, None], 16, 1), precision="fp16x3")[:, 0]
    assert relmax(y3[z["rows"]], z["y"]) <= 2e-6


@pytest.mark.gpu
def test_gpu_cfg5_mini_reaction(dbie):
    z = load("cfg5_mini")
    say("golden scaled config 5: screened volume RHS, solve, reaction matrix vs fixture")
    g = dbie.single_dielectric_box3d(*MG.CFG5)
    assert g.n_points == int(z["N"][0])
    vss = []
    for a in range(8):
        s0, s1 = z["grp"][a], z["grp"][a + 1]
        vss.append(dbie.VolumeSource(z["pos"][s0:s1], z["w"][s0:s1], np.ones(s1 - s0)))
    F = dbie.rhs_dielectric_box3d_fmm3d(g, vss, 1e-12)
    assert relmax(F, z["F"]) <= 1e-13
    op = dbie.BatchedDielectricOperator(g, corrected=False)
    S, st = op.solve(F, rtol=1e-13)
    assert st["solved"]
    assert relmax(S, z["sigma"]) <= 1e-11
    R, _ = dbie.reaction_matrix(g, vss, z["sigma"], lhs_tol=1e-10)
    assert relmax(R, z["R"]) <= 1e-10
