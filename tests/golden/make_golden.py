"""Golden fixtures for the parity tests, generated by the CPU oracle (oracle/, test infrastructure).

    python tests/golden/make_golden.py

Writes tests/golden/<name>.npz (plain float/int arrays; numpy.load reads them with allow_pickle=False)
and tests/golden/golden_meta.json (the git commit of the tree that generated them, the seeds and the
sizes).  The reference ships no stored vectors; these are the oracle's, so a later change of the oracle
or of the library shows up as a difference against a recorded state.

  cfg1       config 1 (box 1.2 x 0.8 x 0.6, n_quad 4, l_ec 0.2, eps 4 / 1; reference README.md:29-31):
             node SoA (x y z nx ny nz w, 7 x N), the dense-LU solution sigma for one unit charge at
             (0.1, 0.1, 0.1) with eps_src = 4, and its total induced charge dot(w, sigma) + 1 / eps_in.
  tiny       two-panel interfaces: parallel unit panels 0.2 apart (an upsampled pair) and two panels
             meeting at a right-angle edge (a touching, adaptive pair); per interface the neighbor
             decisions (src, trg, p_up; adaptive pairs with p_up = -1) and the correction blocks [b, t, c].
  cfg2_rows  config 2 (box 2.4 x 1.6 x 1.2, n_quad 4, l_ec 0.05, eps 10 / 1): 1024 rows (drawn with
             numpy default_rng(8), PCG64) of the corrected operator applied to
             x = default_rng(7).standard_normal(N).
  cfg5_mini  scaled-down config 5 (box 1.2 x 0.8 x 0.6, n_quad 5, l_ec 0.2, eps 4 / 1): eight unit
             Gaussian blobs (sigma 0.02, 6^3 Gauss-Hermite nodes) at the centres of a 2 x 2 x 2 lattice;
             their nodes and weights, the box-screened volume RHS, the dense-LU densities and the 8 x 8
             reaction matrix (near-corrected target potential, hcubature atol 1e-10).
"""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

CFG1 = (1.2, 0.8, 0.6, 4, 0.2, 4.0, 1.0)
CFG2 = (2.4, 1.6, 1.2, 4, 0.05, 10.0, 1.0)
CFG5 = (1.2, 0.8, 0.6, 5, 0.2, 4.0, 1.0)


def cfg1():
    f = O.Interface.box(*CFG1)
    nd = f.nodes()
    F = f.rhs_points(np.array([[0.1, 0.1, 0.1]]), np.array([1.0]), 4.0)[:, 0]
    sigma = O.solve_lu(f.lhs_dense(), F)
    charge = float(nd[6] @ sigma + 1.0 / 4.0)
    return dict(nodes=nd, F=F, sigma=sigma, induced_charge=np.array([charge]))


def tiny_interfaces():
    """(corners [P, 4, 3], normals [P, 3]) of the two two-panel interfaces."""
    par = (np.array([[(-0.5, -0.5, 0.0), (0.5, -0.5, 0.0), (0.5, 0.5, 0.0), (-0.5, 0.5, 0.0)],
                     [(-0.5, -0.5, 0.2), (0.5, -0.5, 0.2), (0.5, 0.5, 0.2), (-0.5, 0.5, 0.2)]]),
           np.array([[0, 0, 1.0], [0, 0, 1.0]]))
    edge = (np.array([[(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (1.0, 1.0, 0.0), (0.0, 1.0, 0.0)],
                      [(0.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 1.0, 1.0), (0.0, 0.0, 1.0)]]),
            np.array([[0, 0, -1.0], [-1.0, 0, 0]]))
    return {"parallel": par, "edge": edge}


def tiny():
    out = {}
    for name, (corners, normals) in tiny_interfaces().items():
        f = O.Interface.from_panels(corners, normals, 4, 2.0, 1.0)
        nl = f.neighbor_list(8, 1e-9, True)
        blocks = nl.corrections()
        up = sorted(nl.upsample.items())
        adp = sorted(nl.adaptive)
        dec = np.array([[i, j, p] for (i, j), p in up] + [[i, j, -1] for (i, j) in adp], dtype=np.int64).reshape(-1, 3)
        src, trg = nl.pairs()
        out[f"{name}_corners"] = corners
        out[f"{name}_normals"] = normals
        out[f"{name}_decisions"] = dec
        out[f"{name}_pairs"] = np.stack([src, trg]).astype(np.int64)
        out[f"{name}_blocks"] = np.ascontiguousarray(blocks)
    return out


def cfg2_rows():
    f = O.Interface.box(*CFG2)
    nl = f.neighbor_list(8, 1e-9, True)
    nl.corrections()
    x = np.random.default_rng(7).standard_normal(f.N)
    rows = np.sort(np.random.default_rng(8).choice(f.N, 1024, replace=False)).astype(np.int32)
    y = f.apply_DT_rows(rows, x) + nl.apply(x)[rows] + f.diag()[rows] * x[rows]
    return dict(rows=rows, y=y, N=np.array([f.N]))


def blob(center, sigma, n=6):
    """Unit Gaussian charge on an n^3 tensor Gauss-Hermite cloud (physicists' nodes x sqrt(2) sigma)."""
    t, wt = np.polynomial.hermite.hermgauss(n)
    pts = center + np.sqrt(2) * sigma * np.stack(np.meshgrid(t, t, t, indexing="ij"), -1).reshape(-1, 3)
    w = (wt[:, None, None] * wt[None, :, None] * wt[None, None, :]).ravel() / np.pi ** 1.5
    return pts, w


def lattice_centers(Lx, Ly, Lz, nx, ny, nz):
    xs = -Lx / 2 + (np.arange(nx) + 0.5) * Lx / nx
    ys = -Ly / 2 + (np.arange(ny) + 0.5) * Ly / ny
    zs = -Lz / 2 + (np.arange(nz) + 0.5) * Lz / nz
    return np.array([[x, y, z] for x in xs for y in ys for z in zs])


def cfg5_mini():
    Lx, Ly, Lz = CFG5[:3]
    f = O.Interface.box(*CFG5)
    C = lattice_centers(Lx, Ly, Lz, 2, 2, 2)
    clouds = [blob(c, 0.02) for c in C]
    pos = np.concatenate([p for p, _ in clouds])
    w = np.concatenate([ww for _, ww in clouds])
    grp = np.concatenate([[0], np.cumsum([len(ww) for _, ww in clouds])]).astype(np.int64)
    F = np.stack([f.rhs_volume(p, ww, f.screen_sharp(p, np.ones(len(ww))), 1.0) for p, ww in clouds], 1)
    sigma = O.solve_lu(f.lhs_dense(), F)
    tm = O.TargetMap(f, pos, 1e-10)
    R = tm.reaction(grp, w, sigma)
    return dict(centers=C, pos=pos, w=w, grp=grp, F=F, sigma=sigma, R=R, N=np.array([f.N]))


def main():
    O.build()
    commit = subprocess.run(["git", "rev-parse", "HEAD"], cwd=ROOT, capture_output=True, text=True).stdout.strip()
    dirty = bool(subprocess.run(["git", "status", "--porcelain", "oracle"], cwd=ROOT, capture_output=True,
                                text=True).stdout.strip())
    meta = {"generated_by": "tests/golden/make_golden.py", "commit": commit, "oracle_modified": dirty,
            "seeds": {"cfg2_rows": {"x": "default_rng(7).standard_normal(N)", "rows": "default_rng(8).choice(N, 1024)"}},
            "fixtures": {}}
    for name, fn in (("cfg1", cfg1), ("tiny", tiny), ("cfg2_rows", cfg2_rows), ("cfg5_mini", cfg5_mini)):
        arrs = fn()
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **arrs)
        meta["fixtures"][name] = {k: list(np.shape(v)) for k, v in arrs.items()}
        print(f"wrote {path} ({os.path.getsize(path)} bytes)", flush=True)
    with open(os.path.join(HERE, "golden_meta.json"), "w") as fh:
        json.dump(meta, fh, indent=1)


if __name__ == "__main__":
    main()
