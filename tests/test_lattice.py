"""Batch results on disk (src/solver/batch_io.jl) and the lattice-batch helpers
(src/solver/lattice_batch.jl:154-184); CPU only (the GPU solve is in test_gpu_multibox.py)."""
import os

import numpy as np
import pytest


def make_batch(dbie, n=40, K=3, seed=0):
    rng = np.random.default_rng(seed)
    gidx = np.stack(np.meshgrid(np.arange(4), np.arange(5), np.arange(2), indexing="ij"), -1).reshape(-1, 3)[:n]
    pos = (gidx * 0.1 - 0.2).T.astype(float)
    dens = rng.standard_normal((n, K))
    return dbie.LatticeBatch(np.array([(i, i + 1) for i in range(K)]), gidx, pos, np.full(n, 1e-3), dens)


def test_envelope_and_sources(dbie):
    b = make_batch(dbie)
    env = dbie.envelope_volume_source(b)
    assert np.allclose(env.density, np.sqrt((b.densities ** 2).sum(axis=1)), rtol=1e-15)
    vss = dbie.batch_volume_sources(b)
    assert len(vss) == 3 and all(np.array_equal(v.positions, b.positions) for v in vss)
    assert np.array_equal(vss[2].density, b.densities[:, 2])
    with pytest.raises(ValueError):
        dbie.LatticeBatch(b.pair_ids[:2], b.gidx, b.positions, b.weights, b.densities)


def test_batch_result_roundtrip(dbie, tmp_path):
    g = dbie.multi_dielectric_box3d(2, 0.5, [((0.0, 0.0, 0.0), 1.0, 1.0, 1.0), ((1.0, 0.0, 0.0), 1.0, 1.0, 1.0)],
                                    [2.0, 4.0], 1.0)
    b = make_batch(dbie)
    sigma = np.random.default_rng(1).standard_normal((g.n_points, 3))
    br = dbie.batch_result(7, b, {"sigma": sigma, "interface": g, "stats": {"niter": 12}}, {"t_total": 1.5})
    path = os.path.join(str(tmp_path), "sub", "batch_0007.npz")
    assert not dbie.is_complete_batch(path)
    assert dbie.save_batch_result(path, br) == path
    assert sorted(os.listdir(os.path.dirname(path))) == ["batch_0007.npz"]  # no temporary left behind
    assert dbie.is_complete_batch(path)
    r = dbie.load_batch_result(path)
    assert r.version == dbie.BATCH_FORMAT_VERSION and r.batch_id == 7
    assert np.array_equal(r.pair_ids, b.pair_ids) and np.array_equal(r.gidx, b.gidx)
    assert np.array_equal(r.weights, b.weights) and np.array_equal(r.densities, b.densities)
    assert np.array_equal(r.sigma, sigma)
    assert r.stats == {"niter": 12, "dof": g.n_points, "n_support": 40, "K": 3, "t_total": 1.5}
    assert np.array_equal(r.interface.nodes, g.nodes)
    assert np.array_equal(r.interface.eps_in, g.eps_in) and np.array_equal(r.interface.eps_out, g.eps_out)
    # overwriting replaces the file whole
    br.batch_id = 8
    dbie.save_batch_result(path, br)
    assert dbie.load_batch_result(path).batch_id == 8


def test_batch_result_rejects_bad_files(dbie, tmp_path):
    g = dbie.single_dielectric_box3d(1.0, 1.0, 1.0, 2, 0.5, 4.0, 1.0)
    b = make_batch(dbie)
    br = dbie.batch_result(1, b, {"sigma": np.zeros((g.n_points, 3)), "interface": g, "stats": {}})
    good = str(tmp_path / "good.npz")
    dbie.save_batch_result(good, br)
    # a truncated file (a write that died half way) is not complete
    raw = open(good, "rb").read()
    cut = str(tmp_path / "cut.npz")
    open(cut, "wb").write(raw[: len(raw) // 2])
    assert not dbie.is_complete_batch(cut)
    with pytest.raises(ValueError):
        dbie.load_batch_result(cut)
    # another format version
    with np.load(good, allow_pickle=False) as z:
        f = {k: z[k] for k in z.files}
    old = str(tmp_path / "old.npz")
    np.savez(old, **{**f, "version": np.int64(0)})
    assert not dbie.is_complete_batch(old)
    with pytest.raises(ValueError, match="format version"):
        dbie.load_batch_result(old)
    # inconsistent shapes
    for key, bad in (("weights", f["weights"][:-1]), ("pair_ids", f["pair_ids"][:-1]), ("gidx", f["gidx"][:-1]),
                     ("sigma", f["sigma"][:, :-1])):
        p = str(tmp_path / f"bad_{key}.npz")
        np.savez(p, **{**f, key: bad})
        with pytest.raises(ValueError, match="mismatch"):
            dbie.load_batch_result(p)
    assert not dbie.is_complete_batch(str(tmp_path / "missing.npz"))
    assert not dbie.is_complete_batch(str(tmp_path))
