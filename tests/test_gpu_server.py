"""The request server on the GPU: results are the same alone and in a mix, failures stay with
their request, memory admission and cancellation (run with -m gpu)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"\n[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


BOX_A = (1.0, 0.8, 0.6, 4, 0.1, 10.0, 1.0)


def charges(seed, K, L):
    rng = np.random.default_rng(seed)
    half = np.array(L) / 2 - 0.05
    return rng.uniform(-half, half, (K, 3)), rng.uniform(-1, 1, K)


def test_alone_equals_mix(dbie):
    pa, qa = charges(1, 5, BOX_A[:3])
    say("server: request A alone (FP64 and split-FP16)")
    with dbie.Server(solve_slots=2) as srv:
        ia = srv.submit_box(*BOX_A, pa, qa, precision="fp64", rtol=1e-12)
        ib = srv.submit_box(*BOX_A, pa, qa, precision="fp16x3", rtol=1e-11)
        S64_alone = srv.result(ia)
        Sm_alone = srv.result(ib)
    # reference: the same solve through the operator API
    itf = dbie.single_dielectric_box3d(*BOX_A[:5], BOX_A[5], BOX_A[6])
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(itf, 1e-9, 1e-9, 8)
    F = dbie.rhs_point_charges(itf, pa, qa, BOX_A[5])
    S_ref, _ = op.solve(F, rtol=1e-12)
    assert rel(S64_alone, S_ref) < 1e-10
    for k in range(5):
        assert rel(Sm_alone[:, k], S_ref[:, k]) < 1e-8

    say("server: A inside a mix of 8 requests (3 share its operator, 4 other boxes)")
    # slots are paused until every request is set up, so that A and its companions are fused
    # into one column block, and the other boxes with the same options join that solve as members
    # of a fused Krylov process
    others = [((1.3, 0.7, 0.9, 4, 0.1, 4.0, 1.0), 3), ((0.6, 0.6, 0.6, 4, 0.07, 40.0, 1.0), 7),
              ((1.1, 0.9, 0.5, 4, 0.1, 2.5, 1.0), 1), ((0.9, 0.9, 0.9, 4, 0.1, 77.0, 1.0), 12)]
    with dbie.Server(solve_slots=2, setup_workers=3) as srv:
        srv.pause(True)
        ids = {}
        for j, (box, K) in enumerate(others[:2]):
            p, q = charges(10 + j, K, box[:3])
            ids[f"o{j}"] = srv.submit_box(*box, p, q, precision="fp64", rtol=1e-12)
        for j in range(2):  # companions on A's operator
            p, q = charges(20 + j, 4 + 3 * j, BOX_A[:3])
            ids[f"c{j}"] = srv.submit_box(*BOX_A, p, q, precision="fp64", rtol=1e-12)
        ids["A"] = srv.submit_box(*BOX_A, pa, qa, precision="fp64", rtol=1e-12)
        ids["Am"] = srv.submit_box(*BOX_A, pa, qa, precision="fp16x3", rtol=1e-11)
        for j, (box, K) in enumerate(others[2:]):
            p, q = charges(30 + j, K, box[:3])
            ids[f"p{j}"] = srv.submit_box(*box, p, q, precision="fp16x3", rtol=1e-11)
        import time
        t0 = time.perf_counter()
        while any(srv.info(v)["state"] != "ready" for v in ids.values()):
            assert time.perf_counter() - t0 < 120, [srv.info(v) for v in ids.values()]
            time.sleep(0.01)
        srv.pause(False)
        infos = {k: srv.wait(v) for k, v in ids.items()}
        assert all(i["status"] == 0 and i["state"] == "done" for i in infos.values()), infos
        S64_mix = srv.result(ids["A"])
        Sm_mix = srv.result(ids["Am"])
        st = srv.stats()
    say(f"  stats {st}; A batch columns {infos['A']['batch_K']}")
    assert st["completed"] == len(ids) and st["failed"] == 0
    assert st["cache_hits"] >= 2 and st["operator_builds"] == 5
    assert st["fused_requests"] >= 2 and infos["A"]["batch_K"] >= 5 + 4 + 7  # A shared a block solve
    assert st["multi_geometry_batches"] >= 1  # ... with requests on other geometries
    assert rel(S64_mix, S64_alone) < 1e-10
    for k in range(5):
        assert rel(Sm_mix[:, k], S64_alone[:, k]) < 1e-8
        assert rel(Sm_mix[:, k], Sm_alone[:, k]) < 1e-8


def test_failures_stay_with_their_request(dbie):
    say("server: injected allocation failure, non-convergence, over-budget, cancellation")
    boxes = [(1.0 + 0.1 * j, 0.8, 0.6, 4, 0.1, 10.0, 1.0) for j in range(4)]
    with dbie.Server(inject_oom_every=2, solve_slots=2, setup_workers=1) as srv:
        ids = []
        for j, box in enumerate(boxes):
            p, q = charges(40 + j, 3, box[:3])
            ids.append(srv.submit_box(*box, p, q, precision="fp64"))
        p, q = charges(50, 2, boxes[0][:3])
        nc = srv.submit_box(*boxes[0], p, q, precision="fp64", itmax=1)
        infos = [srv.wait(i) for i in ids]
        oom = [i for i in infos if i["status"] == 3]
        ok = [i for i in infos if i["status"] == 0]
        assert len(oom) == 2 and len(ok) == 2, infos
        for i in ok:
            assert srv.result(i["id"]).shape == (i["N"], 3)
        for i in oom:
            with pytest.raises(Exception):
                srv.result(i["id"])
        inc = srv.wait(nc)
        assert inc["status"] in (5, 3), inc  # not converged (or its operator build was the failing one)
    with dbie.Server(hbm_budget=1e6) as srv:
        p, q = charges(60, 2, boxes[0][:3])
        r = srv.submit_box(*boxes[0], p, q)
        info = srv.wait(r)
        assert info["status"] == 3 and "budget" in info["message"]
    with dbie.Server(setup_workers=1, solve_slots=1) as srv:
        # the solve slots paused while the requests arrive and one is cancelled (queued, in setup or
        # ready, never inside a batch): the cancellation alone decides its fate, the others complete
        srv.pause(True)
        rs = []
        for j in range(4):
            p, q = charges(70 + j, 2, boxes[j][:3])
            rs.append(srv.submit_box(*boxes[j], p, q))
        srv.cancel(rs[-1])
        srv.pause(False)
        last = srv.wait(rs[-1])
        assert last["state"] == "cancelled"
        done = [srv.wait(r) for r in rs[:-1]]
        assert all(i["state"] == "done" for i in done), done


def test_serving_stream_small(dbie):
    """A short seeded stream of the serving workload at its own sizes (sides U(0.5, 2), l_ec in
    {0.1, 0.07, 0.05}, eps_in U(2, 80), up to 16 charges), arriving at a saturating rate: every served
    split-FP16 result is within 1e-8 (per column) of the same request solved alone in FP64."""
    reqs = dbie.serving_requests(10, seed=3, rate=50.0, kmax=16)
    say("server: 10-request stream at config-4 sizes, Poisson arrivals")
    import time
    with dbie.Server(solve_slots=2) as srv:
        t0 = time.perf_counter()
        ids = []
        for r in reqs:
            while time.perf_counter() - t0 < r["t"]:
                time.sleep(0.001)
            ids.append(srv.submit_box(*r["L"], r["n_quad"], r["l_ec"], r["eps_in"], r["eps_out"], r["points"],
                                      r["charges"], precision="fp16x3"))
        infos = [srv.wait(i) for i in ids]
        bad = [(r["L"], r["l_ec"], len(r["charges"]), i) for r, i in zip(reqs, infos) if i["status"] != 0]
        assert not bad, bad
        results = [srv.result(i) for i in ids]
        lat = [i["total_ms"] for i in infos]
        say(f"  p50 {np.percentile(lat, 50):.1f} ms  p99 {np.percentile(lat, 99):.1f} ms  stats {srv.stats()}")
    worst = 0.0
    for r, S in zip(reqs, results):
        itf = dbie.single_dielectric_box3d(*r["L"], r["n_quad"], r["l_ec"], r["eps_in"], r["eps_out"])
        op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(itf, 1e-9, 1e-9, 8)
        F = dbie.rhs_point_charges(itf, r["points"], r["charges"], r["eps_in"])
        S64, _ = op.solve(F, rtol=1e-12)
        err = (np.linalg.norm(S - S64, axis=0) / np.linalg.norm(S64, axis=0)).max()
        worst = max(worst, err)
        assert err < 1e-8, (r["L"], r["l_ec"], err)
    say(f"  max per-column error vs FP64 {worst:.2e}")


def test_admission_uses_full_width(dbie):
    """A request wider than max_batch_rhs is solved with all its columns, so admission reserves
    workspace for all of them: under a budget that holds the operator and 8 columns of workspace but
    not 40, the 40-column request fails with DBIE_OOM (and its neighbour within budget succeeds)."""
    box = (1.0, 0.8, 0.6, 4, 0.1, 10.0, 1.0)
    N = dbie.single_dielectric_box3d(*box).n_points
    npp = 16
    op_est = N * 8 * 24 + 50 * (N // npp) * npp * npp * 16 + 50 * (N // npp) * 8 + 64 * 1024 * 1024
    ws = lambda K: N * K * 8 * 62
    budget = op_est + 1.5 * ws(8)
    say(f"server: admission at full width (N={N}, budget {budget / 1e6:.0f} MB)")
    with dbie.Server(hbm_budget=budget, max_batch_rhs=8, solve_slots=1) as srv:
        p, q = charges(90, 40, box[:3])
        wide = srv.wait(srv.submit_box(*box, p, q, precision="fp64"))
        assert wide["status"] == 3 and "budget" in wide["message"], wide
        p, q = charges(91, 6, box[:3])
        narrow = srv.wait(srv.submit_box(*box, p, q, precision="fp64"))
        assert narrow["status"] == 0, narrow
    with dbie.Server(max_batch_rhs=8, solve_slots=1) as srv:
        p, q = charges(90, 40, box[:3])
        rid = srv.submit_box(*box, p, q, precision="fp64")
        info = srv.wait(rid)
        assert info["status"] == 0 and info["batch_K"] == 40, info
        assert srv.result(rid).shape == (N, 40)


@pytest.mark.parametrize("precision,krylov,pre", [("fp64", "columns", 1), ("fp16x3", "columns", 1), ("fp16x3", "block", 1),
                                                 ("fp64", "columns", 2), ("fp16x3", "columns", 2), ("fp16x3", "columns", 3)])
def test_fused_solve_equals_alone(dbie, precision, krylov, pre):
    """dbie_solve_fused: three boxes of different geometry (N 5k-13k, 1 to 12 columns) in one Krylov
    process give each box's own solution (FP64 within 1e-10, mixed within 1e-8 per column); block GMRES
    is refused (its block orthogonalisation would mix columns of different operators).  pre = 2 / 3: each
    member's near-field preconditioner on its own rows and columns, in fewer iterations than pre = 1."""
    boxes = [((0.9, 0.7, 0.6, 4, 0.1, 10.0, 1.0), 5), ((0.6, 0.6, 0.6, 4, 0.07, 40.0, 1.0), 1),
             ((1.3, 0.7, 0.9, 4, 0.1, 2.5, 1.0), 12)]
    say(f"fused solve of 3 geometries ({precision}, {krylov})")
    ops, Fs, alone = [], [], []
    for j, (box, K) in enumerate(boxes):
        itf = dbie.single_dielectric_box3d(*box)
        op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(itf, 1e-9, 1e-9, 8)
        p, q = charges(80 + j, K, box[:3])
        F = dbie.rhs_point_charges(itf, p, q, box[5])
        S64, _ = op.solve(F, rtol=1e-12)
        ops.append(op)
        Fs.append(F)
        alone.append(S64)
    rtol = 1e-12 if precision == "fp64" else 1e-11
    if krylov == "block":
        with pytest.raises(ValueError, match="krylov"):
            dbie.solve_fused(ops, Fs, rtol=rtol, precision=precision, krylov=krylov)
        return
    Ss, st = dbie.solve_fused(ops, Fs, rtol=rtol, precision=precision, krylov=krylov, precondition=pre)
    say(f"  stats {st}")
    assert st["solved"]
    if pre >= 2:
        _, st1 = dbie.solve_fused(ops, Fs, rtol=rtol, precision=precision, krylov=krylov, precondition=1)
        say(f"  iterations: diagonal scaling {st1['niter']}, near field {st['niter']}")
        assert st["niter"] < st1["niter"]
    bound = 1e-10 if precision == "fp64" else 1e-8
    for S, R in zip(Ss, alone):
        for k in range(S.shape[1]):
            assert rel(S[:, k], R[:, k]) < bound
