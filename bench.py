"""Benchmark: batched many-RHS solves of the single dielectric box on one MI355X.

Workload (the configuration the metric is quoted on): box 2.4 x 1.6 x 1.2, l_ec = 0.05,
n_quad = 4 (N = 64,384 unknowns), eps_in = 4, eps_out = 1, K = 256 point charges placed
uniformly at random inside the box at least 0.05 from every face, q ~ U(-1, 1) (seeded).
Corrected operator with the block-solve defaults (up_tol = 1e-9, max_order = 8,
correct_edges = true), right diagonal preconditioner, rtol = 1e-10.

A step is one block solve of all K right-hand sides, inputs resident in HBM:
  * mixed:  FP64 iterative refinement around GMRES on the split-FP16 MFMA operator
            (near panel pairs taken exactly from the FP32 near blocks);
  * fp64:   the plain FP64 GMRES path (the comparator), timed beside it.
value = K * steps / (time of the timed steps) for the mixed path, in RHS-solves/s.

Prints ONE JSON line.  Extra fields: the FP64 comparator rate (median of warm solves, per-column
and block GMRES), the mixed/FP64 ratio, p50/p99 step latency, per-RHS max relative error of the
mixed solution vs the FP64 solution, the bf16x3 and fp32 mixed forms at the same size, the roofline
of the dominant kernel (HIP events on the operator's stream, in a profiled pass after the timed
one), the serving stream (requests of the config-4 distribution at a saturating arrival rate, mixed
and FP64, p50/p99 latency, error of every served mixed result vs its FP64 one) and the CPU baseline
(the oracle on this host's cores, bounded sample, median of 3).
"""
from __future__ import annotations

import argparse
import json
import subprocess
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dielectricbie.jl_amd"))

METRIC = "RHS-solves/s at ≤1e-8 rel err vs FP64, mixed vs plain FP64 GPU; p50/p99 latency"
BOX = (2.4, 1.6, 1.2)
L_EC, NQ, EPS_IN, EPS_OUT = 0.05, 4, 4.0, 1.0
PEAK = {"fp64": 78.6, "bf16": 2500.0, "hbm": 8000.0}  # TFLOP/s, TFLOP/s, GB/s
PRECONDITIONERS = {1: "diagonal", 2: "near-field", 3: "near-field, two sweeps"}


def charges(K: int, seed: int = 0, margin: float = 0.05):
    rng = np.random.default_rng(seed)
    half = np.array(BOX) / 2 - margin
    return rng.uniform(-half, half, (K, 3)), rng.uniform(-1.0, 1.0, K)


def cross_plane_pairs(iface) -> int:
    """Nonzero Dᵀ pairs: N^2 minus same-plane pairs (the self pair is same-plane)."""
    nd = iface.nodes
    key = np.round(np.vstack([nd[3:6], (nd[3:6] * nd[0:3]).sum(0)]).T, 12)
    _, counts = np.unique(key, axis=0, return_counts=True)
    N = iface.n_points
    return int(N * N - np.sum(counts.astype(np.int64) ** 2))


def host_cpu():
    """Model name and logical CPU count of this host (lscpu / /proc/cpuinfo)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def cpu_baseline(iters_fp64: int, ksample: int = 8, reps: int = 3):
    """The oracle's FP64 corrected matvec (direct-sum far field + near corrections + diagonal) on all
    host cores for a bounded sample: ksample columns, timed reps times (median); the rate assumes the
    CPU solve needs the FP64 GPU path's iteration count and one matvec per iteration, and ignores the
    Krylov vector work (so it favours the CPU)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    f = O.Interface.box(*BOX, NQ, L_EC, EPS_IN, EPS_OUT)
    A = O.CorrectedOperator(f)
    X = np.random.default_rng(1).standard_normal((f.N, ksample))
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        A @ X
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    rate = ksample / (dt * max(iters_fp64, 1))
    threads = int(O.lib().orc_num_threads())
    return {"value": rate, "unit": "RHS-solves/s", "cores": threads, "kind": "port",
            "label": "CPU restatement (direct sum), not the reference (Julia/FMM3D is not installed)",
            "host": dict(host_cpu(), threads_used=threads),
            "sample": f"FP64 corrected matvec (direct-sum far field + near corrections) of {ksample} columns at "
                      f"N={f.N}, median of {reps} runs {dt:.2f} s (runs {', '.join(f'{t:.2f}' for t in times)}); "
                      f"rate = {ksample} / (t x {iters_fp64} GMRES iterations of the FP64 path)"}


def reaction_workload(a, dbie, log):
    """Box 1.2 x 0.8 x 0.6, n_quad = 5, l_ec = 0.03 (N = 65,800), eps 4/1; K = 1024 unit Gaussian
    blobs (width 0.02, 6^3 Gauss-Hermite points) on the 16 x 8 x 8 interior lattice; screened
    volume RHS, mixed-precision and FP64 block solves, and the 1024 x 1024 reaction matrix from each
    (near-corrected target potential at the 221,184 blob points, hcubature atol 1e-9).
    A step is: RHS + mixed block solve + reaction matrix, with the operator and target map built."""
    L = (1.2, 0.8, 0.6)
    t0 = time.perf_counter()
    iface = dbie.single_dielectric_box3d(*L, 5, 0.03, 4.0, 1.0)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(iface, 1e-9, 1e-9, 8)
    t_op = time.perf_counter() - t0
    C = dbie.blob_lattice_centers(*L)[: a.blobs]
    srcs = [dbie.gaussian_blob_gh(c, 0.02, 6) for c in C]
    K = len(srcs)
    pts = np.concatenate([s.positions.T for s in srcs])
    q = np.concatenate([s.weights * s.density for s in srcs])
    grp = np.concatenate([[0], np.cumsum([len(s.weights) for s in srcs])]).astype(np.int64)
    t1 = time.perf_counter()
    tm = dbie.TargetPotentialMap(iface, pts, 1e-9)
    t_tm = time.perf_counter() - t1
    log(f"N={iface.n_points} K={K} M={len(q)} operator {t_op:.1f}s target map {t_tm:.1f}s "
        f"(refined {tm.n_refined}, cubatures {tm.n_pairs}, capped {tm.n_capped}; phases ms {tm.setup_ms()})")
    N = iface.n_points
    dS = dbie.DeviceMatrix(N, K)

    # both paths with the near-field preconditioner; the mixed solve ends on its refined residual estimate
    # (as the FP64 GMRES ends on its Krylov estimate)
    def step(precision):
        ts = time.perf_counter()
        F = dbie.rhs_dielectric_box3d_fmm3d(iface, srcs, 1e-9)
        tr = time.perf_counter()
        dF = dbie.DeviceMatrix.from_host(F)
        st = op.solve_device(dF, dS, rtol=1e-10, precision=precision, inner_rtol=a.inner_rtol,
                             check_residual=False, precondition=2, verify=False)
        tsv = time.perf_counter()
        R = tm.reaction(grp, q, dS)
        te = time.perf_counter()
        return R, st, {"rhs_ms": (tr - ts) * 1e3, "solve_ms": (tsv - tr) * 1e3, "reaction_ms": (te - tsv) * 1e3,
                       "total_ms": (te - ts) * 1e3}

    R64, st64, t64 = step("fp64")
    log(f"fp64: iters={st64['niter']} {t64}")
    step(a.precision)  # warm-up
    lat, out = [], None
    for _ in range(max(a.steps, 1)):
        out = step(a.precision)
        lat.append(out[2]["total_ms"])
    Rm, stm, tmix = out
    err = float(np.abs(Rm - R64).max() / np.abs(R64).max())
    log(f"{a.precision}: iters={stm['niter']} refine={stm['nrefine']} {tmix} max entry err {err:.2e}")
    print(json.dumps({
        "workload": "reaction", "metric": "RHS-solves/s (RHS + block solve + reaction matrix), mixed precision",
        "value": K / (float(np.median(lat)) * 1e-3), "unit": "RHS-solves/s", "dtype": a.precision,
        "fp64_value": K / (t64["total_ms"] * 1e-3), "config": {"N": N, "K": K, "M": int(len(q)),
        "N_refined": tm.n_refined, "cubatures": tm.n_pairs},
        "setup_s": {"operator": t_op, "target_map": t_tm}, "target_map_phases_ms": tm.setup_ms(), "mixed_ms": tmix, "fp64_ms": t64,
        "mixed_iters": stm["niter"], "mixed_refinements": stm["nrefine"], "fp64_iters": st64["niter"],
        "max_entry_err_vs_fp64_rel_to_max": err, "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99))}), flush=True)


def single_workload(a, dbie, log):
    """Box 2.4 x 1.6 x 1.2, l_ec = 0.05, n_quad = 4, eps_in = 10, eps_out = 1 (N = 64,384); one unit point
    charge at a seeded uniform-random interior point; plain FP64 GMRES to a relative residual of 1e-10
    (matrix-free far field, near corrections).  A step is one solve with the operator built; the
    single-column far field runs on the FP64 VALU (one thread per target, sources staged in LDS)."""
    t0 = time.perf_counter()
    iface = dbie.single_dielectric_box3d(*BOX, NQ, L_EC, 10.0, EPS_OUT)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(iface, 1e-9, 1e-9, 8)
    t_op = time.perf_counter() - t0
    rng = np.random.default_rng(a.seed)
    src = rng.uniform(-0.5, 0.5, (1, 3)) * np.array(BOX)
    F = dbie.rhs_point_charges(iface, src, np.ones(1), 10.0)
    N = iface.n_points
    dF = dbie.DeviceMatrix.from_host(F.reshape(N, 1))
    dS = dbie.DeviceMatrix(N, 1)
    # the reference's diagonal scaling, then the near-field preconditioner (the value: the faster)
    diag_lat = []
    for _ in range(a.warmup):
        op.solve_device(dF, dS, rtol=1e-10, precision="fp64", precondition=1)
    for _ in range(max(a.steps, 1)):
        t1 = time.perf_counter()
        st1 = op.solve_device(dF, dS, rtol=1e-10, precision="fp64", precondition=1)
        diag_lat.append((time.perf_counter() - t1) * 1e3)
    for _ in range(a.warmup):
        op.solve_device(dF, dS, rtol=1e-10, precision="fp64", precondition=2)
    op.profile(True)
    op.profile_read()
    lat, st = [], None
    for _ in range(max(a.steps, 1)):
        t1 = time.perf_counter()
        st = op.solve_device(dF, dS, rtol=1e-10, precision="fp64", precondition=2)
        lat.append((time.perf_counter() - t1) * 1e3)
    prof = op.profile_read()
    op.profile(False)
    log(f"diagonal scaling: iters={st1['niter']} {np.median(diag_lat):.1f} ms; near-field: iters={st['niter']}")
    pairs = cross_plane_pairs(iface)
    n_far, far_ms = prof.get("far_fp64", (0, 0.0))
    avg = far_ms / max(n_far, 1)
    log(f"single FP64 solve: N={N} iters={st['niter']} residual {st['max_rel_residual']:.2e} "
        f"{np.median(lat):.1f} ms; far field {avg:.3f} ms per launch ({n_far} launches), prof={prof}")
    print(json.dumps({
        "workload": "single", "metric": "FP64 solve time, one point charge, relative residual 1e-10",
        "value": float(np.median(lat)), "unit": "ms", "higher_is_better": False, "dtype": "f64",
        "config": {"N": N, "eps_in": 10.0, "source": src.ravel().tolist(), "precondition": "near-field"},
        "iters": st["niter"], "max_rel_residual": st["max_rel_residual"], "setup_s": t_op,
        "diagonal_scaling": {"ms": float(np.median(diag_lat)), "iters": st1["niter"],
                             "max_rel_residual": st1["max_rel_residual"]},
        "far_field": {"avg_launch_ms": avg, "pair_evaluations": pairs,
                      "pairs_per_s": pairs / (avg * 1e-3) if avg > 0 else None},
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99))}), flush=True)


def serve_stream(dbie, reqs, precision, inner_rtol, slots, log, setup_workers=2, verify=False):
    """Submit the requests at their arrival times to a fresh request server; wait for all.
    Returns (infos, results, wall seconds, server stats).  Every request runs with the near-field
    preconditioner; mixed-precision requests end on their refined residual estimate (verify=False),
    as the FP64 ones end on their Krylov estimate; each served result is compared with the FP64 one
    afterwards."""
    infos, results = [], []
    with dbie.Server(solve_slots=slots, setup_workers=setup_workers) as srv:
        t0 = time.perf_counter()
        ids = []
        for r in reqs:
            while time.perf_counter() - t0 < r["t"]:
                time.sleep(0.0005)
            ids.append(srv.submit_box(*r["L"], r["n_quad"], r["l_ec"], r["eps_in"], r["eps_out"], r["points"],
                                      r["charges"], precision=precision, inner_rtol=inner_rtol, verify=verify,
                                      precondition=2))
        for i in ids:
            infos.append(srv.wait(i))
            results.append(srv.result(i) if infos[-1]["status"] == 0 else None)
        wall = time.perf_counter() - t0
        stats = srv.stats()
    ok = sum(1 for i in infos if i["status"] == 0)
    log(f"serve {precision}: {ok}/{len(infos)} ok, wall {wall:.2f}s, stats {stats}")
    return infos, results, wall, stats


def serve_summary(infos, wall):
    ok = [i for i in infos if i["status"] == 0]
    lat = np.array([i["total_ms"] for i in ok]) if ok else np.zeros(1)
    K_total = sum(i["K"] for i in ok)
    return {"value": K_total / wall, "unit": "RHS-solves/s", "completed": len(ok), "rhs_total": K_total,
            "wall_s": wall, "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_setup_ms": float(np.mean([i["setup_ms"] for i in ok])) if ok else None,
            "mean_solve_ms": float(np.mean([i["solve_ms"] for i in ok])) if ok else None,
            "mean_queue_ms": float(np.mean([i["queue_ms"] for i in ok])) if ok else None}


def serve_both(a, dbie, log, requests, rate):
    """The serving stream in mixed precision and in FP64 (same requests, fresh servers); the error of
    each served mixed result against the served FP64 result of the same request."""
    reqs = dbie.serving_requests(requests, seed=a.seed, rate=rate)
    im, rm, wm, sm = serve_stream(dbie, reqs, a.precision, a.inner_rtol, a.slots, log, a.setup_workers)
    i64, r64, w64, s64 = serve_stream(dbie, reqs, "fp64", a.inner_rtol, a.slots, log, a.setup_workers)
    errs = [float(np.max(np.linalg.norm(S - R, axis=0) / np.linalg.norm(R, axis=0)))
            for S, R in zip(rm, r64) if S is not None and R is not None]
    mixed, fp64 = serve_summary(im, wm), serve_summary(i64, w64)
    ok = [i for i in im if i["status"] == 0]
    return {"requests": len(reqs), "arrival_rate": rate, "slots": a.slots, "setup_workers": a.setup_workers,
            "N_range": [int(min(i["N"] for i in ok)), int(max(i["N"] for i in ok))] if ok else None,
            "mixed": dict(mixed, server=sm), "fp64": dict(fp64, server=s64),
            "speedup_vs_fp64": mixed["value"] / fp64["value"] if fp64["value"] > 0 else None,
            "max_rel_err_vs_fp64": max(errs) if errs else None, "results_compared": len(errs)}


def serve_child(a, log, timeout_s=240):
    """The serving stream of the default line, run as `bench.py --workload serve` in a child process
    with a time limit, so that the headline measurement never depends on the request server
    finishing; returns its JSON line (or an error record)."""
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", "serve", "--requests", str(a.serve_requests),
           "--rate", str(a.serve_rate), "--slots", str(a.slots), "--setup-workers", str(a.setup_workers),
           "--precision", a.precision, "--inner-rtol", str(a.inner_rtol), "--seed", str(a.seed)]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        log(f"serve: no result within {timeout_s} s")
        return {"error": f"serving stream did not finish within {timeout_s} s"}
    lines = [ln for ln in r.stdout.strip().splitlines() if ln.startswith("{")]
    if r.returncode != 0 or not lines:
        log(f"serve: child exited {r.returncode}: {r.stderr[-2000:]}")
        return {"error": f"serving stream exited with status {r.returncode}"}
    out = json.loads(lines[-1])
    for k in ("workload", "metric", "value", "unit", "dtype", "p50_ms", "p99_ms"):
        out.pop(k, None)
    return out


def requests_workload(a, dbie, log):
    """Single serving-size requests solved alone on the device (no server, no fusion): per shape the
    FP64 and the split-FP16 solve times (best of 3 after 3 warm-ups), iterations, and the far- and
    near-field time per apply; the operator keeps no stored kernel matrix, as served operators do."""
    rng = np.random.default_rng(a.seed)
    shapes = [((1.0, 0.8, 0.6), 0.1, 4), ((1.0, 0.8, 0.6), 0.1, 32), ((1.5, 1.2, 1.0), 0.07, 8),
              ((1.5, 1.2, 1.0), 0.07, 48), ((1.8, 1.6, 1.4), 0.05, 4), ((1.8, 1.6, 1.4), 0.05, 64)]
    rows = []
    for L, lec, K in shapes:
        g = dbie.single_dielectric_box3d(*L, 4, lec, 20.0, 1.0)
        op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(g, 1e-9, 1e-9, 8)
        op.stored_far(False)
        half = np.asarray(L) / 2 - 0.05
        F = dbie.rhs_point_charges(g, rng.uniform(-half, half, (K, 3)), rng.uniform(-1, 1, K), 20.0)
        dF, dS = dbie.DeviceMatrix.from_host(F), dbie.DeviceMatrix(g.n_points, K)
        row = {"sides": list(L), "l_ec": lec, "N": g.n_points, "K": K}
        for name, kw in (("fp64", dict(precision="fp64")),
                         ("mixed", dict(precision=a.precision, inner_rtol=a.inner_rtol, verify=False))):
            for _ in range(3):
                op.solve_device(dF, dS, rtol=1e-10, precondition=2, check_residual=False, **kw)
            op.synchronize()
            ts, st = [], None
            for _ in range(3):
                t0 = time.perf_counter()
                st = op.solve_device(dF, dS, rtol=1e-10, precondition=2, check_residual=False, **kw)
                op.synchronize()
                ts.append((time.perf_counter() - t0) * 1e3)
            op.profile(True)
            op.profile_read()
            op.apply(F, kw["precision"])
            prof = op.profile_read()
            op.profile(False)
            row[name] = {"ms": min(ts), "iters": st["niter"], "refinements": st["nrefine"],
                         "apply_ms": {k: v[1] / max(v[0], 1) for k, v in prof.items() if v[0]}}
        row["speedup"] = row["fp64"]["ms"] / row["mixed"]["ms"]
        log(f"requests: {row}")
        rows.append(row)
    print(json.dumps({"workload": "requests", "precision": a.precision, "rows": rows}), flush=True)


def serve_workload(a, dbie, log):
    """Seeded Poisson stream of single-box requests (sides ~U(0.5, 2), l_ec in {0.1, 0.07, 0.05},
    n_quad 4, eps_in ~U(2, 80), 1-64 charges) through the request server, in mixed precision and in
    FP64; latency = arrival -> result on the host; error of each mixed result vs the FP64 one."""
    out = serve_both(a, dbie, log, a.requests, a.rate)
    print(json.dumps(dict({"workload": "serve", "metric": "served RHS-solves/s (arrival -> result), mixed precision",
                           "value": out["mixed"]["value"], "unit": "RHS-solves/s", "dtype": a.precision,
                           "p50_ms": out["mixed"]["p50_ms"], "p99_ms": out["mixed"]["p99_ms"]}, **out)), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--K", type=int, default=256)
    ap.add_argument("--precision", default="fp16x3", choices=["fp32", "bf16", "bf16x3", "fp16x3"])
    ap.add_argument("--inner-rtol", type=float, default=1e-5)
    ap.add_argument("--fp64-steps", type=int, default=5, help="timed FP64 solves per Krylov variant (median)")
    ap.add_argument("--prof-steps", type=int, default=3, help="profiled (untimed) mixed solves for the roofline")
    ap.add_argument("--variant-steps", type=int, default=2,
                    help="timed solves of the other mixed precisions (bf16x3, fp32) beside the headline")
    ap.add_argument("--no-serve", action="store_true", help="leave the serving stream out of the default line")
    ap.add_argument("--serve-requests", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=64.0, help="arrival rate of the serving stream in the "
                                                                   "default line (requests/s; saturating)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--mode", default="solve", choices=["solve", "apply", "headline"],
                    help="apply: time only operator applies (one per precision) for kernel profiling")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--krylov", default="columns", choices=["columns", "block"],
                    help="headline mode: GMRES per column (the headline form) or block GMRES")
    ap.add_argument("--precondition", type=int, default=2, choices=[1, 2, 3],
                    help="headline mode: the preconditioner (1 diagonal, 2 near field, 3 near field, two sweeps)")
    ap.add_argument("--shape", default="", help="apply mode: pin the low-precision far-field launch shape "
                                                "'splits,variant' instead of tuning it")
    ap.add_argument("--blobs", type=int, default=1024, help="reaction workload: first B lattice blobs")
    ap.add_argument("--workload", default="box", choices=["box", "reaction", "serve", "single", "requests"],
                    help="reaction: the volume-source reaction-matrix workload (1024 Gauss-Hermite blobs); "
                         "serve: a seeded Poisson stream of box requests through the request server; "
                         "requests: serving-size requests solved alone, FP64 and mixed; "
                         "single: one point charge, eps_in 10, FP64 GMRES to 1e-10 (solve time)")
    ap.add_argument("--requests", type=int, default=32, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=64.0, help="serve: arrival rate (requests/s; 64 saturates one MI355X)")
    # 4 solve slots and 3 setup workers: served throughput of the config-4 stream levels off there
    # (32 requests at 64/s, mixed: 2/2 666-769, 3/2 795, 4/3 835, 6/3 851, 8/4 843 RHS-solves/s)
    ap.add_argument("--slots", type=int, default=4, help="serve: concurrent solve slots")
    ap.add_argument("--setup-workers", type=int, default=3, help="serve: concurrent request setups")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one MI355X; the path shards by RHS "
              "columns or requests across GPUs (no collectives) and is not launched at N > 1", file=sys.stderr)
        sys.exit(2)

    import dbie

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    if a.workload == "reaction":
        reaction_workload(a, dbie, log)
        return
    if a.workload == "serve":
        serve_workload(a, dbie, log)
        return
    if a.workload == "single":
        single_workload(a, dbie, log)
        return
    if a.workload == "requests":
        requests_workload(a, dbie, log)
        return

    t0 = time.perf_counter()
    iface = dbie.single_dielectric_box3d(*BOX, NQ, L_EC, EPS_IN, EPS_OUT)
    op = dbie.batched_lhs_dielectric_box3d_fmm3d_corrected(iface, 1e-9, 1e-9, 8)
    t_setup = time.perf_counter() - t0
    pts, q = charges(a.K, a.seed)
    F = dbie.rhs_point_charges(iface, pts, q, EPS_IN)
    N = iface.n_points
    log(f"setup N={N} blocks={op.n_blocks} {t_setup:.2f}s {op.setup_ms()}")
    if a.mode == "apply":
        X = np.random.default_rng(2).standard_normal((N, a.K))
        res = {}
        for prec in ("fp64", a.precision):
            # let the launch-shape tuner settle (it tries each candidate twice), then time; or pin
            # the low-precision shape given by --shape
            if a.shape and prec == a.precision:
                sp, var = (int(v) for v in a.shape.split(","))
                op.set_tuned_shape(a.K, prec, sp, var)
            for _ in range(96):
                op.apply(X, prec)
                if op.tuned_shape(a.K, prec) is not None:
                    break
            op.apply(X, prec)
            op.profile(True)
            op.profile_read()
            for _ in range(a.reps):
                op.apply(X, prec)
            prof = op.profile_read()
            op.profile(False)
            res[prec] = {k: v[1] / max(v[0], 1) for k, v in prof.items() if v[0]}
            log(f"apply {prec}: per-launch ms {res[prec]} tuned shape {op.tuned_shape(a.K, prec)}")
        # calibration for HBM-traffic counters: one copy kernel of known size (read 1 GiB, write 1 GiB)
        src = dbie.DeviceMatrix(2 ** 24, 8)
        dst = dbie.DeviceMatrix(2 ** 24, 8)
        dst.copy_from(src)
        print(json.dumps({"mode": "apply", "N": N, "K": a.K, "per_launch_ms": res,
                          "shapes": {p: op.tuned_shape(a.K, p) for p in ("fp64", a.precision)},
                          "cross_plane_pairs": cross_plane_pairs(iface), "calibration_copy_bytes": 2 ** 30}),
              flush=True)
        return

    # inputs resident in HBM for every timed solve; results stay on the device
    dF = dbie.DeviceMatrix.from_host(F)
    dS = dbie.DeviceMatrix(N, a.K)
    if a.mode == "headline":  # only the headline form (for kernel timelines): tune, warm up, time
        for prec in ("fp64", a.precision):
            for _ in range(96):
                op.apply(F, prec)
                if op.tuned_shape(a.K, prec) is not None:
                    break
        kw = dict(rtol=1e-10, precision=a.precision, inner_rtol=a.inner_rtol, krylov=a.krylov, recycle=True,
                  verify=False, check_residual=False, precondition=a.precondition)
        for _ in range(a.warmup):
            op.solve_device(dF, dS, **kw)
        op.synchronize()
        lat, st = [], None
        for _ in range(max(a.steps, 1)):
            t1 = time.perf_counter()
            st = op.solve_device(dF, dS, **kw)
            lat.append((time.perf_counter() - t1) * 1e3)
        print(json.dumps({"mode": "headline", "N": N, "K": a.K, "ms": lat, "inner_iters": st["niter"],
                          "refinements": st["nrefine"], "shape": op.tuned_shape(a.K, a.precision)}), flush=True)
        return
    # untimed: let the launch-shape tuner settle for every precision measured (it tries one candidate
    # per apply of a shape class), so that every timed solve runs the chosen shapes
    others = [p for p in ("bf16x3", "fp32") if p != a.precision]
    for prec in ["fp64", a.precision] + others:
        for _ in range(96):
            op.apply(F, prec)  # host result: each apply completes, so its timing is read at the next
            if op.tuned_shape(a.K, prec) is not None:
                break
    log(f"tuned shapes: { {p: op.tuned_shape(a.K, p) for p in ['fp64', a.precision] + others} }")

    # Timed solves run the production path: HIP graphs on, kernel profiling off.  The per-kernel
    # breakdown and the roofline come from a separate profiled pass (eager launches bracketed by
    # HIP events on the operator's stream) after the timed ones.
    def timed(solve_kw, steps, warmup):
        for _ in range(warmup):
            op.solve_device(dF, dS, **solve_kw)
        lat, st = [], None
        op.synchronize()
        t_start = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            st = op.solve_device(dF, dS, **solve_kw)
            lat.append((time.perf_counter() - t1) * 1e3)
        return lat, time.perf_counter() - t_start, st

    # plain FP64 comparator in two Krylov variants: GMRES per column on shared block matvecs, and block
    # GMRES (one block Krylov space, as Krylov.block_gmres); the mixed speed-up is stated against the
    # faster one (median of fp64_steps warm solves)
    # (each with the reference's diagonal scaling and with the near-field preconditioner in one and two
    # sweeps, precondition=2 and 3)
    fp64, S64 = {}, None
    for name, kr, pc in (("columns", "columns", 1), ("block", "block", 1), ("columns_near", "columns", 2),
                         ("block_near", "block", 2), ("columns_near2", "columns", 3), ("block_near2", "block", 3)):
        lat, wall, st64 = timed(dict(rtol=1e-10, precision="fp64", check_residual=False, krylov=kr, precondition=pc),
                                max(a.fp64_steps, 1), 1)
        log(f"fp64 ({name}): iters={st64['niter']} ms {[round(v, 1) for v in lat]}")
        if S64 is None:
            S64 = dS.to_host()
        fp64[name] = {"ms_per_step": float(np.median(lat)), "ms_all": lat, "iters": int(st64["niter"]),
                      "precondition": PRECONDITIONERS[pc]}
    fp64_kr = min(fp64, key=lambda k: fp64[k]["ms_per_step"])
    fp64_ms = fp64[fp64_kr]["ms_per_step"]
    fp64_diag_ms = min(fp64["columns"]["ms_per_step"], fp64["block"]["ms_per_step"])
    iters64 = fp64["columns"]["iters"]

    def col_err(S):
        return np.linalg.norm(S - S64, axis=0) / np.linalg.norm(S64, axis=0)

    # the mixed solve in its Krylov variants (per column with and without recycling the first inner
    # solve's Krylov space in the later refinements, and block GMRES); the fastest is the headline
    # ("unverified": the solve may end on its refined residual estimate, verify=False, as both FP64
    # variants end on their Krylov estimates; "verified": it ends only on an FP64 residual.)  Every
    # variant's true FP64 residual F - A S is evaluated after its timed solves (untimed), and only a
    # variant whose every column meets the tolerance can be the headline.
    def true_residual(S):
        R = F - op.apply(S, "fp64")
        return float((np.linalg.norm(R, axis=0) / np.linalg.norm(F, axis=0)).max())

    mixed = {}
    for name, kr, rec, ver, pc in (("columns", "columns", True, False, 2), ("columns_verified", "columns", True, True, 2),
                                   ("columns_restarted", "columns", False, False, 2), ("block", "block", True, False, 2),
                                   ("columns_diagonal", "columns", True, False, 1),
                                   ("columns_near2", "columns", True, False, 3), ("block_near2", "block", True, False, 3)):
        kw = dict(rtol=1e-10, precision=a.precision, inner_rtol=a.inner_rtol, krylov=kr, recycle=rec, verify=ver,
                  check_residual=False, precondition=pc)
        lat, wall, st = timed(kw, a.steps, a.warmup)
        Sm = dS.to_host()
        mixed[name] = {"lat": lat, "wall": wall, "st": st, "err": col_err(Sm), "kw": kw, "res": true_residual(Sm)}
        log(f"mixed {a.precision} ({name}): iters={st['niter']} refine={st['nrefine']} "
            f"{np.median(lat):.1f} ms/step, true residual {mixed[name]['res']:.2e}")
    ok = [k for k in mixed if mixed[k]["res"] <= 1e-10] or list(mixed)
    mixed_kr = min(ok, key=lambda k: mixed[k]["wall"])
    lat, wall, st, err = (mixed[mixed_kr][k] for k in ("lat", "wall", "st", "err"))
    value = a.K * a.steps / wall
    fp64_value = a.K / (fp64_ms * 1e-3)

    # profiled pass of the headline variant: per-kernel device time with HIP events
    op.profile(True)
    op.profile_read()
    plat, _, _ = timed(mixed[mixed_kr]["kw"], max(a.prof_steps, 1), 0)
    prof = op.profile_read()
    op.profile(False)
    log(f"profiled pass ({mixed_kr}): {np.median(plat):.1f} ms/step; prof={prof}")

    # the other low precisions at the same size (per-column Krylov with the recycled space)
    variants = {}
    for prec in others:
        kw = dict(rtol=1e-10, precision=prec, inner_rtol=a.inner_rtol, krylov="columns", recycle=True, verify=False,
                  check_residual=False, precondition=2)
        vlat, vwall, vst = timed(kw, max(a.variant_steps, 1), 1)
        Sv = dS.to_host()
        verr = col_err(Sv)
        variants[prec] = {"value": a.K * len(vlat) / vwall, "ms_per_step": float(np.median(vlat)),
                          "inner_iters": vst["niter"], "refinements": vst["nrefine"],
                          "max_rel_err_vs_fp64": float(verr.max()), "true_max_rel_residual": true_residual(Sv),
                          "speedup_vs_fp64": a.K * len(vlat) / vwall / fp64_value}
        log(f"mixed {prec}: {variants[prec]}")

    # roofline of the dominant kernel (largest device time in the profiled pass)
    pairs = cross_plane_pairs(iface)
    name, (launches, total_ms) = max(prof.items(), key=lambda kv: kv[1][1])
    avg_ms = total_ms / max(launches, 1)
    if name == "near_combine":
        nb, bd = op.n_blocks, op.block_dim
        bytes_ = nb * bd * bd * 4 + 8 * N * a.K * 3 + 4 * nb  # FP32 blocks, X read twice-ish, Y write
        achieved = bytes_ / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": PEAK["hbm"], "unit": "GB/s",
                "frac": achieved / PEAK["hbm"], "traffic": None, "kernel": name}
    else:
        flops = 2.0 * pairs * a.K
        achieved = flops / (avg_ms * 1e-3) / 1e12
        peak = PEAK["fp64"] if name == "far_fp64" else PEAK["bf16"]
        roof = {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": None, "kernel": ("far_mfma_f64" if name == "far_fp64" else f"far_mfma_{a.precision}"),
                "avg_launch_ms": avg_ms, "launches": launches, "algorithmic_flops_per_launch": flops}
        shp = op.tuned_shape(a.K, a.precision)
        if name == "far_lowprec" and a.precision == "fp16x3" and shp and shp[1] >= 2 and op.stored_far() > 0:
            # the GEMM over the stored kernel matrix streams 4 bytes per cross-plane pair from HBM: its
            # algorithmic bytes are the stored matrix, the split operands (hi/lo fp16 of w.*X) and the
            # FP64 partial sums of its source splits
            nbytes = op.stored_far() + 2 * 2 * N * a.K + 8 * shp[0] * N * a.K
            roof.update({"bound": "hbm", "achieved": nbytes / (avg_ms * 1e-3) / 1e9, "peak": PEAK["hbm"],
                         "unit": "GB/s", "frac": nbytes / (avg_ms * 1e-3) / 1e9 / PEAK["hbm"],
                         "algorithmic_bytes_per_launch": nbytes, "mfma_tflops": achieved,
                         "mfma_frac": achieved / peak, "stored_matrix_bytes": op.stored_far()})
    # HBM bytes per launch of the dominant kernel, from the committed FETCH_SIZE/WRITE_SIZE passes
    # (split-FP16 far field: one entry per launch shape, s<splits>v<variant>; the entry of the shape
    # the tuner chose in this run, if that shape was profiled)
    traffic_file = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        tr = json.load(open(traffic_file)).get(roof.get("kernel"))
        if tr and "bytes" not in tr:
            shp = op.tuned_shape(a.K, "fp64" if roof.get("kernel") == "far_mfma_f64" else a.precision)
            roof["traffic_shape"] = f"s{shp[0]}v{shp[1]}" if shp else None
            tr = tr.get(roof["traffic_shape"]) if shp else None
        if tr and a.K == 256:
            roof["traffic"] = tr["bytes"]
            roof["traffic_fetch_write"] = [tr["fetch_bytes"], tr["write_bytes"]]
    roof["kernels_ms"] = {k: v[1] for k, v in prof.items()}
    roof["kernel_launches"] = {k: v[0] for k, v in prof.items()}
    roof["profiled_ms_per_step"] = float(np.median(plat))

    serve = None
    if not a.no_serve:
        serve = serve_child(a, log)
        log(f"serve: {serve}")

    cpu = None if a.no_cpu else cpu_baseline(iters64)
    out = {
        "metric": METRIC, "value": value, "unit": "RHS-solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": float(np.mean(lat)), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": a.precision, "data": "synthetic",
        "config": {"workload": "single box 2.4x1.6x1.2, l_ec=0.05, n_quad=4, eps_in=4, K=256 point charges, "
                               "block solve to rtol 1e-10 (mixed-precision IR)",
                   "N": N, "K": a.K, "near_blocks": op.n_blocks, "inner_rtol": a.inner_rtol,
                   "parallelism": "single GPU"},
        "timing": {"value": "timed solves with HIP graphs on and kernel profiling off (the product default)",
                   "roofline": f"separate profiled pass of {max(a.prof_steps, 1)} solves (eager launches, HIP events "
                               "around each far-field contraction and near combine on the operator's stream)",
                   "fp64": f"median of {max(a.fp64_steps, 1)} warm solves per Krylov variant"},
        "fp64_value": fp64_value, "fp64_ms_per_step": fp64_ms, "fp64_iters": fp64[fp64_kr]["iters"],
        "fp64_krylov": fp64_kr, "fp64_variants": fp64, "mixed_krylov": mixed_kr,
        "mixed_variants": {k: {"ms_per_step": float(np.mean(v["lat"])), "inner_iters": v["st"]["niter"],
                               "refinements": v["st"]["nrefine"], "value": a.K * a.steps / v["wall"],
                               "max_rel_err_vs_fp64": float(v["err"].max()), "true_max_rel_residual": v["res"],
                               "verify": bool(v["kw"]["verify"]),
                               "precondition": PRECONDITIONERS[v["kw"]["precondition"]]}
                           for k, v in mixed.items()},
        "speedup_vs_fp64_diagonal": value / (a.K / (fp64_diag_ms * 1e-3)),
        "other_precisions": variants,
        "speedup_vs_fp64": value / fp64_value,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "max_rel_err_vs_fp64": float(err.max()), "median_rel_err_vs_fp64": float(np.median(err)),
        "mixed_inner_iters": st["niter"], "mixed_refinements": st["nrefine"],
        "mixed_max_rel_residual": mixed[mixed_kr]["res"], "mixed_residual_note": "true FP64 residual max_k "
        "||F_k - A S_k|| / ||F_k|| of the headline variant's result, evaluated after the timed solves",
        "setup_s": t_setup, "setup_ms": op.setup_ms(),
        "roofline": roof, "cpu_baseline": cpu, "serve": serve,
        "tuned_shapes": {p: op.tuned_shape(a.K, p) for p in ["fp64", a.precision] + others},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
