#!/usr/bin/env python3
"""Benchmark: Jacobi-PCG solves of the 2236 x 2236 LSCM system on one MI355X.

Workload (the large single-system configuration): LSCM normal equations of a
2236 x 2236 height-field grid (n = 9,999,388 unknowns, nnz = 139,919,888,
FP64 CRS), right-hand side from two pinned corners, x0 = 0, stopping rule
|Ax-b| <= 1e-10 |b| or 5000 iterations (the iteration cap of the 317 x 317
configuration of the same construction).  One step = one complete solve with
the inputs resident in HBM.  Synthetic data from a fixed seed.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_PROBE_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_cg_spmv_lscm2236_traffic.md")


def measured_traffic(workload: str):
    """HBM bytes per k_cg_spmv launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a plain copy), when
    it covers this workload; None otherwise."""
    try:
        with open(TRAFFIC_FILE) as f:
            kv = dict(line.split(":", 1) for line in f if line.startswith("- ") and ":" in line)
    except OSError:
        return None
    kv = {k[2:].strip(): v.strip() for k, v in kv.items()}
    if kv.get("workload") != workload:
        return None
    try:
        return float(kv["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(n: int, nnz: int, sv: int = 8, sx: int = 8) -> float:
    """Algorithmic bytes of y = A x: values + colind + rowptr + x once + y."""
    return nnz * (sv + 4) + (n + 1) * 4 + n * sx + n * sx


def cg_direction_bytes(n: int, s: int = 8) -> float:
    """x -= alpha d; d = beta d + h: x, d and h read, x and d written."""
    return 5 * n * s


def cg_update_bytes(n: int, s: int = 8) -> float:
    """r -= alpha q; h = D^-1 r: r, q, dinv read; r, h written."""
    return 5 * n * s


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def lscm_rows_arrays(mesh):
    """Rows of the LSCM least-squares problem, as the context receives them:
    per valid triangle a real and an imaginary row over (u0, v0, u1, v1, u2, v2)."""
    coeffs, ok = mesh.lscm_rows()
    tri = mesh.triangles[ok]
    vars6 = (2 * tri[:, [0, 0, 1, 1, 2, 2]] + np.array([0, 1, 0, 1, 0, 1])).astype(np.uint32)
    c = coeffs[ok]
    var = np.stack([vars6, vars6], axis=1).reshape(-1)
    coef = np.stack([c[:, :6], c[:, 6:]], axis=1).reshape(-1)
    rp = (6 * np.arange(2 * len(tri) + 1, dtype=np.uint64)).astype(np.uint32)
    return rp, var, coef


def assembly_measurement(args, Sw, capi, W):
    k = args.k
    mesh = W.grid_mesh(k, seed=args.seed)
    pins, uv = (0, k * k - 1), (0.0, 0.0, 1.0, 1.0)     # as lscm_grid
    rp, var, coef = lscm_rows_arrays(mesh)
    nrows = len(rp) - 1
    nvars = 2 * mesh.nv
    locked = np.zeros(nvars, dtype=np.uint8)
    values = np.zeros(nvars)
    for p, (u, v) in zip(pins, [(uv[0], uv[1]), (uv[2], uv[3])]):
        locked[2 * p] = locked[2 * p + 1] = 1
        values[2 * p], values[2 * p + 1] = u, v
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        Sa, ba = capi.System.least_squares(nvars, rp, var, coef, locked=locked, values=values)
        times.append(time.perf_counter() - t0)
        if _ == 0:
            rpa, cia, vaa = Sa.crs()
            same = (np.array_equal(rpa, Sw.rowptr) and np.array_equal(cia, Sw.colind)
                    and np.array_equal(vaa.view(np.uint64), Sw.val.view(np.uint64))
                    and np.array_equal(ba[0].view(np.uint64), Sw.b.view(np.uint64)))
            del rpa, cia, vaa
        Sa.close()
    t_dev = min(times)
    # the whole LSCM system from the mesh on the device (rows computed there)
    times_m = []
    for _ in range(2):
        t0 = time.perf_counter()
        Sm, bm = capi.System.lscm(mesh.points, mesh.triangles, pins, uv)
        times_m.append(time.perf_counter() - t0)
        if _ == 0:
            rpa, cia, vaa = Sm.crs()
            same_m = (np.array_equal(rpa, Sw.rowptr) and np.array_equal(cia, Sw.colind)
                      and np.array_equal(vaa.view(np.uint64), Sw.val.view(np.uint64))
                      and np.array_equal(bm.view(np.uint64), Sw.b.view(np.uint64)))
            del rpa, cia, vaa
        Sm.close()
    t0 = time.perf_counter()
    W.lscm_system(mesh, pins=pins, pin_uv=uv)
    t_gather = time.perf_counter() - t0
    res = {"rows": nrows, "entries": int(rp[-1]), "n": Sw.n, "nnz": Sw.nnz,
           "device_s": t_dev, "bitwise_equal_host_assembly": bool(same),
           "device_from_mesh_s": min(times_m), "from_mesh_bitwise_equal": bool(same_m),
           "host_gather_s": t_gather,
           "host_gather_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    log(f"assembly: device {t_dev:.3f}s (rows in from host, CRS resident), bitwise = host "
        f"assembly: {same}; from the mesh {min(times_m):.3f}s ({same_m}); host gather "
        f"{t_gather:.2f}s")
    if args.cpu_iters > 0 and args.assembly_sample_rows > 0:
        from oracle import oracle as O
        ns = min(nrows, args.assembly_sample_rows)
        ctx = O.Context(nvars, least_squares=True)
        for p, (u, v) in zip(pins, [(uv[0], uv[1]), (uv[2], uv[3])]):
            ctx.set_variable(2 * p, u); ctx.set_variable(2 * p + 1, v)
            ctx.lock_variable(2 * p); ctx.lock_variable(2 * p + 1)
        ctx.begin_matrix()
        t0 = time.perf_counter()
        ctx.rows(rp[:ns + 1], var[:int(rp[ns])], coef[:int(rp[ns])])
        t_s = time.perf_counter() - t0
        del ctx
        res["cpu_oracle_s"] = t_s * nrows / ns
        res["cpu_oracle_sample"] = (f"first {ns} of {nrows} rows through the oracle's "
                                    f"least-squares context (OpenNL's row-by-row "
                                    f"accumulation, one core): {t_s:.2f}s, scaled by rows; "
                                    f"the compression to CRS is not counted")
        log(f"assembly cpu (oracle context, 1 core): {t_s:.2f}s for {ns} rows -> "
            f"{res['cpu_oracle_s']:.1f}s for {nrows}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--k", type=int, default=2236, help="grid side")
    ap.add_argument("--max-iter", type=int, default=5000)
    ap.add_argument("--eps", type=float, default=1e-10)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-iters", type=int, default=40,
                    help="oracle PCG iterations timed for the CPU baseline (0: skip)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--workgroups", type=int, default=0,
                    help="SpMV workgroups (with --rows-per-wave: a fixed launch shape)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape instead of the autotuner's")
    ap.add_argument("--no-assembly", action="store_true",
                    help="skip the least-squares assembly measurement")
    ap.add_argument("--assembly-sample-rows", type=int, default=2_000_000,
                    help="rows of the CPU (oracle) assembly sample")
    ap.add_argument("--rows-per-wave", type=int, default=0,
                    help="SpMV engine code for the C2 system (0: the library default)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only launch the kernels whose HBM counters are read under "
                         "rocprofv3 --pmc: a 1 GiB plain copy (x3) and 10 PCG iterations")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the "
              "solver has no multi-GPU path in this build", file=sys.stderr)
        sys.exit(2)

    from graphitethree_amd import capi
    from graphitethree_amd import workloads as W

    L = capi.lib()
    if not L.nlInitExtension(b"HIP"):
        raise SystemExit("nlInitExtension(\"HIP\") failed: " + capi.last_error())
    dev_name, cu, hbm = capi.device_info()
    log(f"device: {dev_name}, {cu} CUs, {hbm / 2**30:.0f} GiB")

    t0 = time.time()
    Sw = W.lscm_grid(args.k, seed=args.seed)
    n, nnz = Sw.n, Sw.nnz
    log(f"LSCM grid {args.k}^2: n={n} nnz={nnz} built in {time.time() - t0:.1f}s")
    t0 = time.time()
    S = capi.System.from_crs(Sw, fp32=not args.no_mixed)
    tuned = None
    if args.rows_per_wave or args.workgroups:
        S.set_launch(args.rows_per_wave or S.launch()[0], args.workgroups or S.launch()[1])
    elif not args.no_autotune and not args.traffic_probe:
        tuned = S.autotune(iters=20)
        log(f"autotuned: engine {tuned[0]}, {tuned[1]} workgroups, "
            f"{tuned[2] * 1e3:.3f} ms per CG iteration")
    lpr, wg = S.launch()
    log(f"uploaded in {time.time() - t0:.1f}s; launch rows/wavefront={lpr} workgroups={wg}")

    if args.traffic_probe:
        t_copy = capi.bench_copy(COPY_PROBE_BYTES, reps=3)
        t_dir, t_a, t_upd = capi.profile_cg(S, iters=10)
        print(json.dumps({"probe": "traffic", "copy_bytes": COPY_PROBE_BYTES, "copy_s": t_copy,
                          "k_cg_spmv_s": t_a, "spmv_bytes": spmv_bytes(n, nnz)}), flush=True)
        return

    # ---- plain copy: the bandwidth this device sustains for a streaming kernel
    t_copy = capi.bench_copy(2 << 30, reps=20)
    copy_gbs = 2 * (2 << 30) / t_copy / 1e9
    log(f"copy 2 GiB: {t_copy * 1e3:.3f} ms  {copy_gbs:.0f} GB/s (read + write)")

    # ---- standalone SpMV bandwidth (FP64, and FP32 values)
    t_spmv = S.bench_spmv(fp32=False, reps=50)
    spmv_gbs = spmv_bytes(n, nnz) / t_spmv / 1e9
    log(f"SpMV fp64: {t_spmv * 1e3:.3f} ms  {spmv_gbs:.0f} GB/s")
    t_spmv32 = spmv32_gbs = None
    if not args.no_mixed:
        t_spmv32 = S.bench_spmv(fp32=True, reps=50)
        spmv32_gbs = spmv_bytes(n, nnz, 4, 4) / t_spmv32 / 1e9
        log(f"SpMV fp32: {t_spmv32 * 1e3:.3f} ms  {spmv32_gbs:.0f} GB/s")

    # ---- warmup + timed solves (FP64 Jacobi-PCG, the plain GPU CG)
    b = Sw.b
    params = dict(eps=args.eps, max_iter=args.max_iter)
    for w in range(args.warmup):
        x, st = S.solve(b, **params)
        log(f"warmup {w}: {st.iterations} it, rel res {st.rel_residual:.3e}, "
            f"{st.device_seconds:.3f}s")
    lat = []
    stats = None
    t_start = time.perf_counter()
    for k in range(args.steps):
        x, st = S.solve(b, **params)
        lat.append(st.device_seconds)
        stats = st
        log(f"step {k}: {st.iterations} it, rel res {st.rel_residual:.3e}, "
            f"dev {st.device_seconds:.3f}s wall {st.wall_seconds:.3f}s")
    t_total = time.perf_counter() - t_start
    ms_per_step = 1e3 * float(np.mean(lat))
    solves_per_s = 1e3 / ms_per_step
    iters = int(stats.iterations)

    # ---- per-kernel timing of one PCG iteration (HIP events, system stream)
    t_dir, t_a, t_upd = capi.profile_cg(S, iters=50)
    bytes_a = spmv_bytes(n, nnz)          # q = A d (+ d.q from the same d)
    achieved = bytes_a / t_a / 1e9
    log(f"cg direction {t_dir * 1e3:.3f} ms ({cg_direction_bytes(n) / t_dir / 1e9:.0f} GB/s), "
        f"spmv {t_a * 1e3:.3f} ms ({achieved:.0f} GB/s), update {t_upd * 1e3:.3f} ms "
        f"({cg_update_bytes(n) / t_upd / 1e9:.0f} GB/s)")
    iter_bytes = bytes_a + cg_direction_bytes(n) + cg_update_bytes(n)
    iter_gbs = iters * iter_bytes / (ms_per_step * 1e-3) / 1e9 if iters else 0.0

    result_extra = {}
    # ---- mixed precision (FP32 inner PCG inside FP64 refinement), same budget
    if not args.no_mixed:
        xm, sm = S.solve(b, precision="mixed", eps=args.eps, max_iter=args.max_iter,
                         inner_eps=1e-4)
        result_extra["mixed"] = {
            "solves_per_s": 1.0 / sm.device_seconds, "iterations": int(sm.iterations),
            "outer": int(sm.outer_iterations), "rel_residual": sm.rel_residual,
            "rel_diff_vs_fp64": float(np.linalg.norm(xm - x) / np.linalg.norm(x))}
        log(f"mixed: {sm.iterations} inner it / {sm.outer_iterations} outer, rel res "
            f"{sm.rel_residual:.3e}, {sm.device_seconds:.3f}s")
        del xm
    del x

    # ---- CPU baseline: the oracle's OpenNL-order PCG on the same matrix
    cpu = None
    if args.cpu_iters > 0:
        from oracle import oracle as O
        O.build() if not os.path.exists(O._LIB_PATH) else None
        A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        secs = O.pcg_timed_iters(A, b, args.cpu_iters, variant=1)
        per_iter = secs / args.cpu_iters
        cpu_solve_s = per_iter * max(iters, 1)
        secs_all = O.pcg_timed_iters(A, b, args.cpu_iters, variant=2)
        cpu = {"value": 1.0 / cpu_solve_s, "unit": "solves/s", "cores": threads,
               "kind": "port",
               "sample": f"{args.cpu_iters} Jacobi-PCG iterations of the oracle (OpenMP "
                         f"SpMV, serial BLAS-1, OpenNL operation order) on this system, "
                         f"{secs:.1f}s, scaled to the {iters}-iteration solve",
               "all_openmp_value": 1.0 / (secs_all / args.cpu_iters * max(iters, 1)),
               "cpu_model": cpu_model(), "nproc": os.cpu_count(),
               "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "")}
        log(f"cpu baseline: {per_iter * 1e3:.1f} ms/iter on {threads} threads -> "
            f"{cpu['value']:.4f} solves/s (all-OpenMP {secs_all / args.cpu_iters * 1e3:.1f} "
            f"ms/iter)")
        del A

    # ---- C1 (317 x 317, cache-resident) to full convergence, no cap
    if not args.no_c1:
        S1w = W.lscm_grid(317, seed=args.seed)
        S1 = capi.System.from_crs(S1w)
        x1, s1 = S1.solve(S1w.b, eps=1e-10, max_iter=100000)
        l1 = []
        for _ in range(3):
            x1, s1 = S1.solve(S1w.b, eps=1e-10, max_iter=100000)
            l1.append(s1.device_seconds)
        x1c, s1c = S1.solve(S1w.b, eps=1e-10, max_iter=5000)
        result_extra["lscm317"] = {
            "n": S1w.n, "nnz": S1w.nnz, "iterations_to_1e-10": int(s1.iterations),
            "rel_residual": s1.rel_residual, "solves_per_s_to_1e-10": 1.0 / float(np.median(l1)),
            "capped_5000": {"solves_per_s": 1.0 / s1c.device_seconds,
                            "rel_residual": s1c.rel_residual}}
        log(f"317^2: {s1.iterations} it to {s1.rel_residual:.2e} in "
            f"{np.median(l1):.3f}s; capped 5000: {s1c.device_seconds:.3f}s")
        S1.close()

    # ---- least-squares assembly of the same system from its LSCM rows:
    # device (nlhipLeastSquaresCreate, host rows in, resident CRS out) against
    # the oracle's row-by-row context (OpenNL's algorithm, one core) on a
    # sample of the rows, and the multithreaded host gather that built Sw
    if not args.no_assembly:
        result_extra["assembly"] = assembly_measurement(args, Sw, capi, W)

    out = {
        "metric": "CG solves/s + SpMV HBM GB/s at 1e-10 rel. residual; p50/p99 solve latency",
        "value": solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded LSCM height-field grid)",
        "config": {"workload": f"lscm_grid_{args.k}", "n": n, "nnz": nnz,
                   "solver": "Jacobi-PCG fp64", "eps": args.eps, "max_iter": args.max_iter,
                   "iterations": iters, "converged": bool(stats.converged),
                   "rel_residual": stats.rel_residual, "rows_per_wave": lpr,
                   "workgroups": wg, "autotuned": tuned is not None,
                   "step": "one Jacobi-PCG solve from x0 = 0 with the stopping rule "
                           "|Ax - b| <= eps |b| or max_iter iterations (the 317 x 317 "
                           "configuration's cap; this system needs far more to reach eps, "
                           "so every step runs the full max_iter iterations)"},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "spmv_gbs": {"fp64": spmv_gbs, "fp32": spmv32_gbs,
                     "bytes_fp64": spmv_bytes(n, nnz)},
        "iteration_gbs": iter_gbs,
        "roofline": {"bound": "hbm", "kernel": "k_cg_spmv", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"lscm_grid_{args.k}"),
                     "bytes_per_launch": bytes_a, "launch_ms": t_a * 1e3,
                     "copy_gbs": copy_gbs},
        "kernels_ms": {"k_cg_dir": t_dir * 1e3, "k_cg_spmv": t_a * 1e3,
                       "k_cg_update": t_upd * 1e3},
        "cpu_baseline": cpu,
        "device": dev_name,
        **result_extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
