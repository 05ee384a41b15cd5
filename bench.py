#!/usr/bin/env python3
"""Benchmark: Jacobi-PCG solves of the 2236 x 2236 LSCM system on one MI355X.

Workload (the large single-system configuration): LSCM normal equations of a
2236 x 2236 height-field grid (n = 9,999,388 unknowns, nnz = 139,919,888,
FP64 CRS), right-hand side from two pinned corners, x0 = 0, solved with
FP64 Jacobi-PCG to a true relative residual |b - Ax| / |b| <= 1e-10
(recomputed in FP64; the solve confirms it, restarting from its iterate when
rounding has made the recursive residual drift below the true one).  One
such solve takes about 700k iterations (minutes), so one step is a slice of
it: the K timed steps are consecutive slices of ceil(max_iter / K)
iterations of ONE solve, which finishes inside them, and `value` is that
solve's rate (1 / the timed seconds).  Warmup steps are slices of 2000
iterations of a separate solve that is thrown away.  The inputs are resident
in HBM before the timed region (the solve is opened, b staged and the
initial residual formed, before it; x is read back after it).  Synthetic
data from a fixed seed.

Per result it also reports the true residual recomputed by the oracle's SpMV
on the host, and for the 317 x 317 system of the same construction: the GPU
solve, the oracle's converged solve on the host cores (timed), their relative
difference and the forward error of both against a sparse direct solve; and
the batched configuration: 2000 LSCM charts solved to 1e-10 as one batch (FP64,
and FP32-inner refinement) against one GPU solve per chart.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_PROBE_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_cg_spmv_lscm2236_traffic.md")


def measured_traffic(workload: str, engine: int, workgroups: int):
    """HBM bytes per k_cg_spmv launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the same counters on a plain copy), for
    the launch shape this run uses (engine code, workgroups), when the summary
    covers this workload and that shape; None otherwise (a summary of another
    shape says nothing of this one).  Engine code 64 is the library's default
    stream engine, which runs as code 73 on a matrix with the 2x2 block form
    (every LSCM grid)."""
    try:
        with open(TRAFFIC_FILE) as f:
            kv = dict(line.split(":", 1) for line in f if line.startswith("- ") and ":" in line)
    except OSError:
        return None
    kv = {k[2:].strip(): v.strip() for k, v in kv.items()}
    if kv.get("workload") != workload:
        return None
    code = 73 if engine == 64 else engine
    try:
        return float(kv[f"traffic[{code},{workgroups}]"])
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


def atlas_measurement(args, capi, W):
    """The batched configuration beside the headline: 2000 disk-like LSCM
    charts (N ~ logU[50, 50000] vertices, seed 1) solved to 1e-10 as one
    batch in FP64 and with FP32 inner PCG inside FP64 refinement, against
    the plain FP64 GPU PCG one chart at a time (every 10th chart timed,
    scaled by 10)."""
    t0 = time.perf_counter()
    A = W.atlas(2000, seed=1, nmin=50, nmax=50000)
    t_build = time.perf_counter() - t0
    B = capi.Batch(packed=(A.row_off, A.rowptr, A.colind, A.val), fp32=True)
    wg, passes, _ = B.autotune(A.b, iters=50)
    B.solve(A.b, eps=1e-10, max_iter=50000)
    x64, p64, t64 = B.solve(A.b, eps=1e-10, max_iter=50000)
    xm, pm, tm = B.solve(A.b, precision="mixed", eps=1e-10, max_iter=200000, inner_eps=1e-4)
    xs = B.split(x64)
    t_plain, worst, stride = 0.0, 0.0, 10
    for c in range(0, A.count, stride):
        S = A.chart(c)
        Sc = capi.System.from_crs(S)
        Sc.solve(S.b, eps=1e-10, max_iter=50000)
        xc, st = Sc.solve(S.b, eps=1e-10, max_iter=50000)
        t_plain += st.device_seconds
        worst = max(worst, float(np.linalg.norm(xc - xs[c]) / np.linalg.norm(xc)))
        Sc.close()
    t_plain *= stride
    B.close()
    res = {"charts": A.count, "n": A.n, "nnz": A.nnz, "host_build_s": t_build,
           "workgroups": wg, "spmv_passes": passes,
           "batched_fp64": {"seconds": t64.device_seconds, "solves_per_s": A.count / t64.device_seconds,
                            "converged": int(sum(p.converged for p in p64)),
                            "max_iterations": int(t64.iterations)},
           "batched_mixed": {"seconds": tm.device_seconds, "solves_per_s": A.count / tm.device_seconds,
                             "converged": int(sum(p.converged for p in pm)),
                             "max_inner_iterations": int(tm.iterations)},
           "per_chart_fp64": {"seconds": t_plain, "solves_per_s": A.count / t_plain,
                              "sample": f"every {stride}th chart timed, scaled by {stride}",
                              "max_rel_diff_vs_batched": worst},
           "speedup_batched_fp64": t_plain / t64.device_seconds,
           "speedup_batched_mixed": t_plain / tm.device_seconds}
    log(f"atlas: batched fp64 {t64.device_seconds:.2f}s, batched mixed {tm.device_seconds:.2f}s, "
        f"per chart fp64 {t_plain:.1f}s -> {res['speedup_batched_fp64']:.2f}x / "
        f"{res['speedup_batched_mixed']:.2f}x")
    return res


def cpu_leg_main(args):
    """The host-core baseline, run in a child process of its own so that the
    OpenMP runtime starts with the thread binding under test (it reads
    OMP_PROC_BIND once).  No GPU call.  Prints one JSON line."""
    from graphitethree_amd import workloads as W
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {"threads": threads, "omp_proc_bind": os.environ.get("OMP_PROC_BIND", "")}
    if args.cpu_leg == "c2":
        Sw = W.lscm_grid(args.k, seed=args.seed)
        A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
        it1 = args.cpu_iters
        out["per_iter_s"] = O.pcg_timed_iters(A, Sw.b, it1, variant=1) / it1
        it2 = max(it1 // 2, 1)
        out["per_iter_all_openmp_s"] = O.pcg_timed_iters(A, Sw.b, it2, variant=2) / it2
        out["iters"] = [it1, it2]
    else:
        Sw = W.lscm_grid(317, seed=args.seed)
        A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
        t0 = time.perf_counter()
        xc, itc, _ = O.pcg(A, Sw.b, eps=1e-10, max_iter=100000, variant=1)
        out["seconds"] = time.perf_counter() - t0
        out["iterations"] = int(itc)
        np.save(args.cpu_out, xc)
    print(json.dumps(out), flush=True)


def run_cpu_leg(args, leg: str, bind: str | None, out_path: str = "") -> dict:
    """bench.py --cpu-leg in a child process with OMP_PROC_BIND = bind (None:
    unset, the OpenMP runtime's default)."""
    env = dict(os.environ)
    env.pop("OMP_PROC_BIND", None)
    if bind is not None:
        env["OMP_PROC_BIND"] = bind
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", leg, "--k", str(args.k),
           "--seed", str(args.seed), "--cpu-iters", str(args.cpu_iters)]
    if out_path:
        cmd += ["--cpu-out", out_path]
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, text=True, timeout=900)
    if r.returncode != 0:
        raise SystemExit(f"CPU baseline leg {leg} failed with status {r.returncode}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def c1_measurement(args, capi, W, O, bind):
    """The 317 x 317 system of the same construction, whole solves on both
    sides: GPU FP64 Jacobi-PCG (latency over repeated solves), the oracle's
    converged solve on the host cores, a sparse direct solve, and the
    per-result parity fields."""
    import scipy.sparse.linalg as sla
    Sw = W.lscm_grid(317, seed=args.seed)
    A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
    nb = float(np.linalg.norm(Sw.b))
    S = capi.System.from_crs(Sw)
    S.solve(Sw.b, eps=1e-10, max_iter=100000)
    lat = []
    for _ in range(10):
        x, st = S.solve(Sw.b, eps=1e-10, max_iter=100000)
        lat.append(st.device_seconds)
    S.close()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "x_cpu.npy")
        leg = run_cpu_leg(args, "c1", bind, path)
        xc = np.load(path)
    itc, t_cpu = leg["iterations"], leg["seconds"]
    t0 = time.perf_counter()
    xd = sla.spsolve(A.to_scipy().tocsc(), Sw.b)
    t_direct = time.perf_counter() - t0
    nd = float(np.linalg.norm(xd))
    res = {
        "n": Sw.n, "nnz": Sw.nnz, "iterations": int(st.iterations),
        "iterations_opennl_rule": int(st.iterations - st.extra_iterations),
        "converged": bool(st.converged),
        "true_rel_residual": float(np.linalg.norm(Sw.b - O.spmv(A, x, parallel=True)) / nb),
        "solves_per_s": 1.0 / float(np.median(lat)),
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "rel_diff_vs_cpu": float(np.linalg.norm(x - xc) / np.linalg.norm(xc)),
        "forward_error": float(np.linalg.norm(x - xd) / nd),
        "forward_error_reference": "scipy SuperLU direct solve of the same matrix",
        "cpu": {"iterations": itc, "seconds": t_cpu,
                "true_rel_residual": float(np.linalg.norm(Sw.b - O.spmv(A, xc, parallel=True)) / nb),
                "forward_error": float(np.linalg.norm(xc - xd) / nd),
                "threads": leg["threads"], "omp_proc_bind": leg["omp_proc_bind"] or "unset",
                "what": "the oracle's Jacobi-PCG (OpenMP SpMV, serial BLAS-1, OpenNL order) "
                        "from x0 = 0 to its recursive 1e-10, whole solve"},
        "direct_seconds": t_direct,
    }
    log(f"317^2: GPU {st.iterations} it ({res['iterations_opennl_rule']} to the recursive test), "
        f"true res {res['true_rel_residual']:.3e}, {np.median(lat) * 1e3:.1f} ms; CPU {itc} it "
        f"in {t_cpu:.1f}s; rel diff {res['rel_diff_vs_cpu']:.2e}; forward error GPU "
        f"{res['forward_error']:.2e} CPU {res['cpu']['forward_error']:.2e}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--k", type=int, default=2236, help="grid side")
    ap.add_argument("--max-iter", type=int, default=1_000_000,
                    help="iteration budget of the solve (the timed slices cover it)")
    ap.add_argument("--eps", type=float, default=1e-10)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--warmup-iters", type=int, default=2000,
                    help="iterations per warmup slice")
    ap.add_argument("--time-budget", type=float, default=450.0,
                    help="seconds after which the timed solve is cut short and reported "
                         "as not converged")
    ap.add_argument("--wall-budget", type=float, default=540.0,
                    help="seconds of the whole run within which the optional legs (CPU "
                         "baseline, 317 x 317, assembly, atlas) must fit; one that would "
                         "not is skipped")
    ap.add_argument("--cpu-iters", type=int, default=200,
                    help="oracle PCG iterations timed for the CPU baseline (0: skip)")
    ap.add_argument("--cpu-leg", choices=["c2", "c1"], default=None,
                    help=argparse.SUPPRESS)   # child process of the CPU baseline
    ap.add_argument("--cpu-out", default="", help=argparse.SUPPRESS)
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--workgroups", type=int, default=0,
                    help="SpMV workgroups (with --rows-per-wave: a fixed launch shape)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape instead of the autotuner's")
    ap.add_argument("--no-atlas", action="store_true",
                    help="skip the batched 2000-chart atlas measurement")
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

    if args.cpu_leg:
        cpu_leg_main(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the "
              "solver has no multi-GPU path in this build", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        raise SystemExit("--steps must be at least 1")

    from graphitethree_amd import capi
    from graphitethree_amd import workloads as W

    L = capi.lib()
    if not L.nlInitExtension(b"HIP"):
        raise SystemExit("nlInitExtension(\"HIP\") failed: " + capi.last_error())
    dev_name, cu, hbm = capi.device_info()
    log(f"device: {dev_name}, {cu} CUs, {hbm / 2**30:.0f} GiB")
    t_begin = time.time()

    t0 = time.time()
    Sw = W.lscm_grid(args.k, seed=args.seed)
    n, nnz = Sw.n, Sw.nnz
    log(f"LSCM grid {args.k}^2: n={n} nnz={nnz} built in {time.time() - t0:.1f}s")
    t0 = time.time()
    S = capi.System.from_crs(Sw, fp32=not args.traffic_probe)
    tuned = None
    if args.rows_per_wave or args.workgroups:
        S.set_launch(args.rows_per_wave or S.launch()[0], args.workgroups or S.launch()[1])
    elif not args.no_autotune:
        tuned = S.autotune(iters=20)
        log(f"autotuned: engine {tuned[0]}, {tuned[1]} workgroups, "
            f"{tuned[2] * 1e3:.3f} ms per CG iteration")
    lpr, wg = S.launch()
    log(f"uploaded in {time.time() - t0:.1f}s; launch rows/wavefront={lpr} workgroups={wg}")

    if args.traffic_probe:
        t_copy = capi.bench_copy(COPY_PROBE_BYTES, reps=3)
        t_dir, t_a, t_upd = capi.profile_cg(S, iters=10)
        print(json.dumps({"probe": "traffic", "copy_bytes": COPY_PROBE_BYTES, "copy_s": t_copy,
                          "k_cg_spmv_s": t_a, "spmv_bytes": spmv_bytes(n, nnz),
                          "engine": lpr, "workgroups": wg}), flush=True)
        return

    # ---- plain copy: the bandwidth this device sustains for a streaming kernel
    t_copy = capi.bench_copy(2 << 30, reps=20)
    copy_gbs = 2 * (2 << 30) / t_copy / 1e9
    log(f"copy 2 GiB: {t_copy * 1e3:.3f} ms  {copy_gbs:.0f} GB/s (read + write)")
    t_read = capi.bench_read(4 << 30, reps=20)
    read_gbs = (4 << 30) / t_read / 1e9
    log(f"read 4 GiB: {t_read * 1e3:.3f} ms  {read_gbs:.0f} GB/s")
    # the 2x2-block SpMV's traffic without its arithmetic: this system's
    # block count and blocks per row pair, a 16-byte read and write per pair
    mix_blocks = nnz // 4
    mix_bpr = max(1, round(mix_blocks / max(1, n // 2)))
    t_mix = capi.bench_stream_mix(mix_blocks, mix_bpr, writes=True, reps=20)
    t_mix_ro = capi.bench_stream_mix(mix_blocks, mix_bpr, writes=False, reps=20)
    mix_bytes = 34.0 * mix_blocks + 32.0 * (mix_blocks // mix_bpr)
    log(f"SpMV-shaped stream ({mix_blocks} blocks, {mix_bpr} per row pair): "
        f"{t_mix * 1e3:.3f} ms  {mix_bytes / t_mix / 1e9:.0f} GB/s; without its writes "
        f"{t_mix_ro * 1e3:.3f} ms")

    # ---- standalone SpMV bandwidth (FP64, and FP32 values)
    t_spmv = S.bench_spmv(fp32=False, reps=50)
    spmv_gbs = spmv_bytes(n, nnz) / t_spmv / 1e9
    log(f"SpMV fp64: {t_spmv * 1e3:.3f} ms  {spmv_gbs:.0f} GB/s")
    t_spmv32 = S.bench_spmv(fp32=True, reps=50)
    spmv32_gbs = spmv_bytes(n, nnz, 4, 4) / t_spmv32 / 1e9
    log(f"SpMV fp32: {t_spmv32 * 1e3:.3f} ms  {spmv32_gbs:.0f} GB/s")

    # ---- warmup: slices of a solve that is thrown away
    b = Sw.b
    params = dict(eps=args.eps, max_iter=args.max_iter)
    if args.warmup > 0:
        S.solve_begin(b, **params)
        for w in range(args.warmup):
            _, st = S.solve_advance(args.warmup_iters)
            log(f"warmup {w}: {st.iterations} it, recursive rel res {st.cg_rel_residual:.3e}")
        S.solve_end()

    # ---- timed: K consecutive slices of one solve from x0 = 0
    K = args.steps
    slice_iters = -(-args.max_iter // K)
    sub = 20000                      # iterations per call: a progress line each ~10 s
    S.solve_begin(b, **params)       # b staged, initial residual formed
    step_s = []
    finished, st = False, None
    cut_short = False
    t_start = time.perf_counter()
    done_iters = 0
    for k in range(K):
        ts = time.perf_counter()
        end = min((k + 1) * slice_iters, args.max_iter)
        while not finished and done_iters < end:
            if time.time() - t_begin > args.time_budget:
                cut_short = True
                break
            finished, st = S.solve_advance(min(sub, end - done_iters))
            if int(st.iterations) <= done_iters and not finished:
                raise SystemExit("the solve made no progress")
            done_iters = int(st.iterations)
            log(f"step {k}: {st.iterations} it, recursive rel res {st.cg_rel_residual:.3e}, "
                f"true {st.rel_residual:.3e}, {time.perf_counter() - t_start:.1f}s")
        step_s.append(time.perf_counter() - ts)
        if cut_short:
            break
    t_total = time.perf_counter() - t_start
    x, st = S.solve_end()
    ms_per_step = 1e3 * t_total / K
    converged = bool(st.converged) and not cut_short
    solves_per_s = 1.0 / t_total if converged else 0.0
    iters = int(st.iterations)
    log(f"solve: {iters} it ({iters - st.extra_iterations} to the recursive test, "
        f"{st.outer_iterations} restarts), true rel res {st.rel_residual:.3e}, "
        f"converged {converged}, {t_total:.1f}s (device {st.device_seconds:.1f}s)")

    from oracle import oracle as O
    O.build() if not os.path.exists(O._LIB_PATH) else None
    A = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
    true_oracle = float(np.linalg.norm(b - O.spmv(A, x, parallel=True)) / np.linalg.norm(b))
    log(f"true rel res by the oracle's SpMV: {true_oracle:.3e}")
    del x

    # ---- per-kernel timing of one PCG iteration (HIP events, system stream)
    t_dir, t_a, t_upd = capi.profile_cg(S, iters=50)
    bytes_a = spmv_bytes(n, nnz)          # q = A d (+ d.q from the same d)
    achieved = bytes_a / t_a / 1e9
    log(f"cg direction {t_dir * 1e3:.3f} ms ({cg_direction_bytes(n) / t_dir / 1e9:.0f} GB/s), "
        f"spmv {t_a * 1e3:.3f} ms ({achieved:.0f} GB/s), update {t_upd * 1e3:.3f} ms "
        f"({cg_update_bytes(n) / t_upd / 1e9:.0f} GB/s)")
    iter_bytes = bytes_a + cg_direction_bytes(n) + cg_update_bytes(n)
    iter_gbs = iters * iter_bytes / t_total / 1e9 if iters else 0.0
    S.close()

    result_extra = {}
    # ---- CPU baseline: the oracle's OpenNL-order PCG on the same matrix, in
    # child processes with the threads bound close and unbound; the faster
    # binding is the baseline
    cpu, bind = None, None
    skipped = []

    def fits(name, seconds):
        """Whether an optional leg of about `seconds` still fits the run's
        wall-clock allowance (--wall-budget); a leg that does not is skipped
        and named in the JSON."""
        if time.time() - t_begin + seconds <= args.wall_budget:
            return True
        skipped.append(name)
        log(f"skipping {name}: {time.time() - t_begin:.0f}s spent, ~{seconds}s more would "
            f"pass the {args.wall_budget:.0f}s allowance")
        return False

    if args.cpu_iters > 0 and fits("cpu_baseline", 40):
        legs = {}
        for b_ in ("close", None):
            legs[b_] = run_cpu_leg(args, "c2", b_)
            log(f"cpu leg (OMP_PROC_BIND={b_ or 'unset'}): "
                f"{legs[b_]['per_iter_s'] * 1e3:.1f} ms/iter (all-OpenMP "
                f"{legs[b_]['per_iter_all_openmp_s'] * 1e3:.1f})")
        bind = min(legs, key=lambda k_: legs[k_]["per_iter_s"])
        best = legs[bind]
        per_iter = best["per_iter_s"]
        cpu = {"value": 1.0 / (per_iter * max(iters, 1)), "unit": "solves/s",
               "cores": best["threads"], "kind": "port",
               "sample": f"{args.cpu_iters} Jacobi-PCG iterations of the oracle (OpenMP SpMV, "
                         f"serial BLAS-1, OpenNL operation order) on this system, "
                         f"{per_iter * args.cpu_iters:.1f}s, scaled to the GPU solve's {iters} "
                         f"iterations; the 317 x 317 system's whole solve is timed in "
                         f"lscm317.cpu",
               "omp_proc_bind": bind or "unset",
               "all_openmp_value": 1.0 / (best["per_iter_all_openmp_s"] * max(iters, 1)),
               "per_iter_ms": {("close" if b_ else "unset"): {
                   "opennl_order": legs[b_]["per_iter_s"] * 1e3,
                   "all_openmp": legs[b_]["per_iter_all_openmp_s"] * 1e3} for b_ in legs},
               "cpu_model": cpu_model(), "nproc": os.cpu_count()}
        log(f"cpu baseline: {per_iter * 1e3:.1f} ms/iter on {best['threads']} threads "
            f"(OMP_PROC_BIND={bind or 'unset'}) -> {cpu['value']:.2e} solves/s")
    del A

    # ---- C1 (317 x 317): whole solves on the GPU and on the host cores
    if not args.no_c1 and fits("lscm317", 45):
        result_extra["lscm317"] = c1_measurement(args, capi, W, O, bind)

    # ---- least-squares assembly of the same system from its LSCM rows:
    # device (nlhipLeastSquaresCreate, host rows in, resident CRS out) against
    # the oracle's row-by-row context (OpenNL's algorithm, one core) on a
    # sample of the rows, and the multithreaded host gather that built Sw
    if not args.no_assembly and fits("assembly", 15):
        result_extra["assembly"] = assembly_measurement(args, Sw, capi, W)
    del Sw

    # ---- the batched atlas (2000 LSCM charts) beside it
    if not args.no_atlas and fits("atlas", 70):
        result_extra["atlas"] = atlas_measurement(args, capi, W)

    out = {
        "metric": "CG solves/s + SpMV HBM GB/s at 1e-10 rel. residual; p50/p99 solve latency",
        "value": solves_per_s,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded LSCM height-field grid)",
        "config": {"workload": f"lscm_grid_{args.k}", "n": n, "nnz": nnz,
                   "solver": "Jacobi-PCG fp64", "eps": args.eps, "max_iter": args.max_iter,
                   "iterations": iters,
                   "iterations_opennl_rule": int(iters - st.extra_iterations),
                   "restarts": int(st.outer_iterations),
                   "converged": converged, "cut_short": cut_short,
                   "rel_residual": st.rel_residual,
                   "true_rel_residual_oracle_spmv": true_oracle,
                   "recursive_rel_residual": st.cg_rel_residual,
                   "rows_per_wave": lpr, "workgroups": wg, "autotuned": tuned is not None,
                   "slice_iterations": slice_iters,
                   "step": f"a slice of {slice_iters} iterations of ONE FP64 Jacobi-PCG solve "
                           f"from x0 = 0 to a true relative residual <= {args.eps:g}; the "
                           f"{K} timed slices hold the whole solve (it finished in them: "
                           f"{converged}), value = 1 / their total seconds"},
        "latency_s": {"p50": t_total, "p99": t_total,
                      "what": "one whole solve (the timed region); the 317 x 317 system's "
                              "latency over repeated solves is in lscm317"},
        "parity": {"rel_diff_vs_cpu": None,
                   "why_not": "the oracle would need ~8 h for this solve on the host cores; "
                              "rel. diff and forward error are measured on the 317 x 317 "
                              "system (lscm317), and here the true residual of the GPU "
                              "result is recomputed by the oracle's SpMV",
                   "true_rel_residual": true_oracle},
        "spmv_gbs": {"fp64": spmv_gbs, "fp32": spmv32_gbs,
                     "bytes_fp64": spmv_bytes(n, nnz)},
        "iteration_gbs": iter_gbs,
        "roofline": {"bound": "hbm", "kernel": "k_cg_spmv", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"lscm_grid_{args.k}", lpr, wg),
                     "bytes_per_launch": bytes_a, "launch_ms": t_a * 1e3,
                     "copy_gbs": copy_gbs, "read_gbs": read_gbs,
                     "frac_of_read": achieved / read_gbs,
                     "stream_mix": {"what": "the block SpMV's traffic without its arithmetic "
                                            "(nlhipBenchStreamMix): 34 bytes per 2x2 block "
                                            "streamed, an L2 gather per block, a 16-byte read "
                                            "and write per row pair",
                                    "blocks": mix_blocks, "blocks_per_row_pair": mix_bpr,
                                    "bytes": mix_bytes, "ms": t_mix * 1e3,
                                    "gbs": mix_bytes / t_mix / 1e9,
                                    "ms_without_writes": t_mix_ro * 1e3,
                                    "spmv_time_frac": t_mix / t_a}},
        "kernels_ms": {"k_cg_dir": t_dir * 1e3, "k_cg_spmv": t_a * 1e3,
                       "k_cg_update": t_upd * 1e3},
        "cpu_baseline": cpu,
        "device": dev_name,
        "bench_seconds": time.time() - t_begin,
        "skipped_for_time": skipped,
        **result_extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
