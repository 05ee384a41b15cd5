"""Measurements of the configurations beside bench.py's single large system.

    python -m graphitethree_amd.benchmarks atlas [--count 2000] [--seed 1]
    python -m graphitethree_amd.benchmarks multirhs [--grid 1500] [--ks 3 4 8 16 32]
    python -m graphitethree_amd.benchmarks serving [--count 120] [--rate 10]
    python -m graphitethree_amd.benchmarks eigen [--grid 700] [--k 10]
    python -m graphitethree_amd.benchmarks spmv-patterns [--k 2236]
    python -m graphitethree_amd.benchmarks mixed [--k 2236] [--seconds 306]

atlas: `count` disk-like LSCM charts (N ~ logU[50, 50000] vertices) solved
  (1) as one batch in FP64, (2) as one batch with FP32 inner PCG inside FP64
  refinement, and (3) chart by chart with the plain FP64 GPU PCG
  (nlhipSolve per chart, the way an atlas builder calls the solver once per
  chart).  Reports solves/s of each, the speed-up of (2) and (1) over (3),
  convergence counts, and the largest relative difference between the
  batched and per-chart solutions.

multirhs: the (I + 1e-3 L) cotangent-Laplacian system of a k x k noisy grid
  (default 1500 x 1500, n = 2.25 M) with 3..32 N(0,1) right-hand sides,
  solved as one multi-right-hand-side PCG (FP64, and FP32 inner with FP64
  refinement) against one FP64 PCG per right-hand side.

eigen: the k smallest eigenpairs of a grid mesh's cotangent Laplacian with
  lumped mass (manifold harmonics), nlhipEigenSolve against scipy's
  shift-invert ARPACK with a SuperLU factorization on the host.

mixed: the FP32-inner / FP64-refined PCG on the 2236 x 2236 LSCM system,
  warm-started calls of a fixed number of inner iterations, until 1e-10 or
  the FP64 solve's device time: the residual it reaches at equal time.

spmv-patterns: the CG SpMV step on the 2236 x 2236 LSCM system against two
  synthetic 2x2-block systems of the same size and row length whose x
  gathers are (a) a band of 7 block columns around the diagonal (every
  gather a cache hit) and (b) the grid stencil of the LSCM system (rows
  i +- 1, i +- w) with random values, and (c) the LSCM system with its
  vertices in random order (32-bit block columns, gathers anywhere), with
  and without the block form: how much of the SpMV step's time the gather
  pattern costs.

serving: a stream of requests with Poisson arrivals (rate --rate per second):
  60% LSCM charts with N ~ logU[1e3, 1e6] vertices, 40% smoothing solves
  (I + 1e-3 L) x = b with 3 right-hand sides (x, y, z) on grids of similar
  size; submitted to the solve server (200 GB HBM budget) at their arrival
  times.  Reports p50 / p99 latency (arrival to completion), solves/s and the
  server's admission figures.

Each subcommand prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import capi
from . import workloads as W


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def init_device():
    L = capi.lib()
    if not L.nlInitExtension(b"HIP"):
        raise SystemExit("nlInitExtension(\"HIP\") failed: " + capi.last_error())
    return capi.device_info()


def run_atlas(args) -> dict:
    dev = init_device()
    t0 = time.time()
    A = W.atlas(args.count, seed=args.seed, nmin=args.nmin, nmax=args.nmax)
    t_host_build = time.time() - t0
    log(f"atlas: {A.count} charts, n={A.n}, nnz={A.nnz}, built in {t_host_build:.1f}s")
    assembly = None if args.no_device_assembly else atlas_device_assembly(args, A)
    t0 = time.perf_counter()
    B = capi.Batch(packed=(A.row_off, A.rowptr, A.colind, A.val), fp32=True)
    t_packed_create = time.perf_counter() - t0
    if args.direct_launches:
        B.set_graphs(False)
    params = dict(eps=args.eps, max_iter=args.max_iter)

    tuned = B.autotune(A.b, iters=args.tune_iters) if args.tune_iters else None
    if tuned:
        log(f"autotuned: {tuned[0]} workgroups, SpMV passes {tuned[1] or 'by active tiles'}, "
            f"{tuned[2] * 1e3:.3f} ms per iteration over the first {args.tune_iters}")
    x64, p64, t64 = B.solve(A.b, **params)            # warm-up (graph capture)
    x64, p64, t64 = B.solve(A.b, **params)
    log(f"batched fp64: {t64.device_seconds:.3f}s, max it {t64.iterations}, "
        f"converged {sum(p.converged for p in p64)}/{A.count}")
    # the same FP64 iterations as one persistent cooperative launch per run
    B.set_persistent(True)
    B.solve(A.b, **params)
    xp, pp, tp = B.solve(A.b, **params)
    B.set_persistent(False)
    persistent = {"seconds": tp.device_seconds, "solves_per_s": A.count / tp.device_seconds,
                  "bitwise_equal_to_graphs": bool(np.array_equal(xp.view(np.uint64),
                                                                 x64.view(np.uint64)))}
    log(f"batched fp64, persistent launches: {tp.device_seconds:.3f}s, same bits "
        f"{persistent['bitwise_equal_to_graphs']}")
    del xp
    xm, pm, tm = B.solve(A.b, precision="mixed", inner_eps=args.inner_eps,
                         max_iter=args.max_iter_mixed, eps=args.eps)
    xm, pm, tm = B.solve(A.b, precision="mixed", inner_eps=args.inner_eps,
                         max_iter=args.max_iter_mixed, eps=args.eps)
    log(f"batched mixed: {tm.device_seconds:.3f}s, converged "
        f"{sum(p.converged for p in pm)}/{A.count}")
    fb_params = dict(precision="mixed-fallback", inner_eps=args.inner_eps,
                     inner_max_iter=args.fallback_inner_iters,
                     max_iter=args.max_iter_mixed, eps=args.eps)
    B.solve(A.b, **fb_params)
    xf, pf, tf = B.solve(A.b, **fb_params)
    nfb = B.fallbacks()
    log(f"batched mixed, FP64 fallback after {args.fallback_inner_iters} inner iterations: "
        f"{tf.device_seconds:.3f}s, {nfb} of {A.count} charts finished in FP64, converged "
        f"{sum(p.converged for p in pf)}/{A.count}")

    # plain FP64 GPU CG, one nlhipSolve per chart (matrices resident)
    charts = list(range(0, A.count, args.baseline_stride))
    xs64 = B.split(x64)
    t_plain = 0.0
    worst_diff = 0.0
    its_plain = []
    per_chart = []   # (n, fp64 s, fp64 it, mixed s, mixed inner it) per sampled chart
    mp = dict(precision="mixed", inner_eps=args.inner_eps, max_iter=args.max_iter_mixed,
              eps=args.eps)
    for k, c in enumerate(charts):
        S = A.chart(c)
        sysc = capi.System.from_crs(S, fp32=args.per_chart_mixed)
        sysc.solve(S.b, **params)                      # warm-up
        x, st = sysc.solve(S.b, **params)
        t_plain += st.device_seconds
        its_plain.append(int(st.iterations))
        worst_diff = max(worst_diff, float(np.linalg.norm(x - xs64[c]) /
                                           max(np.linalg.norm(xs64[c]), 1e-300)))
        row = [S.n, st.device_seconds, int(st.iterations), None, None]
        if args.per_chart_mixed:
            sysc.solve(S.b, **mp)
            _, sm = sysc.solve(S.b, **mp)
            row[3], row[4] = sm.device_seconds, int(sm.iterations)
        per_chart.append(row)
        sysc.close()
        if k % 200 == 0:
            log(f"  per-chart baseline {k}/{len(charts)} ({t_plain:.2f}s so far)")
    scale = A.count / len(charts)
    t_plain_all = t_plain * scale
    res = {
        "metric": "atlas solves/s (charts solved per second) at 1e-10 rel. residual",
        "workload": f"atlas_{A.count}_charts_N{int(args.nmin)}-{int(args.nmax)}",
        "charts": A.count, "n": A.n, "nnz": A.nnz, "eps": args.eps,
        "device": dev[0],
        "autotuned": None if not tuned else {"workgroups": tuned[0], "spmv_passes": tuned[1],
                                             "ms_per_iteration_first_iters": tuned[2] * 1e3},
        "batched_fp64": {"seconds": t64.device_seconds, "solves_per_s": A.count / t64.device_seconds,
                         "max_iterations": int(t64.iterations), "max_iter": args.max_iter,
                         "converged": int(sum(p.converged for p in p64)),
                         "worst_rel_residual": float(t64.rel_residual)},
        "batched_mixed": {"seconds": tm.device_seconds, "solves_per_s": A.count / tm.device_seconds,
                          "max_inner_iterations": int(tm.iterations),
                          "max_iter": args.max_iter_mixed, "inner_eps": args.inner_eps,
                          "converged": int(sum(p.converged for p in pm)),
                          "worst_rel_residual": float(tm.rel_residual),
                          "max_rel_diff_vs_fp64": float(max(
                              np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
                              for a, b in zip(B.split(xm), xs64)))},
        "per_chart_fp64": {"seconds": t_plain_all, "solves_per_s": A.count / t_plain_all,
                           "sampled_charts": len(charts),
                           "sample": f"every {args.baseline_stride}th chart, time scaled by "
                                     f"{scale:g}",
                           "max_rel_diff_vs_batched": worst_diff,
                           "iterations_mean": float(np.mean(its_plain))},
    }
    res["batched_mixed_fp64_fallback"] = {
        "seconds": tf.device_seconds, "solves_per_s": A.count / tf.device_seconds,
        "inner_max_iter": args.fallback_inner_iters, "charts_finished_in_fp64": nfb,
        "converged": int(sum(p.converged for p in pf)),
        "worst_rel_residual": float(tf.rel_residual),
        "max_rel_diff_vs_fp64": float(max(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
                                          for a, b in zip(B.split(xf), xs64))),
        "policy": "FP32 inner PCG in FP64 refinement; a chart whose FP32 inner solve uses "
                  "inner_max_iter without reaching inner_eps finishes with FP64 CG (DSGESV's "
                  "fallback)"}
    res["speedup_mixed_fallback_batched_vs_per_chart_fp64"] = t_plain_all / tf.device_seconds
    res["create"] = {"host_build_seconds": t_host_build,
                     "host_build": "chart meshes, per-chart LSCM assembly and packing on "
                                   "the host (OpenMP), then nlhipBatchCreate",
                     "batch_create_from_packed_seconds": t_packed_create,
                     "device_assembly": assembly}
    res["batched_fp64_persistent"] = persistent
    res["speedup_mixed_batched_vs_per_chart_fp64"] = t_plain_all / tm.device_seconds
    # per chart-size band: iterations of FP32-inner refinement against FP64
    # PCG (from the batched solves: each chart's counts are its own) and the
    # per-chart solve times of both
    edges = [0, 1000, 10000, 10 ** 9]
    bands = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = [i for i, c in enumerate(charts) if lo <= per_chart[i][0] < hi]
        allc = [c for c in range(A.count) if lo <= int(A.row_off[c + 1] - A.row_off[c]) < hi]
        if not allc:
            continue
        it64 = np.array([p64[c].iterations for c in allc], dtype=float)
        itm = np.array([pm[c].iterations for c in allc], dtype=float)
        band = {"unknowns": [lo, hi if hi < 10 ** 9 else None], "charts": len(allc),
                "fp64_iterations_mean": float(it64.mean()),
                "mixed_inner_iterations_mean": float(itm.mean()),
                "inner_iteration_inflation": float((itm / np.maximum(it64, 1)).mean())}
        if sel:
            band["per_chart_fp64_s"] = float(sum(per_chart[i][1] for i in sel) * scale)
            if args.per_chart_mixed:
                band["per_chart_mixed_s"] = float(sum(per_chart[i][3] for i in sel) * scale)
                band["mixed_over_fp64_time"] = band["per_chart_mixed_s"] / band["per_chart_fp64_s"]
        bands.append(band)
    res["bands"] = bands
    res["speedup_fp64_batched_vs_per_chart_fp64"] = t_plain_all / t64.device_seconds
    return res


def run_atlas_trace(args) -> dict:
    """The batched FP64 atlas solve with every kernel launched directly (no
    graph replay), for kernel tracing: per-kernel time and HBM counters of
    the batch kernels.  One warm-up solve, one timed solve."""
    dev = init_device()
    A = W.atlas(args.count, seed=args.seed, nmin=args.nmin, nmax=args.nmax)
    log(f"atlas: {A.count} charts, n={A.n}, nnz={A.nnz}")
    B = capi.Batch(packed=(A.row_off, A.rowptr, A.colind, A.val))
    B.set_graphs(False)
    B.solve(A.b, eps=args.eps, max_iter=args.max_iter)
    x, per, tot = B.solve(A.b, eps=args.eps, max_iter=args.max_iter)
    log(f"batched fp64, direct launches: {tot.device_seconds:.3f}s, max it {tot.iterations}")
    res = {"workload": f"atlas_{A.count}_charts_N{int(args.nmin)}-{int(args.nmax)}",
           "device": dev[0], "n": A.n, "nnz": A.nnz, "launches": "direct",
           "seconds": tot.device_seconds, "max_iterations": int(tot.iterations),
           "converged": int(sum(p.converged for p in per)),
           "spmv_bytes_all_active": 8.5 * A.nnz + 16.0 * A.n}
    del B, x
    if not args.no_one_system:
        # the whole block-diagonal matrix as ONE system: the single-system CG
        # kernels on the same nonzeros (every row active), for comparison
        # with the batched all-active launches
        S = capi.System(A.rowptr, A.colind, A.val)
        tuned = S.autotune(iters=20)
        td, ts, tu = capi.profile_cg(S, iters=50)
        log(f"one system: engine {tuned[0]}, {tuned[1]} workgroups; CG SpMV step "
            f"{ts * 1e6:.1f} us, direction {td * 1e6:.1f} us, update {tu * 1e6:.1f} us")
        res["one_system"] = {"engine": tuned[0], "workgroups": tuned[1],
                             "spmv_us": ts * 1e6, "dir_us": td * 1e6, "update_us": tu * 1e6}
        S.close()
    return res


def block_band_system(nb: int, offsets, seed: int = 1):
    """CRS of a 2 nb x 2 nb matrix in 2x2 block form: block row I holds the
    blocks (I, I + o) for o in `offsets` that fall inside the matrix, values
    uniform in [-0.1, 0), diagonal entries 10."""
    offs = np.asarray(sorted(offsets))
    J = np.arange(nb)[:, None] + offs[None, :]
    ok = (J >= 0) & (J < nb)
    rowcnt = np.repeat(2 * ok.sum(1), 2)
    rp = np.zeros(2 * nb + 1, dtype=np.int64)
    np.cumsum(rowcnt, out=rp[1:])
    cols = np.stack([2 * J, 2 * J + 1], axis=2).reshape(nb, -1)
    valid = np.repeat(ok, 2, axis=1)
    ci = np.repeat(cols, 2, axis=0)[np.repeat(valid, 2, axis=0)]
    val = -np.random.default_rng(seed).random(ci.size) * 0.1
    val[ci == np.repeat(np.arange(2 * nb), rowcnt)] = 10.0
    return rp.astype(np.uint32), ci.astype(np.uint32), val


def run_spmv_patterns(args) -> dict:
    dev = init_device()
    res = {"device": dev[0], "patterns": []}

    def measure(name, rp, ci, val, blocks=True):
        n, nnz = len(rp) - 1, len(ci)
        S = capi.System(rp, ci, val, blocks=blocks)
        tuned = S.autotune(iters=20)
        td, ts, tu = capi.profile_cg(S, iters=50)
        byts = 8.5 * nnz + 26.0 * n
        log(f"{name}: n={n} nnz={nnz}, engine {tuned[0]} / {tuned[1]} workgroups, CG SpMV "
            f"step {ts * 1e6:.1f} us ({byts / ts / 1e9:.0f} GB/s on 8.5 nnz + 26 n bytes)")
        res["patterns"].append({"pattern": name, "n": n, "nnz": nnz, "engine": tuned[0],
                                "workgroups": tuned[1], "spmv_us": ts * 1e6,
                                "block_form_gbs": byts / ts / 1e9})
        S.close()

    G = W.lscm_grid(args.k, seed=args.seed)
    w = args.k  # block columns between vertically neighbouring vertices
    nb = G.n // 2
    measure(f"lscm_grid_{args.k}", G.rowptr, G.colind, G.val)
    del G
    measure("band of 7 block columns", *block_band_system(nb, range(-3, 4), args.seed))
    measure(f"grid stencil (I, I +- 1, I +- {w}, I +- {w + 1}), random values",
            *block_band_system(nb, (-w - 1, -w, -1, 0, 1, w, w + 1), args.seed))
    if not args.no_shuffled:
        # the LSCM system with its vertices in random order: block columns
        # too far apart for 16-bit offsets (32-bit block columns), gathers
        # anywhere in x; and the same matrix without the block form
        import scipy.sparse as sps
        G = W.lscm_grid(args.k, seed=args.seed)
        p = np.random.default_rng(args.seed).permutation(G.n // 2)
        q = np.empty(G.n, np.int64)
        q[0::2], q[1::2] = 2 * p, 2 * p + 1
        M = sps.csr_matrix((G.val, G.colind.astype(np.int64), G.rowptr.astype(np.int64)),
                           shape=(G.n, G.n))[q][:, q].tocsr()
        del G
        M.sort_indices()
        rp, ci = M.indptr.astype(np.uint32), M.indices.astype(np.uint32)
        measure(f"lscm_grid_{args.k}, vertices in random order", rp, ci, M.data)
        measure(f"lscm_grid_{args.k}, vertices in random order, no block form", rp, ci, M.data,
                blocks=False)
    return res


def run_mixed(args) -> dict:
    """FP32-inner / FP64-refined PCG on the k x k LSCM grid system against a
    time budget (the FP64 solve's time): the refinement runs in calls of
    `chunk_iters` inner iterations, each warm-started from the previous
    call's iterate (an outer-iteration boundary of the refinement), until it
    reaches eps or the budget.  Reports the true residual after each call."""
    dev = init_device()
    G = W.lscm_grid(args.k, seed=args.seed)
    S = capi.System.from_crs(G, fp32=True)
    tuned = S.autotune(iters=20)
    log(f"lscm_grid_{args.k}: n={G.n} nnz={G.nnz}; engine {tuned[0]} / {tuned[1]} workgroups")
    x = np.zeros(G.n)
    seconds, inner, outer, trace = 0.0, 0, 0, []
    converged, stalled, last = False, False, float("inf")
    while seconds < args.seconds and not converged and not stalled:
        x, st = S.solve(G.b, x0=x, precision="mixed", eps=args.eps, max_iter=args.chunk_iters,
                        inner_eps=args.inner_eps)
        seconds += st.device_seconds
        inner += int(st.iterations)
        outer += int(st.outer_iterations)
        converged = bool(st.converged)
        # a call that ends where it started: its refinement undid its only
        # correction (it did not lower the residual), and the next call would
        # repeat it
        stalled = not converged and not (st.rel_residual < last)
        last = min(last, st.rel_residual)
        trace.append({"seconds": seconds, "inner_iterations": inner,
                      "true_rel_residual": st.rel_residual})
        log(f"{seconds:.1f} s: {inner} inner iterations, {outer} outer, true rel res "
            f"{st.rel_residual:.3e}, converged {converged}, stalled {stalled}")
    return {"workload": f"lscm_grid_{args.k}", "device": dev[0], "n": G.n, "nnz": G.nnz,
            "eps": args.eps, "inner_eps": args.inner_eps, "chunk_iterations": args.chunk_iters,
            "budget_s": args.seconds, "seconds": seconds, "inner_iterations": inner,
            "outer_iterations": outer, "converged": converged, "stalled": stalled,
            "true_rel_residual": last, "trace": trace}


def atlas_device_assembly(args, A) -> dict:
    """The same atlas built on the device from its chart meshes
    (nlhipBatchLSCMCreate: LSCM rows, normal equations, block form and tiles);
    checked bitwise against the host-packed CRS."""
    t0 = time.time()
    M = W.atlas_mesh(args.count, seed=args.seed, nmin=args.nmin, nmax=args.nmax)
    t_mesh = time.time() - t0
    times = []
    same = None
    for rep in range(2):                       # the first call also sizes the sort
        t0 = time.perf_counter()
        Bd, rhs = capi.Batch.lscm_atlas(M.vertex_off, M.points, M.triangles, M.pins, M.pin_uv,
                                        fp32=True)
        times.append(time.perf_counter() - t0)
        if rep == 0:
            ro, rp, ci, va = Bd.crs()
            same = bool(np.array_equal(ro, A.row_off) and np.array_equal(rp, A.rowptr) and
                        np.array_equal(ci, A.colind) and
                        np.array_equal(va.view(np.uint64), A.val.view(np.uint64)) and
                        np.array_equal(rhs.view(np.uint64), A.b.view(np.uint64)))
            del ro, rp, ci, va
        Bd.close()
    log(f"device atlas assembly: {times[-1]:.3f}s (first call {times[0]:.3f}s), "
        f"bitwise equal to the host packing: {same}")
    return {"seconds": times[-1], "first_call_seconds": times[0],
            "mesh_generation_seconds": t_mesh, "vertices": int(M.vertex_off[-1]),
            "triangles": int(len(M.triangles)), "bitwise_equal_to_host": same,
            "scope": "host mesh arrays in, batch resident out: upload, LSCM rows, "
                     "normal equations, right-hand sides, block form, tiles"}


def run_multirhs(args) -> dict:
    dev = init_device()
    t0 = time.time()
    mesh = W.grid_mesh(args.grid, seed=args.seed)
    Sw = W.cotan_system(mesh, t=args.t, nrhs=max(args.ks), rhs="gaussian", seed=args.seed)
    del mesh
    log(f"cotan {args.grid}^2: n={Sw.n} nnz={Sw.nnz} built in {time.time() - t0:.1f}s")
    S = capi.System.from_crs(Sw, fp32=True)
    rows = []
    for k in args.ks:
        Bk = Sw.rhs[:k]
        S.solve_multi(Bk, eps=args.eps, max_iter=args.max_iter)          # warm-up
        X, per, tot = S.solve_multi(Bk, eps=args.eps, max_iter=args.max_iter)
        S.solve_multi(Bk, precision="mixed", eps=args.eps, max_iter=args.max_iter,
                      inner_eps=args.inner_eps)
        Xm, perm, totm = S.solve_multi(Bk, precision="mixed", eps=args.eps,
                                       max_iter=args.max_iter, inner_eps=args.inner_eps)
        t_single, diff = 0.0, 0.0
        for j in range(k):
            S.solve(Bk[j], eps=args.eps, max_iter=args.max_iter)
            xj, st = S.solve(Bk[j], eps=args.eps, max_iter=args.max_iter)
            t_single += st.device_seconds
            diff = max(diff, float(np.linalg.norm(xj - X[j]) / np.linalg.norm(xj)))
        row = {"k": k,
               "block_fp64": {"seconds": tot.device_seconds,
                              "rhs_per_s": k / tot.device_seconds,
                              "iterations": int(tot.iterations),
                              "worst_rel_residual": tot.rel_residual},
               "block_mixed": {"seconds": totm.device_seconds,
                               "rhs_per_s": k / totm.device_seconds,
                               "inner_iterations": int(totm.iterations),
                               "worst_rel_residual": totm.rel_residual,
                               "max_rel_diff_vs_fp64": float(max(
                                   np.linalg.norm(Xm[j] - X[j]) / np.linalg.norm(X[j])
                                   for j in range(k)))},
               "per_rhs_fp64": {"seconds": t_single, "rhs_per_s": k / t_single,
                                "max_rel_diff_vs_block": diff},
               "speedup_block_fp64": t_single / tot.device_seconds,
               "speedup_block_mixed": t_single / totm.device_seconds}
        log(json.dumps(row))
        rows.append(row)
    return {"metric": "right-hand sides solved per second at 1e-10 rel. residual",
            "workload": f"cotan_{args.grid}_t{args.t:g}", "n": Sw.n, "nnz": Sw.nnz,
            "device": dev[0], "eps": args.eps, "rows": rows}


def make_requests(count: int, seed: int, nmin: float, nmax: float):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        N = int(round(float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))))
        if rng.random() < 0.6:
            S = W.lscm_system(W.chart_mesh(N, seed=seed * 100003 + i))
            reqs.append(("lscm", S, S.b[None, :]))
        else:
            g = max(3, int(round(np.sqrt(N))))
            S = W.cotan_system(W.grid_mesh(g, seed=seed * 100003 + i), t=1e-3, nrhs=3,
                               rhs="xyz")
            reqs.append(("smooth", S, S.rhs))
    return reqs


def run_serving(args) -> dict:
    """Poisson arrivals of mixed requests through the solve server.  With
    --resident, the smoothing requests re-solve a pool of meshes whose
    systems stay resident on the device (nlhipServerSubmitSystem: only the
    right-hand sides travel), the way a modelling session smooths the same
    surfaces again and again; otherwise every request uploads its matrix."""
    import threading
    dev = init_device()
    t0 = time.time()
    reqs = make_requests(args.count, args.seed, args.nmin, args.nmax)
    log(f"serving: {len(reqs)} requests generated in {time.time() - t0:.1f}s "
        f"({sum(r[0] == 'lscm' for r in reqs)} LSCM, {sum(r[0] == 'smooth' for r in reqs)} "
        f"smoothing x3)")
    pool = []
    if args.resident:
        smooth = [i for i, r in enumerate(reqs) if r[0] == "smooth"]
        for i in smooth[:args.resident]:
            pool.append((i, capi.System.from_crs(reqs[i][1])))
        log(f"resident pool: {len(pool)} smoothing systems")
    rng = np.random.default_rng(args.seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, size=len(reqs)))
    srv = capi.Server(hbm_budget_bytes=int(args.budget_gb * (1 << 30)),
                      max_active=args.max_active)
    ids = [None] * len(reqs)
    results = [None] * len(reqs)
    target = [None] * len(reqs)   # (kind, CRS, B) each request actually solved
    submitted = threading.Semaphore(0)

    def waiter():
        for i in range(len(reqs)):
            submitted.acquire()
            try:
                X, res = srv.wait(ids[i])
                results[i] = (0, res, X)
            except capi.NLHIPError as e:
                results[i] = (e.code, getattr(e, "result", None), None)
            if i % 20 == 0:
                log(f"  completed {i + 1}/{len(reqs)}")

    th = threading.Thread(target=waiter)
    th.start()
    start = time.perf_counter()
    n_smooth = 0
    for i, (kind, S, B) in enumerate(reqs):
        delay = arrivals[i] - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        if kind == "smooth" and pool:
            j, Sys = pool[n_smooth % len(pool)]
            n_smooth += 1
            Sj, Bj = reqs[j][1], reqs[j][2]
            target[i] = (kind, Sj, Bj)
            ids[i] = srv.submit_system(Sys, Bj, eps=args.eps, max_iter=args.max_iter)
        else:
            target[i] = (kind, S, B)
            ids[i] = srv.submit(S, B, eps=args.eps, max_iter=args.max_iter)
        submitted.release()
    th.join()
    info = srv.info()
    ok = [r for r in results if r[0] == 0]
    lat = np.array([r[1].t_done - r[1].t_submit for r in ok])
    wait = np.array([r[1].t_admit - r[1].t_submit for r in ok])
    t_first = min(r[1].t_submit for r in ok)
    t_last = max(r[1].t_done for r in ok)
    conv_idx = [i for i, r in enumerate(results) if r[0] == 0 and int(r[1].stats.converged)]
    lat_conv = np.array([results[i][1].t_done - results[i][1].t_submit for i in conv_idx])
    not_conv = [{"request": i, "kind": target[i][0], "n": int(target[i][1].n),
                 "iterations": int(results[i][1].stats.iterations),
                 "rel_residual": float(results[i][1].stats.rel_residual)}
                for i, r in enumerate(results)
                if r[0] == 0 and not int(r[1].stats.converged)]
    # spot check: a few requests solved alone give the same bits as in the mix
    same = True
    for i in (0, len(reqs) // 2, len(reqs) - 1):
        kind, S, B = target[i]
        alone = capi.Server(max_active=1)
        Xa, _ = alone.wait(alone.submit(S, B, eps=args.eps, max_iter=args.max_iter))
        alone.close()
        same &= bool(np.array_equal(Xa, results[i][2]))
    srv.close()
    for _, Sys in pool:
        Sys.close()
    return {
        "metric": "serving p50/p99 solve latency and solves/s",
        "workload": f"serving_{len(reqs)}_requests_rate{args.rate:g}"
                    + (f"_resident{len(pool)}" if pool else ""),
        "device": dev[0], "requests": len(reqs), "rate_per_s": args.rate,
        "lscm": int(sum(r[0] == "lscm" for r in reqs)),
        "smoothing_x3": int(sum(r[0] == "smooth" for r in reqs)),
        "resident_systems": len(pool),
        "eps": args.eps, "max_iter": args.max_iter,
        "completed": int(info.completed), "failed": int(info.failed),
        "converged_to_eps": len(conv_idx),
        "not_converged": not_conv,
        "latency_s": {"p50": float(np.percentile(lat_conv, 50)),
                      "p99": float(np.percentile(lat_conv, 99)),
                      "mean": float(lat_conv.mean()), "max": float(lat_conv.max()),
                      "what": "requests that reached eps"},
        "latency_all_s": {"p50": float(np.percentile(lat, 50)),
                          "p99": float(np.percentile(lat, 99))},
        "queue_wait_s": {"p50": float(np.percentile(wait, 50)),
                         "p99": float(np.percentile(wait, 99))},
        "solves_per_s": float(len(conv_idx) / (t_last - t_first)),
        "solves_per_s_what": "requests that reached eps per second of the makespan",
        "makespan_s": float(t_last - t_first),
        "budget_gb": info.budget_bytes / 2**30,
        "peak_in_use_gb": info.peak_in_use_bytes / 2**30,
        "peak_running": int(info.peak_running), "max_active": args.max_active,
        "alone_equals_mix_spot_check": same,
    }


def run_ssor(args) -> dict:
    """SSOR-preconditioned CG through nlSolveSystemIterative (the wrapper's
    solve_iterative(..., "CG", "SSOR")) on a cotangent smoothing system,
    against Jacobi-CG on the device and the oracle's SSOR-CG on one host
    core."""
    import scipy.sparse as sp
    from oracle import oracle as O
    dev = init_device()
    L = capi.lib()
    mesh = W.grid_mesh(args.grid, seed=args.seed)
    Sw = W.cotan_system(mesh, t=args.t, nrhs=1)
    del mesh
    n = Sw.n
    log(f"ssor: cotan {args.grid}^2 t={args.t:g}, n={n}, nnz={Sw.nnz}")
    M = L.nlCRSMatrixNewFromArrays(n, n, capi.ptr(Sw.rowptr), capi.ptr(Sw.colind),
                                   capi.ptr(Sw.val))
    b = Sw.b.copy()
    t0 = time.perf_counter()
    P = L.nlNewSSORPreconditioner(M, 1.5)
    t_create = time.perf_counter() - t0
    if not P:
        raise RuntimeError("nlNewSSORPreconditioner failed")
    rows = {}
    for name, prec in (("ssor", P), ("jacobi", None)):
        if name == "jacobi":
            prec = L.nlNewJacobiPreconditioner(M)
        for rep in range(2):
            x = np.zeros(n)
            t0 = time.perf_counter()
            its = L.nlSolveSystemIterative(L.nlHostBlas(), M, prec, capi.ptr(b), capi.ptr(x),
                                           capi.NL_CG, args.eps, args.max_iter, 5)
            dt = time.perf_counter() - t0
        A = sp.csr_matrix((Sw.val, Sw.colind, Sw.rowptr), shape=(n, n))
        rr = float(np.linalg.norm(A @ x - b) / np.linalg.norm(b))
        rows[name] = {"seconds": dt, "iterations": int(its), "rel_residual": rr,
                      "ms_per_iteration": 1e3 * dt / max(int(its), 1)}
        log(f"{name}-CG: {its} iterations, {dt:.3f}s, residual {rr:.2e}")
        if name == "jacobi":
            L.nlDeleteMatrix(prec)
    L.nlDeleteMatrix(P)
    L.nlDeleteMatrix(M)
    out = {"metric": "SSOR-preconditioned CG solve time (nlSolveSystemIterative)",
           "workload": f"cotan_{args.grid}_t{args.t:g}", "n": n, "nnz": Sw.nnz,
           "device": dev[0], "eps": args.eps, "omega": 1.5,
           "ssor_create_seconds": t_create, "ssor_cg": rows["ssor"],
           "jacobi_cg_device": rows["jacobi"]}
    if not args.no_cpu:
        C = O.CRS(Sw.rowptr, Sw.colind, Sw.val)
        t0 = time.perf_counter()
        xc, itc, relc = O.solve_ssor(C, b, "cg", eps=args.eps, max_iter=args.max_iter)
        dt = time.perf_counter() - t0
        out["cpu_ssor_cg"] = {"seconds": dt, "iterations": itc, "rel_residual": relc,
                              "cores": 1, "kind": "port"}
        out["speedup_vs_cpu_ssor_cg"] = dt / rows["ssor"]["seconds"]
        log(f"cpu ssor-CG: {itc} iterations, {dt:.3f}s")
    return out


def run_eigen(args) -> dict:
    """k smallest eigenpairs of a mesh Laplacian with lumped mass (manifold
    harmonics): the device solver against scipy's shift-invert ARPACK with a
    SuperLU factorization (the reference's ARPACK-extension configuration)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    dev = init_device()
    mesh = W.grid_mesh(args.grid, seed=args.seed)
    L, mass = W.laplacian(mesh)
    log(f"eigen: grid {args.grid}^2, n={L.n}, nnz={L.nnz}, k={args.k}")
    S = capi.System.from_crs(L)
    vals, vecs, st = S.eigen(args.k, mass=mass, shift=args.shift, eps=args.eps)
    t_gpu = st.seconds
    vals, vecs, st = S.eigen(args.k, mass=mass, shift=args.shift, eps=args.eps)
    t_gpu = min(t_gpu, st.seconds)
    log(f"  device: {t_gpu:.3f}s, {st.outer_iterations} outer / {st.inner_iterations} filter SpMMs "
        f"iterations, {st.converged}/{args.k} below {args.eps:g}")
    res = {"metric": "eigenpairs: seconds for k smallest of L x = lambda M x",
           "workload": f"laplacian_grid_{args.grid}_k{args.k}", "device": dev[0],
           "n": L.n, "nnz": L.nnz, "k": args.k, "shift": args.shift, "eps": args.eps,
           "device_s": t_gpu, "outer_iterations": int(st.outer_iterations),
           "filter_spmms": int(st.inner_iterations), "converged": int(st.converged),
           "max_rel_residual": st.max_rel_residual, "eigenvalues": [float(v) for v in vals]}
    if not args.no_cpu:
        A = L.to_scipy().tocsc()
        M = sp.diags(mass).tocsc()
        t0 = time.perf_counter()
        ref = np.sort(spla.eigsh(A, k=args.k, M=M, sigma=args.shift, which="LM",
                                 tol=args.eps)[0])
        t_cpu = time.perf_counter() - t0
        res["cpu_arpack_superlu_s"] = t_cpu
        res["max_rel_diff_eigenvalues"] = float(np.max(np.abs(vals - ref) /
                                                       np.maximum(np.abs(ref), 1e-12)))
        res["speedup"] = t_cpu / t_gpu
        log(f"  cpu (scipy eigsh shift-invert, SuperLU): {t_cpu:.2f}s, speedup "
            f"{t_cpu / t_gpu:.1f}x")
    S.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("atlas")
    a.add_argument("--count", type=int, default=2000)
    a.add_argument("--seed", type=int, default=1)
    a.add_argument("--nmin", type=float, default=50)
    a.add_argument("--nmax", type=float, default=50000)
    a.add_argument("--eps", type=float, default=1e-10)
    a.add_argument("--max-iter", type=int, default=50000)
    a.add_argument("--max-iter-mixed", type=int, default=200000)
    a.add_argument("--inner-eps", type=float, default=1e-4)
    a.add_argument("--baseline-stride", type=int, default=1)
    a.add_argument("--no-device-assembly", action="store_true")
    a.add_argument("--fallback-inner-iters", type=int, default=300,
                   help="inner_max_iter of the mixed solve with FP64 fallback")
    a.add_argument("--direct-launches", action="store_true",
                   help="no graph replay (kernel tracing); same results")
    a.add_argument("--per-chart-mixed", action="store_true",
                   help="also time one FP32-inner refined solve per sampled chart")
    at = sub.add_parser("atlas-trace",
                        help="the batched FP64 atlas solve alone, launched directly")
    at.add_argument("--count", type=int, default=2000)
    at.add_argument("--seed", type=int, default=1)
    at.add_argument("--nmin", type=float, default=50)
    at.add_argument("--nmax", type=float, default=50000)
    at.add_argument("--eps", type=float, default=1e-10)
    at.add_argument("--max-iter", type=int, default=50000)
    at.add_argument("--no-one-system", action="store_true",
                    help="skip the single-system CG kernels on the same matrix")
    a.add_argument("--tune-iters", type=int, default=50,
                   help="iterations per candidate of the batch autotuner (0: no tuning)")
    mr = sub.add_parser("multirhs")
    mr.add_argument("--grid", type=int, default=1500)
    mr.add_argument("--t", type=float, default=1e-3)
    mr.add_argument("--ks", type=int, nargs="+", default=[3, 4, 8, 16, 32])
    mr.add_argument("--seed", type=int, default=1)
    mr.add_argument("--eps", type=float, default=1e-10)
    mr.add_argument("--max-iter", type=int, default=5000)
    mr.add_argument("--inner-eps", type=float, default=1e-4)
    ev = sub.add_parser("eigen")
    ev.add_argument("--grid", type=int, default=300)
    ev.add_argument("--k", type=int, default=10)
    ev.add_argument("--shift", type=float, default=-1e-2)
    ev.add_argument("--eps", type=float, default=1e-8)
    ev.add_argument("--seed", type=int, default=1)
    ev.add_argument("--no-cpu", action="store_true")
    so = sub.add_parser("ssor")
    so.add_argument("--grid", type=int, default=300)
    so.add_argument("--t", type=float, default=1.0)
    so.add_argument("--seed", type=int, default=1)
    so.add_argument("--eps", type=float, default=1e-10)
    so.add_argument("--max-iter", type=int, default=5000)
    so.add_argument("--no-cpu", action="store_true")
    pa = sub.add_parser("spmv-patterns")
    pa.add_argument("--k", type=int, default=2236)
    pa.add_argument("--seed", type=int, default=1)
    pa.add_argument("--no-shuffled", action="store_true",
                    help="skip the system with its vertices in random order")
    mx = sub.add_parser("mixed")
    mx.add_argument("--k", type=int, default=2236)
    mx.add_argument("--seed", type=int, default=1)
    mx.add_argument("--eps", type=float, default=1e-10)
    mx.add_argument("--inner-eps", type=float, default=1e-4)
    mx.add_argument("--chunk-iters", type=int, default=150000)
    mx.add_argument("--seconds", type=float, default=306.0,
                    help="device-time budget (the FP64 solve of the same system)")
    sv = sub.add_parser("serving")
    sv.add_argument("--count", type=int, default=120)
    sv.add_argument("--rate", type=float, default=10.0)
    sv.add_argument("--seed", type=int, default=1)
    sv.add_argument("--nmin", type=float, default=1e3)
    sv.add_argument("--nmax", type=float, default=1e6)
    sv.add_argument("--eps", type=float, default=1e-10)
    sv.add_argument("--max-iter", type=int, default=500000,
                    help="iteration cap per request (LSCM charts of 1e6 vertices need "
                         "~100k iterations to reach 1e-10)")
    sv.add_argument("--resident", type=int, default=0,
                    help="smoothing requests re-solve this many resident systems")
    sv.add_argument("--budget-gb", type=float, default=200.0)
    sv.add_argument("--max-active", type=int, default=8)
    args = ap.parse_args(argv)
    if args.cmd == "serving":
        out = run_serving(args)
    elif args.cmd == "atlas":
        out = run_atlas(args)
    elif args.cmd == "atlas-trace":
        out = run_atlas_trace(args)
    elif args.cmd == "multirhs":
        out = run_multirhs(args)
    elif args.cmd == "eigen":
        out = run_eigen(args)
    elif args.cmd == "ssor":
        out = run_ssor(args)
    elif args.cmd == "spmv-patterns":
        out = run_spmv_patterns(args)
    elif args.cmd == "mixed":
        out = run_mixed(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
