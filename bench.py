"""Benchmark: Lloyd iterations of the K-means hot path on one MI355X.

Workload (default): N=10,000,000 points, D=128, K=1024 centres, Gaussian
mixture (1024 means U[-5,5]^128, sigma U[0.5,2.0], seed 1), K sampled
initial centres; points generated on the GPU so they are resident in HBM
before the timed region.  One step = one full Lloyd iteration (assign,
re-resolve for reduced precision, canonical FP64 update, compactness).

Prints one JSON line: value = N*K*steps / elapsed (point-centre pairs per
second), plus a roofline object for the assign kernel and the CPU baseline
(the oracle on a bounded sample of the same points, timed on this host).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp32": 157.3e12, "bf16": 2516.6e12, "fp16": 2516.6e12}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", default="bf16", choices=["fp32", "bf16", "fp16", "fp64"])
    ap.add_argument("--n", type=int, default=None, help="override N (smaller runs)")
    ap.add_argument("--cpu-sample", type=int, default=0, help="points in the CPU baseline sample (0: all)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32-ref", action="store_true")
    ap.add_argument("--full-gemm", action="store_true",
                    help="fp32 only: assign through rocBLAS SGEMM + row argmin (the reference's "
                         "doFullSGEMM = 1 algorithm), a comparison point for the fused kernels")
    ap.add_argument("--calib", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy (PMC byte-counter calibration)")
    ap.add_argument("--proposer", default=None, choices=["none", "mlp", "centroid"],
                    help="candidate proposer (default: mlp for c5, none otherwise)")
    ap.add_argument("--groups", type=int, default=128)
    ap.add_argument("--top-groups", type=int, default=4)
    ap.add_argument("--train-iters", type=int, default=3)
    ap.add_argument("--hidden", type=int, default=256)
    ap.add_argument("--tile-rows", type=int, default=0,
                    help="stream the points from pinned host memory through HBM in tiles of this "
                         "many rows (a multiple of 2^23), the reference's mini-batch tiling; the "
                         "timed iterations then include every host-to-device copy")
    args = ap.parse_args()
    if args.proposer is None:
        args.proposer = "mlp" if args.workload == "c5" else "none"
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus > 1 is not supported "
                                   "(the multi-GPU sharding is a design note, not built)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import kmeans_amd as km
    from kmeans_amd import datagen

    torch.cuda.init()
    if args.workload == "c4":
        print(json.dumps(bench_batch(args, km, datagen, np, torch)))
        return
    X, C0, spec = datagen.generate_on_gpu(args.workload, n=args.n, torch=torch)
    n, d, k = spec["n"], spec["d"], spec["k"]

    if args.full_gemm and args.precision not in ("fp32", "fp64"):
        ap.error("--full-gemm needs --precision fp32 or fp64")
    if args.precision == "fp64":
        print(json.dumps(bench_fp64(args, X, C0, spec, km, np)))
        return
    if args.tile_rows > 0:
        print(json.dumps(bench_streamed(args, X, C0, spec, km, np, torch)))
        return

    prep_ms = {}

    def run(precision, steps, warmup, full_gemm=False, timing=True):
        e = km.KMeans(n, d, k, precision=precision, timing=True, full_gemm=full_gemm)
        e.set_data(device_ptr=X.data_ptr())
        e.set_centers(C0)
        if warmup:
            e.run(warmup)
            e.synchronize()
            e.set_centers(C0)
        # fresh stats for the timed run; its per-data-set preparation (the
        # reduced-precision copy of the points and their norm bounds) is timed
        # on its own and is not part of the iterations
        e2 = km.KMeans(n, d, k, precision=precision, timing=timing, full_gemm=full_gemm)
        e.close()
        torch.cuda.synchronize()
        tp = time.perf_counter()
        e2.set_data(device_ptr=X.data_ptr())
        e2.set_centers(C0)
        e2.synchronize()
        prep_ms[precision] = (time.perf_counter() - tp) * 1e3
        if not timing and warmup:
            # graph launches: the engine's iteration graphs (one per update
            # parity and reuse state) are captured and instantiated by a
            # warm-up on the timed engine itself, as a caller's first run would
            e2.run(max(warmup, 3))
            e2.set_centers(C0)
            e2.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e2.run(steps)
        e2.synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return e2, dt

    if args.proposer != "none":
        print(json.dumps(bench_proposer(args, X, C0, spec, km, np, torch)))
        return

    # FP32 over one canonical block: the timed iterations are graph launches
    # (timing off); the per-phase times come from an eagerly launched rerun
    graphs = args.precision == "fp32" and not args.full_gemm and n <= km.KM_BLOCK
    eng, dt = run(args.precision, args.steps, args.warmup, args.full_gemm, timing=not graphs)
    graph_info = None
    if graphs:
        gl = eng.stats()["graph_launches"]
        eng.close()
        eng, dt_eager = run(args.precision, args.steps, 0)
        graph_info = {"graph_launches_in_timed_run": gl - max(args.warmup, 3) if args.warmup else gl,
                      "eager_ms_per_step": dt_eager * 1e3 / args.steps,
                      "note": "ms_per_step: each iteration one hipGraph launch (assign, update, "
                              "compactness, slot commit), graphs instantiated in the warm-up; "
                              "phases_ms_per_step and the roofline from an eagerly launched "
                              "rerun of the same iterations with HIP events"}
    st = eng.stats()
    pairs = float(n) * k
    value = pairs * args.steps / dt
    assign_ms = st["ms_assign"] / args.steps
    flops = 2.0 * d * pairs
    if args.precision != "fp32" and st["assign_launches"]:
        # the reduced-precision assign kernel itself: flops of the points it
        # scored (each against all K centres) over its own event-timed
        # launches; the drift bounds and the pair test leave the rest unscored
        kernel_ms = st["ms_assign_kernel"] / st["assign_launches"]
        kernel_flops = 2.0 * d * k * st["scored"] / st["assign_launches"]
    else:
        # the FP32 assign phase is the one kernel, over every point
        kernel_ms, kernel_flops = assign_ms, flops
    achieved = kernel_flops / (kernel_ms * 1e-3)
    prof = measured_profile(args.workload, args.precision, args.n)
    out = {
        "metric": "assign pts*ctrs/s per Lloyd iter",
        "value": value,
        "unit": "pts*ctrs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision if args.precision != "fp32" else "f32",
        "data": "synthetic (Gaussian mixture generated on the GPU)",
        "config": {"workload": f"{args.workload}: N={n} D={d} K={k} Lloyd, {args.precision} assign"
                               + (" by library GEMM + row argmin" if args.full_gemm else ""),
                   "N": n, "D": d, "K": k, "precision": args.precision},
        "phases_ms_per_step": {p: st[f"ms_{p}"] / args.steps for p in
                               ("assign", "reresolve", "update", "compactness")},
        "assign_pairs_per_s": pairs / (assign_ms * 1e-3),
        "reresolved_per_step": st["reresolved"] / args.steps,
        "rescored_per_step": st["rescored"] / args.steps,
        "deep_per_step": st["deep"] / args.steps,
        "bounded_per_step": st["bounded"] / args.steps,
        "paired_per_step": st["paired"] / args.steps,
        "prep_ms_excluded": prep_ms.get(args.precision),
        **({"graphs": graph_info} if graph_info else {}),
        "timing_scope": "ms_per_step covers the Lloyd iterations only (assign, re-resolve, update, "
                        "compactness) with the points resident in HBM; the per-data-set "
                        "preparation (reduced-precision copy of the points and their norm bounds, "
                        "prep_ms_excluded) runs once before the timed region",
        "assign_tiles_frac": (st["tiles"] / st["tiles_total"]) if st["tiles_total"] else None,
        "roofline": {"bound": "mfma",
                     "achieved": achieved / 1e12, "peak": PEAK[args.precision] / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK[args.precision],
                     "traffic": None if args.full_gemm or prof is None else prof["bytes_per_launch"],
                     "mfma_busy": None if args.full_gemm or prof is None else prof.get("mfma_util"),
                     "kernel": ("rocBLAS SGEMM + k_row_transform_min" if args.full_gemm else
                                "k_assign_mfma" if args.precision != "fp32" else "k_assign_f32_mfma"),
                     "launches": st["assign_launches"] if args.precision != "fp32" else args.steps,
                     "kernel_ms_per_launch": kernel_ms,
                     "flops_per_launch": kernel_flops,
                     "algorithmic": ("2*D*K flops per point the kernel scores (against all K "
                                     "centres), averaged over its launches in the timed "
                                     "iterations, over its own HIP-event-timed durations on "
                                     "the engine's stream" if args.precision != "fp32" else
                                     "2*D*N*K flops per launch over the assign phase's "
                                     "event-timed duration (the phase is this one kernel)"),
                     "effective_dense": {
                         "achieved": flops / (assign_ms * 1e-3) / 1e12, "unit": "TFLOP/s",
                         "frac": flops / (assign_ms * 1e-3) / PEAK[args.precision],
                         "note": "dense equivalent: 2*D*N*K per iteration over the whole assign "
                                 "phase (drift-bound filter, pair decisions and the kernel)"}},
    }
    labels = eng.labels()
    centers = eng.centers()
    eng.close()

    if args.precision != "fp32" and not args.no_fp32_ref:
        ref, dt32 = run("fp32", args.steps, 0)
        st32 = ref.stats()
        lab32 = ref.labels()
        out["fp32_value"] = pairs * args.steps / dt32
        out["fp32_assign_ms"] = st32["ms_assign"] / args.steps
        out["speedup_assign_vs_fp32"] = out["fp32_assign_ms"] / (
            (st["ms_assign"] + st["ms_reresolve"]) / args.steps)
        out["label_mismatches_vs_fp32"] = int((lab32 != labels).sum())
        out["centers_equal_fp32"] = bool(np.array_equal(ref.centers(), centers))
        ref.close()

    if not args.no_cpu:
        # CPU baseline: one full Lloyd iteration of the oracle (assign, update,
        # compactness) on the first `s` points, and the same iteration on the
        # GPU's plain FP32 path for a full-size bit-exact parity check.
        from oracle import oracle as orc
        orc.use_native()
        s = min(args.cpu_sample or n, n)
        Xs = X[:s].cpu().numpy()
        t0 = time.perf_counter()
        lab_cpu = orc.assign(Xs, C0)
        C1, _, _, _ = orc.update(Xs, lab_cpu, C0)
        comp_cpu = orc.compactness(Xs, lab_cpu, C1)
        t_cpu = time.perf_counter() - t0
        e1 = km.KMeans(s, d, k, precision="fp32")
        e1.set_data(device_ptr=X.data_ptr())
        e1.set_centers(C0)
        e1.run(1)
        parity = (bool(np.array_equal(lab_cpu, e1.labels())) and bool(np.array_equal(C1, e1.centers()))
                  and e1.stats()["compactness"] == comp_cpu)
        e1.close()
        # error of the reduced-precision run's final centres and compactness
        # against FP64 means and FP64 compactness of its own final labels
        Xh = X.cpu().numpy()
        M64, cnt64 = orc.means_fp64(Xh, labels, k)
        nz = cnt64 > 0
        rel_c = float(np.max(np.abs(centers[nz].astype(np.float64) - M64[nz]) /
                             np.maximum(np.abs(M64[nz]), 1e-30)))
        comp64 = orc.compactness_fp64(Xh, labels, centers)
        out["error_vs_fp64"] = {
            "centers_max_rel": rel_c,
            "centers_max_abs_over_max_abs": float(np.max(np.abs(centers[nz] - M64[nz])) /
                                                  np.max(np.abs(M64[nz]))),
            "compactness_rel": abs(st["compactness"] - comp64) / abs(comp64),
            "reference": "FP64 means and FP64 compactness (CPU oracle) of the final labels",
        }
        del Xh
        out["cpu_baseline"] = {
            "value": s * float(k) / t_cpu, "unit": "pts*ctrs/s", "cores": orc.num_threads(),
            "kind": "port", **cpu_info(orc),
            "sample": f"one full Lloyd iteration (assign + canonical update + compactness) of the "
                      f"CPU oracle on the first {s} points of the same workload, same initial centres",
            "seconds": t_cpu,
            "gpu_fp32_iteration_bit_exact": parity,
        }
    if args.calib:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        out["calibration_copy_bytes"] = 1 << 30
    print(json.dumps(out))


def bench_fp64(args, X, C0, spec, km, np):
    """computeKmeansD on the same workload in double: the points are copied to
    the host as double and the call's wall clock covers its own host-to-device
    copy and set-up, so per-iteration kernel times come from a kernel trace
    (two calls, `--warmup` and `--warmup + --steps` iterations, differenced)."""
    n, d, k = spec["n"], spec["d"], spec["k"]
    Xh = X.double().cpu().numpy()
    C0d = np.ascontiguousarray(C0.astype(np.float64).T)

    def call(iters):
        t0 = time.perf_counter()
        st, res = km.compute_kmeans_d(Xh, k, -1.0, iters, result=C0d.copy(), init_from_result=True,
                                      full_gemm=args.full_gemm)
        return time.perf_counter() - t0, st, res

    t_w, _, _ = call(max(args.warmup, 1))
    t_all, st, res = call(max(args.warmup, 1) + args.steps)
    dt = (t_all - t_w) / args.steps
    return {
        "metric": "assign pts*ctrs/s per Lloyd iter", "value": float(n) * k / dt,
        "unit": "pts*ctrs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (Gaussian mixture generated on the GPU, as double)",
        "config": {"workload": f"{args.workload}: N={n} D={d} K={k} Lloyd in double (computeKmeansD"
                               + (", doFullSGEMM = 1)" if args.full_gemm else ")"),
                   "N": n, "D": d, "K": k, "precision": "fp64"},
        "timing": "difference of two computeKmeansD calls (set-up and copies cancel)",
        "status": int(st),
    }


def cpu_info(orc):
    """The host the CPU baseline ran on: model, logical CPUs available to the
    process, OpenMP threads used and the oracle build's -march."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    return {"cpu_model": model, "nproc": nproc, "omp_threads": orc.num_threads(),
            "march": orc.march()}


def bench_streamed(args, X, C0, spec, km, np, torch):
    """The points in pinned host memory, streamed through HBM in tiles of
    --tile-rows rows every iteration (copies on their own stream, double
    buffered, overlapped with the previous tile's assign and sums).  The
    timed iterations include every copy."""
    n, d, k = spec["n"], spec["d"], spec["k"]
    Xp = torch.empty((n, d), dtype=torch.float32, pin_memory=True)
    Xp.copy_(X)
    del X
    torch.cuda.empty_cache()
    e = km.KMeans(n, d, k, precision=args.precision, timing=True, tile_rows=args.tile_rows)
    e.set_data(host_ptr=Xp.data_ptr())
    e.set_centers(C0)
    if args.warmup:
        e.run(args.warmup)
        e.synchronize()
        e.set_centers(C0)
    e.close()
    e = km.KMeans(n, d, k, precision=args.precision, timing=True, tile_rows=args.tile_rows)
    e.set_data(host_ptr=Xp.data_ptr())
    e.set_centers(C0)
    e.synchronize()
    t0 = time.perf_counter()
    e.run(args.steps)
    e.synchronize()
    dt = time.perf_counter() - t0
    st = e.stats()
    e.close()
    pairs = float(n) * k
    xbytes = float(n) * d * 4
    return {
        "metric": "assign pts*ctrs/s per Lloyd iter", "value": pairs * args.steps / dt,
        "unit": "pts*ctrs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic (Gaussian mixture generated on the GPU, then held in pinned host memory)",
        "config": {"workload": f"{args.workload}: N={n} D={d} K={k} Lloyd, {args.precision} assign, "
                               f"points streamed from pinned host memory in tiles of "
                               f"{args.tile_rows} rows",
                   "N": n, "D": d, "K": k, "precision": args.precision, "tile_rows": args.tile_rows},
        "host_to_device_GBps": xbytes * args.steps / dt / 1e9,
        "phases_ms_per_step": {p: st[f"ms_{p}"] / args.steps for p in
                               ("assign", "reresolve", "update", "compactness")},
        "timing_scope": "wall clock of the iterations, every tile's host-to-device copy included; "
                        "host_to_device_GBps = bytes of the points per iteration / iteration time",
    }


def measured_profile(workload, precision, n_override):
    """Counters of the dominant kernel from the committed PMC summary
    (profiles/traffic.json): HBM bytes per launch (FETCH_SIZE and WRITE_SIZE
    runs of this workload, scaled by the calibration copy) and the MFMA
    utilisation (SQ_VALU_MFMA_BUSY_CYCLES over the SIMD-cycles of the launch),
    averaged over the launches of the timed iterations; None when not
    measured."""
    if n_override is not None:
        return None
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        tab = json.load(open(path))
    except (OSError, ValueError):
        return None
    key = f"{workload}:{precision}:" + ("k_assign_mfma" if precision != "fp32" else "k_assign_f32_mfma")
    return tab.get(key)


def bench_batch(args, km, datagen, np, torch):
    """2048 independent clusterings (D=16, K=256, lognormal N, first-K initial
    centres) as one batched engine: every iteration is one persistent assign
    launch and one persistent update launch, the run one hipGraph.  Compared
    with running the single-problem engine once per instance (timed on a
    sample of instances, scaled by their share of the points)."""
    spec = datagen.C4
    d, k = spec["d"], spec["k"]
    X, n, seeds = datagen.generate_c4_on_gpu(torch, nb=args.n)
    nb = len(n)
    total = int(n.sum())
    steps = args.steps
    prec = args.precision if args.precision in ("fp32", "bf16") else "bf16"
    b = km.KMeansBatch(n, d, k, precision=prec)
    b.set_data(device_ptr=X.data_ptr())
    b.init_first_k()
    if args.warmup:
        b.run(args.warmup)  # also autotunes the tile size
        b.init_first_k()
    st0 = b.stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st = b.stats()
    for f in ("kept", "paired", "scored", "exact_searched"):  # the timed run's own counts
        st[f] -= st0[f]
    # the same trajectory once more on an engine with timing on (enqueued
    # eagerly, HIP events around every assign and update launch) for the
    # kernels' own durations
    bt = km.KMeansBatch(n, d, k, precision=prec, timing=True)
    bt.set_data(device_ptr=X.data_ptr())
    bt.set_tile_rows(st["tile_rows"])
    bt.init_first_k()
    bt.run(1)  # first touch of the engine's buffers, as the warm-up did for the graph run
    bt.init_first_k()
    bt.run(steps)
    stt = bt.stats()
    bt.close()
    st["ms_assign"], st["ms_update"] = stt["ms_assign"], stt["ms_update"]
    pairs = float(total) * k
    out = {
        "metric": "assign pts*ctrs/s per Lloyd iter",
        "value": pairs * steps / dt,
        "unit": "pts*ctrs/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 (labels certified against f32)" if prec == "bf16" else "f32",
        "data": "synthetic (per-instance Gaussian mixtures generated on the GPU)",
        "config": {"workload": f"c4: {nb} instances, D={d} K={k}, lognormal N (median 20000, "
                               f"sigma 1, clipped [1000, 200000]), sum N={total}, first-K init, "
                               f"{prec} assign, one hipGraph per run",
                   "instances": nb, "D": d, "K": k, "sum_N": total, "precision": prec},
        "graph_ms_per_step": st["ms_graph"] / steps,
        "phases_ms_per_step": {"assign": st["ms_assign"] / steps, "update": st["ms_update"] / steps,
                               "note": "kernel durations from a second, eagerly enqueued run of "
                                       "the same iterations with HIP events"},
        "batch_stats": st,
    }
    # the dominant kernel, the batched assign, on the HBM roofline (D=16 puts
    # its intensity far below the BF16 ridge): the bytes it must move per
    # launch -- the FP32 rows of the points it scores or decides between two
    # candidates, and per point its label and drift bounds (read, and written
    # back) -- over its event-timed launches in the graph
    b0 = total * (4 + 16 + 16)
    rows_read = st["scored"] + st["paired"]
    if prec == "fp32":
        kbytes = (4.0 * d + 4.0) * total * steps
    else:
        kbytes = float(b0) * steps + 4.0 * d * rows_read
    out["roofline"] = {
        "bound": "hbm", "achieved": kbytes / (st["ms_assign"] * 1e-3) / 1e9, "peak": 8000.0,
        "unit": "GB/s", "frac": kbytes / (st["ms_assign"] * 1e-3) / 8e12, "traffic": None,
        "kernel": "k_batch_assign_bf16" if prec != "fp32" else "k_batch_assign_f32x",
        "kernel_ms_per_launch": st["ms_assign"] / steps,
        "bytes_per_launch": kbytes / steps,
        "algorithmic": ("per point 36 B (label, drift bounds read and written) plus 4*D B of FP32 "
                        "row for each point scored or decided between two candidates (counts from "
                        "the stats of the timed run); over the assign launches' durations (HIP "
                        "events around each launch in an eagerly enqueued rerun of the same "
                        "iterations)" if prec != "fp32" else
                        "(4*D + 4) B per point per launch (rows read, labels written)"),
        "mfma_flops_per_s": 2.0 * d * k * (st["scored"] if prec != "fp32" else total * steps) /
                            (st["ms_assign"] * 1e-3)}
    labels = b.labels()
    centers_b = b.centers()
    b.close()
    if prec != "fp32":
        # the plain FP32 batched engine from the same centres: labels must match
        b32 = km.KMeansBatch(n, d, k, precision="fp32")
        b32.set_data(device_ptr=X.data_ptr())
        b32.init_first_k()
        b32.run(1)
        b32.init_first_k()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b32.run(steps)
        torch.cuda.synchronize()
        out["fp32_batch_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
        out["label_mismatches_vs_fp32"] = int((b32.labels() != labels).sum())
        out["centers_equal_fp32"] = bool(np.array_equal(b32.centers(), centers_b))
        b32.close()
    # the method without batching: one single-problem engine per instance
    sample = np.linspace(0, nb - 1, min(nb, 48)).astype(int)
    off = np.concatenate([[0], np.cumsum(n)])
    t_seq = 0.0
    mism = 0
    for i in sample:
        Xi = X[off[i]:off[i + 1]]
        e = km.KMeans(int(n[i]), d, k)
        e.set_data(device_ptr=Xi.data_ptr())
        e.set_centers(Xi[:k].cpu().numpy())
        e.synchronize()
        ta = time.perf_counter()
        e.run(steps)
        e.synchronize()
        t_seq += time.perf_counter() - ta
        mism += int((e.labels() != labels[off[i]:off[i + 1]]).sum())
        e.close()
    share = float(n[sample].sum()) / total
    out["per_instance_engine"] = {
        "sample_instances": len(sample), "sample_points": int(n[sample].sum()),
        "ms_per_step_extrapolated": t_seq / share * 1e3 / steps,
        "label_mismatches_vs_batch": mism,
    }
    out["speedup_vs_per_instance_engine"] = out["per_instance_engine"]["ms_per_step_extrapolated"] / out["ms_per_step"]
    if not args.no_cpu:
        from oracle import oracle as orc
        orc.use_native()
        cs = sample[:8]
        t_cpu = 0.0
        for i in cs:
            Xi = X[off[i]:off[i + 1]].cpu().numpy()
            ta = time.perf_counter()
            orc.lloyd(Xi, Xi[:k].copy(), 1)
            t_cpu += time.perf_counter() - ta
        pts = float(n[cs].sum())
        out["cpu_baseline"] = {"value": pts * k / t_cpu, "unit": "pts*ctrs/s",
                               "cores": orc.num_threads(), "kind": "port", **cpu_info(orc),
                               "sample": f"one full Lloyd iteration of the CPU oracle on {len(cs)} "
                                         f"instances ({int(pts)} points), first-K centres"}
    return out


def bench_proposer(args, X, C0, spec, km, np, torch):
    """Lloyd with the candidate proposer: train_iters plain iterations, grouping
    and MLP training, then proposer iterations (exact FP32 scores within the
    proposed groups, certified, full FP32 search for the rest).  Compared with
    full-search runs (FP32 and the reduced-precision path) from the same
    initial centres; labels must equal the FP32 run's."""
    n, d, k = spec["n"], spec["d"], spec["k"]
    steps, ti = args.steps, args.train_iters
    assert steps > ti, "steps must exceed train_iters"
    kw = dict(proposer=args.proposer, groups=args.groups, top_groups=args.top_groups,
              train_iters=ti, hidden=args.hidden)

    def fresh(precision, **extra):
        e = km.KMeans(n, d, k, precision=precision, timing=True, **extra)
        e.set_data(device_ptr=X.data_ptr())
        e.set_centers(C0)
        e.synchronize()
        torch.cuda.synchronize()
        return e

    if args.warmup:
        w = fresh(args.precision)
        w.run(args.warmup)
        w.synchronize()
        w.close()
    e = fresh(args.precision, **kw)
    t0 = time.perf_counter()
    e.run(ti)
    e.synchronize()
    t1 = time.perf_counter()
    st_plain = e.stats()
    e.run(steps - ti)
    e.synchronize()
    t2 = time.perf_counter()
    st = e.stats()
    labels = e.labels()
    centers = e.centers()
    e.close()
    nl = steps - ti
    learned_assign_ms = ((st["ms_assign"] - st_plain["ms_assign"]) +
                         (st["ms_reresolve"] - st_plain["ms_reresolve"])) / nl
    learned_iter_ms = (st["ms_total"] - st_plain["ms_total"]) / nl
    pairs = float(n) * k
    total = t2 - t0
    out = {
        "metric": "assign pts*ctrs/s per Lloyd iter",
        "value": pairs * steps / total,
        "unit": "pts*ctrs/s",
        "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": total * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (exact candidate scores) + bf16 MLP",
        "data": "synthetic (Gaussian mixture generated on the GPU)",
        "config": {"workload": f"{args.workload}: N={n} D={d} K={k} Lloyd, {args.proposer} proposer, "
                               f"{args.groups} groups, top {args.top_groups}, trained on "
                               f"{ti} plain iterations",
                   "N": n, "D": d, "K": k, "precision": args.precision, "proposer": args.proposer},
        "plain_iter_ms": st_plain["ms_total"] / ti,
        "learned_iter_ms": learned_iter_ms,
        "learned_assign_ms": learned_assign_ms,
        "learned_assign_pairs_per_s": pairs / (learned_assign_ms * 1e-3),
        "train_ms": st["ms_train"],
        "train_loss": [st["train_loss_first"], st["train_loss_last"]],
        "proposer_stats": {f: st[f] for f in ("proposed", "proposal_hits", "certified",
                                              "certified_group", "fallback", "candidate_pairs")},
        "wall_ms": {"plain_iters": (t1 - t0) * 1e3, "train_and_learned_iters": (t2 - t1) * 1e3},
    }
    out["proposer_stats"]["candidate_pairs_fraction_of_full"] = st["candidate_pairs"] / (pairs * nl)
    for prec in ([] if args.no_fp32_ref else ["fp32", "bf16"]):
        r = fresh(prec)
        ta = time.perf_counter()
        r.run(steps)
        r.synchronize()
        tb = time.perf_counter()
        sr = r.stats()
        out[f"full_{prec}"] = {"iter_ms": (tb - ta) * 1e3 / steps,
                               "assign_ms": (sr["ms_assign"] + sr["ms_reresolve"]) / steps}
        if prec == "fp32":
            lab32 = r.labels()
            out["label_mismatches_vs_fp32"] = int((lab32 != labels).sum())
            out["centers_equal_fp32"] = bool(np.array_equal(r.centers(), centers))
        r.close()
    if "full_fp32" in out:
        out["speedup_assign_vs_fp32"] = out["full_fp32"]["assign_ms"] / learned_assign_ms
    return out


if __name__ == "__main__":
    main()
