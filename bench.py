#!/usr/bin/env python3
"""Hash + match throughput on one MI355X.

One step = hash B synthetic 512x512 u8 images resident in HBM (fused kernel:
tent compression, DCT, quality, median quantisation, 8 dihedral hashes), then
match the 8B dihedral hashes against a 10M-row database resident in HBM at
threshold 31 (exact (query, index, distance) triples into a device buffer),
then read the hit count back -- one library call per step
(yume_pdq_amd_hash_match_batch), the count stored into page-locked host
memory by the GPU.  The database carries 50,000 planted
near-duplicates (U{0..40} distinct bit flips) of the 8192 dihedral hashes of
1024 images; a batch of B images is the first B of them.

Prints one JSON line.  `value` is images/s at the headline batch (128 images
= 1024 queries per pass, the largest batch whose step p99 stays under 2 ms);
the other batch sizes, the plain unbatched f32 path, step latency
percentiles, the roofline of the step's dominant kernel and of the others
(each timed as 20-50 back-to-back launches between two HIP events on the
launch stream), the HBM traffic measured by rocprofv3 counters (read from
profiles/hbm_traffic_counters.md) and a CPU baseline (the reference's CPU
kernel + matcher restated in C, timed on this host's cores) ride along.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "yume-pdq_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

METRIC = "hash+match images/s @10M DB thr31; p50/p99 latency; % HBM roofline"
HBM_PEAK_GBS = 8000.0
PLANT_POOL_IMAGES = 1024
MFMA_FP4_PEAK_TFLOPS = 10100.0  # MI355X dense fp4 matrix peak
COUNTER_SUMMARY = os.path.join(REPO, "profiles", "hbm_traffic_counters.md")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def percentile(xs, p):
    import numpy as np
    return float(np.percentile(np.asarray(xs), p))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--db-rows", type=int, default=10_000_000)
    ap.add_argument("--plants", type=int, default=50_000)
    ap.add_argument("--batches", default="1,16,128,1024")
    ap.add_argument("--headline-batch", type=int, default=128)
    ap.add_argument("--in-flight", type=int, default=2,
                    help="steps in flight in the throughput loop, each on its own stream")
    ap.add_argument("--latency-steps", type=int, default=2000,
                    help="steps of the separate pass the headline p50/p99 come from")
    ap.add_argument("--ramp-ms", type=float, default=400.0,
                    help="untimed steps run first so the GPU reaches its sustained clock")
    ap.add_argument("--threshold", type=int, default=31)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-images", type=int, default=8192)
    ap.add_argument("--verify", action="store_true", help="check B=1 hits against numpy (slow)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop sections")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batching section")
    ap.add_argument("--no-hash-precisions", action="store_true",
                    help="skip the batched-hashing section (4096 images, every DCT precision)")
    ap.add_argument("--hash-images", type=int, default=4096)
    ap.add_argument("--no-pipe", action="store_true", help="skip the stream-ingestion section")
    ap.add_argument("--pipe-frames", type=int, default=4096)
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-offered-10m", default="30000,50000,67000,80000,90000,100000,150000",
                    help="offered images/s of the 10M-row serving runs (comma-separated)")
    ap.add_argument("--serve-offered-100m", default="8000,16000,20000",
                    help="offered images/s of the 100M-row serving runs (comma-separated)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X and has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import yume_pdq
    from yume_pdq import synthetic

    yume_pdq.lib()  # fails loudly if the HIP library is not built
    dev = torch.device("cuda:0")
    info = yume_pdq.device_info(0)
    batches = [int(b) for b in args.batches.split(",")]
    bmax = max(batches + [args.headline_batch])
    thr = args.threshold

    # Plants are drawn from the dihedral hashes of 1024 images whatever batches
    # run (smaller batches use prefixes), so every run sees the same database.
    n_imgs = max(bmax, PLANT_POOL_IMAGES)
    log(f"[bench] device {info}; generating {n_imgs} images on the GPU")
    planes_u8 = synthetic.images_u8_device(n_imgs, dev)
    planes_f32 = planes_u8[:1].to(torch.float32).contiguous()
    torch.cuda.synchronize()

    # queries of the whole image set (used to plant the database)
    h_all, _, _ = yume_pdq.hash_batch(planes_u8[:PLANT_POOL_IMAGES])
    queries_all = h_all.reshape(-1, 32).cpu().numpy()
    log(f"[bench] building {args.db_rows} row database with {args.plants} plants")
    db = synthetic.random_db(args.db_rows, seed=7)
    synthetic.plant_fast(db, queries_all, args.plants, seed=11)
    d_db = torch.from_numpy(db).to(dev)
    # The database's prefix plane (the first 16 bytes of every row, packed) is
    # built once with the database, as part of loading it into HBM.
    d_plane = yume_pdq.build_prefix_plane(d_db)
    torch.cuda.synchronize()

    cap = 1 << 22
    d_hits = torch.empty((cap, 16), dtype=torch.uint8, device=dev)
    d_count = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    # one stream per in-flight step (default priority, as every stream here)
    in_flight = max(1, args.in_flight)
    streams = [stream] + [torch.cuda.Stream(device=dev) for _ in range(in_flight - 1)]

    hm_cache = {}

    def hash_match_for(planes, B, plane, slot=0):
        """The step's one-call object for (planes, B, plane); slot k runs on
        stream k."""
        key = (planes.data_ptr(), planes.dtype, B, plane is not None, slot)
        if key not in hm_cache:
            hm_cache[key] = yume_pdq.HashMatch(planes[:B], d_db, thr, plane=plane, capacity=1 << 20,
                                               stream=streams[slot])
        return hm_cache[key]

    def step(planes, B, ev=None, plane=d_plane):
        """One batch through the one-call path (yume_pdq_amd_hash_match_batch:
        hash kernel, counter reset, matcher back to back, the hit count stored
        into page-locked host memory); the step ends when the count is on the
        host."""
        hm = hash_match_for(planes, B, plane)
        if ev is not None:
            ev[0].record(stream)
        hm.run(sync=False)
        if ev is not None:
            ev[1].record(stream)
        stream.synchronize()
        return int(hm.h_count[0])

    out_bufs = {}
    for B in set(batches + [args.headline_batch, 1]):
        out_bufs[B] = (torch.empty((B, 8, 32), dtype=torch.uint8, device=dev),
                       torch.empty((B, 8), dtype=torch.float32, device=dev),
                       torch.empty((B, 8), dtype=torch.float32, device=dev), None)

    def run(planes, B, plane=d_plane):
        for _ in range(args.warmup):
            step(planes, B, plane=plane)
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(args.steps)]
        lat = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hits = 0
        for k in range(args.steps):
            s0 = time.perf_counter()
            hits = step(planes, B, evs[k], plane=plane)
            lat.append((time.perf_counter() - s0) * 1e3)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        t_gpu = [e[0].elapsed_time(e[1]) for e in evs]
        sync_s = t1 - t0
        # The same K steps with F in flight (default two; three and four
        # measured slower), each slot on its own stream: step k + 1 is
        # enqueued before the host waits for step k's hit count, so the
        # host's wait, the count read and the next call's enqueue overlap the
        # GPU work, and the GPU may start step k + 1's kernels while step k's
        # last blocks finish (each step has its own outputs and record area).
        # This is the throughput; the latency pass below runs steps one at a
        # time.
        F = in_flight
        hms = [hash_match_for(planes, B, plane, slot=k) for k in range(F)]
        done = [torch.cuda.Event() for _ in range(F)]
        counts = []
        for rep in range(2):  # an untimed round, then the timed one
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(args.steps):
                if k >= F:  # step k - F, on this slot, is done: its count is read first
                    done[k % F].synchronize()
                    counts.append(int(hms[k % F].h_count[0]))
                hms[k % F].run(sync=False)
                done[k % F].record(streams[k % F])
            for k in range(max(0, args.steps - F), args.steps):
                done[k % F].synchronize()
                counts.append(int(hms[k % F].h_count[0]))
            torch.cuda.synchronize()
            t1 = time.perf_counter()
        assert all(c == hits for c in counts), "pipelined steps found different hit counts"
        total_s = t1 - t0
        # the step's two kernels timed apart, through the separate entry points
        queries = hash_match_for(planes, B, plane).hashes.reshape(-1, 32)
        t_hash = kernel_ms(lambda: yume_pdq.hash_batch(planes[:B], stream=stream, out=out_bufs[B]), k=10)
        t_match = kernel_ms(lambda: match_appending(queries, plane), k=10)
        return {
            "batch_images": B,
            "queries_per_pass": 8 * B,
            "images_per_s": B * args.steps / total_s,
            "ms_per_step": total_s * 1e3 / args.steps,
            "images_per_s_one_step_at_a_time": B * args.steps / sync_s,
            "ms_per_step_one_step_at_a_time": sync_s * 1e3 / args.steps,
            "latency_ms": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
            "step_gpu_ms": float(np.mean(t_gpu)),
            "hash_kernel_ms": t_hash,
            "match_kernel_ms": t_match,
            "hits": hits,
        }

    # Kernel durations (kernel_ms).  The match calls append to the hit
    # counter, so no reset kernel runs between the timed launches.
    def kernel_ms(fn, k=20):
        """Average GPU duration of one launch of fn on `stream`: k launches
        back to back between one pair of events.  About 20 ms of full-size
        matching is queued first (40 launches at 1024 queries, a few ms of
        host time), so the host has queued every timed launch before the GPU
        reaches them: they run back to back at the sustained clock, as in a
        kernel trace, and no host-side gap (tens of us of Python and ctypes
        per call, longer than a 25 us scan) falls between the events."""
        for _ in range(10):
            fn()
        torch.cuda.synchronize()
        r0, r1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(40):
            match_appending(q_head, d_plane)
        r0.record(stream)
        for _ in range(k):
            fn()
        r1.record(stream)
        r1.synchronize()
        return r0.elapsed_time(r1) / k

    q_head = out_bufs[args.headline_batch][0].reshape(-1, 32)
    q_one = out_bufs[1][0].reshape(-1, 32)

    def match_appending(q, plane):
        yume_pdq.match(d_db, q, thr, hits=d_hits, count=d_count, sync=False, stream=stream, plane=plane,
                       append=True)

    # The GPU raises its clock over the first tens of milliseconds of
    # sustained load; untimed steps first, so the timed ones see the
    # sustained rate.
    t_ramp = time.perf_counter()
    n_ramp = 0
    while (time.perf_counter() - t_ramp) * 1e3 < args.ramp_ms:
        step(planes_u8, args.headline_batch)
        n_ramp += 1
    log(f"[bench] clock ramp: {n_ramp} untimed steps in {args.ramp_ms:.0f} ms")

    results = {}
    for B in sorted(set(batches + [args.headline_batch])):
        log(f"[bench] batch {B}: {args.warmup} warmup + {args.steps} timed steps")
        results[B] = run(planes_u8, B)
        log(f"[bench]   {results[B]}")
    log("[bench] plain unbatched f32 path (B=1, f32 image, raw database rows)")
    plain = run(planes_f32, 1, plane=None)
    log(f"[bench]   {plain}")

    if args.verify:
        log("[bench] verifying B=1 hit count against numpy")
        h1, _, _ = yume_pdq.hash_batch(planes_u8[:1])
        exp = synthetic.brute_force_hits(db, h1.reshape(-1, 32).cpu().numpy(), thr)
        assert exp.shape[0] == results.get(1, plain)["hits"], "hit count mismatch"

    head = results[args.headline_batch]
    n_rows = args.db_rows
    B = args.headline_batch

    # Step latency from a pass of its own, long enough for a p99 that is not
    # just the maximum of the timed steps.
    log(f"[bench] latency pass: {args.latency_steps} steps at batch {B}")
    lat = []
    for _ in range(args.latency_steps):
        s0 = time.perf_counter()
        step(planes_u8, B)
        lat.append((time.perf_counter() - s0) * 1e3)
    latency = {"p50": round(percentile(lat, 50), 4), "p99": round(percentile(lat, 99), 4),
               "max": round(max(lat), 4), "steps": len(lat),
               "timed_steps": head["latency_ms"]}

    log("[bench] back-to-back kernel timing")
    d_count.zero_()
    yume_pdq.match(d_db, q_head, thr, hits=d_hits, count=d_count, sync=False, stream=stream, plane=d_plane)
    hits_head = int(d_count.item())
    t_match_head = kernel_ms(lambda: match_appending(q_head, d_plane))
    yume_pdq.match(d_db, q_one, thr, hits=d_hits, count=d_count, sync=False, stream=stream, plane=d_plane)
    hits_8 = int(d_count.item())
    t_scan8 = kernel_ms(lambda: match_appending(q_one, d_plane), k=50)
    t_scan8_raw = kernel_ms(lambda: match_appending(q_one, None), k=50)
    t_hash_head = kernel_ms(lambda: yume_pdq.hash_batch(planes_u8[:B], stream=stream, out=out_bufs[B]))
    t_hash_max = kernel_ms(lambda: yume_pdq.hash_batch(planes_u8[:bmax], stream=stream, out=out_bufs[bmax]))
    log(f"[bench]   match {8 * B}q {t_match_head:.4f} ms, scan 8q {t_scan8:.4f} ms "
        f"(raw rows {t_scan8_raw:.4f} ms), "
        f"hash B={B} {t_hash_head:.4f} ms, hash B={bmax} {t_hash_max:.4f} ms")

    counters = load_counter_summary()
    nq_head = 8 * B
    mfma_path = yume_pdq.match_uses_matrix_cores(nq_head, thr)
    # Matcher algorithmic bytes over the prefix plane: 16 B per database row,
    # plus, per hit, the rest of its row (16 B) and the hit record (16 B).
    # Over raw rows: 32 B per row + 16 B per hit.
    bytes_head = 16.0 * n_rows + 32.0 * hits_head
    bytes_8 = 16.0 * n_rows + 32.0 * hits_8
    bytes_8_raw = 32.0 * n_rows + 16.0 * hits_8
    scan8_gbs = bytes_8 / (t_scan8 * 1e-3) / 1e9
    scan8_raw_gbs = bytes_8_raw / (t_scan8_raw * 1e-3) / 1e9
    roofline_scan8 = {
        "kernel": f"pdq_match_plane_kernel (vector ALU, prefix plane): {n_rows} rows x 8 queries, thr {thr}",
        "bound": "hbm", "achieved": round(scan8_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(scan8_gbs / HBM_PEAK_GBS, 4),
        "traffic": counters.get("match_plane_x8", {}).get("traffic_bytes"),
        "algorithmic_bytes_per_launch": bytes_8, "avg_launch_ms": round(t_scan8, 5),
        "rows_per_s": n_rows / (t_scan8 * 1e-3),
    }
    roofline_scan8_raw = {
        "kernel": f"pdq_match_kernel (vector ALU, 32-byte rows): {n_rows} rows x 8 queries, thr {thr}",
        "bound": "hbm", "achieved": round(scan8_raw_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(scan8_raw_gbs / HBM_PEAK_GBS, 4),
        "traffic": counters.get("match_x8", {}).get("traffic_bytes"),
        "algorithmic_bytes_per_launch": bytes_8_raw, "avg_launch_ms": round(t_scan8_raw, 5),
        "rows_per_s": n_rows / (t_scan8_raw * 1e-3),
    }
    if mfma_path:
        # 128-bit prefix dot product of +-1 fp4 operands on the matrix cores:
        # 128 multiply-adds = 256 ops per (row, query) pair.
        ops = 256.0 * n_rows * nq_head
        achieved = ops / (t_match_head * 1e-3) / 1e12
        roofline_match = {
            "kernel": (f"matrix-core matcher call over the prefix plane: pdq_match_mfma_kernel (fp4 MFMA "
                       f"16x16x128 prefix filter, waves sharing rows) + pdq_match_resolve_kernel (exact "
                       f"256-bit recheck of candidate groups): {n_rows} rows x {nq_head} queries, thr {thr}"),
            "bound": "mfma", "achieved": round(achieved, 1), "peak": MFMA_FP4_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / MFMA_FP4_PEAK_TFLOPS, 4),
            "traffic": counters.get(f"match_plane_x{nq_head}", {}).get("traffic_bytes"),
            "algorithmic_ops_per_launch": ops, "avg_launch_ms": round(t_match_head, 5),
            "pairs_per_s": n_rows * nq_head / (t_match_head * 1e-3),
            "hbm_gbs": round(bytes_head / (t_match_head * 1e-3) / 1e9, 1),
        }
    else:
        gbs = bytes_head / (t_match_head * 1e-3) / 1e9
        roofline_match = {
            "kernel": (f"pdq_match_plane_kernel (vector ALU, prefix plane): {n_rows} rows x "
                       f"{nq_head} queries, thr {thr}"),
            "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": counters.get(f"match_plane_x{nq_head}", {}).get("traffic_bytes"),
            "algorithmic_bytes_per_launch": bytes_head, "avg_launch_ms": round(t_match_head, 5),
        }
    # Hash algorithmic bytes: the u8 plane (262,144 B) + 8 x (32 + 4 + 4) B out.
    hash_bytes = (262144.0 + 320.0) * B
    hash_gbs = hash_bytes / (t_hash_head * 1e-3) / 1e9
    # the automatic layout: a wave per image past 1024 images, sixteen waves
    # while the images fit one per CU, else eight
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    layout_head = ("pdq_hash_kernel<u8> (one wave per image)" if B > 1024 else
                   "pdq_hash_block_kernel<u8> (sixteen waves per image)" if B <= cus else
                   "pdq_hash_block_kernel<u8> (eight waves per image)")
    roofline_hash = {
        "kernel": f"{layout_head}: {B} images per launch",
        "bound": "hbm", "achieved": round(hash_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(hash_gbs / HBM_PEAK_GBS, 4),
        "traffic": counters.get(f"hash_u8_b{B}", {}).get("traffic_bytes"),
        "algorithmic_bytes_per_launch": hash_bytes, "avg_launch_ms": round(t_hash_head, 5),
        "images_per_s_kernel_only": B / (t_hash_head * 1e-3),
        f"images_per_s_kernel_only_b{bmax}": bmax / (t_hash_max * 1e-3),
    }
    dom_is_match = t_match_head >= t_hash_head
    roofline = roofline_match if dom_is_match else roofline_hash

    log("[bench] drop-in FFI: yume_pdq_hash_smart_kernel on a host f32 image")
    ffi_us = drop_in_ffi_us(yume_pdq, planes_u8[0].to(torch.float32).cpu().numpy())

    pool_u8 = planes_u8[:256].cpu().numpy()
    pool_hashes = h_all[:256].reshape(-1, 32).cpu().numpy()
    serving = None
    serving_summary = None
    if not args.no_serving:
        serving = [
            serving_section(args, yume_pdq, torch, dev, synthetic, d_db, None, pool_u8, thr,
                            offered=off, label=f"10M rows (the headline database), offered {off:.0f} images/s",
                            max_batch=128, max_contexts=4)
            for off in [float(x) for x in args.serve_offered_10m.split(",")]
        ]
        ok = [r for r in serving if r["latency_ms"]["p99"] < 2.0 and r["errors"] == 0]
        best = max(ok, key=lambda r: r["served_images_per_s"]) if ok else None
        serving_summary = {
            "best_served_at_p99_under_2ms": None if best is None else {
                "offered_images_per_s": best["offered_images_per_s"],
                "served_images_per_s": best["served_images_per_s"], "latency_ms": best["latency_ms"],
                "served_vs_plain": best["served_vs_plain"]},
            "note": "the highest served rate among the 10M-row runs whose p99 latency is under 2 ms",
        }
        del d_plane
        torch.cuda.empty_cache()
        # config 4's database: 100M rows (3.2 GB), 0.01 % planted, built once
        log("[serving] building 100000000 rows, 0.01 % planted")
        db100 = synthetic.random_db(100_000_000, seed=17)
        synthetic.plant_fast(db100, pool_hashes, 10_000, seed=18)
        d_db100 = torch.from_numpy(db100).to(dev)
        del db100
        runs100 = [
            serving_section(args, yume_pdq, torch, dev, synthetic, d_db100, None, pool_u8, thr, offered=off,
                            label=f"100M rows, 0.01 % planted, offered {off:.0f} images/s")
            for off in [float(x) for x in args.serve_offered_100m.split(",")]
        ]
        serving.extend(runs100)
        # the highest rate the server kept up with (served >= 95 % of offered)
        kept = [r for r in runs100 if r["served_images_per_s"] >= 0.95 * r["offered_images_per_s"]]
        top100 = max(kept or runs100, key=lambda r: r["served_images_per_s"])
        serving_summary["100m_rows_best_served"] = {
            "offered_images_per_s": top100["offered_images_per_s"],
            "served_images_per_s": top100["served_images_per_s"], "latency_ms": top100["latency_ms"],
            "served_vs_plain": top100["served_vs_plain"],
            "note": ("the highest rate served at >= 95 % of offered on the 100M-row database of the serving "
                     "configuration; one 64-image request there needs about 2 ms of matching, so its p99 stays "
                     "above 2 ms")}
        del d_db100
        torch.cuda.empty_cache()
    uneven = None
    if not args.no_uneven:
        uneven = uneven_section(args, yume_pdq, torch, dev, synthetic, planes_u8[:256], pool_hashes, thr)

    precisions = None
    if not args.no_hash_precisions:
        precisions = hash_precision_section(args, yume_pdq, torch, dev, synthetic, planes_u8)

    pipe_stats = None
    if not args.no_pipe:
        pipe_stats = pipe_section(args, yume_pdq, torch, planes_u8)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, db, thr, n_rows)

    line = {
        "metric": METRIC,
        "value": round(head["images_per_s"], 1),
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 hash / u32 match",
        "data": "synthetic",
        "config": {
            "workload": (f"hash {args.headline_batch} u8 512x512 images (fused kernel, 8 dihedral "
                         f"hashes each) + exact match of their {8 * args.headline_batch} hashes "
                         f"against a {n_rows}-row HBM-resident DB, threshold {thr}"),
            "db_rows": n_rows,
            "planted": args.plants,
            "batch_images": args.headline_batch,
            "queries_per_pass": 8 * args.headline_batch,
            "threshold": thr,
            "input": "u8 planes resident in HBM",
            "db_layout": ("32-byte rows + their 16-byte prefixes packed in a prefix plane, both "
                          "resident in HBM; the plane is built once with the database"),
            "dihedral_mode": ("avx2: the reference's x86 visit_dihedrals (column negation over "
                              "columns 0..7, src/kernel/x86.rs:330-376)"),
            "dct_precision": "exact fp32 in the reference's operation order (bit-identical hashes)",
            "device": info["name"],
        },
        "latency_ms": latency,
        "dominant_kernel": "match" if dom_is_match else "hash",
        "roofline": roofline,
        "roofline_hbm_scan_8_queries": roofline_scan8,
        "roofline_hbm_scan_8_queries_raw_rows": roofline_scan8_raw,
        "roofline_match": roofline_match,
        "roofline_hash": roofline_hash,
        "counter_source": counters.get("_source"),
        "batches": [results[b] for b in sorted(results)],
        "plain_unbatched_f32": plain,
        "speedup_vs_plain_unbatched": round(head["images_per_s"] / plain["images_per_s"], 2),
        "cpu_baseline": cpu,
        "serving": serving,
        "serving_summary": serving_summary,
        "uneven_batching": uneven,
        "batched_hashing": precisions,
        "stream_ingestion": pipe_stats,
        "drop_in_ffi": {
            "us_per_hash": round(ffi_us, 2),
            "note": ("PCIe-inclusive per-image rate of the reference ABI: 1 MiB f32 host image "
                     "copied in, fused kernel, hash/threshold/PDQF/buffers copied out, per call "
                     "through ctypes with preallocated buffers (as integration/hash.py)"),
        },
    }
    print(json.dumps(line), flush=True)


def plain_host_rate(yume_pdq, torch, dev, d_db, pool_u8, thr, n=60):
    """The plain unbatched GPU path from host memory: per image, upload the u8
    plane (from the same page-locked pool), hash it (one launch), match its 8
    hashes against the raw rows, read the hit count.  Returns images/s."""
    hits = torch.empty((1 << 16, 16), dtype=torch.uint8, device=dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    for i in range(5):
        h, _, _ = yume_pdq.hash_batch(pool_u8[i:i + 1].to(dev))
        yume_pdq.match(d_db, h.reshape(-1, 32), thr, hits=hits, count=count, sync=False)
        count.item()
    t0 = time.perf_counter()
    for i in range(n):
        h, _, _ = yume_pdq.hash_batch(pool_u8[i % pool_u8.shape[0]][None].to(dev))
        yume_pdq.match(d_db, h.reshape(-1, 32), thr, hits=hits, count=count, sync=False)
        count.item()
    return n / (time.perf_counter() - t0)


def serving_section(args, yume_pdq, torch, dev, synthetic, d_db, db_rows, pool_u8, thr, offered, label,
                    pool_hashes=None, clients=8, max_batch=64, max_contexts=4, in_order_compute=True,
                    scan_chunk_rows=0):
    """Open-loop Poisson arrivals at `offered` images/s of requests carrying
    U{1..64} u8 512x512 images from host memory, served by the server (hash +
    8-dihedral match against every shard; continuous batching, or with
    scan_chunk_rows > 0 the shared circular scan); latency is submit ->
    results on the host, PCIe included."""
    import threading

    import numpy as np
    if d_db is None:
        log(f"[serving] building {db_rows} rows, 0.01 % planted")
        db = synthetic.random_db(db_rows, seed=17)
        synthetic.plant_fast(db, pool_hashes, max(1, db_rows // 10_000), seed=18)
        d_db = torch.from_numpy(db).to(dev)
        del db
    n_rows = d_db.shape[0]
    shards = yume_pdq.Shards()
    shards.add(d_db)  # the rows stay where they are; the set adds the prefix plane
    srv = yume_pdq.Server(shards, max_batch_images=max_batch, max_contexts=max_contexts, hbm_budget_bytes=16 << 30,
                          threshold=thr, in_order_compute=in_order_compute, scan_chunk_rows=scan_chunk_rows)
    rng = np.random.default_rng(23)
    qps = offered / 32.5
    gaps = rng.exponential(1.0 / qps, size=int(qps * args.serve_seconds * 1.2) + 10)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < args.serve_seconds]
    sizes = rng.integers(1, 65, size=arrivals.size)
    starts = rng.integers(0, pool_u8.shape[0], size=arrivals.size)
    # request bodies sit in page-locked host memory (as a network stack's
    # receive buffers would) and are submitted without a copy
    pool2 = torch.from_numpy(np.concatenate([pool_u8, pool_u8])).pin_memory()
    log(f"[serving] {label}: warm-up")
    # untimed requests for 300 ms that read every image of the pool: the GPU
    # reaches its sustained clock, and the first DMA from each region of the
    # page-locked pool (which blocks the copy call for several ms) happens
    # before the measured arrivals
    t_w = time.perf_counter()
    j = 0
    n_pool = pool2.shape[0]
    while j * 64 < n_pool or time.perf_counter() - t_w < 0.3:
        k = (j * 64) % n_pool
        srv.wait(srv.submit_pinned(pool2[k:k + 64]))
        j += 1
    log(f"[serving] {label}: {arrivals.size} requests over {args.serve_seconds} s "
        f"(offered {offered:.0f} images/s), {clients} client threads")
    ids = [0] * arrivals.size
    lag = np.zeros(arrivals.size)  # submit later than the scheduled arrival (client behind)
    t0 = time.perf_counter() + 0.05

    def client(k):
        for j in range(k, arrivals.size, clients):
            delay = t0 + arrivals[j] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            lag[j] = max(0.0, time.perf_counter() - (t0 + arrivals[j]))
            ids[j] = srv.submit_pinned(pool2[starts[j]:starts[j] + sizes[j]])

    th = [threading.Thread(target=client, args=(k,)) for k in range(clients)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    res = [srv.wait(i) for i in ids]
    # open-loop latency: from the scheduled arrival, not from the actual submit
    lat = np.array([r["latency_ms"] for r in res]) + lag * 1e3
    done = np.array([arrivals[j] + lat[j] / 1e3 for j in range(arrivals.size)])
    images = int(sizes.sum())
    span = max(float(done.max()), float(arrivals.max()))
    st = srv.stats()
    bad = sum(1 for r in res if r["status"] != 0)
    srv.close()
    shards.close()
    log(f"[serving] {label}: plain unbatched path from host memory")
    plain = plain_host_rate(yume_pdq, torch, dev, d_db, pool2, thr)
    served = images / span
    out = {
        "database": label,
        "mode": f"shared scan, {scan_chunk_rows}-row chunks" if scan_chunk_rows else "batches",
        "db_rows": int(n_rows),
        "requests": int(arrivals.size),
        "images": images,
        "offered_images_per_s": round(offered, 1),
        "served_images_per_s": round(served, 1),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3)},
        "batches": int(st["batches"]),
        "mean_batch_images": round(float(np.mean([r["batch_images"] for r in res])), 1),
        "contexts": st["contexts"],
        "errors": bad,
        "plain_unbatched_host_images_per_s": round(plain, 1),
        "served_vs_plain": round(served / plain, 2),
        "client_lag_ms_p99": round(float(np.percentile(lag, 99) * 1e3), 3),
        "server_latency_ms": {"p50": round(float(np.percentile([r["latency_ms"] for r in res], 50)), 3),
                              "p99": round(float(np.percentile([r["latency_ms"] for r in res], 99)), 3)},
        "note": ("latency from the scheduled arrival to results on the host; request planes in "
                 "page-locked host memory, uploaded by the server (PCIe included); plain = one image per "
                 "call from the same memory: upload, hash launch, 8-query scan of the raw rows, count read"),
    }
    log(f"[serving]   {out}")
    return out


def uneven_section(args, yume_pdq, torch, dev, synthetic, planes_u8, pool_hashes, thr, n_batches=40):
    """Uneven batches of U{1..256} images hashed and matched against 6 shards
    of 10k, 100k, 1M, 5M, 20M and 50M rows (1 % planted) by the batch
    context: one graph replay per batch, matcher per batch size chosen by the
    autotuner."""
    import numpy as np
    sizes = [10_000, 100_000, 1_000_000, 5_000_000, 20_000_000, 50_000_000]
    shards = yume_pdq.Shards()
    keep = []
    for k, n in enumerate(sizes):
        log(f"[uneven] shard {k}: {n} rows, 1 % planted")
        db = synthetic.random_db(n, seed=300 + k)
        synthetic.plant_fast(db, pool_hashes, max(1, n // 100), seed=400 + k)
        d = torch.from_numpy(db).to(dev)
        del db
        keep.append(d)
        shards.add(d)
    b = yume_pdq.Batcher(shards, 256, threshold=thr, hit_capacity=1 << 22)
    log("[uneven] autotuning the matcher per batch size")
    tune = b.autotune(iterations=3)
    choices = {}
    for r in tune:
        best = min((x for x in tune if x["n_queries"] == r["n_queries"]), key=lambda x: x["ms"])
        choices[r["n_queries"]] = f"{best['path']}" + (f" x{best['query_tiles']}" if best["path"] == "mfma" else "")
    rng = np.random.default_rng(29)
    ns = rng.integers(1, 257, size=n_batches)

    def one(n):
        # planes in, one graph replay, the hit count back on the host (the
        # hashes and hits stay in the context's device buffers)
        b.run(planes_u8[:n], sync=False)
        with torch.cuda.stream(b.stream):
            c = b.count.clone()
        b.stream.synchronize()
        return int(c.item())

    for n in ns[:3]:
        one(n)
    ms, hits = [], []
    t0 = time.perf_counter()
    for n in ns:
        s0 = time.perf_counter()
        hits.append(one(n))
        ms.append((time.perf_counter() - s0) * 1e3)
    total = time.perf_counter() - t0
    b.close()
    shards.close()
    del keep
    torch.cuda.empty_cache()
    out = {
        "shards": sizes,
        "batches": int(n_batches),
        "images": int(ns.sum()),
        "images_per_s": round(float(ns.sum()) / total, 1),
        "batch_ms": {"mean": round(float(np.mean(ms)), 3), "p50": round(float(np.percentile(ms, 50)), 3),
                     "p99": round(float(np.percentile(ms, 99)), 3)},
        "autotuned_matcher_by_queries": choices,
        "mean_hits_per_batch": round(float(np.mean(hits)), 1),
        "note": ("per batch: planes copied into the context (device to device), one graph replay, the hit "
                 "count read back; queries are hashes of the images the 1 % plants were made from"),
    }
    log(f"[uneven]   {out}")
    return out


def dct64_matrix(torch, dev):
    """D[i][j] = sqrt(2/127) cos(pi (i+1)(2j+1) / 254), i < 16, j < 127, in FP64."""
    i = torch.arange(1, 17, dtype=torch.float64, device=dev)[:, None]
    j = torch.arange(127, dtype=torch.float64, device=dev)[None, :]
    return torch.sqrt(torch.tensor(2.0 / 127.0, dtype=torch.float64, device=dev)) * \
        torch.cos(torch.pi / 254.0 * i * (2.0 * j + 1.0))


def hash_precision_section(args, yume_pdq, torch, dev, synthetic, planes_u8):
    """Batched hashing of n synthetic images (u8 planes and the same images
    as f32) in every kernel layout and DCT precision: kernel rate against the
    HBM roofline, and for the matrix-core precisions the error of each PDQF
    coefficient against the FP64 DCT of the same (exact) compressed plane --
    computed on the GPU in FP64 as a verification oracle -- and the hash bits
    that differ from the exact kernel's (= the reference's)."""
    n = args.hash_images
    log(f"[bench] batched hashing: {n} images, every layout and DCT precision")
    have = planes_u8.shape[0]
    if have < n:
        extra = synthetic.images_u8_device(n - have, dev, start=have, seed=199)
        u8 = torch.cat([planes_u8, extra]).contiguous()
    else:
        u8 = planes_u8[:n].contiguous()
    f32 = u8.to(torch.float32)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    variants = [("exact", yume_pdq.DCT_EXACT, yume_pdq.LAYOUT_AUTO),
                ("exact_wave", yume_pdq.DCT_EXACT, yume_pdq.LAYOUT_WAVE),
                ("exact_block8", yume_pdq.DCT_EXACT, yume_pdq.LAYOUT_BLOCK8),
                ("bf16", yume_pdq.DCT_BF16, yume_pdq.LAYOUT_BLOCK8),
                ("f16", yume_pdq.DCT_F16, yume_pdq.LAYOUT_BLOCK8),
                ("bf16x3", yume_pdq.DCT_BF16X3, yume_pdq.LAYOUT_BLOCK8),
                ("f16x3", yume_pdq.DCT_F16X3, yume_pdq.LAYOUT_BLOCK8)]
    outs = (torch.empty((n, 8, 32), dtype=torch.uint8, device=dev),
            torch.empty((n, 8), dtype=torch.float32, device=dev),
            torch.empty((n, 8), dtype=torch.float32, device=dev), None)
    rates = {}
    for kind, x, plane_bytes in (("u8", u8, 262144.0), ("f32", f32, 1048576.0)):
        for name, dct, layout in variants:
            def launch():
                yume_pdq.hash_batch(x, stream=stream, out=outs, dct=dct, layout=layout)
            for _ in range(2):
                launch()
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(10):
                launch()
            b.record(stream)
            b.synchronize()
            ms = a.elapsed_time(b) / 10
            gbs = n * (plane_bytes + 320.0) / (ms * 1e-3) / 1e9
            rates[f"{kind}_{name}"] = {"ms_per_launch": round(ms, 4), "images_per_s": round(n / (ms * 1e-3), 1),
                                       "achieved_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    # the precise ReferenceKernel<f32> (box-filter resampling), u8 planes
    n_ref = min(n, 1024)
    ws = torch.empty(int(yume_pdq.lib().yume_pdq_amd_hash_ref_workspace_bytes(n_ref)) // 4, dtype=torch.float32,
                     device=dev)
    ref_out = (torch.empty((n_ref, 32), dtype=torch.uint8, device=dev),
               torch.empty(n_ref, dtype=torch.float32, device=dev), torch.empty(n_ref, dtype=torch.float32, device=dev))
    import ctypes

    def launch_ref():
        rc = yume_pdq.lib().yume_pdq_amd_hash_ref_batch(
            ctypes.c_void_p(u8.data_ptr()), yume_pdq.INPUT_U8, 512 * 512, n_ref, ctypes.c_void_p(ws.data_ptr()),
            ws.numel() * 4, ctypes.c_void_p(ref_out[0].data_ptr()), ctypes.c_void_p(ref_out[1].data_ptr()),
            ctypes.c_void_p(ref_out[2].data_ptr()), None, ctypes.c_void_p(stream.cuda_stream))
        assert rc == 0
    launch_ref()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(5):
        launch_ref()
    b.record(stream)
    b.synchronize()
    ms = a.elapsed_time(b) / 5
    rates["u8_reference_kernel"] = {"ms_per_launch": round(ms, 4), "images": n_ref,
                                    "images_per_s": round(n_ref / (ms * 1e-3), 1),
                                    "note": ("ReferenceKernel<f32>: four sequential box passes through "
                                             "2 MiB of scratch per image, NORMAL orientation only")}
    del ws
    log(f"[bench]   rates {rates}")
    # errors (u8 input; the f32 copy of the same images gives the same planes)
    D = dct64_matrix(torch, dev)
    absD = D.abs()
    ex_h, _, _, ex_p, ex_b = yume_pdq.hash_batch(u8, want_pdqf=True, want_buf1=True)
    C = ex_b[:, :127, :127].to(torch.float64)
    P64 = torch.matmul(torch.matmul(D, C), D.T)
    M = torch.matmul(torch.matmul(absD, C.abs()), absD.T)
    cmax = C.abs().amax(dim=(1, 2), keepdim=True)

    def err_stats(p):
        e = (p.to(torch.float64) - P64).abs()
        return {"max_abs": float(e.max()), "rms_abs": float(e.pow(2).mean().sqrt()),
                "max_rel_to_abs_product": float((e / M).max()),
                "max_abs_per_unit_input": float((e / cmax).max())}

    errors = {"exact": dict(err_stats(ex_p), hash_bits_vs_exact_mean=0.0, hash_bits_vs_exact_max=0,
                            identical_hash_fraction=1.0)}
    popc = torch.tensor([bin(v).count("1") for v in range(256)], dtype=torch.int32, device=dev)
    for name, dct, layout in variants[3:]:
        h, _, _, p = yume_pdq.hash_batch(u8, want_pdqf=True, dct=dct, layout=layout)
        bits = popc[(h ^ ex_h).to(torch.int64)].sum(-1)
        errors[name] = dict(err_stats(p), hash_bits_vs_exact_mean=round(float(bits.float().mean()), 4),
                            hash_bits_vs_exact_max=int(bits.max()),
                            identical_hash_fraction=round(float((bits == 0).float().mean()), 5))
    log(f"[bench]   errors {errors}")
    return {
        "images": n,
        "input": "u8 planes and the same images as f32 planes, resident in HBM",
        "rates": rates,
        "dct_error_vs_fp64": errors,
        "reference_tolerances": {
            "dct_abs_per_unit_input": "32 * f32 epsilon = 3.81e-06 (src/kernel/mod.rs:1239-1246, inputs in [-1, 1])",
            "hash_bits": "16 bits per hash vs the pdqhash library (src/lib.rs:570-574)",
        },
        "note": ("errors: PDQF of every orientation 0 (NORMAL) against D64 C D64^T of the exact "
                 "compressed plane C, FP64 on the GPU; max_abs_per_unit_input divides by max|C| of "
                 "each image, the scale of the reference's DCT tolerance test; hash bits over all 8 "
                 "orientations of every image against the exact kernel, which equals the reference's "
                 "AVX2 kernel bit for bit"),
    }


def pipe_section(args, yume_pdq, torch, planes_u8):
    """Stream ingestion (the reference's `pipe`): a file of raw luma8 frames
    in the page cache, hashed in GPU batches and written as text records to
    /dev/null; frames/s of the whole call (reader thread, uploads, kernel,
    formatting, writes).  Each line also reports the steady rate: the extra
    frames of the full file over a quarter of it, divided by the extra time,
    so that filling and draining the batch pipeline (a larger share of a
    fixed file at larger batches) cancels."""
    import tempfile
    n = args.pipe_frames
    pool = planes_u8[:64].cpu().numpy().tobytes()
    reps = (n + 63) // 64
    reps_short = max(1, reps // 4)
    paths = []
    try:
        for r in (reps, reps_short):
            fd, path = tempfile.mkstemp(prefix="pipe_frames_", suffix=".raw")
            paths.append(path)
            with os.fdopen(fd, "wb") as f:
                for _ in range(r):
                    f.write(pool)
        frames, frames_short = reps * 64, reps_short * 64
        lines = (("hex+lf", "hex+lf", False, 64), ("hex+lf_batch256", "hex+lf", False, 256),
                 ("q+hex+lf_dihedrals", "q+hex+lf", True, 64), ("bin", "bin", False, 64))
        # two rounds over all lines, each line keeping its best times: the
        # host's rate drifts over the section, which would otherwise favour
        # the lines measured first
        best = {label: [float("inf"), float("inf")] for label, *_ in lines}
        for _ in range(2):
            for label, fmt, dih, batch in lines:
                dn = os.open(os.devnull, os.O_WRONLY)
                try:
                    yume_pdq.pipe(paths[0], dn, fmt=fmt, dihedrals=dih, batch=batch)  # warm-up
                    for k, (path, want) in enumerate(((paths[0], frames), (paths[1], frames_short))):
                        for _ in range(3):  # the best of three runs
                            t0 = time.perf_counter()
                            got = yume_pdq.pipe(path, dn, fmt=fmt, dihedrals=dih, batch=batch)
                            best[label][k] = min(best[label][k], time.perf_counter() - t0)
                        assert got == want
                finally:
                    os.close(dn)
        out = {}
        for label, fmt, dih, batch in lines:
            dt, dt_short = best[label]
            out[label] = {"frames_per_s": round(frames / dt, 1), "seconds": round(dt, 4), "batch": batch,
                          "records_per_frame": 8 if dih else 1,
                          "steady_frames_per_s": (round((frames - frames_short) / (dt - dt_short), 1)
                                                  if dt > dt_short else None)}
        log(f"[bench]   pipe {out}")
        return {"frames": frames, "frames_short": frames_short,
                "input": "luma8 512x512 frames from a file in the page cache",
                "output": "/dev/null", "rates": out,
                "note": ("the best of three yume_pdq_amd_pipe_run calls per line and file in each of two rounds "
                         "over the lines, after a warm-up call of the same shape whose page-locked ring and device "
                         "buffers the timed calls reuse; "
                         "steady_frames_per_s = (frames - frames_short) / (seconds - seconds of the short "
                         "file); the reference's pipe hashes on two CPU threads")}
    finally:
        for path in paths:
            os.unlink(path)
        yume_pdq.pipe_trim()


def drop_in_ffi_us(yume_pdq, image, n=300):
    """Mean microseconds per yume_pdq_hash_smart_kernel call on a host image."""
    import ctypes as C

    import numpy as np
    L = yume_pdq.lib()
    img = np.ascontiguousarray(image, np.float32)
    thr = np.zeros(1, np.float32)
    out = np.zeros(32, np.uint8)
    buf1 = np.zeros((128, 128), np.float32)
    tmp = np.zeros(128, np.float32)
    pdqf = np.zeros((16, 16), np.float32)
    args = [x.ctypes.data_as(C.c_void_p) for x in (img, thr, out, buf1, tmp, pdqf)]
    for _ in range(20):
        L.yume_pdq_hash_smart_kernel(*args)
    t0 = time.perf_counter()
    for _ in range(n):
        L.yume_pdq_hash_smart_kernel(*args)
    return (time.perf_counter() - t0) * 1e6 / n


def load_counter_summary():
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (scaled by the same counters on a copy of known size), from the JSON block
    of the committed counter summary; {} when absent."""
    try:
        text = open(COUNTER_SUMMARY).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        data = json.loads(block)
        data["_source"] = os.path.relpath(COUNTER_SUMMARY, REPO)
        return data
    except (OSError, IndexError, ValueError):
        return {}


def cpu_baseline(args, db, thr, n_rows):
    """The reference's CPU kernel + matcher (C restatement of Avx2F32Kernel,
    visit_dihedrals and CpuMatcher) on this host's cores over a bounded sample."""
    import numpy as np

    import oracle_lib
    from yume_pdq import synthetic

    n_img = args.cpu_images
    threads = max(1, min(16, os.cpu_count() or 1))
    log(f"[bench] CPU baseline: {n_img} images, {8 * n_img} queries x {n_rows} rows, {threads} threads")
    pool = synthetic.images_u8(min(n_img, 64)).astype(np.float32)
    use512 = bool(oracle_lib.lib().oracle_cpu_has_vpopcntdq())
    hits, _, t_hash, t_match = oracle_lib.cpu_baseline(pool, db, thr, threads, use_avx512=use512,
                                                       n_images=n_img)
    total = t_hash + t_match
    # one thread, as the reference's own single-core figures (README.md:191-195, bench.csv)
    n1 = 64
    log(f"[bench] CPU baseline, one thread: {n1} images, {8 * n1} queries x {n_rows} rows")
    _, _, t1_hash, t1_match = oracle_lib.cpu_baseline(pool, db, thr, 1, use_avx512=use512, n_images=n1)
    return {
        "value": round(n_img / total, 2),
        "unit": "images/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{n_img} f32 images (cycling through {pool.shape[0]} distinct synthetic images) "
                   f"hashed (AVX2 kernel restated with intrinsics) and their "
                   f"{8 * n_img} dihedral hashes matched against all {n_rows} rows "
                   f"({'AVX-512 VPOPCNTDQ' if use512 else 'scalar u64 popcount'} CpuMatcher), "
                   f"{threads} threads"),
        "hash_s": round(t_hash, 4),
        "match_s": round(t_match, 4),
        "hash_images_per_s": round(n_img / t_hash, 1),
        "match_images_per_s": round(n_img / t_match, 1),
        "hits": hits,
        "one_thread": {"images": n1, "hash_images_per_s": round(n1 / t1_hash, 1),
                       "match_images_per_s": round(n1 / t1_match, 1),
                       "hash_match_images_per_s": round(n1 / (t1_hash + t1_match), 2)},
    }


if __name__ == "__main__":
    main()
