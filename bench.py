#!/usr/bin/env python3
"""Hash + match throughput on one MI355X.

One step = hash B synthetic 512x512 u8 images resident in HBM (fused kernel:
tent compression, DCT, quality, median quantisation, 8 dihedral hashes), then
match the 8B dihedral hashes against a 10M-row database resident in HBM at
threshold 31 (exact (query, index, distance) triples into a device buffer),
then read the hit count back.  The database carries 50,000 planted
near-duplicates of the batch's queries at distances U{0..40}.

Prints one JSON line (rank 0).  `value` is images/s at the headline batch;
the other batch sizes, the unbatched f32 path, step latency percentiles, the
roofline of the dominant kernels (timed with HIP events on the stream they
run on) and a CPU baseline (reference CPU kernel + matcher restated in C,
timed on this host) ride along.
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
    ap.add_argument("--threshold", type=int, default=31)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-images", type=int, default=512)
    ap.add_argument("--verify", action="store_true", help="check B=1 hits against numpy (slow)")
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

    log(f"[bench] device {info}; generating {bmax} images on the GPU")
    planes_u8 = synthetic.images_u8_device(bmax, dev)
    planes_f32 = planes_u8[:1].to(torch.float32).contiguous()
    torch.cuda.synchronize()

    # queries of the whole image set (used to plant the database)
    h_all, _, _ = yume_pdq.hash_batch(planes_u8)
    queries_all = h_all.reshape(-1, 32).cpu().numpy()
    log(f"[bench] building {args.db_rows} row database with {args.plants} plants")
    db = synthetic.random_db(args.db_rows, seed=7)
    synthetic.plant_fast(db, queries_all, args.plants, seed=11)
    d_db = torch.from_numpy(db).to(dev)
    torch.cuda.synchronize()

    cap = 1 << 22
    d_hits = torch.empty((cap, 16), dtype=torch.uint8, device=dev)
    d_count = torch.zeros(1, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    def step(planes, B, ev=None):
        if ev is not None:
            ev[0].record(stream)
        h, _, _ = yume_pdq.hash_batch(planes[:B], stream=stream, out=out_bufs[B])
        if ev is not None:
            ev[1].record(stream)
        yume_pdq.match(d_db, h.reshape(-1, 32), thr, hits=d_hits, count=d_count, sync=False,
                       stream=stream)
        if ev is not None:
            ev[2].record(stream)
        return int(d_count.item())  # the step ends when its hit count is on the host

    out_bufs = {}
    for B in set(batches + [args.headline_batch, 1]):
        out_bufs[B] = (torch.empty((B, 8, 32), dtype=torch.uint8, device=dev),
                       torch.empty((B, 8), dtype=torch.float32, device=dev),
                       torch.empty((B, 8), dtype=torch.float32, device=dev), None)

    def run(planes, B):
        for _ in range(args.warmup):
            step(planes, B)
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
        lat = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hits = 0
        for k in range(args.steps):
            s0 = time.perf_counter()
            hits = step(planes, B, evs[k])
            lat.append((time.perf_counter() - s0) * 1e3)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        t_hash = [e[0].elapsed_time(e[1]) for e in evs]
        t_match = [e[1].elapsed_time(e[2]) for e in evs]
        total_s = t1 - t0
        return {
            "batch_images": B,
            "queries_per_pass": 8 * B,
            "images_per_s": B * args.steps / total_s,
            "ms_per_step": total_s * 1e3 / args.steps,
            "latency_ms": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
            "hash_kernel_ms": float(np.mean(t_hash)),
            "match_kernel_ms": float(np.mean(t_match)),
            "hits": hits,
        }

    results = {}
    for B in sorted(set(batches + [args.headline_batch])):
        log(f"[bench] batch {B}: {args.warmup} warmup + {args.steps} timed steps")
        results[B] = run(planes_u8, B)
        log(f"[bench]   {results[B]}")
    log("[bench] plain unbatched f32 path (B=1, f32 plane)")
    out_bufs[1] = out_bufs[1]
    plain = run(planes_f32, 1)
    log(f"[bench]   {plain}")

    if args.verify:
        log("[bench] verifying B=1 hit count against numpy")
        h1, _, _ = yume_pdq.hash_batch(planes_u8[:1])
        exp = synthetic.brute_force_hits(db, h1.reshape(-1, 32).cpu().numpy(), thr)
        assert exp.shape[0] == results.get(1, plain)["hits"], "hit count mismatch"

    head = results[args.headline_batch]
    b1 = results.get(1, plain)
    n_rows = args.db_rows
    # algorithmic bytes: 32 B per row scanned (+16 B per hit written)
    match_bytes = 32.0 * n_rows + 16.0 * b1["hits"]
    match_gbs = match_bytes / (b1["match_kernel_ms"] * 1e-3) / 1e9
    hb = results[bmax]
    hash_bytes = 262144.0 * bmax + 320.0 * bmax
    hash_gbs = hash_bytes / (hb["hash_kernel_ms"] * 1e-3) / 1e9
    # the headline step's dominant kernel
    dom_is_match = head["match_kernel_ms"] >= head["hash_kernel_ms"]

    roofline = {
        "kernel": "pdq_match_kernel: 10M rows x 8 queries (one image's dihedrals), thr 31",
        "bound": "hbm",
        "achieved": round(match_gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(match_gbs / HBM_PEAK_GBS, 4),
        "traffic": None,
        "algorithmic_bytes_per_launch": match_bytes,
        "avg_launch_ms": b1["match_kernel_ms"],
    }
    roofline_hash = {
        "kernel": f"pdq_hash_kernel<u8>: {bmax} images per launch",
        "bound": "hbm",
        "achieved": round(hash_gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(hash_gbs / HBM_PEAK_GBS, 4),
        "algorithmic_bytes_per_launch": hash_bytes,
        "avg_launch_ms": hb["hash_kernel_ms"],
        "images_per_s_kernel_only": bmax / (hb["hash_kernel_ms"] * 1e-3),
    }

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
            "device": info["name"],
        },
        "latency_ms": head["latency_ms"],
        "dominant_kernel": "match" if dom_is_match else "hash",
        "roofline": roofline,
        "roofline_hash": roofline_hash,
        "batches": [results[b] for b in sorted(results)],
        "plain_unbatched_f32": plain,
        "speedup_vs_plain_unbatched": round(head["images_per_s"] / plain["images_per_s"], 2),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(args, db, thr, n_rows):
    """The reference's CPU kernel + matcher (C restatement of Avx2F32Kernel,
    visit_dihedrals and CpuMatcher) on this host's cores over a bounded sample."""
    import numpy as np

    import oracle_lib
    from yume_pdq import synthetic

    n_img = args.cpu_images
    threads = max(1, min(16, os.cpu_count() or 1))
    log(f"[bench] CPU baseline: {n_img} images, {8 * n_img} queries x {n_rows} rows, {threads} threads")
    base = synthetic.images_u8(min(n_img, 32)).astype(np.float32)
    imgs = np.concatenate([base] * ((n_img + base.shape[0] - 1) // base.shape[0]))[:n_img]
    use512 = bool(oracle_lib.lib().oracle_cpu_has_vpopcntdq())
    hits, _, t_hash, t_match = oracle_lib.cpu_baseline(imgs, db, thr, threads, use_avx512=use512)
    total = t_hash + t_match
    return {
        "value": round(n_img / total, 2),
        "unit": "images/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{n_img} f32 images hashed (AVX2 kernel restated with intrinsics) and their "
                   f"{8 * n_img} dihedral hashes matched against all {n_rows} rows "
                   f"({'AVX-512 VPOPCNTDQ' if use512 else 'scalar u64 popcount'} CpuMatcher), "
                   f"{threads} threads"),
        "hash_s": round(t_hash, 4),
        "match_s": round(t_match, 4),
        "hits": hits,
    }


if __name__ == "__main__":
    main()
