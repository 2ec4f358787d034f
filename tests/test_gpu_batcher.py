"""GPU parity of the batching layer: shard sets (persistent multi-shard
matchers, file loader, in-place replacement) and batch contexts (one graph
per batch: hash -> match against every shard, uneven batch sizes,
autotuner).  Every result is compared with the single-database entry points,
which the other GPU tests pin to the oracle, and with numpy brute force.
"""
import sys

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu

SHARD_SIZES = [10_000, 100_000, 1_000_000, 37, 250_000]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def triples(hits):
    if hits.size == 0:
        return np.zeros((0, 3), np.int64)
    return np.stack([hits["query"].astype(np.int64), hits["index"].astype(np.int64),
                     hits["distance"].astype(np.int64)], axis=1)


@pytest.fixture(scope="module")
def db_parts():
    """Shards of uneven sizes with 1 % planted near-duplicates of 2048 query
    hashes (256 images' worth), global indices contiguous."""
    q = synthetic.random_db(2048, seed=90)
    parts = []
    for k, n in enumerate(SHARD_SIZES):
        d = synthetic.random_db(n, seed=100 + k)
        synthetic.plant_fast(d, q, max(1, n // 100), seed=200 + k)
        parts.append(d)
    return parts, q


@pytest.fixture(scope="module")
def shard_set(gpu, db_parts):
    parts, _ = db_parts
    s = gpu.Shards()
    for d in parts:
        s.add(d)  # host rows, copied into HBM
    return s


@pytest.mark.parametrize("nq,path", [(8, None), (24, "valu"), (64, "mfma"), (1024, None), (2048, "valu")])
def test_shards_match_equals_concatenated(gpu, dev, db_parts, shard_set, nq, path):
    """One persistent launch over 5 uneven shards == the single-database
    matcher over their concatenation (global indices)."""
    import torch
    parts, q = db_parts
    full = np.concatenate(parts)
    d_q = torch.from_numpy(q[:nq]).to(dev)
    say(f"[batcher] shards match nq={nq} path={path}")
    n1, h1 = shard_set.match(d_q, 31, capacity=1 << 20, path=path)
    n2, h2 = gpu.match(torch.from_numpy(full).to(dev), d_q, 31, capacity=1 << 20)
    assert n1 == n2 and n1 > 0
    assert np.array_equal(triples(h1), triples(h2))
    if nq <= 24:
        assert np.array_equal(triples(h1), synthetic.brute_force_hits(full, q[:nq], 31))


def test_shard_info(shard_set):
    info = shard_set.info()
    assert info["shards"] == len(SHARD_SIZES)
    assert info["rows"] == sum(SHARD_SIZES)
    # 32-byte rows + 16-byte prefix plane (padded to 16-row tiles) per shard
    assert info["device_bytes"] == sum(32 * n + (n + 15) // 16 * 256 for n in SHARD_SIZES)


def test_batcher_uneven_batches(gpu, dev, db_parts, shard_set):
    """Uneven batch sizes replay the same graphs (batch size read from device
    memory); each batch equals hash_batch + match on the concatenated
    database.  Eager work and synchronisation sit between the replays."""
    import torch
    parts, _ = db_parts
    full = torch.from_numpy(np.concatenate(parts)).to(dev)
    imgs = torch.from_numpy(synthetic.images_u8(64)).to(dev)
    b = gpu.Batcher(shard_set, 64, threshold=31, hit_capacity=1 << 20)
    try:
        for n in (1, 3, 17, 64, 5, 2, 64, 1):
            say(f"[batcher] graph batch n={n}")
            planes = imgs[(n * 7) % 64:][:n] if n < 64 else imgs
            n = planes.shape[0]
            h, qv, t, total, hits = b.run(planes)
            eh, eq, et = gpu.hash_batch(planes)
            assert torch.equal(h, eh)
            assert torch.equal(qv.view(torch.int32), eq.view(torch.int32))
            assert torch.equal(t.view(torch.int32), et.view(torch.int32))
            en, ehits = gpu.match(full, eh.reshape(-1, 32), 31, capacity=1 << 20)
            assert total == en
            assert np.array_equal(triples(hits), triples(ehits))
            # the same batch without the graph
            h2, _, _, total2, hits2 = b.run(planes, use_graph=False)
            assert torch.equal(h2, eh) and total2 == en and np.array_equal(triples(hits2), triples(ehits))
    finally:
        b.close()


def test_batcher_large_u8_context(gpu, dev, db_parts, shard_set):
    """A batch context of at least 16 u8 images per CU hashes in 16-wave
    blocks with the batch size read from device memory: a full batch and a
    small one through the same graphs equal the eight-wave block kernel's
    hashes, and their hit counts equal match on the concatenated database."""
    import torch
    parts, _ = db_parts
    full = torch.from_numpy(np.concatenate(parts)).to(dev)
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    cap = 16 * cus + 8
    g = torch.Generator(device=dev)
    g.manual_seed(77)
    imgs = torch.randint(0, 256, (cap, 512, 512), dtype=torch.uint8, device=dev, generator=g)
    imgs[:64] = torch.from_numpy(synthetic.images_u8(64)).to(dev)
    b = gpu.Batcher(shard_set, cap, threshold=31, hit_capacity=1 << 20)
    try:
        for n in (cap - 5, 5, cap):
            say(f"[batcher] large u8 context: n={n} of {cap}")
            planes = imgs[:n]
            h, qv, t, total, _ = b.run(planes)
            eh, eq, et = gpu.hash_batch(planes, layout=gpu.LAYOUT_BLOCK8)
            assert torch.equal(h, eh)
            assert torch.equal(qv.view(torch.int32), eq.view(torch.int32))
            assert torch.equal(t.view(torch.int32), et.view(torch.int32))
            en, _ = gpu.match(full, eh.reshape(-1, 32), 31, capacity=1 << 20)
            assert total == en
    finally:
        b.close()


def test_batcher_planted_hits_found(gpu, dev, db_parts, shard_set):
    """Queries that are the hashes of real images: plant near-duplicates of
    them into a shard through in-place replacement and find every one."""
    import torch
    imgs = torch.from_numpy(synthetic.images_u8(8)).to(dev)
    b = gpu.Batcher(shard_set, 8, threshold=31)
    try:
        h, _, _, _, _ = b.run(imgs)
        hq = h.reshape(-1, 32).cpu().numpy()
        rng = np.random.default_rng(5)
        rows = rng.choice(100_000, size=64, replace=False).astype(np.int64)
        new = np.repeat(hq[:8], 8, axis=0).copy()
        dists = []
        for i in range(64):  # distinct flips -> exact distance
            k = int(rng.integers(0, 41))
            bits = rng.choice(256, size=k, replace=False)
            for bit in bits:
                new[i, bit >> 3] ^= np.uint8(1 << (bit & 7))
            dists.append(k)
        say("[batcher] replace 64 rows of shard 1 and rerun")
        bad = shard_set.replace(1, torch.from_numpy(rows).to(dev), torch.from_numpy(new).to(dev))
        assert bad == 0
        torch.cuda.synchronize()
        _, _, _, total, hits = b.run(imgs)
        t = triples(hits)
        base = SHARD_SIZES[0]
        for i in range(64):
            if dists[i] <= 31:
                want = (i // 8, base + rows[i], dists[i])
                assert ((t[:, 0] == want[0]) & (t[:, 1] == want[1]) & (t[:, 2] == want[2])).any(), want
        # an out-of-range index is skipped and counted
        bad = shard_set.replace(3, torch.tensor([36, 37], dtype=torch.int64, device=dev),
                                torch.zeros((2, 32), dtype=torch.uint8, device=dev))
        assert bad == 1
    finally:
        b.close()


def test_batcher_autotune(gpu, dev, shard_set):
    """The autotuner times every candidate per query bucket and the batcher
    keeps producing the same results with its choices."""
    import torch
    imgs = torch.from_numpy(synthetic.images_u8(16)).to(dev)
    b = gpu.Batcher(shard_set, 16, threshold=31)
    try:
        before = b.run(imgs)
        say("[batcher] autotune")
        recs = b.autotune(iterations=2)
        assert {r["n_queries"] for r in recs} == {8, 16, 32, 64, 128}
        assert all(r["ms"] > 0 for r in recs)
        for nq in (8, 16, 32, 64, 128):
            best = min((r for r in recs if r["n_queries"] == nq), key=lambda r: r["ms"])
            assert b.choice(nq) == (best["path"], best["query_tiles"])
        after = b.run(imgs)
        assert torch.equal(before[0], after[0]) and before[3] == after[3]
        assert np.array_equal(triples(before[4]), triples(after[4]))
    finally:
        b.close()


def test_tuning_shared_and_saved(gpu, dev, shard_set, tmp_path):
    """Autotuner choices are kept per (GPU, threshold, shard-set shape) for
    the whole process: a second batcher on the same set uses them without
    tuning, clearing the cache brings back the defaults, a saved file brings
    the choices back, and another threshold is another key."""
    buckets = (8, 16, 32, 64, 128)
    gpu.tune_clear()
    b1 = gpu.Batcher(shard_set, 16, threshold=31)
    b2 = b3 = None
    try:
        defaults = {nq: b1.choice(nq) for nq in buckets}
        say("[batcher] autotune for the cache")
        b1.autotune(iterations=2)
        tuned = {nq: b1.choice(nq) for nq in buckets}
        path = tmp_path / "tune.txt"
        assert gpu.tune_save(str(path)) == 1
        b2 = gpu.Batcher(shard_set, 16, threshold=31)
        assert {nq: b2.choice(nq) for nq in buckets} == tuned
        assert gpu.tune_clear() == 1
        assert {nq: b2.choice(nq) for nq in buckets} == defaults
        assert gpu.tune_load(str(path)) == 1
        assert {nq: b2.choice(nq) for nq in buckets} == tuned
        assert {nq: b1.choice(nq) for nq in buckets} == tuned
        b3 = gpu.Batcher(shard_set, 16, threshold=20)
        assert {nq: b3.choice(nq) for nq in buckets} == defaults
    finally:
        for b in (b1, b2, b3):
            if b is not None:
                b.close()
        gpu.tune_clear()


def test_query_tiles_follow_padding(gpu, dev, shard_set, tmp_path):
    """A bucket tuned to 16 query tiles per wave (256 queries per group) at
    its first query count runs 320 queries at 8 tiles (384 computed instead of
    512) and keeps 16 where the groups are full; the hits do not change."""
    import torch
    gpu.tune_clear()
    b = gpu.Batcher(shard_set, 64, threshold=31)
    try:
        b.autotune(iterations=1)
        path = tmp_path / "tune.txt"
        assert gpu.tune_save(str(path)) == 1
        line = [x for x in path.read_text().splitlines() if not x.startswith("#")][0]
        key = line.split(" :")[0]
        path.write_text(f"{key} : 8=valu 32=mfma:2 256=mfma:16\n")
        gpu.tune_clear()
        assert gpu.tune_load(str(path)) == 1
        assert b.choice(256) == ("mfma", 16)
        assert b.choice(512) == ("mfma", 16)
        assert b.choice(320) == ("mfma", 8)
        assert b.choice(264) == ("mfma", 8)
        assert b.choice(40) == ("mfma", 2)
        imgs = torch.from_numpy(synthetic.images_u8(40, start=70)).to(dev)
        got = b.run(imgs)
        gpu.tune_clear()
        ref = b.run(imgs)
        assert got[3] == ref[3]
        assert sorted(map(tuple, triples(got[4]))) == sorted(map(tuple, triples(ref[4])))
    finally:
        b.close()
        gpu.tune_clear()


def test_server_autotunes_once(gpu, dev, db_parts, shard_set):
    """A server started with autotune_iterations tunes when the cache has no
    choices for its shard set and stores them for everyone; its results equal
    a batcher's."""
    import torch
    gpu.tune_clear()
    imgs = synthetic.images_u8(6, start=40)
    srv = gpu.Server(shard_set, max_batch_images=16, hbm_budget_bytes=4 << 30, max_contexts=2,
                     autotune_iterations=1)
    try:
        rid = srv.submit(imgs)
        r = srv.wait(rid, timeout_ms=60_000)
    finally:
        srv.close()
    assert r["status"] == 0
    b = gpu.Batcher(shard_set, 16, threshold=31)
    try:
        h, q, t, total, hits = b.run(torch.from_numpy(imgs).to(dev))
    finally:
        b.close()
        saved = gpu.tune_clear()
    assert saved == 1
    assert np.array_equal(np.asarray(r["hashes"]).reshape(-1), h.cpu().numpy().reshape(-1))
    assert sorted(map(tuple, triples(r["hits"]))) == sorted(map(tuple, triples(hits)))


def test_shards_load_raw(gpu, dev, tmp_path):
    """A raw N x 32-byte file split into shards of 3000 rows equals one
    database of the same rows."""
    import torch
    db = synthetic.random_db(10_007, seed=77)
    q = synthetic.random_db(40, seed=78)
    synthetic.plant(db, q, 100, seed=79)
    path = tmp_path / "db.raw"
    db.tofile(path)
    s = gpu.Shards()
    try:
        say("[batcher] load raw file")
        assert s.load_raw(path, 3000) == 4
        assert s.info()["rows"] == 10_007
        d_q = torch.from_numpy(q).to(dev)
        n1, h1 = s.match(d_q, 31)
        n2, h2 = gpu.match(torch.from_numpy(db).to(dev), d_q, 31)
        assert n1 == n2 and np.array_equal(triples(h1), triples(h2))
    finally:
        s.close()


def test_batcher_shared_rows_planted(gpu, dev):
    """Batches of 4 or more query groups (64, 100 and 128 images: 512, 800
    and 1024 queries) replay the shared-rows graphs: a block's waves stream
    the same rows, candidates go through records to the recheck kernel.
    Ragged shards (5003, 1, 70,000 and 300,000 rows) hold near-duplicates of
    the batch's own hashes; each batch equals the single-database matcher
    over the concatenation, with and without the graph."""
    import torch
    imgs = torch.from_numpy(synthetic.images_u8(128)).to(dev)
    hq = gpu.hash_batch(imgs)[0].reshape(-1, 32).cpu().numpy()
    parts = []
    for k, n in enumerate((5003, 1, 70_000, 300_000)):
        d = synthetic.random_db(n, seed=300 + k)
        synthetic.plant_fast(d, hq, max(1, n // 50), seed=400 + k)
        parts.append(d)
    s = gpu.Shards()
    for d in parts:
        s.add(d)
    full = torch.from_numpy(np.concatenate(parts)).to(dev)
    b = gpu.Batcher(s, 128, threshold=31, hit_capacity=1 << 20)
    try:
        for n in (128, 64, 100, 128):
            say(f"[batcher] shared rows, batch of {n}")
            planes = imgs[:n]
            h, _, _, total, hits = b.run(planes)
            en, ehits = gpu.match(full, h.reshape(-1, 32), 31, capacity=1 << 20)
            assert total == en and en > 0
            assert np.array_equal(triples(hits), triples(ehits))
            _, _, _, total2, hits2 = b.run(planes, use_graph=False)
            assert total2 == en and np.array_equal(triples(hits2), triples(ehits))
    finally:
        b.close()
        s.close()


def test_batcher_shared_rows_record_overflow(gpu, dev, record_cap_4096):
    """256 images (2048 queries) at threshold 63 over 1M rows in two shards:
    every group of tiles is a candidate, 977 sub-chunks x 16 query groups,
    more records than the 4096 allowed here (set_match_record_cap,
    read when the graph is captured), so the captured recheck kernel matches
    every item again per wave.  The set equals the single-database
    vector-ALU matcher's."""
    import torch
    imgs = torch.from_numpy(synthetic.images_u8(256)).to(dev)
    hq = gpu.hash_batch(imgs)[0].reshape(-1, 32).cpu().numpy()
    parts = [synthetic.random_db(n, seed=500 + k) for k, n in enumerate((600_000, 400_001))]
    for k, d in enumerate(parts):
        synthetic.plant_fast(d, hq, 5000, seed=600 + k)
    s = gpu.Shards()
    for d in parts:
        s.add(d)
    full = torch.from_numpy(np.concatenate(parts)).to(dev)
    b = gpu.Batcher(s, 256, threshold=63, hit_capacity=1 << 20)
    try:
        say("[batcher] record overflow, graph")
        h, _, _, total, hits = b.run(imgs)
        en, ehits = gpu.match(full, h.reshape(-1, 32), 63, capacity=1 << 20, path="valu")
        assert total == en and en >= 9000
        assert np.array_equal(triples(hits), triples(ehits))
    finally:
        b.close()
        s.close()
