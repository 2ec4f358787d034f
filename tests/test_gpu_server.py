"""GPU tests of the serving loop: each request's result is the same alone and
in a mix of concurrent requests (continuous batching), equals the direct
hash + match entry points, survives a one-context HBM budget and a hit
buffer overflow."""
import sys
import threading

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def triples(hits):
    if hits.size == 0:
        return np.zeros((0, 3), np.int64)
    return np.stack([hits["query"].astype(np.int64), hits["index"].astype(np.int64),
                     hits["distance"].astype(np.int64)], axis=1)


@pytest.fixture(scope="module")
def pool():
    return synthetic.images_u8(96)


@pytest.fixture(scope="module")
def served_db(gpu, dev, pool):
    """Two shards (600k + 150k rows) with near-duplicates of the pool's hashes."""
    import torch
    h, _, _ = gpu.hash_batch(torch.from_numpy(pool).to(dev))
    q = h.reshape(-1, 32).cpu().numpy()
    a = synthetic.random_db(600_000, seed=61)
    b = synthetic.random_db(150_000, seed=62)
    synthetic.plant_fast(a, q, 3000, seed=63)
    synthetic.plant_fast(b, q, 800, seed=64)
    s = gpu.Shards()
    s.add(a)
    s.add(b)
    return s, np.concatenate([a, b])


def requests(n_req, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_req):
        n = int(rng.integers(1, 65))
        out.append(rng.choice(96, size=n, replace=True))
    return out


def direct(gpu, dev, pool, db, idx):
    import torch
    planes = torch.from_numpy(pool[idx]).to(dev)
    h, q, t = gpu.hash_batch(planes)
    total, hits = gpu.match(db, h.reshape(-1, 32), 31, capacity=1 << 20)
    return h.cpu().numpy(), q.cpu().numpy(), t.cpu().numpy(), hits


def same(a, b):
    return (np.array_equal(a["hashes"], b["hashes"]) and np.array_equal(a["quality"].view(np.uint32),
                                                                      b["quality"].view(np.uint32))
            and np.array_equal(a["threshold"].view(np.uint32), b["threshold"].view(np.uint32))
            and np.array_equal(triples(a["hits"]), triples(b["hits"])))


@pytest.mark.parametrize("in_order", [True, False])
def test_alone_equals_mix(gpu, dev, pool, served_db, in_order):
    """Each request's result is the same alone and sharing batches, in both
    launch policies (in-order compute with upload on admission, and a batch
    per free context).  The second policy starts a batch on any free context
    at once, so it runs with one context: requests then queue behind the
    batch in flight and share the next one."""
    import torch
    shards, full = served_db
    d_full = torch.from_numpy(full).to(dev)
    reqs = requests(40, seed=1)
    srv = gpu.Server(shards, max_batch_images=128, max_contexts=3 if in_order else 1, threshold=31,
                     in_order_compute=in_order)
    try:
        say("[server] requests alone")
        alone = []
        for idx in reqs:
            alone.append(srv.wait(srv.submit(pool[idx])))
        batches_alone = srv.stats()["batches"]
        assert batches_alone == len(reqs)
        say("[server] the same requests from 4 threads at once")
        ids = [None] * len(reqs)

        def client(k):
            for j in range(k, len(reqs), 4):
                ids[j] = srv.submit(pool[reqs[j]])

        th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        mix = [srv.wait(i) for i in ids]
        st = srv.stats()
        say(f"[server] stats {st}")
        assert st["batches"] - batches_alone < len(reqs)  # requests shared batches
        assert any(m["batch_images"] > len(idx) for m, idx in zip(mix, reqs))
        for idx, a, m in zip(reqs, alone, mix):
            assert a["status"] == 0 and m["status"] == 0
            assert same(a, m)
        say("[server] against the direct entry points")
        for idx, a in list(zip(reqs, alone))[:8]:
            h, q, t, hits = direct(gpu, dev, pool, d_full, idx)
            assert np.array_equal(a["hashes"], h)
            assert np.array_equal(a["quality"].view(np.uint32), q.view(np.uint32))
            assert np.array_equal(a["threshold"].view(np.uint32), t.view(np.uint32))
            assert np.array_equal(triples(a["hits"]), triples(hits))
    finally:
        srv.close()


def test_pinned_submit_equals_staged(gpu, dev, pool, served_db):
    """Zero-copy submission from page-locked memory gives the same results."""
    import torch
    shards, _ = served_db
    pinned = torch.from_numpy(pool).pin_memory()
    srv = gpu.Server(shards, max_batch_images=64, max_contexts=2)
    try:
        say("[server] pinned submit")
        reqs = requests(12, seed=3)
        ids = [srv.submit(pool[r]) for r in reqs]
        # contiguous slices of the pinned pool, as the pinned form requires
        starts = [int(r[0]) % 32 for r in reqs]
        ids_p = [srv.submit_pinned(pinned[s0:s0 + len(r)]) for s0, r in zip(starts, reqs)]
        for r, s0, i, j in zip(reqs, starts, ids, ids_p):
            a = srv.wait(i)
            b = srv.wait(j)
            c = srv.wait(srv.submit(pool[s0:s0 + len(r)]))
            assert same(b, c)
            assert a["status"] == 0
    finally:
        srv.close()


def test_one_context_budget(gpu, dev, pool, served_db):
    """A budget that holds one batch context serves one batch at a time,
    with the same results."""
    shards, _ = served_db
    probe = gpu.Server(shards, max_batch_images=32, max_contexts=1)
    one = probe.stats()["device_bytes"]
    probe.close()
    srv = gpu.Server(shards, max_batch_images=32, hbm_budget_bytes=one + one // 2, max_contexts=4)
    big = gpu.Server(shards, max_batch_images=32, hbm_budget_bytes=4 * one, max_contexts=4)
    try:
        assert srv.stats()["contexts"] == 1
        assert big.stats()["contexts"] == 4
        say("[server] one-context budget")
        reqs = requests(10, seed=2)
        reqs = [r[:32] for r in reqs]
        ids = [srv.submit(pool[r]) for r in reqs]
        ids_big = [big.submit(pool[r]) for r in reqs]
        for i, j in zip(ids, ids_big):
            assert same(srv.wait(i), big.wait(j))
    finally:
        srv.close()
        big.close()


def test_budget_too_small_is_an_error(gpu, served_db):
    shards, _ = served_db
    with pytest.raises(gpu.YumePdqError):
        gpu.Server(shards, max_batch_images=64, hbm_budget_bytes=1 << 20)


def test_hit_overflow_rematched(gpu, dev, pool):
    """Threshold 256 on a 50-row shard: every pair is a hit, far more than the
    10-record batch buffer; the server re-matches into a larger buffer."""
    s = gpu.Shards()
    s.add(synthetic.random_db(50, seed=5))
    srv = gpu.Server(s, max_batch_images=16, threshold=256, hit_capacity=10)
    try:
        say("[server] overflow")
        r = srv.wait(srv.submit(pool[:3]))
        assert r["status"] == 0
        assert r["hits"].size == 3 * 8 * 50
        assert set(zip(r["hits"]["query"].tolist(), r["hits"]["index"].tolist())) == \
            {(q, i) for q in range(24) for i in range(50)}
    finally:
        srv.close()
        s.close()


def test_hits_past_readback_prefix(gpu, dev, pool):
    """Threshold 256 on a 4000-row shard: 3 images give 24 x 4000 = 96,000
    hits, more than the 65,536 the batch's readback kernel stores with its
    hashes; the rest come back by a second copy, and the set is whole."""
    s = gpu.Shards()
    s.add(synthetic.random_db(4000, seed=6))
    srv = gpu.Server(s, max_batch_images=16, threshold=256, hit_capacity=1 << 17)
    try:
        say("[server] hits past the readback prefix")
        r = srv.wait(srv.submit(pool[:3]))
        assert r["status"] == 0
        assert r["hits"].size == 24 * 4000
        q = r["hits"]["query"].astype(np.int64)
        i = r["hits"]["index"].astype(np.int64)
        assert np.array_equal(np.sort(q * 4000 + i), np.arange(24 * 4000))
        # and a batch under the prefix afterwards reads back alone
        r1 = srv.wait(srv.submit(pool[3:4]))
        assert r1["status"] == 0 and r1["hits"].size == 8 * 4000
    finally:
        srv.close()
        s.close()


def test_submit_validation(gpu, pool, served_db):
    shards, _ = served_db
    srv = gpu.Server(shards, max_batch_images=8)
    try:
        with pytest.raises(gpu.YumePdqError):
            srv.submit(pool[:9])
        with pytest.raises(gpu.YumePdqError):
            srv.wait(12345, timeout_ms=10)
    finally:
        srv.close()


def test_scan_alone_mix_and_direct(gpu, dev, pool, served_db):
    """The shared circular scan (200,704-row chunks: three of the 600k-row
    shard, one of the 150k-row shard): each request's result is the same
    alone and joining a scan already under way with other requests, and
    equals the direct hash + match entry points over the whole database."""
    import torch
    shards, full = served_db
    d_full = torch.from_numpy(full).to(dev)
    reqs = requests(40, seed=11)
    srv = gpu.Server(shards, max_batch_images=64, max_contexts=3, threshold=31, scan_chunk_rows=200_000)
    try:
        say("[server] scan: requests alone")
        alone = [srv.wait(srv.submit(pool[idx])) for idx in reqs]
        say("[server] scan: the same requests from 4 threads at once")
        ids = [None] * len(reqs)

        def client(k):
            for j in range(k, len(reqs), 4):
                ids[j] = srv.submit(pool[reqs[j]])

        th = [threading.Thread(target=client, args=(k,)) for k in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        mix = [srv.wait(i) for i in ids]
        say(f"[server] scan stats {srv.stats()}")
        assert any(m["batch_images"] > len(idx) for m, idx in zip(mix, reqs))  # shared launches
        for idx, a, m in zip(reqs, alone, mix):
            assert a["status"] == 0 and m["status"] == 0
            assert same(a, m)
        say("[server] scan: against the direct entry points")
        for idx, a in zip(reqs, alone):
            h, q, t, hits = direct(gpu, dev, pool, d_full, idx)
            assert np.array_equal(a["hashes"], h)
            assert np.array_equal(a["quality"].view(np.uint32), q.view(np.uint32))
            assert np.array_equal(a["threshold"].view(np.uint32), t.view(np.uint32))
            assert np.array_equal(triples(a["hits"]), triples(hits))
    finally:
        srv.close()


def test_scan_pinned_and_small_area(gpu, dev, pool, served_db):
    """Scan mode with an active area of 64 images (one context of 64): more
    requests than fit wait for slots in arrival order; zero-copy submissions
    give the same results as staged ones."""
    import torch
    shards, _ = served_db
    pinned = torch.from_numpy(pool).pin_memory()
    srv = gpu.Server(shards, max_batch_images=64, max_contexts=1, scan_chunk_rows=100_000)
    try:
        say("[server] scan: pinned submit, 64-image area")
        reqs = requests(16, seed=13)
        starts = [int(r[0]) % 32 for r in reqs]
        ids = [srv.submit(pool[s0:s0 + len(r)]) for s0, r in zip(starts, reqs)]
        ids_p = [srv.submit_pinned(pinned[s0:s0 + len(r)]) for s0, r in zip(starts, reqs)]
        for i, j in zip(ids, ids_p):
            a, b = srv.wait(i), srv.wait(j)
            assert a["status"] == 0 and b["status"] == 0
            assert same(a, b)
    finally:
        srv.close()


@pytest.mark.parametrize("rows,cap,chunk", [(50, 10, 1024), (4000, 1 << 17, 4096)])
def test_scan_hit_overflow_and_prefix(gpu, dev, pool, rows, cap, chunk):
    """Scan mode at threshold 256 (every pair a hit): a launch with more hits
    than its buffer is matched again into a larger one (50 rows, 10-hit
    buffer), and hits past the 65,536 the readback kernel stores come back
    by a second copy (3 images x 8 x 4000 rows = 96,000 hits in one chunk);
    the sets are whole."""
    s = gpu.Shards()
    s.add(synthetic.random_db(rows, seed=7))
    srv = gpu.Server(s, max_batch_images=16, threshold=256, hit_capacity=cap, scan_chunk_rows=chunk)
    try:
        say(f"[server] scan overflow rows={rows} cap={cap}")
        r = srv.wait(srv.submit(pool[:3]))
        assert r["status"] == 0
        q = r["hits"]["query"].astype(np.int64)
        i = r["hits"]["index"].astype(np.int64)
        assert np.array_equal(np.sort(q * rows + i), np.arange(24 * rows))
    finally:
        srv.close()
        s.close()
