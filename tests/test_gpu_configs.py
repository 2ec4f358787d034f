"""GPU: each benchmark configuration at its stated size, checked against the
oracle -- the reference's CPU kernel + visit_dihedrals (oracle mode 1, the x86
smart kernel) for hashes and the brute-force matcher (oracle_match_all on
host threads) for hit sets -- or, where a full brute force would take
minutes, against it on whole shards and 1M-row windows plus the planted
controls of src/matching/mod.rs:460-538 (every planted row within the
threshold found with its exact distance).

* config 0: 64 images -> 512 dihedral queries -> 1M rows with 1000 plants,
  threshold 31, end to end through the one-call hash + match;
* config 1: 10M rows with 50,000 plants, 8 x {1, 16, 128, 1024} queries, each
  batch's full triple set against the brute force over all 10M rows;
* config 3: shards of 10k, 100k, 1M, 5M, 20M and 50M rows (1 % planted)
  through the graph-captured batcher, uneven batches;
* config 4: the server on a 100M-row shard (0.01 % planted), U{1..64}-image
  requests from concurrent clients: alone == in the mix, plants found,
  1M-row windows and two whole requests against the brute force.
"""
import sys
import threading

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu

THR = 31


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def triples(hits):
    if hits.size == 0:
        return np.zeros((0, 3), np.int64)
    t = np.stack([hits["query"].astype(np.int64), hits["index"].astype(np.int64),
                  hits["distance"].astype(np.int64)], axis=1)
    return t[np.lexsort((t[:, 0], t[:, 1]))]


def brute(oracle, db, queries, thr=THR, index_base=0):
    """Every (query, index, distance) of db against queries, sorted by
    (index, query), from the oracle's threaded brute force."""
    n, t = oracle.match_all_mt(db, queries, thr, cap=1 << 24)
    assert n == t.shape[0], "oracle cap too small"
    t = t.copy()
    t[:, 1] += index_base
    return t[np.lexsort((t[:, 0], t[:, 1]))]


def plant_masks(db, pool, n_plants, seed, max_dist=40):
    """n_plants distinct rows of db become pool queries XOR a mask of
    k ~ U{0..max_dist} bit flips drawn with replacement (as the reference's
    planted controls flip bits, src/matching/mod.rs:489-493), so the planted
    distance is popcount(mask) <= k.  Returns (rows, pool ids, distances)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rows = rng.choice(db.shape[0], size=n_plants, replace=False)
    qids = rng.integers(0, pool.shape[0], size=n_plants)
    k = rng.integers(0, max_dist + 1, size=n_plants)
    pos = rng.integers(0, 256, size=(n_plants, max_dist))
    bits = np.zeros((n_plants, 256), np.uint8)
    use = np.arange(max_dist)[None, :] < k[:, None]
    bits[np.repeat(np.arange(n_plants), max_dist)[use.reshape(-1)], pos[use]] = 1
    masks = np.packbits(bits, axis=1, bitorder="little")
    db[rows] = pool[qids] ^ masks
    return rows, qids, bits.sum(axis=1).astype(np.int64)


def keys(t):
    """(query, index, distance) rows as one int64 key each."""
    return (t[:, 1] << 24) | (t[:, 0] << 8) | t[:, 2]


def assert_plants_found(t, rows, qids, dists, query_of_pool, label, require_some=True):
    """Every plant whose pool query is among this call's queries
    (query_of_pool maps a pool id to the call's query id, -1 if absent) and
    whose distance is within the threshold is a hit with that distance."""
    want = np.stack([query_of_pool[qids], rows, dists], axis=1).astype(np.int64)
    want = want[(want[:, 0] >= 0) & (want[:, 2] <= THR)]
    assert want.shape[0] > 0 or not require_some, label
    found = np.isin(keys(want), keys(t))
    assert found.all(), f"{label}: {int((~found).sum())} of {want.shape[0]} plants missing, e.g. {want[~found][:3]}"


# --------------------------------------------------------------------------- config 0

def test_config0_end_to_end(gpu, dev, oracle):
    """64 seeded images (u8 planes) -> the reference's 8 hashes each (oracle)
    -> 512 queries against 1M random rows with 1000 plants (distinct flips,
    k ~ U{0..40}) at threshold 31: the one-call hash + match returns the
    oracle's hashes, qualities, thresholds and exact triple set, over the
    prefix plane and over raw rows."""
    import torch
    imgs = synthetic.images_u8(64)
    H = np.zeros((64, 8, 32), np.uint8)
    Q = np.zeros((64, 8), np.float32)
    T = np.zeros((64, 8), np.float32)
    for i in range(64):
        H[i], Q[i], T[i], _ = oracle.hash_dihedrals(imgs[i].astype(np.float32), 1)
    queries = H.reshape(-1, 32)
    db = synthetic.random_db(1_000_000, seed=7)
    rows, qids, dists = synthetic.plant_fast(db, queries, 1000, seed=11)
    exp = brute(oracle, db, queries)
    assert_plants_found(exp, rows, qids, dists, np.arange(512), "oracle")
    d_db = torch.from_numpy(db).to(dev)
    planes = torch.from_numpy(imgs).to(dev)
    for plane in (gpu.build_prefix_plane(d_db), None):
        say(f"[config0] hash_match plane={plane is not None}")
        h, q, t, total, hits = gpu.hash_match(planes, d_db, THR, plane=plane, capacity=1 << 16)
        assert np.array_equal(h.cpu().numpy(), H)
        assert np.array_equal(q.cpu().numpy().view(np.uint32), Q.view(np.uint32))
        assert np.array_equal(t.cpu().numpy().view(np.uint32), T.view(np.uint32))
        assert total == exp.shape[0]
        assert np.array_equal(triples(hits), exp)


# --------------------------------------------------------------------------- config 1

@pytest.fixture(scope="module")
def config1(gpu, dev):
    """10M rows, 50,000 plants (distinct flips, U{0..40}) from the 8192
    dihedral hashes of 1024 images; the rows, their prefix plane on the GPU."""
    import torch
    planes = synthetic.images_u8_device(1024, dev)
    h, _, _ = gpu.hash_batch(planes)
    queries = h.reshape(-1, 32).cpu().numpy()
    db = synthetic.random_db(10_000_000, seed=7)
    rows, qids, dists = synthetic.plant_fast(db, queries, 50_000, seed=11)
    d_db = torch.from_numpy(db).to(dev)
    return db, d_db, gpu.build_prefix_plane(d_db), queries, (rows, qids, dists)


@pytest.mark.parametrize("images", [1, 16, 128, 1024])
def test_config1_full_brute_force(gpu, dev, oracle, config1, images):
    """8 x {1, 16, 128, 1024} queries (prefixes of the plant pool) against all
    10M rows: the matcher's triples (whichever kernel it picks, over the
    prefix plane) equal the brute force over every row."""
    import torch
    db, d_db, d_plane, queries, (rows, qids, dists) = config1
    nq = 8 * images
    q = queries[:nq]
    say(f"[config1] 10M x {nq}: GPU")
    total, hits = gpu.match(d_db, torch.from_numpy(q).to(dev), THR, capacity=1 << 20, plane=d_plane)
    say(f"[config1] 10M x {nq}: oracle")
    exp = brute(oracle, db, q)
    assert total == exp.shape[0] > 0
    assert np.array_equal(triples(hits), exp)
    pool_to_query = np.where(np.arange(queries.shape[0]) < nq, np.arange(queries.shape[0]), -1)
    assert_plants_found(exp, rows, qids, dists, pool_to_query, f"10M x {nq}")


# --------------------------------------------------------------------------- config 3

SHARDS = [10_000, 100_000, 1_000_000, 5_000_000, 20_000_000, 50_000_000]


def test_config3_full_shard_set(gpu, dev, oracle):
    """Six shards of 10k...50M rows (76.11M rows, 2.4 GB), 1 % of each planted
    from the hashes of a 256-image pool, matched by the batcher's captured
    hash -> match graph for uneven batches of U{1..256} contiguous pool
    images.  Per batch: the batch's hashes are hash_batch's; every plant of
    its queries within the threshold is found on every shard with its exact
    distance; the hits on the 10k, 100k and 1M shards and in a 1M-row window
    of each larger shard equal the brute force there."""
    import torch
    pool = synthetic.images_u8_device(256, dev, start=300)
    ph, _, _ = gpu.hash_batch(pool)
    pool_q = ph.reshape(-1, 32).cpu().numpy()
    parts, plants, bases = [], [], []
    base = 0
    for k, n in enumerate(SHARDS):
        d = synthetic.random_db(n, seed=1000 + k)
        plants.append(plant_masks(d, pool_q, n // 100, seed=2000 + k))
        parts.append(d)
        bases.append(base)
        base += n
    shards = gpu.Shards()
    for d in parts:
        shards.add(d)
    say(f"[config3] {base} rows in {len(SHARDS)} shards on the GPU")
    b = gpu.Batcher(shards, max_images=256, hit_capacity=1 << 21)
    rng = np.random.default_rng(3)
    try:
        for batch in range(4):
            n = int(rng.integers(1, 257)) if batch else 256
            a = int(rng.integers(0, 256 - n + 1))
            say(f"[config3] batch {batch}: images {a}..{a + n}")
            h, _, _, total, hits = b.run(pool[a:a + n])
            assert np.array_equal(h.cpu().numpy().reshape(-1, 32), pool_q[8 * a:8 * (a + n)])
            assert total == hits.shape[0], "hit buffer overflow"
            t = triples(hits)
            q = pool_q[8 * a:8 * (a + n)]
            pool_to_query = np.full(pool_q.shape[0], -1)
            pool_to_query[8 * a:8 * (a + n)] = np.arange(8 * n)
            for k, (rows, qids, dists) in enumerate(plants):
                assert_plants_found(t, rows + bases[k], qids, dists, pool_to_query, f"shard {SHARDS[k]}",
                                    require_some=SHARDS[k] >= 1_000_000)
                lo = 0 if SHARDS[k] <= 1_000_000 else int(rng.integers(0, SHARDS[k] - 1_000_000))
                hi = min(SHARDS[k], lo + 1_000_000)
                exp = brute(oracle, parts[k][lo:hi], q, index_base=bases[k] + lo)
                sel = t[(t[:, 1] >= bases[k] + lo) & (t[:, 1] < bases[k] + hi)]
                assert np.array_equal(sel, exp), f"shard {SHARDS[k]} rows {lo}..{hi}"
    finally:
        b.close()
        shards.close()


# --------------------------------------------------------------------------- config 4

def test_config4_server_100m(gpu, dev, oracle):
    """The serving loop over one 100M-row shard (3.2 GB, 10,000 plants from
    a 96-image pool): 24 requests of U{1..64} u8 images, each served alone,
    then all of them from four client threads at once.  Each request's
    result is the same alone and in the mix; its hashes are hash_batch's;
    every plant of its images within the threshold is found with its exact
    distance; its hits in a 1M-row window equal the brute force there; two
    whole requests equal the brute force over all 100M rows."""
    import torch
    imgs = synthetic.images_u8(96, start=400)
    ph, pq, pt = gpu.hash_batch(torch.from_numpy(imgs).to(dev))
    ph, pq, pt = ph.cpu().numpy(), pq.cpu().numpy(), pt.cpu().numpy()
    pool_q = ph.reshape(-1, 32)
    say("[config4] building 100M rows")
    db = synthetic.random_db(100_000_000, seed=17)
    rows, qids, dists = plant_masks(db, pool_q, 10_000, seed=18)
    shards = gpu.Shards()
    shards.add(db)
    rng = np.random.default_rng(19)
    reqs = [rng.choice(96, size=int(rng.integers(1, 65)), replace=False) for _ in range(24)]
    srv = gpu.Server(shards, max_batch_images=64, hbm_budget_bytes=8 << 30, max_contexts=4, threshold=THR)
    try:
        say("[config4] alone")
        alone = []
        for idx in reqs:
            rid = srv.submit(imgs[idx])
            alone.append(srv.wait(rid, timeout_ms=60_000))
        say("[config4] mix of four clients")
        mixed = [None] * len(reqs)

        def client(c):
            for i in range(c, len(reqs), 4):
                rid = srv.submit(imgs[reqs[i]])
                mixed[i] = srv.wait(rid, timeout_ms=60_000)

        th = [threading.Thread(target=client, args=(c,)) for c in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        shared = sum(1 for i, r in enumerate(mixed) if r["batch_images"] != len(reqs[i]))
        say(f"[config4] requests that shared a batch: {shared}")
    finally:
        srv.close()
        shards.close()
    for i, idx in enumerate(reqs):
        a, m = alone[i], mixed[i]
        assert a["status"] == 0 and m["status"] == 0
        assert np.array_equal(a["hashes"], ph[idx]) and np.array_equal(m["hashes"], ph[idx])
        assert np.array_equal(a["quality"].view(np.uint32), pq[idx].view(np.uint32))
        assert np.array_equal(a["threshold"].view(np.uint32), pt[idx].view(np.uint32))
        ta, tm = triples(a["hits"]), triples(m["hits"])
        assert np.array_equal(ta, tm), f"request {i}: alone != mix"
        # pool query 8 * image + d is request query 8 * position + d
        pool_to_query = np.full(pool_q.shape[0], -1)
        for pos, im in enumerate(idx):
            pool_to_query[8 * im:8 * im + 8] = 8 * pos + np.arange(8)
        assert_plants_found(ta, rows, qids, dists, pool_to_query, f"request {i}")
        q = ph[idx].reshape(-1, 32)
        lo = int(rng.integers(0, 99_000_000))
        exp = brute(oracle, db[lo:lo + 1_000_000], q, index_base=lo)
        sel = ta[(ta[:, 1] >= lo) & (ta[:, 1] < lo + 1_000_000)]
        assert np.array_equal(sel, exp), f"request {i}: window {lo}"
    for i in sorted(range(len(reqs)), key=lambda i: len(reqs[i]))[:2]:
        say(f"[config4] request {i} ({len(reqs[i])} images) against all 100M rows")
        exp = brute(oracle, db, ph[reqs[i]].reshape(-1, 32))
        assert np.array_equal(triples(alone[i]["hits"]), exp)
