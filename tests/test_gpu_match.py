"""GPU parity of the Hamming matcher: exact (query, index, distance) sets.

Oracles: the C brute force (oracle/pdq_oracle.c) at small sizes, numpy
bitwise_count brute force at larger sizes, and the reference's planted-control
protocol (src/matching/mod.rs:460-538).
"""
import sys

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def as_triples(hits):
    return np.stack([hits["query"].astype(np.int64), hits["index"].astype(np.int64),
                     hits["distance"].astype(np.int64)], axis=1) if hits.size else np.zeros((0, 3), np.int64)


def sort_triples(t):
    if t.shape[0] == 0:
        return t
    return t[np.lexsort((t[:, 0], t[:, 1]))]


@pytest.mark.parametrize("nq", [1, 8, 24, 64])
def test_planted_controls(gpu, oracle, dev, nq):
    """Reference protocol: plant a control with 31 flips (with replacement, so
    distance <= 31) at each position for each needle; find must report it."""
    import torch
    rng = np.random.default_rng(int.from_bytes(b"yume-pdq", "big") % (2 ** 63))
    hay = rng.integers(0, 256, (2048, 32), dtype=np.uint8)
    needles = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    d_hay = torch.from_numpy(hay).to(dev)
    say(f"[match] planted controls nq={nq}: negative control")
    total, hits = gpu.match(d_hay, torch.from_numpy(needles).to(dev), 31)
    assert total == 0
    assert gpu.match_first(d_hay, torch.from_numpy(needles).to(dev), 31) is None
    positions = rng.choice(2048, size=64, replace=False)
    for pos in positions:
        i = int(rng.integers(0, nq))
        nd = needles.copy()
        nd[i] = hay[pos]
        for _ in range(31):
            b = int(rng.integers(0, 32))
            nd[i, b] ^= np.uint8(1 << int(rng.integers(0, 8)))
        d_nd = torch.from_numpy(nd).to(dev)
        total, hits = gpu.match(d_hay, d_nd, 31)
        exp_n, exp = oracle.match_all(hay, nd, 31)
        assert total == exp_n >= 1
        assert np.array_equal(sort_triples(as_triples(hits)), exp)
        first = gpu.match_first(d_hay, d_nd, 31)
        assert first == oracle.match_first(hay, nd, 31)
        assert first == (int(pos), i)


THRESHOLD_CASES = [(t, p) for t in (0, 1, 31, 63, 64, 128, 200, 255, 256)
                   for p in ("valu", "mfma", "valu_plane", "mfma_plane")
                   if not (p.startswith("mfma") and t >= 64)]


def run_path(gpu, d_db, d_q, threshold, path, **kw):
    """path: valu / mfma over the raw rows, valu_plane / mfma_plane over the
    prefix plane."""
    if path.endswith("_plane"):
        return gpu.match(d_db, d_q, threshold, path=path[:-6], plane=gpu.build_prefix_plane(d_db), **kw)
    return gpu.match(d_db, d_q, threshold, path=path, **kw)


@pytest.mark.parametrize("threshold,path", THRESHOLD_CASES)
def test_thresholds(gpu, oracle, dev, threshold, path):
    import torch
    db = synthetic.random_db(3000, seed=3)
    q = synthetic.random_db(40, seed=4)
    synthetic.plant(db, q, 200, seed=5, max_dist=70)
    say(f"[match] threshold={threshold} path={path}")
    exp_n, exp = oracle.match_all(db, q, threshold, cap=3000 * 40)
    total, hits = run_path(gpu, torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev), threshold, path,
                           capacity=3000 * 40)
    assert total == exp_n
    assert np.array_equal(sort_triples(as_triples(hits)), exp)


@pytest.mark.parametrize("plane", [False, True])
@pytest.mark.parametrize("nq", [1, 15, 16, 17, 33, 100, 257, 385, 513, 1000, 2048])
def test_mfma_query_counts(gpu, oracle, dev, nq, plane):
    """Ragged query counts on the matrix-core path (partial 16-query tiles,
    partial query groups) against the oracle, ragged row counts too (5003
    rows: a partial last tile and a partial last wave), raw rows and plane.
    From 4 query groups the block's waves share their rows (385: a last wave
    with one query; 513: a second group of waves with one active; 1000, 2048:
    full and partial groups of waves)."""
    import torch
    db = synthetic.random_db(5003, seed=40 + nq)
    q = synthetic.random_db(nq, seed=41)
    synthetic.plant(db, q, 300, seed=42, max_dist=45)
    say(f"[match] mfma nq={nq} plane={plane}")
    exp_n, exp = oracle.match_all(db, q, 31)
    total, hits = run_path(gpu, torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev), 31,
                           "mfma_plane" if plane else "mfma")
    assert total == exp_n
    assert np.array_equal(sort_triples(as_triples(hits)), exp)


@pytest.mark.parametrize("plane", [False, True])
def test_mfma_dense_candidates(gpu, oracle, dev, plane):
    """Threshold 63 puts the prefix bound at 2, which about half of all
    random pairs reach: every group of tiles of every wave is a candidate and
    goes to the second pass, pair by pair.  150,003 rows x 2048 queries (147
    shared sub-chunks x 16 query groups, a partial last sub-chunk) against the
    oracle."""
    import torch
    db = synthetic.random_db(150_003, seed=90)
    q = synthetic.random_db(2048, seed=91)
    synthetic.plant(db, q, 2000, seed=92, max_dist=70)
    say(f"[match] mfma dense candidates plane={plane}")
    exp_n, exp = oracle.match_all(db, q, 63)
    total, hits = run_path(gpu, torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev), 63,
                           "mfma_plane" if plane else "mfma")
    assert total == exp_n
    assert np.array_equal(sort_triples(as_triples(hits)), exp)


@pytest.mark.parametrize("threshold", [0, 20, 27, 28, 29, 31, 45, 62, 63])
@pytest.mark.parametrize("plane", [False, True])
def test_mfma_sixteen_tiles_thresholds(gpu, oracle, dev, threshold, plane):
    """2048 queries (sixteen query tiles per wave, rows shared) at thresholds
    on both sides of where three products share an accumulator (28), with
    pairs built against its flag fields: row p sits exactly at the bound
    from query 16 + p (query tile 1, the second product of an accumulator),
    while queries p and 32 + p (tiles 0 and 2, the same accumulator and
    lane) are near-complements of its prefix, whose fields borrow; the same
    again for tiles 8-10.  Plus 300 plants at distances up to 70 among 4000
    random rows.  Every set equals the oracle's."""
    import torch
    rng = np.random.default_rng(500 + threshold)
    q = synthetic.random_db(2048, seed=501)
    rows = []
    for base in (0, 128):  # query tiles 0-2 and 8-10
        for p in range(16):
            target = q[base + 16 + p].copy()
            row = target.copy()
            for b in rng.choice(128, size=threshold, replace=False):  # prefix bits only: d256 = threshold
                row[b // 8] ^= np.uint8(1 << (b % 8))
            rows.append(row)
            for off in (0, 32):  # prefix distance 128 - k from the row: far past 64 + t
                comp = row.copy()
                comp[:16] ^= np.uint8(0xFF)
                for b in rng.choice(128, size=int(rng.integers(0, 4)), replace=False):
                    comp[b // 8] ^= np.uint8(1 << (b % 8))
                q[base + off + p] = comp
    db = np.concatenate([synthetic.random_db(4000, seed=502), np.stack(rows)])
    synthetic.plant(db[:4000], q, 300, seed=503, max_dist=70)
    say(f"[match] sixteen tiles threshold={threshold} plane={plane}")
    exp_n, exp = oracle.match_all(db, q, threshold, cap=1 << 20)
    total, hits = run_path(gpu, torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev), threshold,
                           "mfma_plane" if plane else "mfma", capacity=1 << 20)
    assert total == exp_n and exp_n >= 32  # the 32 rows built at the bound
    assert np.array_equal(sort_triples(as_triples(hits)), exp)


def near(base, n, max_flips, seed):
    """n copies of the 32-byte hash `base`, each with up to max_flips bits flipped."""
    rng = np.random.default_rng(seed)
    out = np.repeat(base[None, :], n, axis=0)
    for i in range(n):
        for b in rng.choice(256, size=int(rng.integers(0, max_flips + 1)), replace=False):
            out[i, b // 8] ^= np.uint8(1 << (b % 8))
    return out


@pytest.mark.parametrize("plane", [False, True])
def test_mfma_every_pair_a_hit(gpu, oracle, dev, plane):
    """2048 queries and 300 rows, all within 8 bits of one hash: every pair is
    a hit at threshold 31 (614,400 hits from 16 candidate records, far more
    than a block of the second pass gathers before it appends directly).  The
    set equals the oracle's; with a 1000-hit buffer the count stays exact and
    the stored hits are 1000 distinct pairs of the set."""
    import torch
    base = synthetic.random_db(1, seed=96)[0]
    q = near(base, 2048, 8, seed=97)
    db = near(base, 300, 8, seed=98)
    d_db, d_q = torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev)
    say(f"[match] every pair a hit plane={plane}")
    exp_n, exp = oracle.match_all(db, q, 31, cap=300 * 2048)
    assert exp_n == 300 * 2048
    total, hits = run_path(gpu, d_db, d_q, 31, "mfma_plane" if plane else "mfma", capacity=1 << 20)
    assert total == exp_n
    assert np.array_equal(sort_triples(as_triples(hits)), exp)
    say(f"[match] every pair a hit, 1000-hit buffer plane={plane}")
    total, hits = run_path(gpu, d_db, d_q, 31, "mfma_plane" if plane else "mfma", capacity=1000)
    assert total == exp_n and hits.size == 1000
    t = as_triples(hits)
    assert len(set(map(tuple, t.tolist()))) == 1000
    assert set(map(tuple, t.tolist())) <= set(map(tuple, exp.tolist()))


def test_mfma_candidate_records_overflow(gpu, dev, record_cap_4096):
    """1,000,003 rows x 2048 queries at threshold 63: 977 sub-chunks x 16
    query groups, every one a candidate, are more candidate records than the
    4096 the first pass may keep here (set_match_record_cap;
    the default is 1,048,576), so the second pass matches every work item again
    with per-wave candidate checks.  The set equals the vector-ALU kernel's,
    and holds the planted pairs."""
    import torch
    db = synthetic.random_db(1_000_003, seed=93)
    q = synthetic.random_db(2048, seed=94)
    synthetic.plant_fast(db, q, 20_000, seed=95)
    d_db, d_q = torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev)
    for plane in (None, gpu.build_prefix_plane(d_db)):
        say(f"[match] record overflow: mfma plane={plane is not None}")
        n1, h1 = gpu.match(d_db, d_q, 63, capacity=1 << 20, path="mfma", plane=plane)
        say("[match] record overflow: valu")
        n2, h2 = gpu.match(d_db, d_q, 63, capacity=1 << 20, path="valu")
        assert n1 == n2 and np.array_equal(as_triples(h1), as_triples(h2))
        assert n1 >= 0.9 * 20_000


def test_prefix_plane_layout(gpu, dev):
    """The plane holds the first 16 bytes of every row, packed, zero-padded to
    whole 16-row tiles (checked on the host against numpy)."""
    import torch
    for n in (1, 15, 16, 17, 1000):
        db = synthetic.random_db(n, seed=n)
        say(f"[match] prefix plane layout n={n}")
        plane = gpu.build_prefix_plane(torch.from_numpy(db).to(dev)).cpu().numpy()
        padded = (n + 15) // 16 * 16
        exp = np.zeros((padded, 16), np.uint8)
        exp[:n] = db[:, :16]
        assert np.array_equal(plane.reshape(-1, 16), exp)


def test_empty_and_tiny(gpu, dev):
    import torch
    q = torch.from_numpy(synthetic.random_db(8, seed=1)).to(dev)
    say("[match] empty db")
    empty = torch.empty((0, 32), dtype=torch.uint8, device=dev)
    assert gpu.match(empty, q, 31)[0] == 0
    assert gpu.match_first(empty, q, 31) is None
    say("[match] db of 1 row equal to a query")
    one = q[3:4].clone()
    total, hits = gpu.match(one, q, 0)
    assert total >= 1 and (hits["query"] == 3).any() and (hits["distance"][hits["query"] == 3] == 0).all()


def test_overflow_reports_exact_count(gpu, oracle, dev):
    """More hits than capacity: the count stays exact, the stored hits are a subset."""
    import torch
    db = synthetic.random_db(5000, seed=21)
    q = synthetic.random_db(4, seed=22)
    say("[match] overflow, threshold 256 (every pair)")
    total, hits = gpu.match(torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev), 256, capacity=1000)
    assert total == 5000 * 4
    assert hits.size == 1000
    assert len(set(zip(hits["query"].tolist(), hits["index"].tolist()))) == 1000


@pytest.mark.parametrize("nq", [64, 1024])
def test_counter_reset_without_append(gpu, oracle, dev, nq):
    """Without the append flag the matrix-core matcher clears the caller's
    counter itself (in the kernel that clears its candidate records): a
    counter holding garbage reads the exact count afterwards, also for an
    empty database, and with the append flag it keeps adding.  64 queries
    run the per-wave layout, 1024 the shared-rows one."""
    import torch
    db = synthetic.random_db(5003, seed=70)
    q = synthetic.random_db(nq, seed=71)
    synthetic.plant(db, q, 200, seed=72)
    d_db, d_q = torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev)
    exp_n, exp = oracle.match_all(db, q, 31)
    count = torch.full((1,), 123456789, dtype=torch.int64, device=dev)
    say(f"[match] counter reset nq={nq}")
    total, hits = gpu.match(d_db, d_q, 31, count=count, path="mfma")
    assert total == exp_n and np.array_equal(sort_triples(as_triples(hits)), exp)
    total, _ = gpu.match(d_db, d_q, 31, count=count, path="mfma", append=True)
    assert total == 2 * exp_n
    empty = torch.empty((0, 32), dtype=torch.uint8, device=dev)
    assert gpu.match(empty, d_q, 31, count=count, path="mfma")[0] == 0


def test_record_count_slots_alternate(gpu, oracle, dev):
    """Shared-rows calls on one stream take the record area's two count slots
    in turn: each call's streaming kernel clears the slot of the stream's
    next call and, without the append flag, the caller's hit counter, so only
    a stream's first call (or one after a failed launch) needs a clearing
    launch.  A sequence of shared-rows calls mixed with per-wave calls (64
    queries), an append, garbage counters and a call whose records overflow
    the cap stays exact call after call."""
    import torch
    db = synthetic.random_db(20011, seed=80)
    q = synthetic.random_db(1024, seed=81)
    synthetic.plant(db, q, 300, seed=82)
    d_db, d_q = torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev)
    exp_n, exp = oracle.match_all(db, q, 31)
    exp64_n, exp64 = oracle.match_all(db, q[:64], 31)
    plane = gpu.build_prefix_plane(d_db)
    stream = torch.cuda.Stream()
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    steps = ["shared", "shared", "append", "per-wave", "shared", "overflow", "shared", "shared", "per-wave", "shared"]
    with torch.cuda.stream(stream):
        for i, kind in enumerate(steps):
            say(f"[match] count slots: step {i} {kind}")
            append = kind == "append"
            if not append:
                count.fill_(987654321)
            if kind == "overflow":
                gpu.set_match_record_cap(8)
            try:
                queries = d_q[:64] if kind == "per-wave" else d_q
                total, hits = gpu.match(d_db, queries, 31, capacity=1 << 16, count=count, path="mfma",
                                        plane=plane if i % 2 else None, append=append, stream=stream)
            finally:
                if kind == "overflow":
                    gpu.set_match_record_cap(0)
            if kind == "per-wave":
                assert total == exp64_n and np.array_equal(sort_triples(as_triples(hits)), exp64)
            elif append:
                assert total == 2 * exp_n
            else:
                assert total == exp_n and np.array_equal(sort_triples(as_triples(hits)), exp)


def test_explicit_stream_and_per_thread_streams(gpu, oracle, dev):
    """match() on a stream that is not torch's current one waits for that
    stream before it reads the count; and two host threads matching at once
    on hipStreamPerThread (one handle, a different stream in each thread) get
    record areas of their own, so both stay exact call after call."""
    import threading

    import torch
    db = synthetic.random_db(30011, seed=90)
    qa = synthetic.random_db(1024, seed=91)
    qb = synthetic.random_db(1024, seed=92)
    synthetic.plant(db, qa, 200, seed=93)
    synthetic.plant(db, qb, 200, seed=94)
    d_db = torch.from_numpy(db).to(dev)
    d_qa, d_qb = torch.from_numpy(qa).to(dev), torch.from_numpy(qb).to(dev)
    exp = {0: oracle.match_all(db, qa, 31), 1: oracle.match_all(db, qb, 31)}
    say("[match] explicit stream, read outside its context")
    s = torch.cuda.Stream()
    total, hits = gpu.match(d_db, d_qa, 31, path="mfma", stream=s)
    assert total == exp[0][0] and np.array_equal(sort_triples(as_triples(hits)), exp[0][1])
    torch.cuda.synchronize()
    per_thread = torch.cuda.ExternalStream(2)  # hipStreamPerThread
    results = {0: [], 1: []}

    def worker(k):
        for _ in range(4):
            n, h = gpu.match(d_db, d_qa if k == 0 else d_qb, 31, path="mfma", stream=per_thread)
            results[k].append((n, sort_triples(as_triples(h))))

    say("[match] two threads on hipStreamPerThread")
    threads = [threading.Thread(target=worker, args=(k,)) for k in (0, 1)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=60)
    for k in (0, 1):
        assert len(results[k]) == 4
        for n, t in results[k]:
            assert n == exp[k][0] and np.array_equal(t, exp[k][1])


def test_index_base_and_append(gpu, oracle, dev):
    """Two shards into one hit buffer with global indices (sharded DB)."""
    import torch
    db = synthetic.random_db(4000, seed=31)
    q = synthetic.random_db(8, seed=32)
    synthetic.plant(db, q, 100, seed=33)
    d_db = torch.from_numpy(db).to(dev)
    d_q = torch.from_numpy(q).to(dev)
    say("[match] two shards, append")
    hits, count = gpu.match(d_db[:1500], d_q, 31, capacity=4096, sync=False)
    gpu.match(d_db[1500:], d_q, 31, index_base=1500, hits=hits, count=count, append=True, sync=False)
    total = int(count.item())
    arr = np.sort(hits[:total].cpu().numpy().copy().view(gpu.HIT_DTYPE).reshape(-1), order=("index", "query"))
    exp_n, exp = oracle.match_all(db, q, 31)
    assert total == exp_n
    assert np.array_equal(as_triples(arr), exp)


@pytest.mark.parametrize("n_rows,nq", [(1_000_000, 8), (1_000_000, 128), (10_000_000, 8)])
def test_scale_exact_set(gpu, dev, n_rows, nq):
    """Config-sized DBs with planted near-duplicates: exact set vs numpy brute
    force, over the raw rows and over the prefix plane."""
    import torch
    db = synthetic.random_db(n_rows, seed=7)
    q = synthetic.random_db(nq, seed=8)
    synthetic.plant(db, q, max(1000, n_rows // 200), seed=12)
    d_db, d_q = torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev)
    say(f"[match] scale n_rows={n_rows} nq={nq}")
    total, hits = gpu.match(d_db, d_q, 31, capacity=1 << 20)
    exp = synthetic.brute_force_hits(db, q, 31)
    assert total == exp.shape[0]
    assert np.array_equal(as_triples(hits), exp)
    say(f"[match] scale n_rows={n_rows} nq={nq} prefix plane")
    total, hits = gpu.match(d_db, d_q, 31, capacity=1 << 20, plane=gpu.build_prefix_plane(d_db))
    assert total == exp.shape[0]
    assert np.array_equal(as_triples(hits), exp)


def test_scale_10m_x_1024_paths_agree(gpu, dev):
    """10M rows x 1024 queries (8 x 128 images' worth), 50k plants: the
    matrix-core and vector-ALU kernels return the identical set, and the
    first 500k rows agree with a numpy brute force."""
    import torch
    db = synthetic.random_db(10_000_000, seed=7)
    q = synthetic.random_db(1024, seed=8)
    synthetic.plant_fast(db, q, 50_000, seed=12)
    d_db, d_q = torch.from_numpy(db).to(dev), torch.from_numpy(q).to(dev)
    say("[match] 10M x 1024 mfma")
    n1, h1 = gpu.match(d_db, d_q, 31, capacity=1 << 20, path="mfma")
    say("[match] 10M x 1024 valu")
    n2, h2 = gpu.match(d_db, d_q, 31, capacity=1 << 20, path="valu")
    assert n1 == n2 and np.array_equal(as_triples(h1), as_triples(h2))
    say("[match] 10M x 1024 mfma over the prefix plane")
    n3, h3 = gpu.match(d_db, d_q, 31, capacity=1 << 20, path="mfma", plane=gpu.build_prefix_plane(d_db))
    assert n3 == n1 and np.array_equal(as_triples(h3), as_triples(h1))
    assert n1 >= 0.9 * 50_000 * 32 / 41  # plants at d <= 31 are hits
    sub = synthetic.brute_force_hits(db[:500_000], q, 31)
    t1 = as_triples(h1)
    assert np.array_equal(t1[t1[:, 1] < 500_000], sub)
