"""GPU: yume_pdq_amd_match_noalloc, the matcher entry point that allocates
nothing (its candidate records live in the library's static per-GPU scratch
area), called through the C ABI, returns exactly what the allocating entry
points (yume_pdq_amd_match_with_plane, yume_pdq_amd_match) return -- on two
calls in a row, the second smaller, with and without the prefix plane, on
every matcher path; it refuses a call whose worst case exceeds the area
before enqueueing anything; and it can be captured into a HIP graph, whose
replays (with eager work between them) equal the eager call."""
import ctypes as C
import sys

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu

ESCRATCH = -9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def vp(t):
    return C.c_void_p(t.data_ptr() if t is not None else 0)


def call(gpu, fn, d_db, plane, d_q, thr, hits, count, flags=0, stream=None):
    """One call of a match entry point through the C ABI; returns its status."""
    import torch
    s = C.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)
    L = gpu.lib()
    if fn == "noalloc":
        return L.yume_pdq_amd_match_noalloc(vp(d_db), vp(plane), d_db.shape[0], vp(d_q), d_q.shape[0], thr,
                                            vp(hits), hits.shape[0], vp(count), 0, flags, s)
    if plane is not None:
        return L.yume_pdq_amd_match_with_plane(vp(d_db), vp(plane), d_db.shape[0], vp(d_q), d_q.shape[0], thr,
                                               vp(hits), hits.shape[0], vp(count), 0, flags, s)
    return L.yume_pdq_amd_match(vp(d_db), d_db.shape[0], vp(d_q), d_q.shape[0], thr, vp(hits), hits.shape[0],
                                vp(count), 0, flags, s)


def result(hits, count):
    import torch
    torch.cuda.synchronize()
    n = int(count.item())
    k = min(n, hits.shape[0])
    arr = hits[:k].cpu().numpy().copy().view(np.dtype([("query", "<u4"), ("distance", "<u4"),
                                                        ("index", "<u8")])).reshape(-1)
    t = np.stack([arr["query"].astype(np.int64), arr["index"].astype(np.int64),
                  arr["distance"].astype(np.int64)], axis=1) if k else np.zeros((0, 3), np.int64)
    return n, t[np.lexsort((t[:, 0], t[:, 1]))] if k else t


@pytest.fixture(scope="module")
def data(gpu, dev):
    import torch
    q = synthetic.random_db(1024, seed=701)
    db = synthetic.random_db(1_000_000, seed=702)
    synthetic.plant_fast(db, q, 20_000, seed=703)
    d_db = torch.from_numpy(db).to(dev)
    return d_db, gpu.build_prefix_plane(d_db), torch.from_numpy(q).to(dev)


def buffers(dev, cap=1 << 18):
    import torch
    return (torch.empty((cap, 16), dtype=torch.uint8, device=dev),
            torch.full((1,), 12345, dtype=torch.int64, device=dev))


@pytest.mark.parametrize("use_plane", [True, False])
def test_two_calls_equal_the_allocating_entry_point(gpu, dev, data, use_plane):
    """1M rows x 1024 queries (the shared-rows layout: records in the static
    area), then 300k rows x 600 queries on the same area: each equals the
    allocating entry point, which the other GPU tests pin to the oracle."""
    d_db, d_plane, d_q = data
    cases = [(d_db, d_plane, d_q), (d_db[:300_000], gpu.build_prefix_plane(d_db[:300_000]), d_q[:600])]
    for k, (db, plane, q) in enumerate(cases):
        plane = plane if use_plane else None
        say(f"[noalloc] call {k}: {db.shape[0]} rows x {q.shape[0]} queries, plane={use_plane}")
        assert gpu.match_uses_matrix_cores(q.shape[0], 31)
        h1, c1 = buffers(dev)
        assert call(gpu, "noalloc", db, plane, q, 31, h1, c1) == 0
        n1, t1 = result(h1, c1)
        h2, c2 = buffers(dev)
        assert call(gpu, "alloc", db, plane, q, 31, h2, c2) == 0
        n2, t2 = result(h2, c2)
        assert n1 == n2 > 0
        assert np.array_equal(t1, t2)


@pytest.mark.parametrize("nq,flags", [(8, 0), (24, 2), (64, 0), (256, 4), (2048, 0)])
def test_every_path(gpu, dev, data, nq, flags):
    """Vector-ALU (8, and 24 forced), the per-wave matrix-core layout (64,
    256 forced: no records) and the shared-rows layout (2048) over 200k rows
    at threshold 40, each equal to the allocating entry point; the counter
    starts from garbage and is reset by the call."""
    d_db, d_plane, d_q = data
    db, q = d_db[:200_000], d_q[:nq]
    plane = gpu.build_prefix_plane(db)
    say(f"[noalloc] path nq={nq} flags={flags}")
    h1, c1 = buffers(dev)
    assert call(gpu, "noalloc", db, plane, q, 40, h1, c1, flags) == 0
    h2, c2 = buffers(dev)
    assert call(gpu, "alloc", db, plane, q, 40, h2, c2, flags) == 0
    n1, t1 = result(h1, c1)
    n2, t2 = result(h2, c2)
    assert n1 == n2 > 0 and np.array_equal(t1, t2)


def test_append_flag(gpu, dev, data):
    """MATCH_APPEND keeps the count: two appended calls hold the same set as
    the allocating entry point's two appended calls."""
    d_db, d_plane, d_q = data
    got = []
    for fn in ("noalloc", "alloc"):
        h, c = buffers(dev)
        c.zero_()
        assert call(gpu, fn, d_db, d_plane, d_q[:512], 31, h, c, 1) == 0
        assert call(gpu, fn, d_db, d_plane, d_q[512:], 31, h, c, 1) == 0
        got.append(result(h, c))
    assert got[0][0] == got[1][0] > 0
    assert np.array_equal(got[0][1], got[1][1])


def test_too_large_is_refused_before_enqueueing(gpu, dev):
    """10M rows x 32,768 queries needs up to 1,250,304 records (one per
    1024 rows and 256-query group), more than the 2^20 the area holds:
    ESCRATCH, and the counter is left untouched (nothing ran).  26,624
    queries (1,015,872 records) fit and run."""
    import torch
    d_db = torch.randint(0, 256, (10_000_000, 32), dtype=torch.uint8, device=dev)
    d_q = torch.randint(0, 256, (32768, 32), dtype=torch.uint8, device=dev)
    plane = gpu.build_prefix_plane(d_db)
    h, c = buffers(dev, cap=1 << 10)
    say("[noalloc] 10M x 32768: refused")
    assert call(gpu, "noalloc", d_db, plane, d_q, 31, h, c) == ESCRATCH
    torch.cuda.synchronize()
    assert int(c.item()) == 12345
    assert gpu.lib().yume_pdq_amd_strerror(ESCRATCH).decode().startswith("the call needs more")
    say("[noalloc] 10M x 26624: runs")
    assert call(gpu, "noalloc", d_db, plane, d_q[:26624], 31, h, c) == 0
    n1, _ = result(h, c)
    h2, c2 = buffers(dev, cap=1 << 10)
    assert call(gpu, "alloc", d_db, plane, d_q[:26624], 31, h2, c2) == 0
    n2, _ = result(h2, c2)
    assert n1 == n2


def test_graph_capture(gpu, dev, data):
    """The no-allocation call captured into a HIP graph (through torch): three
    replays, with eager matcher calls on the same area between them, each
    equal to the eager result.  The allocating entry point, under capture,
    falls back to the per-wave layout; this one keeps the shared rows."""
    import torch
    d_db, d_plane, d_q = data
    h, c = buffers(dev)
    assert call(gpu, "noalloc", d_db, d_plane, d_q, 31, h, c) == 0  # eager (looks the area up)
    n0, t0 = result(h, c)
    hg, cg = buffers(dev)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            rc = call(gpu, "noalloc", d_db, d_plane, d_q, 31, hg, cg, stream=s)
    assert rc == 0
    torch.cuda.current_stream().wait_stream(s)
    for r in range(3):
        say(f"[noalloc] graph replay {r}")
        cg.fill_(777)
        g.replay()
        n, t = result(hg, cg)
        assert n == n0 and np.array_equal(t, t0)
        # eager work on the same static area between replays
        he, ce = buffers(dev)
        assert call(gpu, "noalloc", d_db[:100_000], None, d_q[:700], 31, he, ce) == 0
        torch.cuda.synchronize()
