"""GPU: yume_pdq_amd_hash_match_batch, the one-call hash + match of a batch,
against the same batch through the separate entry points (hash_batch, then
match / match_with_plane): identical hashes, qualities, thresholds and hit
sets, and the hit count stored into page-locked host memory.  The separate
entry points are themselves pinned to the oracle by test_gpu_hash.py and
test_gpu_match.py."""
import sys

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def triples(hits):
    return sorted(zip(hits["query"].tolist(), hits["index"].tolist(), hits["distance"].tolist()))


@pytest.fixture(scope="module")
def batch(gpu, dev):
    import torch
    planes = synthetic.images_u8_device(40, dev, start=7)
    h, _, _ = gpu.hash_batch(planes)
    queries = h.reshape(-1, 32).cpu().numpy()
    db = synthetic.random_db(200_000, seed=21)
    synthetic.plant(db, queries, 3000, seed=22)
    d_db = torch.from_numpy(db).to(dev)
    return planes, d_db, gpu.build_prefix_plane(d_db)


@pytest.mark.parametrize("n", [1, 3, 40])
@pytest.mark.parametrize("use_plane", [True, False])
def test_equals_separate_calls(gpu, dev, batch, n, use_plane):
    planes, d_db, d_plane = batch
    plane = d_plane if use_plane else None
    say(f"[hash_match] n={n} plane={use_plane}")
    h, q, t, total, hits = gpu.hash_match(planes[:n], d_db, 31, plane=plane, capacity=1 << 16)
    h2, q2, t2 = gpu.hash_batch(planes[:n])
    total2, hits2 = gpu.match(d_db, h2.reshape(-1, 32), 31, capacity=1 << 16, plane=plane)
    assert np.array_equal(h.cpu().numpy(), h2.cpu().numpy())
    assert np.array_equal(q.cpu().numpy().view(np.uint32), q2.cpu().numpy().view(np.uint32))
    assert np.array_equal(t.cpu().numpy().view(np.uint32), t2.cpu().numpy().view(np.uint32))
    assert total == total2 and total > 0
    assert triples(hits) == triples(hits2)


def test_repeated_runs_reset_the_count(gpu, dev, batch):
    """The counter is reset by every call: two runs of one batch, then a
    smaller batch, each count the hits of that batch alone."""
    planes, d_db, d_plane = batch
    hm = gpu.HashMatch(planes[:8], d_db, 31, plane=d_plane)
    first = hm.run()
    assert hm.run() == first
    assert int(hm.count.item()) == first
    small = gpu.HashMatch(planes[:1], d_db, 31, plane=d_plane)
    _, want = gpu.match(d_db, gpu.hash_batch(planes[:1])[0].reshape(-1, 32), 31, plane=d_plane)
    assert small.run() == want.size


def test_f32_planes(gpu, dev, batch):
    import torch
    planes, d_db, _ = batch
    f = planes[:2].to(torch.float32)
    h, _, _, total, hits = gpu.hash_match(f, d_db, 31)
    h2, _, _ = gpu.hash_batch(f)
    assert np.array_equal(h.cpu().numpy(), h2.cpu().numpy())
    total2, hits2 = gpu.match(d_db, h2.reshape(-1, 32), 31)
    assert total == total2 and triples(hits) == triples(hits2)


def test_argument_errors(gpu, dev, batch):
    import ctypes as C
    import torch
    planes, d_db, d_plane = batch
    L = gpu.lib()
    hashes = torch.empty((2, 8, 32), dtype=torch.uint8, device=dev)
    qual = torch.empty((2, 8), dtype=torch.float32, device=dev)
    thr = torch.empty((2, 8), dtype=torch.float32, device=dev)
    hits = torch.empty((16, 16), dtype=torch.uint8, device=dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    args = [C.c_void_p(planes.data_ptr()), gpu.INPUT_U8, 512 * 512, 2, 0, C.c_void_p(d_db.data_ptr()),
            C.c_void_p(d_plane.data_ptr()), d_db.shape[0], 31, C.c_void_p(hashes.data_ptr()),
            C.c_void_p(qual.data_ptr()), C.c_void_p(thr.data_ptr()), C.c_void_p(hits.data_ptr()), 16,
            C.c_void_p(count.data_ptr()), None, None]
    assert L.yume_pdq_amd_hash_match_batch(*args) == 0
    torch.cuda.synchronize()
    bad = list(args)
    bad[14] = None  # no device counter
    assert L.yume_pdq_amd_hash_match_batch(*bad) < 0
    bad = list(args)
    bad[5] = C.c_void_p(d_db.data_ptr() + 4)  # misaligned database
    assert L.yume_pdq_amd_hash_match_batch(*bad) < 0
    bad = list(args)
    bad[2] = 512 * 512 - 16  # stride shorter than a plane
    assert L.yume_pdq_amd_hash_match_batch(*bad) < 0
    # an empty batch stores a zero count
    pinned = torch.full((1,), 7, dtype=torch.int64).pin_memory()
    empty = list(args)
    empty[3] = 0
    empty[15] = C.c_void_p(pinned.data_ptr())
    assert L.yume_pdq_amd_hash_match_batch(*empty) == 0
    torch.cuda.synchronize()
    assert int(pinned[0]) == 0 and int(count.item()) == 0
