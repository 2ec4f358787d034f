"""GPU: every batched entry point, called with its default arguments, returns
what the reference's x86 smart kernel returns -- hash_get_threshold then
visit_dihedrals with the AVX2 column negation (src/lib.rs:391-418,
src/kernel/x86.rs:330-376), the oracle's mode 1 -- on config-0 images, in
both kernel layouts; and the hit sets built on those hashes are the oracle's.

The last test records what the two dihedral modes mean for recall: the AVX2
kernel's FLOPPED, ROTATED_180, FLOPPED_ROTATED_270 and ROTATED_90 queries of
an image are far from the hash of the truly transformed image, the exact
mode's are near it."""
import sys

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu

N_IMAGES = 24
REFERENCE_MODE = 1  # the oracle's AVX2 column extent


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def triples(hits):
    if hits.size == 0:
        return np.zeros((0, 3), np.int64)
    t = np.stack([hits["query"].astype(np.int64), hits["index"].astype(np.int64),
                  hits["distance"].astype(np.int64)], axis=1)
    return t[np.lexsort((t[:, 0], t[:, 1]))]


def oracle_triples(oracle, db, queries, thr):
    n, t = oracle.match_all(db, queries, thr)
    assert n == t.shape[0]
    return t[np.lexsort((t[:, 0], t[:, 1]))]


@pytest.fixture(scope="module")
def images():
    return synthetic.images_u8(N_IMAGES)


@pytest.fixture(scope="module")
def expected(oracle, images):
    H = np.zeros((N_IMAGES, 8, 32), np.uint8)
    Q = np.zeros((N_IMAGES, 8), np.float32)
    T = np.zeros((N_IMAGES, 8), np.float32)
    P = np.zeros((N_IMAGES, 16, 16), np.float32)
    for i in range(N_IMAGES):
        H[i], Q[i], T[i], P[i] = oracle.hash_dihedrals(images[i].astype(np.float32), REFERENCE_MODE)
    return H, Q, T, P


@pytest.fixture(scope="module")
def planted_db(expected):
    H = expected[0]
    db = synthetic.random_db(300_000, seed=301)
    synthetic.plant_fast(db, H.reshape(-1, 32), 2000, seed=302)
    return db


def check_hashes(h, q, t, expected, n=N_IMAGES):
    H, Q, T, _ = expected
    h, q, t = (np.asarray(x) for x in (h, q, t))
    assert np.array_equal(h.reshape(n, 8, 32), H[:n]), "hashes differ from the reference's visit_dihedrals"
    assert np.array_equal(q.reshape(n, 8).view(np.uint32), Q[:n].view(np.uint32))
    assert np.array_equal(t.reshape(n, 8).view(np.uint32), T[:n].view(np.uint32))


def test_the_modes_differ_on_these_images(expected, oracle, images):
    """The check below is only meaningful if the AVX2 and exact modes give
    different hashes here: the four orientations whose column negation the
    AVX2 kernel cuts short differ in every image (and ROTATED_270 may too,
    since it is thresholded with the threshold of an orientation before it);
    NORMAL, FLIPPED and ROTATED_90_FLOPPED never differ."""
    H = expected[0]
    for i in range(8):
        He, _, _, _ = oracle.hash_dihedrals(images[i].astype(np.float32), 0)
        differ = {d for d in range(8) if not np.array_equal(He[d], H[i, d])}
        # FLOPPED, ROTATED_180, FLOPPED_ROTATED_270, ROTATED_90
        assert {1, 2, 6, 7} <= differ <= {1, 2, 5, 6, 7}, differ


@pytest.mark.parametrize("layout", ["auto", "wave", "block8", "block16", "block4"])
@pytest.mark.parametrize("kind", ["u8", "f32"])
def test_hash_batch_default(gpu, dev, images, expected, layout, kind):
    import torch
    say(f"[defaults] hash_batch layout={layout} kind={kind}")
    planes = torch.from_numpy(images).to(dev)
    if kind == "f32":
        planes = planes.to(torch.float32)
    lay = {"auto": gpu.LAYOUT_AUTO, "wave": gpu.LAYOUT_WAVE, "block8": gpu.LAYOUT_BLOCK8,
           "block4": gpu.LAYOUT_BLOCK4, "block16": gpu.LAYOUT_BLOCK16}[layout]
    h, q, t, p = gpu.hash_batch(planes, want_pdqf=True, layout=lay)
    check_hashes(h.cpu(), q.cpu(), t.cpu(), expected)
    assert np.array_equal(p.cpu().numpy().view(np.uint32), expected[3].view(np.uint32))


def test_hash_batch_without_options_is_the_reference(gpu, dev, images, expected):
    """yume_pdq_amd_hash_batch_ex with NULL options (the C ABI's default)."""
    import ctypes as C

    import torch
    planes = torch.from_numpy(images).to(dev)
    n = N_IMAGES
    h = torch.empty((n, 8, 32), dtype=torch.uint8, device=dev)
    q = torch.empty((n, 8), dtype=torch.float32, device=dev)
    t = torch.empty((n, 8), dtype=torch.float32, device=dev)
    rc = gpu.lib().yume_pdq_amd_hash_batch_ex(
        C.c_void_p(planes.data_ptr()), gpu.INPUT_U8, 512 * 512, n, None, C.c_void_p(h.data_ptr()),
        C.c_void_p(q.data_ptr()), C.c_void_p(t.data_ptr()), None, None,
        C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    check_hashes(h.cpu(), q.cpu(), t.cpu(), expected)


def test_dihedrals_batch_default(gpu, dev, expected):
    import torch
    H, Q, T, P = expected
    pdqf = torch.from_numpy(P).to(dev)
    thr = torch.from_numpy(np.ascontiguousarray(T[:, 0])).to(dev)
    h, q, t, _ = gpu.dihedrals_batch(pdqf, thr)
    check_hashes(h.cpu(), q.cpu(), t.cpu(), expected)


@pytest.mark.parametrize("use_plane", [True, False])
def test_hash_match_default(gpu, dev, oracle, images, expected, planted_db, use_plane):
    import torch
    say(f"[defaults] hash_match plane={use_plane}")
    d_db = torch.from_numpy(planted_db).to(dev)
    plane = gpu.build_prefix_plane(d_db) if use_plane else None
    h, q, t, total, hits = gpu.hash_match(torch.from_numpy(images).to(dev), d_db, 31, plane=plane,
                                          capacity=1 << 16)
    check_hashes(h.cpu(), q.cpu(), t.cpu(), expected)
    exp = oracle_triples(oracle, planted_db, expected[0].reshape(-1, 32), 31)
    assert total == exp.shape[0] > 0
    assert np.array_equal(triples(hits), exp)


def test_batcher_default(gpu, dev, oracle, images, expected, planted_db):
    import torch
    say("[defaults] batcher")
    shards = gpu.Shards()
    shards.add(planted_db[:120_000])
    shards.add(planted_db[120_000:])
    b = gpu.Batcher(shards, max_images=32)
    try:
        h, q, t, total, hits = b.run(torch.from_numpy(images).to(dev))
    finally:
        b.close()
        shards.close()
    check_hashes(h.cpu(), q.cpu(), t.cpu(), expected)
    exp = oracle_triples(oracle, planted_db, expected[0].reshape(-1, 32), 31)
    assert total == exp.shape[0] > 0
    assert np.array_equal(triples(hits), exp)


def test_server_default(gpu, oracle, images, expected, planted_db):
    say("[defaults] server")
    shards = gpu.Shards()
    shards.add(planted_db)
    srv = gpu.Server(shards, max_batch_images=32, hbm_budget_bytes=2 << 30, max_contexts=2)
    try:
        rid = srv.submit(images)
        r = srv.wait(rid, timeout_ms=60_000)
    finally:
        srv.close()
        shards.close()
    assert r["status"] == 0
    check_hashes(r["hashes"], r["quality"], r["threshold"], expected)
    exp = oracle_triples(oracle, planted_db, expected[0].reshape(-1, 32), 31)
    assert np.array_equal(triples(r["hits"]), exp)


def test_dihedral_recall_in_both_modes(gpu, dev, images):
    """Hash each image and its truly flipped / rotated copies on the GPU.  The
    exact mode's orientation queries of an image lie within a few bits of the
    NORMAL hash of the transformed image; four of the AVX2 mode's lie about
    64 bits away, past any usual threshold.  The other three agree in both
    modes.  (Oracle, 16 images: exact means 5-8 bits, AVX2 means 64-67 bits
    for FLOPPED, ROTATED_180, FLOPPED_ROTATED_270, ROTATED_90.)"""
    import torch
    tf = {1: np.fliplr, 2: lambda a: np.rot90(a, 2), 3: np.flipud, 4: lambda a: a.T,
          5: lambda a: np.rot90(a, 1), 6: lambda a: np.rot90(a, 2).T, 7: lambda a: np.rot90(a, 3)}
    imgs = images[:16]
    moved = np.stack([np.ascontiguousarray(tf[d](im)) for im in imgs for d in range(1, 8)])
    mean = {}
    for mode in (gpu.DIHEDRAL_EXACT, gpu.DIHEDRAL_AVX2):
        h, _, _ = gpu.hash_batch(torch.from_numpy(imgs).to(dev), dihedral_mode=mode)
        hm, _, _ = gpu.hash_batch(torch.from_numpy(moved).to(dev), dihedral_mode=mode)
        h = h.cpu().numpy()
        hn = hm.cpu().numpy()[:, 0].reshape(16, 7, 32)
        d = np.unpackbits(h[:, 1:] ^ hn, axis=-1).sum(-1)  # (16, 7)
        mean[mode] = d.mean(axis=0)
        say(f"[defaults] mode {mode}: mean distance per orientation {np.round(mean[mode], 2).tolist()}")
    ex, av = mean[gpu.DIHEDRAL_EXACT], mean[gpu.DIHEDRAL_AVX2]
    assert np.all(ex < 12)
    cut = [0, 1, 5, 6]  # FLOPPED, ROTATED_180, FLOPPED_ROTATED_270, ROTATED_90 (minus NORMAL)
    assert np.all(av[cut] > 48)
    keep = [2, 3, 4]
    assert np.all(av[keep] < 12)
