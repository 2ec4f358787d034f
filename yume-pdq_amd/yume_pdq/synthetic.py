"""Seeded synthetic workloads for the hash and match paths.

Images: 512x512 grayscale, a smooth 2-D gradient plus 4..16 Gaussian blobs
(amplitude +-U[20,120], sigma U[10,80] px) plus U[-8,8] noise, clamped to
[0, 255] and rounded to integers, so the f32 and u8 forms of an image hold the
same values.  Image k uses numpy PCG64 seed 0x59554D45 + k.

Databases: uniform random 256-bit rows (u64 x 4), with planted near-duplicates
of given query hashes at distances drawn from U{0..40} (distinct bit flips, so
the planted distance is exact before any other coincidence).
"""
import numpy as np

IMAGE_SEED_BASE = 0x59554D45


def image_u8(k):
    rng = np.random.Generator(np.random.PCG64(IMAGE_SEED_BASE + k))
    c0 = rng.uniform(40.0, 215.0)
    cx, cy = rng.uniform(-0.2, 0.2, size=2)
    y, x = np.mgrid[0:512, 0:512].astype(np.float64)
    img = c0 + cx * x + cy * y
    nb = int(rng.integers(4, 17))
    for _ in range(nb):
        bx, by = rng.uniform(0.0, 512.0, size=2)
        amp = rng.uniform(20.0, 120.0) * (1.0 if rng.random() < 0.5 else -1.0)
        sig = rng.uniform(10.0, 80.0)
        img += amp * np.exp(-((x - bx) ** 2 + (y - by) ** 2) / (2.0 * sig * sig))
    img += rng.uniform(-8.0, 8.0, size=(512, 512))
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def images_u8(n, start=0):
    return np.stack([image_u8(start + k) for k in range(n)])


def random_db(n_rows, seed=7):
    """n_rows uniform random 32-byte rows (u8, shape (n_rows, 32))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.frombuffer(rng.bytes(n_rows * 32), dtype=np.uint8).reshape(n_rows, 32).copy()


def flip_bits(row32, k, rng):
    """Copy of a 32-byte row with k distinct random bits flipped."""
    out = np.array(row32, dtype=np.uint8, copy=True)
    pos = rng.choice(256, size=k, replace=False)
    for p in pos:
        out[p >> 3] ^= np.uint8(1 << (p & 7))
    return out


def plant(db_u8, queries_u8, n_plants, seed=11, max_dist=40):
    """Overwrite n_plants distinct random rows of db_u8 (n,32) with copies of
    random queries at distance U{0..max_dist}.  Returns (rows, query_ids, dists)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = db_u8.shape[0]
    rows = rng.choice(n, size=n_plants, replace=False)
    qids = rng.integers(0, queries_u8.shape[0], size=n_plants)
    dists = rng.integers(0, max_dist + 1, size=n_plants)
    for r, q, d in zip(rows, qids, dists):
        db_u8[r] = flip_bits(queries_u8[q], int(d), rng)
    return rows, qids, dists


def brute_force_hits(db_u8, queries_u8, threshold, chunk=1 << 20):
    """Exact (query, index, distance) triples with numpy, sorted by (index, query)."""
    db64 = np.ascontiguousarray(db_u8).view(np.uint64).reshape(-1, 4)
    q64 = np.ascontiguousarray(queries_u8).view(np.uint64).reshape(-1, 4)
    out = []
    for s in range(0, db64.shape[0], chunk):
        blk = db64[s:s + chunk]
        for qi in range(q64.shape[0]):
            d = np.bitwise_count(blk ^ q64[qi]).sum(axis=1, dtype=np.int64)
            idx = np.nonzero(d <= threshold)[0]
            if idx.size:
                out.append(np.stack([np.full(idx.size, qi), s + idx, d[idx]], axis=1))
    if not out:
        return np.zeros((0, 3), np.int64)
    t = np.concatenate(out).astype(np.int64)
    order = np.lexsort((t[:, 0], t[:, 1]))
    return t[order]


def image_params(k):
    """The random parameters image_u8(k) draws, in the same order."""
    rng = np.random.Generator(np.random.PCG64(IMAGE_SEED_BASE + k))
    c0 = rng.uniform(40.0, 215.0)
    cx, cy = rng.uniform(-0.2, 0.2, size=2)
    blobs = []
    for _ in range(int(rng.integers(4, 17))):
        bx, by = rng.uniform(0.0, 512.0, size=2)
        amp = rng.uniform(20.0, 120.0) * (1.0 if rng.random() < 0.5 else -1.0)
        sig = rng.uniform(10.0, 80.0)
        blobs.append((bx, by, amp, sig))
    return c0, cx, cy, blobs


def images_u8_device(n, device, start=0, seed=99):
    """Bulk images of the same family rendered on the GPU (torch), for timing
    workloads: same parameter draws as image_u8, noise from torch's generator,
    so pixels are not bit-identical to image_u8."""
    import torch
    y, x = torch.meshgrid(torch.arange(512, device=device, dtype=torch.float32),
                          torch.arange(512, device=device, dtype=torch.float32), indexing="ij")
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    out = torch.empty((n, 512, 512), dtype=torch.uint8, device=device)
    for i in range(n):
        c0, cx, cy, blobs = image_params(start + i)
        img = c0 + cx * x + cy * y
        for bx, by, amp, sig in blobs:
            img = img + amp * torch.exp(-((x - bx) ** 2 + (y - by) ** 2) / (2.0 * sig * sig))
        img = img + (torch.rand((512, 512), generator=gen, device=device) * 16.0 - 8.0)
        out[i] = torch.clamp(torch.round(img), 0, 255).to(torch.uint8)
    return out


def plant_fast(db_u8, queries_u8, n_plants, seed=11, max_dist=40):
    """Vectorised plant(): distinct rows, random queries, U{0..max_dist} distinct flips."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = db_u8.shape[0]
    rows = rng.choice(n, size=n_plants, replace=False)
    qids = rng.integers(0, queries_u8.shape[0], size=n_plants)
    dists = rng.integers(0, max_dist + 1, size=n_plants)
    order = np.argsort(rng.random((n_plants, 256)), axis=1)
    mask = np.arange(256)[None, :] < dists[:, None]
    bits = np.zeros((n_plants, 256), np.uint8)
    np.put_along_axis(bits, order, mask.astype(np.uint8), axis=1)
    flips = np.packbits(bits, axis=1, bitorder="little")
    db_u8[rows] = queries_u8[qids] ^ flips
    return rows, qids, dists
