"""GPU parity of the precise ReferenceKernel<f32> (src/kernel/mod.rs:691-821,
src/kernel/ref.rs) against the oracle's restatement oracle_hash_ref32 (itself
pinned to the closed form of the box-filter resampler in test_oracle.py):
bit-exact hash, threshold, quality and PDQF, for f32 and u8 planes, uneven
batches and edge planes."""
import sys

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same_f32(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


def check(gpu, oracle, dev, planes_f32, as_u8=False, label=""):
    import torch
    x = torch.from_numpy(planes_f32.astype(np.uint8) if as_u8 else planes_f32).to(dev)
    say(f"[ref] {label} n={planes_f32.shape[0]} u8={as_u8}")
    h, q, t, p = (y.cpu().numpy() for y in gpu.hash_ref_batch(x, want_pdqf=True))
    for i in range(planes_f32.shape[0]):
        H, T, Q, P = oracle.hash_ref32(planes_f32[i])
        assert np.array_equal(h[i], H), f"{label} image {i}: hash differs by " \
            f"{int(np.unpackbits(h[i] ^ H).sum())} bits"
        assert same_f32(p[i], P), f"{label} image {i}: pdqf"
        assert same_f32(t[i], T), f"{label} image {i}: threshold {t[i]} vs {T}"
        assert same_f32(q[i], Q), f"{label} image {i}: quality {q[i]} vs {Q}"


def test_synthetic(gpu, oracle, dev):
    imgs = synthetic.images_u8(9, start=40).astype(np.float32)
    check(gpu, oracle, dev, imgs, label="synthetic f32")
    check(gpu, oracle, dev, imgs, as_u8=True, label="synthetic u8")
    check(gpu, oracle, dev, imgs[2:5], label="sub-batch")
    check(gpu, oracle, dev, imgs[:1], label="one image")


def test_edge_planes(gpu, oracle, dev):
    rng = np.random.default_rng(8)
    planes = np.stack([
        np.zeros((512, 512), np.float32),
        np.full((512, 512), 77.0, np.float32),
        ((np.indices((512, 512)).sum(0) % 2) * 255).astype(np.float32),
        rng.uniform(0, 255, (512, 512)).astype(np.float32),
        rng.normal(0, 1e3, (512, 512)).astype(np.float32),
        (rng.uniform(0, 1, (512, 512)) * 1e-39).astype(np.float32),
        np.tile(np.arange(512, dtype=np.float32) % 256, (512, 1)),
    ])
    check(gpu, oracle, dev, planes, label="edge planes")


def test_against_optimised_kernel(gpu, oracle, dev):
    """The two kernels differ by design (box filter vs tent); the reference's
    own tests print their distance (src/lib.rs:598-658).  Report it."""
    import torch
    imgs = synthetic.images_u8(16, start=60)
    x = torch.from_numpy(imgs).to(dev)
    hr, _, _ = gpu.hash_ref_batch(x)
    ho, _, _ = gpu.hash_batch(x)
    d = np.unpackbits(hr.cpu().numpy() ^ ho[:, 0].cpu().numpy(), axis=1).sum(1)
    say(f"[ref] distance precise vs optimised kernel: mean {d.mean():.2f} max {d.max()}")
    assert d.max() <= 64
