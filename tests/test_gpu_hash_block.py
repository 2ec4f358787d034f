"""GPU tests of the block-per-image hash kernel (four or eight wavefronts per
image) and of its matrix-core DCT precisions.

Exact arithmetic: bit-identical to the wave-per-image kernel (itself checked
against the CPU oracle in test_gpu_hash.py) for every input kind, both
dihedral modes, uneven batches and planes holding NaN/Inf; plus a direct check
against the oracle.

Matrix-core precisions (bf16, f16 and their hi+lo split forms): each PDQF
coefficient against the FP64 DCT of the same compressed plane, within the
bound of the precision's operand rounding and f32 accumulation,

    |P - D64 C D64^T| <= e_p * (|D64| C |D64|^T)     (elementwise, C >= 0)

with e_p = 2 * 3 * u_p + 2 * 128 * 2^-24 (two stages, three rounded factors
per product, 128-term f32 sums), u_p = 2^-8 (bf16), 2^-11 (f16), 2^-15 (split
forms: u^2 of the hi+lo pair and the f32 and subnormal-f16 residues);
and the hash bits against the exact kernel (= the reference's) within the
reference's own tolerance of 16 bits per hash (src/lib.rs:570-574).
"""
import sys

import numpy as np
import pytest

from yume_pdq import synthetic

pytestmark = pytest.mark.gpu

ALL = slice(None)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def same_f32(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


def run(gpu, planes, layout, dct=0, mode=0, luma=None):
    h, q, t, p, b = gpu.hash_batch(planes, dihedral_mode=mode, want_pdqf=True, want_buf1=True,
                                   layout=layout, dct=dct, luma=luma)
    return [x.cpu().numpy() for x in (h, q, t, p, b)]


def assert_same(a, b, label):
    names = ["hashes", "quality", "threshold", "pdqf", "buf1"]
    assert np.array_equal(a[0], b[0]), f"{label}: hashes differ"
    for k in range(1, 5):
        assert same_f32(a[k], b[k]), f"{label}: {names[k]} differ"


@pytest.fixture(scope="module")
def imgs():
    return synthetic.images_u8(21)


def colour(imgs, channels):
    rng = np.random.default_rng(channels + 10)
    base = imgs.astype(np.int16)
    px = np.empty(imgs.shape + (channels,), np.uint8)
    for c in range(channels):
        px[..., c] = np.clip(base + rng.integers(-40, 41, size=(imgs.shape[0], 1, 1)), 0, 255).astype(np.uint8)
    return px


@pytest.mark.parametrize("kind", ["f32", "u8", "rgb", "rgba"])
@pytest.mark.parametrize("waves", [4, 8, 16])
def test_block_exact_equals_wave(gpu, dev, imgs, kind, waves):
    import torch
    if kind == "f32":
        x = torch.from_numpy(imgs.astype(np.float32)).to(dev)
    elif kind == "u8":
        x = torch.from_numpy(imgs).to(dev)
    else:
        x = torch.from_numpy(colour(imgs, 3 if kind == "rgb" else 4)).to(dev)
    layout = {4: gpu.LAYOUT_BLOCK4, 8: gpu.LAYOUT_BLOCK8, 16: gpu.LAYOUT_BLOCK16}[waves]
    for mode in (0, 1):
        say(f"[block] exact {kind} waves={waves} mode={mode} n={x.shape[0]}")
        a = run(gpu, x, gpu.LAYOUT_WAVE, mode=mode)
        b = run(gpu, x, layout, mode=mode)
        assert_same(a, b, f"{kind} waves={waves} mode={mode}")
        # an uneven sub-batch at an offset
        a = run(gpu, x[3:10], gpu.LAYOUT_WAVE, mode=mode)
        b = run(gpu, x[3:10], layout, mode=mode)
        assert_same(a, b, f"{kind} waves={waves} mode={mode} sub-batch")


def test_block_against_oracle(gpu, oracle, dev, imgs):
    import torch
    f = imgs[:4].astype(np.float32)
    h, q, t, p, b = run(gpu, torch.from_numpy(f).to(dev), gpu.LAYOUT_BLOCK4)
    for i in range(4):
        H, Q, T, P = oracle.hash_dihedrals(f[i], 0)
        assert np.array_equal(h[i], H)
        assert same_f32(q[i], Q) and same_f32(t[i], T) and same_f32(p[i], P)
        assert same_f32(b[i], oracle.compress_avx2(f[i]))


def edge_planes():
    rng = np.random.default_rng(5)
    planes = {
        "zeros": np.zeros((512, 512), np.float32),
        "checker": ((np.indices((512, 512)).sum(0) % 2) * 255).astype(np.float32),
        "huge": rng.uniform(-3e37, 3e37, (512, 512)).astype(np.float32),
        "tiny": (rng.uniform(0, 1, (512, 512)) * 1e-39).astype(np.float32),
    }
    for name, (r, c, v) in {"nan_inside": (200, 300, np.nan), "inf_col0": (100, 0, np.inf),
                            "ninf_row511": (511, 100, -np.inf), "inf_inside": (256, 256, np.inf),
                            "nan_row1": (1, 77, np.nan), "inf_col511": (300, 511, np.inf)}.items():
        p = rng.uniform(0, 255, (512, 512)).astype(np.float32)
        p[r, c] = v
        planes[name] = p
    planes["all_nan"] = np.full((512, 512), np.nan, np.float32)
    return planes


NONFINITE = ["nan_inside", "inf_col0", "ninf_row511", "inf_inside", "nan_row1", "inf_col511", "all_nan"]


def test_block_edge_planes(gpu, dev):
    """Finite extremes and NaN/Inf anywhere the reference reads: every
    layout gives the wave kernel's bits; the matrix-core precisions take the
    exact arithmetic for non-finite planes, so they give the same bits too."""
    import torch
    planes = edge_planes()
    names = list(planes)
    x = torch.from_numpy(np.stack([planes[k] for k in names])).to(dev)
    for mode in (0, 1):
        say(f"[block] edge planes mode={mode}")
        a = run(gpu, x, gpu.LAYOUT_WAVE, mode=mode)
        for layout in (gpu.LAYOUT_BLOCK4, gpu.LAYOUT_BLOCK8, gpu.LAYOUT_BLOCK16):
            assert_same(a, run(gpu, x, layout, mode=mode), f"edge planes layout={layout} mode={mode}")
        nf = [names.index(k) for k in NONFINITE]
        xn = x[nf]
        an = run(gpu, xn, gpu.LAYOUT_WAVE, mode=mode)
        for dct in (gpu.DCT_BF16, gpu.DCT_F16X3):
            assert_same(an, run(gpu, xn, gpu.LAYOUT_BLOCK4, dct=dct, mode=mode), f"non-finite dct={dct}")


def test_block_single_and_auto(gpu, dev, imgs):
    """One image, and the automatic layout, equal the wave kernel."""
    import torch
    x = torch.from_numpy(imgs[:1]).to(dev)
    a = run(gpu, x, gpu.LAYOUT_WAVE)
    assert_same(a, run(gpu, x, gpu.LAYOUT_BLOCK8), "single image")
    assert_same(a, run(gpu, x, gpu.LAYOUT_AUTO), "auto")


@pytest.mark.parametrize("mode", [0, 1])
def test_wave_layout_sixteen_wave_blocks(gpu, dev, imgs, mode):
    """A u8 batch of 16 images per CU or more runs the wave-per-image kernel
    in blocks of sixteen waves (one block per CU, four waves per SIMD): 16 x
    CUs + 7 images, the last block partly empty, equal the eight-wave block
    kernel image for image, buf1 included."""
    import torch
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    n = 16 * cus + 7
    g = torch.Generator(device=dev)
    g.manual_seed(1234 + mode)
    x = torch.randint(0, 256, (n, 512, 512), dtype=torch.uint8, device=dev, generator=g)
    x[: imgs.shape[0]] = torch.from_numpy(imgs).to(dev)
    x[-imgs.shape[0]:] = torch.from_numpy(imgs[::-1].copy()).to(dev)
    say(f"[hash] sixteen-wave blocks: n={n} cus={cus} mode={mode}")
    a = run(gpu, x, gpu.LAYOUT_WAVE, mode=mode)
    b = run(gpu, x, gpu.LAYOUT_BLOCK8, mode=mode)
    assert_same(a, b, f"sixteen-wave blocks n={n} mode={mode}")


def dct64_matrix():
    i = np.arange(1, 17, dtype=np.float64)[:, None]
    j = np.arange(127, dtype=np.float64)[None, :]
    return np.sqrt(2.0 / 127.0) * np.cos(np.pi / 254.0 * i * (2.0 * j + 1.0))


BOUND_U = {1: 2.0 ** -8, 2: 2.0 ** -11, 3: 2.0 ** -15, 4: 2.0 ** -15}
NAMES = {1: "bf16", 2: "f16", 3: "bf16x3", 4: "f16x3"}


@pytest.mark.parametrize("dct", [1, 2, 3, 4])
def test_matrix_core_dct_error(gpu, dev, dct):
    """Per-coefficient error of the matrix-core DCT against the FP64 DCT of
    the exact compressed plane, and hash-bit disagreement against the exact
    kernel, over 64 synthetic images (config-0 family) in both input types."""
    import torch
    imgs = synthetic.images_u8(64, start=100)
    D = dct64_matrix()
    absD = np.abs(D)
    e_p = 2 * 3 * BOUND_U[dct] + 2 * 128 * 2.0 ** -24
    for kind in ("f32", "u8"):
        x = torch.from_numpy(imgs.astype(np.float32) if kind == "f32" else imgs).to(dev)
        say(f"[block] dct={NAMES[dct]} {kind}")
        ex = run(gpu, x, gpu.LAYOUT_BLOCK4)
        ap = run(gpu, x, gpu.LAYOUT_BLOCK4, dct=dct)
        assert same_f32(ex[4], ap[4]), "the compression is exact in every precision"
        C = ex[4][:, :127, :127].astype(np.float64)
        P64 = np.einsum("ka,nab,jb->nkj", D, C, D)
        M = np.einsum("ka,nab,jb->nkj", absD, np.abs(C), absD)
        err = np.abs(ap[3].astype(np.float64) - P64)
        rel = float((err / M).max())
        ex_rel = float((np.abs(ex[3].astype(np.float64) - P64) / M).max())
        bits = np.unpackbits(ap[0] ^ ex[0], axis=-1).sum(-1)
        say(f"[block] dct={NAMES[dct]} {kind}: max |err|/M = {rel:.3e} (bound {e_p:.3e}; exact path "
            f"{ex_rel:.3e}), max |err| = {err.max():.3e}, hash bits vs exact: mean {bits.mean():.3f} "
            f"max {bits.max()}")
        assert rel <= e_p
        assert bits.mean() <= 16.0
        if dct in (3, 4):
            assert bits.mean() <= 1.0


def test_matrix_core_layout_8(gpu, dev, imgs):
    """The eight-wave layout computes the matrix-core DCT with one column
    tile per wave: same values as the four-wave layout."""
    import torch
    x = torch.from_numpy(imgs).to(dev)
    for dct in (1, 3):
        assert_same(run(gpu, x, gpu.LAYOUT_BLOCK4, dct=dct), run(gpu, x, gpu.LAYOUT_BLOCK8, dct=dct),
                    f"dct={dct} 4 vs 8 waves")


def test_options_validation(gpu, dev, imgs):
    import torch
    x = torch.from_numpy(imgs[:2]).to(dev)
    with pytest.raises(gpu.YumePdqError):
        gpu.hash_batch(x, layout=gpu.LAYOUT_WAVE, dct=gpu.DCT_BF16)
    with pytest.raises(gpu.YumePdqError):
        gpu.hash_batch(x, dct=9)
