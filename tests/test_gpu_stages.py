"""GPU parity of the staged entry points (one launch per Kernel-trait method,
src/kernel/mod.rs:268-497) against the CPU oracle, and of the C++ host API
(yume-pdq_amd/host/yume_pdq.hpp) through its compiled test program.

Bar: bit-exact for every float and byte (NaN compares as NaN).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from test_gpu_hash import edge_planes, same_f32
from yume_pdq import synthetic

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def planes():
    """7 synthetic images plus the edge planes (non-finite, extreme, flat)."""
    imgs = synthetic.images_u8(7).astype(np.float32)
    edges = edge_planes()
    return np.concatenate([imgs, np.stack([edges[k] for k in sorted(edges)])]).astype(np.float32)


def test_compress_dct_gradient_quantize(gpu, oracle, dev, planes):
    import torch
    n = planes.shape[0]
    say(f"[stages] compress -> dct2d -> gradients -> quantize -> threshold, n={n}")
    buf1 = gpu.stage_compress(torch.from_numpy(planes).to(dev))
    tmp, pdqf = gpu.stage_dct2d(buf1)
    grad = gpu.stage_sum_of_gradients(pdqf)
    thr, h = gpu.stage_quantize(pdqf)
    h2 = gpu.stage_threshold(pdqf, thr * 0.75)
    buf1, tmp, pdqf, grad, thr, h, h2 = (x.cpu().numpy() for x in (buf1, tmp, pdqf, grad, thr, h, h2))
    for i in range(n):
        b = oracle.compress_avx2(planes[i])
        assert same_f32(buf1[i], b), f"image {i}: jarosz_compress"
        p, t = oracle.dct2d_avx2(b)
        assert same_f32(pdqf[i], p), f"image {i}: dct2d"
        assert same_f32(tmp[i], t), f"image {i}: dct2d tmp row"
        assert same_f32(grad[i], oracle.sum_of_gradients(p)), f"image {i}: sum_of_gradients"
        eh, et = oracle.quantize_avx2(p)
        assert same_f32(thr[i], et), f"image {i}: quantize threshold"
        assert np.array_equal(h[i], eh), f"image {i}: quantize bits"
        assert np.array_equal(h2[i], oracle.threshold_2d(p, np.float32(et) * np.float32(0.75))), \
            f"image {i}: threshold_2d"


def test_compress_u8_equals_f32(gpu, dev):
    import torch
    imgs = synthetic.images_u8(5)
    say("[stages] compress u8 vs f32")
    a = gpu.stage_compress(torch.from_numpy(imgs).to(dev)).cpu().numpy()
    b = gpu.stage_compress(torch.from_numpy(imgs.astype(np.float32)).to(dev)).cpu().numpy()
    assert same_f32(a, b)


def test_staged_composition_equals_fused(gpu, dev):
    """hash_get_threshold composed of staged launches == the fused kernel's NORMAL
    orientation (hash, threshold, PDQF), over an uneven batch."""
    import torch
    imgs = torch.from_numpy(synthetic.images_u8(13)).to(dev)
    say("[stages] staged composition vs fused kernel, n=13")
    _, pdqf = gpu.stage_dct2d(gpu.stage_compress(imgs))
    thr, h = gpu.stage_quantize(pdqf)
    q = gpu.stage_sum_of_gradients(pdqf)
    fh, fq, ft, fp = gpu.hash_batch(imgs, dihedral_mode=gpu.DIHEDRAL_AVX2, want_pdqf=True)
    assert torch.equal(h, fh[:, 0])
    assert same_f32(thr.cpu().numpy(), ft[:, 0].cpu().numpy())
    assert same_f32(pdqf.cpu().numpy(), fp.cpu().numpy())
    quality = np.minimum(q.cpu().numpy() / np.float32(180.0), np.float32(1.0))
    assert same_f32(quality, fq[:, 0].cpu().numpy())


TRANSFORMS = {
    0: lambda p: p.transpose(0, 2, 1),
    1: lambda p: p * np.where((np.arange(16) < 8) & (np.arange(16) % 2 == 0), -1, 1).astype(np.float32),
    2: lambda p: p * np.where((np.arange(16) < 8) & (np.arange(16) % 2 == 1), -1, 1).astype(np.float32),
    3: lambda p: p * np.where(np.arange(16) % 2 == 0, -1, 1).astype(np.float32),
    4: lambda p: p * np.where(np.arange(16) % 2 == 1, -1, 1).astype(np.float32),
    5: lambda p: p * np.where(np.arange(16) % 2 == 1, -1, 1).astype(np.float32)[:, None],
    6: lambda p: p * np.where(np.arange(16) % 2 == 0, -1, 1).astype(np.float32)[:, None],
    7: lambda p: p * np.where((np.arange(16)[:, None] + np.arange(16)) % 2 == 1, -1, 1).astype(np.float32),
}


@pytest.mark.parametrize("op", sorted(TRANSFORMS))
def test_transforms(gpu, dev, op):
    """PDQF transforms (mod.rs:345-405; AVX2 column extent x86.rs:330-376;
    AVX2 row parity x86.rs:285-326) against numpy index algebra."""
    import torch
    rng = np.random.default_rng(op)
    p = rng.standard_normal((9, 16, 16)).astype(np.float32)
    t = torch.from_numpy(p.copy()).to(dev)
    say(f"[stages] transform op={op}")
    gpu.stage_transform(t, op)
    assert same_f32(t.cpu().numpy(), TRANSFORMS[op](p))


def test_visit_sequence_equals_dihedrals_batch(gpu, oracle, dev):
    """The reference's visit_dihedrals sequence (src/lib.rs:391-418) run as
    staged launches gives the fused dihedral kernel's hashes and thresholds."""
    import torch
    imgs = torch.from_numpy(synthetic.images_u8(6)).to(dev)
    say("[stages] visit_dihedrals sequence vs dihedrals_batch")
    _, pdqf = gpu.stage_dct2d(gpu.stage_compress(imgs))
    thr, _ = gpu.stage_quantize(pdqf)
    fh, fq, ft, _ = gpu.dihedrals_batch(pdqf, thr, dihedral_mode=gpu.DIHEDRAL_AVX2)
    p = pdqf.clone()
    out_h, out_t = [], []
    p = gpu.stage_transform(p, gpu.T_NEG_COLS_EVEN_AVX2)
    t_cols, h = gpu.stage_quantize(p); out_h.append(h); out_t.append(t_cols)
    p = gpu.stage_transform(p, gpu.T_NEG_ROWS_EVEN)
    t_diag, h = gpu.stage_quantize(p); out_h.append(h); out_t.append(t_diag)
    p = gpu.stage_transform(p, gpu.T_NEG_COLS_EVEN_AVX2)
    t_rows, h = gpu.stage_quantize(p); out_h.append(h); out_t.append(t_rows)
    p = gpu.stage_transform(p, gpu.T_NEG_ROWS_EVEN)
    p = gpu.stage_transform(p, gpu.T_TRANSPOSE)
    out_h.append(gpu.stage_threshold(p, thr)); out_t.append(thr)
    p = gpu.stage_transform(p, gpu.T_NEG_ROWS_EVEN)
    out_h.append(gpu.stage_threshold(p, t_cols)); out_t.append(t_cols)
    p = gpu.stage_transform(p, gpu.T_NEG_COLS_EVEN_AVX2)
    out_h.append(gpu.stage_threshold(p, t_diag)); out_t.append(t_diag)
    p = gpu.stage_transform(p, gpu.T_NEG_ROWS_EVEN)
    out_h.append(gpu.stage_threshold(p, t_rows)); out_t.append(t_rows)
    for d in range(7):
        assert torch.equal(out_h[d], fh[:, d + 1]), f"dihedral {d + 1}"
        assert same_f32(out_t[d].cpu().numpy(), ft[:, d + 1].cpu().numpy()), f"dihedral {d + 1} threshold"


def test_luma(gpu, dev):
    """cvt_rgb8_to_luma8f: trait default R*r + G*g + B*b (mod.rs:303-322);
    cvt_rgba8_to_luma8f: AVX2 override r*R, fma(g,G,.), fma(b,B,.)
    (x86.rs:686-753).  Expected values: float32 products and sums in numpy;
    the fused multiply-adds as the float64 sum (exact: at most 35 significant
    bits here) rounded once to float32."""
    import torch
    rng = np.random.default_rng(5)
    c = np.array(gpu.RGB8_TO_LUMA8_TABLE_ITU, np.float32)
    rgb = rng.integers(0, 256, (3, 512, 3), dtype=np.uint8)
    rgba = rng.integers(0, 256, (3, 512, 4), dtype=np.uint8)
    rgb[0, 0] = (255, 255, 255)
    rgba[0, 0] = (255, 255, 255, 0)
    say("[stages] luma rgb / rgba")
    got = gpu.stage_luma(torch.from_numpy(rgb).to(dev)).cpu().numpy()
    f = rgb.astype(np.float32)
    exp = (c[0] * f[..., 0] + c[1] * f[..., 1]) + c[2] * f[..., 2]
    assert same_f32(got, exp)
    got = gpu.stage_luma(torch.from_numpy(rgba).to(dev)).cpu().numpy()
    f = rgba.astype(np.float32)
    v = f[..., 0] * c[0]
    v = (f[..., 1].astype(np.float64) * np.float64(c[1]) + v.astype(np.float64)).astype(np.float32)
    v = (f[..., 2].astype(np.float64) * np.float64(c[2]) + v.astype(np.float64)).astype(np.float32)
    assert same_f32(got, v)


def test_host_api_program(gpu):
    """The C++ host API (Kernel / PDQMatcher mirrors) through its test program."""
    exe = os.path.join(REPO, "tests", "cpp", "_build", "test_host_api")
    assert os.path.exists(exe), "tests/cpp/_build/test_host_api is not built (run the build step)"
    say("[stages] C++ host API test program")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr, file=sys.stderr)
    assert r.returncode == 0, r.stdout[-2000:]
