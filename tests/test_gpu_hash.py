"""GPU parity of the fused hash kernel against the CPU oracle (AVX2 semantics).

Bar: bit-exact hashes, thresholds, qualities and PDQFs (integer/bit work and an
exact replay of the reference's FP32 rounding sequence).  NaN outputs compare
as NaN.
"""
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


def oracle_batch(oracle, planes_f32, mode):
    n = planes_f32.shape[0]
    H = np.zeros((n, 8, 32), np.uint8)
    Q = np.zeros((n, 8), np.float32)
    T = np.zeros((n, 8), np.float32)
    P = np.zeros((n, 16, 16), np.float32)
    for i in range(n):
        H[i], Q[i], T[i], P[i] = oracle.hash_dihedrals(planes_f32[i], mode)
    return H, Q, T, P


def check_batch(gpu, oracle, planes_f32, torch_planes, mode, label):
    """Both kernel layouts (a wave per image; eight waves per image, the
    automatic choice for small batches) against the oracle."""
    H, Q, T, P = oracle_batch(oracle, planes_f32, mode)
    for layout in (gpu.LAYOUT_WAVE, gpu.LAYOUT_BLOCK8):
        say(f"[hash] {label}: n={planes_f32.shape[0]} dtype={torch_planes.dtype} mode={mode} layout={layout}")
        h, q, t, p = gpu.hash_batch(torch_planes, dihedral_mode=mode, want_pdqf=True, layout=layout)
        h, q, t, p = (x.cpu().numpy() for x in (h, q, t, p))
        check_against(planes_f32, h, q, t, p, H, Q, T, P, f"{label} layout={layout}")


def check_against(planes_f32, h, q, t, p, H, Q, T, P, label):
    for i in range(planes_f32.shape[0]):
        assert np.array_equal(h[i], H[i]), f"{label} image {i}: hash mismatch " \
            f"{[int(np.unpackbits(h[i][d] ^ H[i][d]).sum()) for d in range(8)]}"
        assert same_f32(p[i], P[i]), f"{label} image {i}: pdqf mismatch"
        assert same_f32(t[i], T[i]), f"{label} image {i}: thresholds {t[i]} vs {T[i]}"
        assert same_f32(q[i], Q[i]), f"{label} image {i}: qualities {q[i]} vs {Q[i]}"


@pytest.fixture(scope="module")
def imgs():
    return synthetic.images_u8(12)


@pytest.mark.parametrize("mode", [0, 1])
def test_synthetic_f32(gpu, oracle, dev, imgs, mode):
    import torch
    f = imgs.astype(np.float32)
    check_batch(gpu, oracle, f, torch.from_numpy(f).to(dev), mode, "synthetic f32")


def test_synthetic_u8_equals_f32(gpu, oracle, dev, imgs):
    import torch
    f = imgs.astype(np.float32)
    check_batch(gpu, oracle, f, torch.from_numpy(imgs).to(dev), 0, "synthetic u8")


@pytest.mark.parametrize("n", [1, 3, 5, 9])
def test_uneven_batches(gpu, oracle, dev, imgs, n):
    import torch
    f = imgs[:n].astype(np.float32)
    check_batch(gpu, oracle, f, torch.from_numpy(f).to(dev), 0, f"uneven n={n}")


def test_golden_readme_fm(gpu, oracle, dev):
    """README.md:305-452: <= 2 bits from the 32 expected bytes (the oracle hits 0)."""
    import json
    import os
    import torch
    fm = oracle.readme_fm_image()
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "readme_fm.json")))
    check_batch(gpu, oracle, fm[None], torch.from_numpy(fm[None]).to(dev), 1, "README FM")
    h, q, t = gpu.hash_batch(torch.from_numpy(fm[None]).to(dev))
    bits = int(np.unpackbits(h[0, 0].cpu().numpy() ^ np.array(g["expected_bytes"], np.uint8)).sum())
    assert bits <= g["max_bits_different"]


def test_golden_aaa_orig(gpu, oracle, dev):
    """test-data/aaa-orig.jpg.expect:1-3 whitelist via the ctypes harness preprocessing."""
    import os
    import torch
    here = os.path.dirname(__file__)
    a = np.load(os.path.join(here, "golden", "aaa_orig_512_u8.npy"))
    wl = open(os.path.join(here, "golden", "aaa_orig_expect.txt")).read().split()
    check_batch(gpu, oracle, a[None].astype(np.float32), torch.from_numpy(a[None]).to(dev), 0, "aaa-orig")
    h, _, _ = gpu.hash_batch(torch.from_numpy(a[None]).to(dev))
    assert h[0, 0].cpu().numpy().tobytes().hex() in wl


def edge_planes():
    rng = np.random.default_rng(5)
    planes = {
        "zeros": np.zeros((512, 512), np.float32),
        "constant": np.full((512, 512), 77.0, np.float32),
        "max255": np.full((512, 512), 255.0, np.float32),
        "noise": rng.uniform(0, 255, (512, 512)).astype(np.float32),
        "checker": ((np.indices((512, 512)).sum(0) % 2) * 255).astype(np.float32),
        "negative": rng.normal(0, 1e3, (512, 512)).astype(np.float32),
        "huge": rng.uniform(-3e37, 3e37, (512, 512)).astype(np.float32),
        "tiny": (rng.uniform(0, 1, (512, 512)) * 1e-39).astype(np.float32),
    }
    nan = rng.uniform(0, 255, (512, 512)).astype(np.float32)
    nan[200, 300] = np.nan
    planes["nan_inside"] = nan
    inf = rng.uniform(0, 255, (512, 512)).astype(np.float32)
    inf[100, 0] = np.inf  # column 0: read only through the zero-weight over-read
    planes["inf_col0"] = inf
    inf2 = rng.uniform(0, 255, (512, 512)).astype(np.float32)
    inf2[511, 100] = -np.inf  # row 511: read only by the zero-weight row di=6
    planes["ninf_row511"] = inf2
    inf3 = rng.uniform(0, 255, (512, 512)).astype(np.float32)
    inf3[256, 256] = np.inf
    planes["inf_inside"] = inf3
    allnan = np.full((512, 512), np.nan, np.float32)
    planes["all_nan"] = allnan
    return planes


@pytest.mark.parametrize("name", list(edge_planes().keys()))
def test_edge_planes(gpu, oracle, dev, name):
    import torch
    p = edge_planes()[name][None]
    for mode in (0, 1):
        check_batch(gpu, oracle, p, torch.from_numpy(p).to(dev), mode, f"edge {name}")


def test_dropin_hash_smart_kernel(gpu, oracle, imgs):
    """yume_pdq_hash_smart_kernel: every out-buffer as the reference fills it."""
    for i in range(3):
        f = imgs[i].astype(np.float32)
        say(f"[dropin] hash image {i}")
        out, thr, q, pdqf, buf1, tmp = gpu.hash_smart_kernel(f)
        o_out, o_thr, o_q, o_pdqf, o_buf1, o_tmp = oracle.hash_get_threshold(f)
        assert np.array_equal(out, o_out)
        assert same_f32(thr, o_thr) and same_f32(q, o_q)
        assert same_f32(pdqf, o_pdqf)
        assert same_f32(buf1, o_buf1)
        assert same_f32(tmp, o_tmp)


def test_dropin_visit_dihedrals(gpu, oracle, imgs):
    """yume_pdq_visit_dihedrals_smart_kernel: 7 callbacks in order, AVX2 semantics,
    input threshold passed through (src/lib.rs:223-224), early stop."""
    f = imgs[4].astype(np.float32)
    _, thr, _, pdqf, _, _ = oracle.hash_get_threshold(f)
    codes, th, q, h, pd = oracle.visit_dihedrals(pdqf, thr, 1)
    seen = []

    def cb(code, t, qual, hsh, pdq):
        seen.append((code, t, qual, hsh, pdq))
        return True

    say("[dropin] visit dihedrals")
    assert gpu.visit_dihedrals_smart_kernel(thr, pdqf, cb) is True
    assert [s[0] for s in seen] == [int(c) for c in codes]
    for k, (code, t, qual, hsh, pdq) in enumerate(seen):
        assert same_f32(t, thr)
        assert same_f32(qual, q[k])
        assert np.array_equal(hsh, h[k])
        assert same_f32(pdq, pd[k])
    stop = []
    assert gpu.visit_dihedrals_smart_kernel(thr, pdqf, lambda *a: stop.append(1) or len(stop) < 3) is False
    assert len(stop) == 3


def test_dihedrals_batch(gpu, oracle, dev, imgs):
    import torch
    f = imgs[:4].astype(np.float32)
    for mode in (0, 1):
        say(f"[dihedrals_batch] mode={mode}")
        P, T = [], []
        for i in range(4):
            _, thr, _, pdqf, _, _ = oracle.hash_get_threshold(f[i])
            P.append(pdqf)
            T.append(thr)
        h, q, t, pds = gpu.dihedrals_batch(torch.tensor(np.stack(P)).to(dev),
                                           torch.tensor(np.array(T, np.float32)).to(dev), mode)
        h, q, t, pds = (x.cpu().numpy() for x in (h, q, t, pds))
        for i in range(4):
            codes, th, qq, hh, pd = oracle.visit_dihedrals(P[i], T[i], mode)
            assert np.array_equal(h[i, 1:], hh)
            assert same_f32(q[i, 1:], qq)
            assert same_f32(t[i, 1:], th)
            assert same_f32(pds[i, 1:], pd)
            assert same_f32(pds[i, 0], P[i])


@pytest.mark.parametrize("channels", [3, 4])
@pytest.mark.parametrize("coeff", [None, (0.2126, 0.7152, 0.0722)])
def test_colour_inputs_fused(gpu, dev, channels, coeff):
    """RGB8 / RGBA8 planes hashed with the luma conversion fused into the row
    loader equal converting first (cvt_rgb8_to_luma8f / cvt_rgba8_to_luma8f,
    checked against numpy in test_gpu_stages.test_luma) and hashing the f32
    luma plane -- bit for bit, over an uneven batch."""
    import torch
    rng = np.random.default_rng(channels)
    base = synthetic.images_u8(5).astype(np.int16)
    px = np.empty((5, 512, 512, channels), np.uint8)
    for c in range(channels):
        px[..., c] = np.clip(base + rng.integers(-40, 41, size=(5, 1, 1)) +
                             rng.integers(-6, 7, size=base.shape), 0, 255).astype(np.uint8)
    d_px = torch.from_numpy(px).to(dev)
    say(f"[hash] colour channels={channels} coeff={coeff}")
    lum = gpu.stage_luma(d_px, coeff=coeff or gpu.RGB8_TO_LUMA8_TABLE_ITU)
    for mode in (0, 1):
        fh, fq, ft = gpu.hash_batch(d_px, dihedral_mode=mode, luma=coeff)
        eh, eq, et = gpu.hash_batch(lum.contiguous(), dihedral_mode=mode)
        assert torch.equal(fh, eh)
        assert same_f32(fq.cpu().numpy(), eq.cpu().numpy())
        assert same_f32(ft.cpu().numpy(), et.cpu().numpy())
