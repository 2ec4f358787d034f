"""CPU: pin the oracle against the reference's own golden vectors, tolerances
and invariants, and against closed forms -- before it is trusted as the
checker of the GPU kernels.  No GPU is used here.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
F32_EPS = float(np.finfo(np.float32).eps)


def bits(a, b):
    return int(np.unpackbits(np.asarray(a, np.uint8) ^ np.asarray(b, np.uint8)).sum())


# --- end-to-end known-answer tests held by the reference ---------------------------------

def test_readme_fm_golden(oracle):
    """README.md:305-452: FM-stripe image -> 32 expected bytes (<= 2 bits) and the
    printed PDQF (README.md:397-412, 0.1 precision)."""
    g = json.load(open(os.path.join(GOLD, "readme_fm.json")))
    img = oracle.readme_fm_image()
    h, thr, q, pdqf, _, _ = oracle.hash_get_threshold(img)
    assert bits(h, g["expected_bytes"]) <= g["max_bits_different"]
    # printed with {:5.1}: every value within half a unit of the last digit
    assert np.abs(pdqf - np.array(g["pdqf_printed_1dp"])).max() <= 0.05 + 1e-4
    # quality == quality_easy (README.md:375) -- hash() and hash_get_threshold() agree
    assert q == oracle.hash_get_threshold(img)[2]


def test_aaa_orig_whitelist(oracle):
    """test-data/aaa-orig.jpg with the ctypes harness preprocessing must hit the
    CI whitelist (build.yml:96-128, aaa-orig.jpg.expect:1-3)."""
    a = np.load(os.path.join(GOLD, "aaa_orig_512_u8.npy")).astype(np.float32)
    wl = open(os.path.join(GOLD, "aaa_orig_expect.txt")).read().split()
    h = oracle.hash_get_threshold(a)[0]
    assert h.tobytes().hex() in wl


# --- the two restatements of the AVX2 kernel agree bit for bit ---------------------------

def _images(n):
    from yume_pdq import synthetic
    return synthetic.images_u8(n).astype(np.float32)


def test_lane_restatement_equals_intrinsic_restatement(oracle):
    """pdq_oracle.c (explicit lanes) vs pdq_oracle_avx2.c (the intrinsic
    sequence of x86.rs:100-682): a transposed lane, a wrong accumulator
    alternation or a wrong reduction tree in either would show here."""
    rng = np.random.default_rng(3)
    planes = list(_images(6)) + [rng.uniform(0, 255, (512, 512)).astype(np.float32),
                                 rng.normal(0, 500, (512, 512)).astype(np.float32)]
    for p in planes:
        a = oracle.hash_get_threshold(p)
        b = oracle.avx2_hash(p)
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(np.float32(a[1]).view(np.uint32), np.float32(b[1]).view(np.uint32))
        assert np.array_equal(np.float32(a[2]).view(np.uint32), np.float32(b[2]).view(np.uint32))
        for x, y in zip(a[3:], b[3:]):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


# --- component tolerances of the reference's tests ---------------------------------------

def _ref32_dct(buf127, d32):
    """ReferenceKernel<f32>::dct2d, mod.rs:711-735 (sequential f32 sums)."""
    D = d32[:16 * 127].reshape(16, 127).astype(np.float32)
    out = np.zeros((16, 16), np.float32)
    for i in range(16):
        tmp = np.zeros(127, np.float32)
        for j in range(127):
            s = np.float32(0)
            for k in range(127):
                s = np.float32(s + np.float32(D[i, k] * buf127[k, j]))
            tmp[j] = s
        for j in range(16):
            s = np.float32(0)
            for k in range(127):
                s = np.float32(s + np.float32(tmp[k] * D[j, k]))
            out[i, j] = s
    return out


def test_dct_vs_reference_kernel_tolerance(oracle):
    """mod.rs:1187-1246: smart kernel DCT vs ReferenceKernel<f32> on the
    structured +-random input, |delta| <= 32 * f32::EPSILON."""
    d32, _, _ = oracle.tables()
    rng = np.random.default_rng(17)
    buf = np.zeros((128, 128), np.float32)
    for i in range(127):
        for j in range(127):
            buf[i, j] = rng.uniform(0, 1) if (i + 2 * j) % 20 < 10 else rng.uniform(-1, 0)
    ours, _ = oracle.dct2d_avx2(buf)
    ref = _ref32_dct(buf[:127, :127], d32)
    assert np.abs(ours - ref).max() <= 32 * F32_EPS


def test_dct_avx2_vs_scalar_equivalence(oracle):
    """mod.rs:1248-1350: the AVX2 and scalar DCTs with synthetic matrices,
    relative 4e-7 (part 1) and 8e-7 (part 2).  Uses the oracle's own tables
    for the scalar path with the test's constant / sequential matrices."""
    import ctypes as C
    L = oracle.lib()
    # part 1: DCT matrix all ones, buffer sequential * 1e-5
    buf = np.zeros((128, 128), np.float32)
    for i in range(127):
        for j in range(127):
            buf[i, j] = np.float32((i * 127 + j) * 0.00001)
    ones = np.ones(16 * 127 + 16, np.float32)
    ones[16 * 127:] = 0
    a = _dct_with_matrix(L, "avx2", buf, ones)
    s = _dct_with_matrix(L, "scalar", buf, ones)
    assert np.all(np.abs(s - a) / s < 4e-7)
    # part 2: buffer constant 1e-5, matrix sequential
    buf2 = np.zeros((128, 128), np.float32)
    buf2[:127, :127] = np.float32(0.00001)
    seq = np.arange(16 * 127 + 16, dtype=np.float32)
    seq[16 * 127:] = 0
    a = _dct_with_matrix(L, "avx2", buf2, seq)
    s = _dct_with_matrix(L, "scalar", buf2, seq)
    assert np.all(np.abs(s - a) / s < 8e-7)
    del C


def _dct_with_matrix(L, which, buf, mat):
    """Evaluate the oracle's AVX2 / scalar DCT formulas in numpy with a given
    matrix (the oracle's C functions are bound to the real table)."""
    D = mat
    out = np.zeros((16, 16), np.float32)
    if which == "avx2":
        alt = [1, 0, 0, 1]
        for k in range(16):
            tmp = np.zeros(128, np.float32)
            s = np.zeros((2, 128), np.float32)
            for k2 in range(128):
                t = alt[k2 & 3]
                s[t] = _fma(buf[k2], np.float32(D[k * 127 + k2]), s[t])
            tmp[:] = np.where(np.isnan(s[0]), 0, s[0]).astype(np.float32) + \
                np.where(np.isnan(s[1]), 0, s[1]).astype(np.float32)
            for j in range(16):
                h = np.zeros((2, 8), np.float32)
                for c in range(16):
                    t = alt[c & 3]
                    h[t] = _fma(tmp[8 * c:8 * c + 8], D[j * 127 + 8 * c:j * 127 + 8 * c + 8].astype(np.float32), h[t])
                v = (h[0] + h[1]).astype(np.float32)
                out[k, j] = np.float32(np.float32(np.float32(v[0] + v[1]) + np.float32(v[2] + v[3])) +
                                       np.float32(np.float32(v[4] + v[5]) + np.float32(v[6] + v[7])))
    else:
        for k in range(16):
            tmp = np.zeros(127, np.float32)
            for j in range(127):
                s = np.zeros(4, np.float32)
                for k2 in range(127):
                    s[k2 & 3] = np.float32(s[k2 & 3] + np.float32(np.float32(D[k * 127 + k2]) * buf[k2, j]))
                tmp[j] = np.float32(np.float32(np.float32(s[0] + s[1]) + s[2]) + s[3])
            for j in range(16):
                s = np.zeros(4, np.float32)
                for m in range(127):
                    s[m & 3] = np.float32(s[m & 3] + np.float32(tmp[m] * np.float32(D[j * 127 + m])))
                out[k, j] = np.float32(np.float32(np.float32(s[0] + s[1]) + s[2]) + s[3])
    return out


def _fma(a, b, c):
    """Correctly rounded f32 fused multiply-add, elementwise (exact in f64 for
    these magnitudes: the f32 product fits in 48 bits)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def test_oracle_dct_matches_fp64_closed_form(oracle):
    """The oracle's AVX2 DCT equals D C D^T from the f64 table within f32
    accumulation error, for a random compressed plane."""
    _, d64, _ = oracle.tables()
    D = d64[:16 * 127].reshape(16, 127)
    rng = np.random.default_rng(23)
    buf = np.zeros((128, 128), np.float32)
    buf[:127, :127] = rng.uniform(0, 255, (127, 127)).astype(np.float32)
    ours, _ = oracle.dct2d_avx2(buf)
    exact = D @ buf[:127, :127].astype(np.float64) @ D.T
    assert np.abs(ours - exact).max() <= 2e-5 * np.abs(exact).max()


def test_compress_matches_exact_tent(oracle):
    """Compression equals the exact rational 5x5 tent (weights k/81 at the
    sampling centres 4(i+1)) within f32 rounding -- a wrong offset, weight or
    transposition would be far off."""
    img = _images(1)[0]
    ours = oracle.compress_avx2(img)[:127, :127].astype(np.float64)
    a = np.array([0, 1, 2, 3, 2, 1, 0], np.float64) / 9.0
    W = np.outer(a, a)
    exact = np.zeros((127, 127))
    for i in range(127):
        for j in range(127):
            r, c = 4 * (i + 1) - 3, 4 * (j + 1) - 3
            exact[i, j] = (W * img[r:r + 7, c:c + 7]).sum()
    assert np.abs(ours - exact).max() <= 255 * 8 * F32_EPS
    # scalar DefaultKernel compression agrees to f32 rounding as well
    sc = oracle.compress_scalar(img)[:127, :127].astype(np.float64)
    assert np.abs(sc - exact).max() <= 255 * 8 * F32_EPS


def test_gradient_closed_form(oracle):
    rng = np.random.default_rng(2)
    p = rng.uniform(-50, 50, (16, 16)).astype(np.float32)
    g = oracle.sum_of_gradients(p)
    exact = np.abs(np.diff(p.astype(np.float64), axis=0)).sum() + np.abs(np.diff(p.astype(np.float64), axis=1)).sum()
    assert abs(g - exact) <= 480 * 64 * F32_EPS * exact / 480
    # quality saturates at 1 (mod.rs:611-615)
    assert oracle.lib().oracle_adjust_quality(180.0 * 2) == 1.0
    assert oracle.lib().oracle_adjust_quality(90.0) == pytest.approx(0.5)


# --- quantisation ----------------------------------------------------------------------

def test_quantize_threshold_splits_at_median(oracle):
    """The AVX2 search stops with at most 128 elements on either side of the
    threshold; bits are x > threshold at byte 31 - i/8, bit i%8."""
    rng = np.random.default_rng(4)
    for _ in range(200):
        p = rng.normal(0, 10, (16, 16)).astype(np.float32)
        h, t = oracle.quantize_avx2(p)
        flat = p.reshape(-1)
        assert (flat > t).sum() <= 128 and (flat < t).sum() <= 128
        exp = np.zeros(32, np.uint8)
        for i in range(256):
            if flat[i] > t:
                exp[31 - i // 8] |= 1 << (i % 8)
        assert np.array_equal(h, exp)


def test_rethreshold_reproduces_quantize(oracle):
    """lib.rs:659-687: threshold_2d_f32 with the returned threshold equals quantize."""
    for img in _images(4):
        h, t, _, pdqf, _, _ = oracle.hash_get_threshold(img)
        assert np.array_equal(oracle.threshold_2d(pdqf, t), h)


def test_torben_median_is_lower_median(oracle):
    rng = np.random.default_rng(6)
    for _ in range(100):
        p = rng.normal(0, 5, 256).astype(np.float32)
        assert oracle.lib().oracle_torben_median(p.ctypes.data_as(__import__("ctypes").c_void_p)) == \
            np.sort(p)[127]


def test_avx2_quantize_vs_torben_bits(oracle):
    """The AVX2 search and the scalar Torben median give the same bits on
    distinct-valued inputs (both split 128/128)."""
    rng = np.random.default_rng(8)
    for _ in range(100):
        p = rng.normal(0, 5, (16, 16)).astype(np.float32)
        assert np.array_equal(oracle.quantize_avx2(p)[0], oracle.quantize_torben(p)[0])


# --- dihedrals --------------------------------------------------------------------------

def test_dihedral_sign_algebra_is_exact_for_true_transforms(oracle):
    """Exact mode: each orientation's PDQF equals the f64 DCT of the
    correspondingly transformed 127x127 compressed plane (numpy flips)."""
    _, d64, _ = oracle.tables()
    D = d64[:16 * 127].reshape(16, 127)
    rng = np.random.default_rng(12)
    C = rng.normal(0, 50, (127, 127))
    dct = lambda X: (D @ X @ D.T).astype(np.float32)  # noqa: E731
    P = dct(C)
    codes, th, q, h, pd = oracle.visit_dihedrals(P, 0.0, 0)
    transforms = {  # label -> numpy transform of the compressed plane
        "FLOPPED": np.fliplr(C), "ROTATED_180": np.rot90(C, 2), "FLIPPED": np.flipud(C),
        "ROTATED_90_FLOPPED": C.T, "ROTATED_270": np.rot90(C, 1),
        "FLOPPED_ROTATED_270": np.rot90(C, 2).T, "ROTATED_90": np.rot90(C, 3)}
    order = ["FLOPPED", "ROTATED_180", "FLIPPED", "ROTATED_90_FLOPPED", "ROTATED_270",
             "FLOPPED_ROTATED_270", "ROTATED_90"]
    for k, name in enumerate(order):
        assert codes[k] == oracle.DIHEDRAL_CODES[name]
        assert np.abs(pd[k] - dct(transforms[name])).max() <= 1e-3 * np.abs(P).max(), name


def test_avx2_column_extent_differs_from_exact(oracle):
    """The AVX2 kernel negates columns 0..7 only (x86.rs:330-376): its FLOPPED
    PDQF keeps columns 8..15 unnegated; FLIPPED and the transposed-only
    orientations agree with the exact algebra."""
    rng = np.random.default_rng(13)
    P = rng.normal(0, 10, (16, 16)).astype(np.float32)
    _, _, _, _, ex = oracle.visit_dihedrals(P, 0.0, 0)
    _, _, _, _, av = oracle.visit_dihedrals(P, 0.0, 1)
    assert np.array_equal(av[0][:, 8:], P[:, 8:])          # FLOPPED: cols 8.. untouched
    assert np.array_equal(av[0][:, :8:2], -P[:, :8:2])
    assert np.array_equal(ex[0][:, ::2], -P[:, ::2])
    assert np.array_equal(av[2], ex[2])                     # FLIPPED identical
    assert np.array_equal(av[3], ex[3])                     # ROTATED_90_FLOPPED = P^T
    assert np.array_equal(av[4], ex[4])                     # ROTATED_270


# --- matching ----------------------------------------------------------------------------

def test_matcher_planted_controls(oracle):
    """src/matching/mod.rs:460-538 protocol: positive controls with 31 flips
    (with replacement, distance <= 31) at every tested position; negative
    controls on random data."""
    rng = np.random.default_rng(int.from_bytes(b"yume-pdq", "big") % (2 ** 63))
    for nq in (8, 16, 24):
        needles = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
        hay = rng.integers(0, 256, (2048, 32), dtype=np.uint8)
        assert oracle.match_all(hay, needles, 31)[0] == 0
        assert oracle.match_first(hay, needles, 31) is None
        assert oracle.cpu_match_count(hay, needles, 31, use_avx512=False) == 0
        for pos in rng.choice(2048, size=32, replace=False):
            i = int(rng.integers(0, nq))
            nd = needles.copy()
            nd[i] = hay[pos]
            for _ in range(31):
                nd[i, int(rng.integers(0, 32))] ^= np.uint8(1 << int(rng.integers(0, 8)))
            n, hits = oracle.match_all(hay, nd, 31)
            assert n >= 1 and [pos, i] in hits[:, [1, 0]].tolist()
            assert oracle.match_first(hay, nd, 31) == (int(pos), i)
            assert oracle.cpu_match_count(hay, nd, 31, use_avx512=False) == n
            if oracle.lib().oracle_cpu_has_vpopcntdq():
                assert oracle.cpu_match_count(hay, nd, 31, use_avx512=True) == n


def test_matcher_brute_force_closed_form(oracle):
    """Triples equal an independent numpy brute force (bitwise_count) and the
    distance of every planted row is exact."""
    from yume_pdq import synthetic
    db = synthetic.random_db(20000, seed=1)
    q = synthetic.random_db(16, seed=2)
    rows, qids, dists = synthetic.plant(db, q, 300, seed=3)
    n, hits = oracle.match_all(db, q, 31)
    exp = synthetic.brute_force_hits(db, q, 31)
    assert n == exp.shape[0] and np.array_equal(hits, exp)
    for r, qi, d in zip(rows, qids, dists):
        assert oracle.hamming(db[r], q[qi]) == d
    # inclusive threshold (mod.rs:251,354)
    assert oracle.match_all(db[:1], db[:1], 0)[0] == 1


def test_threaded_matcher_equals_single_threaded(oracle):
    """The threaded brute force gives oracle_match_all's triples in the same
    row-major order, for any thread count, at thresholds where few, some and
    all pairs match, and with a truncating cap (exact total)."""
    from yume_pdq import synthetic
    db = synthetic.random_db(30_011, seed=41)
    q = synthetic.random_db(40, seed=42)
    synthetic.plant(db, q, 500, seed=43)
    for thr in (0, 31, 40, 256):
        n, ref = oracle.match_all(db[:3001] if thr == 256 else db, q, thr, cap=200_000)
        for threads in (1, 3, 8):
            m, got = oracle.match_all_mt(db[:3001] if thr == 256 else db, q, thr, n_threads=threads)
            assert m == n and np.array_equal(got, ref), (thr, threads)
    n, ref = oracle.match_all(db, q, 40)
    m, got = oracle.match_all_mt(db, q, 40, cap=100, n_threads=4)
    assert m == n > 100 and np.array_equal(got, ref[:100])
    assert oracle.match_all_mt(db[:0], q, 31)[0] == 0


# --- precise (reference) kernel and the FP64 oracle -------------------------------------

def _tent_closed_form_f64(img, centres, w):
    """sum_a sum_b w[a] w[b] img[c_i + a - r][c_j + b - r] over the given centres."""
    r = len(w) // 2
    img = img.astype(np.float64)
    out = np.zeros((len(centres), len(centres)))
    W = np.outer(w, w)
    for i, ci in enumerate(centres):
        for j, cj in enumerate(centres):
            out[i, j] = (W * img[ci - r:ci + r + 1, cj - r:cj + r + 1]).sum()
    return out


def test_pdqf_f64_closed_form(oracle):
    """The FP64 oracle (tent at centres 4(i+1), weights k/81, f64 DCT) equals
    the closed form D64 C D64^T evaluated independently in numpy."""
    _, d64, _ = oracle.tables()
    D = d64[:16 * 127].reshape(16, 127)
    img = _images(1)[0]
    C = _tent_closed_form_f64(img, [4 * (i + 1) for i in range(127)],
                              np.array([1, 2, 3, 2, 1], np.float64) / 9.0)
    exact = D @ C @ D.T
    got = oracle.pdqf_f64(img)
    assert np.abs(got - exact).max() <= 1e-12 * np.abs(exact).max()


def test_reference_kernel_closed_form(oracle):
    """ReferenceKernel (mod.rs:790-820, ref.rs:7-194): two Jarosz passes of a
    width-3 box in each direction are the separable (1,2,3,2,1)/9 filter away
    from the border; sampled at floor((2i+1) 512 / 254) every sample is
    interior, so the f64 kernel's PDQF is D64 C D64^T of that closed form to
    f64 rounding, and the f32 kernel's to f32 accumulation error.  The bits
    of both follow the lower-median threshold (mod.rs:764-788)."""
    _, d64, _ = oracle.tables()
    D = d64[:16 * 127].reshape(16, 127)
    for img in _images(2):
        centres = [((2 * i + 1) * 512) // 254 for i in range(127)]
        C = _tent_closed_form_f64(img, centres, np.array([1, 2, 3, 2, 1], np.float64) / 9.0)
        exact = D @ C @ D.T
        h64, t64, q64, p64 = oracle.hash_ref64(img)
        assert np.abs(p64 - exact).max() <= 1e-11 * np.abs(exact).max()
        h32, t32, q32, p32 = oracle.hash_ref32(img)
        assert np.abs(p32 - exact).max() <= 1e-4 * np.abs(exact).max()
        for h, p, t in ((h64, p64, t64), (h32, p32, t32)):
            flat = p.reshape(-1)
            assert t == np.sort(flat)[127]
            exp = np.zeros(32, np.uint8)
            for i in range(256):
                if flat[i] > t:
                    exp[31 - i // 8] |= 1 << (i % 8)
            assert np.array_equal(h, exp)
        assert bits(h32, h64) <= 8


def test_scalar_dct_matches_fp64_closed_form(oracle):
    """DefaultKernel::dct2d (mod.rs:543-574) restated in the oracle equals
    D C D^T from the f64 table within f32 accumulation error, and agrees with
    the AVX2 restatement to the same order."""
    _, d64, _ = oracle.tables()
    D = d64[:16 * 127].reshape(16, 127)
    rng = np.random.default_rng(29)
    buf = np.zeros((128, 128), np.float32)
    buf[:127, :127] = rng.uniform(0, 255, (127, 127)).astype(np.float32)
    exact = D @ buf[:127, :127].astype(np.float64) @ D.T
    sc = oracle.dct2d_scalar(buf)
    av, _ = oracle.dct2d_avx2(buf)
    scale = np.abs(exact).max()
    assert np.abs(sc - exact).max() <= 2e-5 * scale
    assert np.abs(sc - av).max() <= 4e-5 * scale


def test_fp32_path_error_vs_fp64(oracle):
    """The error statement for the FP32 path (which the GPU kernel reproduces
    bit for bit): on synthetic images every PDQF coefficient is within
    1e-6 x max|coefficient| of the FP64 oracle, and the hash equals the hash
    of the FP64 PDQF thresholded at its lower median (0 bits disagree)."""
    for img in _images(16):
        h, _, _, p, _, _ = oracle.hash_get_threshold(img)
        p64 = oracle.pdqf_f64(img)
        assert np.abs(p.astype(np.float64) - p64).max() <= 1e-6 * np.abs(p64).max()
        flat = p64.reshape(-1)
        t = np.sort(flat)[127]
        h64 = np.zeros(32, np.uint8)
        for i in range(256):
            if flat[i] > t:
                h64[31 - i // 8] |= 1 << (i % 8)
        assert bits(h, h64) == 0
