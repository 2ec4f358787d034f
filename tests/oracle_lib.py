"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker, never by the product.
"""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(REPO, "oracle", "liboracle.so")

_lib = None

DIHEDRAL_ORDER = ["NORMAL", "FLOPPED", "ROTATED_180", "FLIPPED", "ROTATED_90_FLOPPED",
                  "ROTATED_270", "FLOPPED_ROTATED_270", "ROTATED_90"]
DIHEDRAL_CODES = {
    "NORMAL": 0x01000001, "FLIPPED": 0x010000FF, "FLOPPED": 0xFF000001,
    "ROTATED_180": 0xFF0000FF, "ROTATED_90": 0x0001FF00, "ROTATED_270": 0x00010100,
    "ROTATED_90_FLOPPED": 0x00FFFF00, "FLOPPED_ROTATED_270": 0x00FF0100,
}


class Hit(C.Structure):
    _fields_ = [("query", C.c_uint32), ("distance", C.c_uint32), ("index", C.c_uint64)]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, f, u32, u64, i = C.c_void_p, C.c_float, C.c_uint32, C.c_uint64, C.c_int
        sig = {
            "oracle_tables": (None, [vp, vp, vp]),
            "oracle_compress_avx2": (None, [vp, vp]),
            "oracle_dct2d_avx2": (None, [vp, vp, vp]),
            "oracle_sum_of_gradients": (f, [vp]),
            "oracle_adjust_quality": (f, [f]),
            "oracle_quantize_avx2": (None, [vp, vp, vp]),
            "oracle_threshold_2d": (None, [vp, vp, f]),
            "oracle_torben_median": (f, [vp]),
            "oracle_quantize_torben": (None, [vp, vp, vp]),
            "oracle_compress_scalar": (None, [vp, vp]),
            "oracle_dct2d_scalar": (None, [vp, vp]),
            "oracle_hash_ref32": (f, [vp, vp, vp, vp]),
            "oracle_hash_ref64": (C.c_double, [vp, vp, vp, vp]),
            "oracle_pdqf_f64": (None, [vp, vp]),
            "oracle_readme_fm_image": (None, [vp]),
            "oracle_hash_get_threshold": (f, [vp, vp, vp, vp, vp, vp]),
            "oracle_avx2_hash": (f, [vp, vp, vp, vp, vp, vp]),
            "oracle_visit_dihedrals": (None, [vp, f, i, vp, vp, vp, vp, vp]),
            "oracle_hash_dihedrals": (None, [vp, i, vp, vp, vp, vp]),
            "oracle_hamming": (u32, [vp, vp]),
            "oracle_match_all": (u64, [vp, u64, vp, u32, u32, vp, u64]),
            "oracle_match_first": (i, [vp, u64, vp, u32, u32, vp, vp]),
            "oracle_match_all_mt": (u64, [vp, u64, vp, u32, u32, vp, u64, u32]),
            "oracle_cpu_has_vpopcntdq": (i, []),
            "oracle_cpu_match_count": (u64, [vp, u64, vp, u32, u32, i]),
            "oracle_cpu_baseline": (u64, [vp, u32, u32, vp, u64, u32, u32, i, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def tables():
    d32 = np.zeros(16 * 127 + 16, np.float32)
    d64 = np.zeros(16 * 127 + 16, np.float64)
    tent = np.zeros(56, np.float32)
    lib().oracle_tables(_p(d32), _p(d64), _p(tent))
    return d32, d64, tent


def _plane(img):
    a = np.ascontiguousarray(img, dtype=np.float32)
    assert a.shape == (512, 512)
    return a


def compress_avx2(img):
    out = np.zeros((128, 128), np.float32)
    lib().oracle_compress_avx2(_p(_plane(img)), _p(out))
    return out


def compress_scalar(img):
    out = np.zeros((128, 128), np.float32)
    lib().oracle_compress_scalar(_p(_plane(img)), _p(out))
    return out


def dct2d_avx2(buf):
    buf = np.ascontiguousarray(buf, np.float32)
    tmp = np.zeros(128, np.float32)
    out = np.zeros((16, 16), np.float32)
    lib().oracle_dct2d_avx2(_p(buf), _p(tmp), _p(out))
    return out, tmp


def dct2d_scalar(buf):
    buf = np.ascontiguousarray(buf, np.float32)
    out = np.zeros((16, 16), np.float32)
    lib().oracle_dct2d_scalar(_p(buf), _p(out))
    return out


def quantize_avx2(pdqf):
    p = np.ascontiguousarray(pdqf, np.float32)
    thr = np.zeros(1, np.float32)
    out = np.zeros(32, np.uint8)
    lib().oracle_quantize_avx2(_p(p), _p(thr), _p(out))
    return out, float(thr[0])


def quantize_torben(pdqf):
    p = np.ascontiguousarray(pdqf, np.float32)
    thr = np.zeros(1, np.float32)
    out = np.zeros(32, np.uint8)
    lib().oracle_quantize_torben(_p(p), _p(thr), _p(out))
    return out, float(thr[0])


def threshold_2d(pdqf, thr):
    p = np.ascontiguousarray(pdqf, np.float32)
    out = np.zeros(32, np.uint8)
    lib().oracle_threshold_2d(_p(p), _p(out), C.c_float(thr))
    return out


def sum_of_gradients(pdqf):
    return lib().oracle_sum_of_gradients(_p(np.ascontiguousarray(pdqf, np.float32)))


def hash_get_threshold(img):
    """lib.rs:427-455 with the AVX2 kernel: (hash[32], threshold, quality, pdqf, buf1, tmp)."""
    a = _plane(img)
    thr = np.zeros(1, np.float32)
    out = np.zeros(32, np.uint8)
    buf1 = np.zeros((128, 128), np.float32)
    tmp = np.zeros(128, np.float32)
    pdqf = np.zeros((16, 16), np.float32)
    q = lib().oracle_hash_get_threshold(_p(a), _p(thr), _p(out), _p(buf1), _p(tmp), _p(pdqf))
    return out, float(thr[0]), float(q), pdqf, buf1, tmp


def avx2_hash(img):
    a = _plane(img)
    thr = np.zeros(1, np.float32)
    out = np.zeros(32, np.uint8)
    buf1 = np.zeros((128, 128), np.float32)
    tmp = np.zeros(128, np.float32)
    pdqf = np.zeros((16, 16), np.float32)
    q = lib().oracle_avx2_hash(_p(a), _p(thr), _p(out), _p(buf1), _p(tmp), _p(pdqf))
    return out, float(thr[0]), float(q), pdqf, buf1, tmp


def hash_ref32(img):
    a = _plane(img)
    thr = np.zeros(1, np.float32)
    out = np.zeros(32, np.uint8)
    pdqf = np.zeros((16, 16), np.float32)
    q = lib().oracle_hash_ref32(_p(a), _p(thr), _p(out), _p(pdqf))
    return out, float(thr[0]), float(q), pdqf


def hash_ref64(img):
    a = _plane(img)
    thr = np.zeros(1, np.float64)
    out = np.zeros(32, np.uint8)
    pdqf = np.zeros((16, 16), np.float64)
    q = lib().oracle_hash_ref64(_p(a), _p(thr), _p(out), _p(pdqf))
    return out, float(thr[0]), float(q), pdqf


def pdqf_f64(img):
    out = np.zeros((16, 16), np.float64)
    lib().oracle_pdqf_f64(_p(_plane(img)), _p(out))
    return out


def visit_dihedrals(pdqf, threshold, mode):
    """lib.rs:358-420.  mode 0 = exact sign algebra, 1 = AVX2 column extent."""
    p = np.ascontiguousarray(pdqf, np.float32)
    codes = np.zeros(7, np.uint32)
    th = np.zeros(7, np.float32)
    q = np.zeros(7, np.float32)
    h = np.zeros((7, 32), np.uint8)
    pd = np.zeros((7, 16, 16), np.float32)
    lib().oracle_visit_dihedrals(_p(p), C.c_float(threshold), mode, _p(codes), _p(th), _p(q),
                                 _p(h), _p(pd))
    return codes, th, q, h, pd


def hash_dihedrals(img, mode=0):
    """8 hashes (NORMAL first, then the reference's visiting order), qualities, thresholds."""
    a = _plane(img)
    h = np.zeros((8, 32), np.uint8)
    q = np.zeros(8, np.float32)
    t = np.zeros(8, np.float32)
    pd = np.zeros((16, 16), np.float32)
    lib().oracle_hash_dihedrals(_p(a), mode, _p(h), _p(q), _p(t), _p(pd))
    return h, q, t, pd


def readme_fm_image():
    out = np.zeros((512, 512), np.float32)
    lib().oracle_readme_fm_image(_p(out))
    return out


def hamming(a, b):
    return int(lib().oracle_hamming(_p(np.ascontiguousarray(a, np.uint8)),
                                    _p(np.ascontiguousarray(b, np.uint8))))


def match_all(db, queries, threshold, cap=None):
    """All (query, index, distance) triples, row-major order; returns (total, array)."""
    db = np.ascontiguousarray(db, np.uint8).reshape(-1, 32)
    q = np.ascontiguousarray(queries, np.uint8).reshape(-1, 32)
    if cap is None:
        cap = 1 << 20
    hits = (Hit * max(cap, 1))()
    n = lib().oracle_match_all(_p(db), db.shape[0], _p(q), q.shape[0], threshold, hits, cap)
    k = min(n, cap)
    arr = np.array([(hits[i].query, hits[i].index, hits[i].distance) for i in range(k)],
                   dtype=np.int64).reshape(-1, 3)
    return int(n), arr


HIT_DTYPE = np.dtype([("query", "<u4"), ("distance", "<u4"), ("index", "<u8")])


def default_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def match_all_mt(db, queries, threshold, cap=1 << 24, n_threads=None):
    """match_all on several threads (same triples, same row-major order):
    returns (total, (k, 3) int64 array of (query, index, distance))."""
    db = np.ascontiguousarray(db, np.uint8).reshape(-1, 32)
    q = np.ascontiguousarray(queries, np.uint8).reshape(-1, 32)
    buf = np.zeros(max(int(cap), 1), HIT_DTYPE)
    n = lib().oracle_match_all_mt(_p(db), db.shape[0], _p(q), q.shape[0], int(threshold), _p(buf), int(cap),
                                  int(n_threads or default_threads()))
    if n == 0xFFFFFFFFFFFFFFFF:
        raise MemoryError("oracle_match_all_mt: a worker ran out of memory")
    k = min(int(n), int(cap))
    b = buf[:k]
    t = np.stack([b["query"].astype(np.int64), b["index"].astype(np.int64), b["distance"].astype(np.int64)], 1) \
        if k else np.zeros((0, 3), np.int64)
    return int(n), t


def match_first(db, queries, threshold):
    db = np.ascontiguousarray(db, np.uint8).reshape(-1, 32)
    q = np.ascontiguousarray(queries, np.uint8).reshape(-1, 32)
    row = C.c_uint64(0)
    qq = C.c_uint32(0)
    ok = lib().oracle_match_first(_p(db), db.shape[0], _p(q), q.shape[0], threshold,
                                  C.byref(row), C.byref(qq))
    return (int(row.value), int(qq.value)) if ok else None


def cpu_match_count(db, queries, threshold, use_avx512=True):
    db = np.ascontiguousarray(db, np.uint8).reshape(-1, 32)
    q = np.ascontiguousarray(queries, np.uint8).reshape(-1, 32)
    return int(lib().oracle_cpu_match_count(_p(db), db.shape[0], _p(q), q.shape[0], threshold,
                                            1 if use_avx512 else 0))


def cpu_baseline(images, db, threshold, n_threads, use_avx512=True, n_images=None):
    """Reference CPU kernel + matcher over n_images images drawn cyclically from
    the pool `images` (n,512,512 f32) and db (m,32 u8)."""
    imgs = np.ascontiguousarray(images, np.float32).reshape(-1, 512, 512)
    db = np.ascontiguousarray(db, np.uint8).reshape(-1, 32)
    n_images = imgs.shape[0] if n_images is None else int(n_images)
    hashes = np.zeros((n_images, 8, 32), np.uint8)
    th = C.c_double(0)
    tm = C.c_double(0)
    hits = lib().oracle_cpu_baseline(_p(imgs), imgs.shape[0], n_images, _p(db), db.shape[0], threshold,
                                     n_threads, 1 if use_avx512 else 0, _p(hashes),
                                     C.byref(th), C.byref(tm))
    return int(hits), hashes, th.value, tm.value
