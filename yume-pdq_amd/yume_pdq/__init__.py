"""Python binding of libyume_pdq.so (MI355X / gfx950 PDQ hash + Hamming matcher).

Mirrors the reference's FFI (include/yume_pdq.h, integration/hash.py) and adds
the batched device API of include/yume_pdq_amd.h over torch device tensors.
Every call runs on the GPU through the HIP library; if the library is missing
the import of the binding raises -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("LIBYUME_PDQ_PATH", os.path.join(PKG_DIR, "libyume_pdq.so"))

INPUT_F32, INPUT_U8, INPUT_RGB8, INPUT_RGBA8 = 0, 1, 2, 3
DIHEDRAL_EXACT, DIHEDRAL_AVX2 = 0, 1
MATCH_APPEND = 1
MATCH_VALU = 2
MATCH_MFMA = 4
# DCT arithmetic and kernel layout of hash_batch (YUME_PDQ_AMD_DCT_*, _LAYOUT_*)
DCT_EXACT, DCT_BF16, DCT_F16, DCT_BF16X3, DCT_F16X3 = 0, 1, 2, 3, 4
LAYOUT_AUTO, LAYOUT_WAVE, LAYOUT_BLOCK4, LAYOUT_BLOCK8, LAYOUT_BLOCK16 = 0, 1, 2, 3, 4

# Orientation order of the 8 hashes per image (YUME_PDQ_AMD_D_*).
DIHEDRALS = ["NORMAL", "FLOPPED", "ROTATED_180", "FLIPPED", "ROTATED_90_FLOPPED",
             "ROTATED_270", "FLOPPED_ROTATED_270", "ROTATED_90"]

DIHEDRAL_CODES = {  # Dihedrals::* packed, src/lib.rs:119-133
    "NORMAL": 0x01000001, "FLIPPED": 0x010000FF, "FLOPPED": 0xFF000001,
    "ROTATED_180": 0xFF0000FF, "ROTATED_90": 0x0001FF00, "ROTATED_270": 0x00010100,
    "ROTATED_90_FLOPPED": 0x00FFFF00, "FLOPPED_ROTATED_270": 0x00FF0100,
}


class YumePdqError(RuntimeError):
    pass


class Hit(C.Structure):
    _fields_ = [("query", C.c_uint32), ("distance", C.c_uint32), ("index", C.c_uint64)]


DihedralCallback = C.CFUNCTYPE(C.c_bool, C.c_void_p, C.c_uint32, C.c_float, C.c_float)


class BatchBuffers(C.Structure):
    _fields_ = [("d_input", C.c_void_p), ("input_stride_bytes", C.c_uint64), ("d_hashes", C.c_void_p),
                ("d_quality", C.c_void_p), ("d_threshold", C.c_void_p), ("d_hits", C.c_void_p),
                ("hit_capacity", C.c_uint64), ("d_count", C.c_void_p), ("stream", C.c_void_p),
                ("max_images", C.c_uint32), ("device_bytes", C.c_uint64)]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch_images", C.c_uint32), ("hbm_budget_bytes", C.c_uint64), ("max_contexts", C.c_uint32),
                ("threshold", C.c_uint32), ("dihedral_mode", C.c_uint32), ("hit_capacity", C.c_uint64),
                ("use_graph", C.c_int), ("busy_poll", C.c_int), ("in_order_compute", C.c_int),
                ("autotune_iterations", C.c_uint32), ("scan_chunk_rows", C.c_uint64)]


class ServerResult(C.Structure):
    _fields_ = [("status", C.c_int), ("n_images", C.c_uint32), ("hashes", C.c_void_p), ("quality", C.c_void_p),
                ("threshold", C.c_void_p), ("hits", C.c_void_p), ("n_hits", C.c_uint64),
                ("latency_ms", C.c_double), ("batch", C.c_uint64), ("batch_images", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("contexts", C.c_uint32), ("device_bytes", C.c_uint64), ("batches", C.c_uint64),
                ("images", C.c_uint64)]


class HashOptions(C.Structure):
    _fields_ = [("dihedral_mode", C.c_uint32), ("dct_precision", C.c_uint32), ("layout", C.c_uint32),
                ("luma_coeff", C.POINTER(C.c_float))]


class PipeConfig(C.Structure):
    _fields_ = [("input_fd", C.c_int), ("output_fd", C.c_int), ("input_kind", C.c_int),
                ("output_format", C.c_char_p), ("dihedrals", C.c_int), ("dihedral_mode", C.c_uint32),
                ("batch_frames", C.c_uint32), ("stats", C.c_int)]


class TuneRecord(C.Structure):
    _fields_ = [("n_queries", C.c_uint32), ("path", C.c_uint32), ("query_tiles", C.c_uint32), ("ms", C.c_float)]

_lib = None


def lib():
    """Load libyume_pdq.so (built by __graft_entry__.build() / make -C yume-pdq_amd)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise YumePdqError(f"libyume_pdq.so not found at {LIB_PATH}; build it first "
                               "(python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        vp, f, u32, u64, i = C.c_void_p, C.c_float, C.c_uint32, C.c_uint64, C.c_int
        L.yume_pdq_hash_smart_kernel.restype = f
        L.yume_pdq_hash_smart_kernel.argtypes = [vp, vp, vp, vp, vp, vp]
        L.yume_pdq_visit_dihedrals_smart_kernel.restype = C.c_bool
        L.yume_pdq_visit_dihedrals_smart_kernel.argtypes = [vp, f, vp, vp, DihedralCallback]
        L.yume_pdq_amd_strerror.restype = C.c_char_p
        L.yume_pdq_amd_strerror.argtypes = [i]
        L.yume_pdq_amd_device_info.restype = i
        L.yume_pdq_amd_device_info.argtypes = [i, C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]
        L.yume_pdq_amd_hash_batch.restype = i
        L.yume_pdq_amd_hash_batch.argtypes = [vp, i, u64, u32, u32, vp, vp, vp, vp, vp]
        L.yume_pdq_amd_hash_ref_workspace_bytes.restype = u64
        L.yume_pdq_amd_hash_ref_workspace_bytes.argtypes = [u32]
        L.yume_pdq_amd_hash_ref_batch.restype = i
        L.yume_pdq_amd_hash_ref_batch.argtypes = [vp, i, u64, u32, vp, u64, vp, vp, vp, vp, vp]
        L.yume_pdq_amd_pipe_run.restype = i
        L.yume_pdq_amd_pipe_run.argtypes = [C.POINTER(PipeConfig), C.POINTER(C.c_uint64)]
        L.yume_pdq_amd_pipe_trim.restype = i
        L.yume_pdq_amd_pipe_trim.argtypes = []
        L.yume_pdq_amd_format_frames.restype = C.c_int64
        L.yume_pdq_amd_format_frames.argtypes = [C.c_char_p, u32, i, vp, vp, vp, vp, vp, u64]
        L.yume_pdq_amd_format_bound.restype = u64
        L.yume_pdq_amd_format_bound.argtypes = [C.c_char_p, u32, i]
        L.yume_pdq_amd_hash_batch_ex.restype = i
        L.yume_pdq_amd_hash_batch_ex.argtypes = [vp, i, u64, u32, C.POINTER(HashOptions), vp, vp, vp, vp, vp, vp]
        L.yume_pdq_amd_hash_batch_luma.restype = i
        L.yume_pdq_amd_hash_batch_luma.argtypes = [vp, i, u64, u32, u32, C.POINTER(C.c_float), vp, vp, vp, vp, vp]
        L.yume_pdq_amd_dihedrals_batch.restype = i
        L.yume_pdq_amd_dihedrals_batch.argtypes = [vp, vp, u32, u32, vp, vp, vp, vp, vp]
        L.yume_pdq_amd_match.restype = i
        L.yume_pdq_amd_match.argtypes = [vp, u64, vp, u32, u32, vp, u64, vp, u64, u32, vp]
        L.yume_pdq_amd_match_first.restype = i
        L.yume_pdq_amd_match_first.argtypes = [vp, u64, vp, u32, u32, vp, vp]
        L.yume_pdq_amd_match_with_plane.restype = i
        L.yume_pdq_amd_match_with_plane.argtypes = [vp, vp, u64, vp, u32, u32, vp, u64, vp, u64, u32, vp]
        L.yume_pdq_amd_match_noalloc.restype = i
        L.yume_pdq_amd_match_noalloc.argtypes = [vp, vp, u64, vp, u32, u32, vp, u64, vp, u64, u32, vp]
        L.yume_pdq_amd_match_set_record_cap.restype = i
        L.yume_pdq_amd_match_set_record_cap.argtypes = [u32]
        L.yume_pdq_amd_hash_match_batch.restype = i
        L.yume_pdq_amd_hash_match_batch.argtypes = [vp, i, u64, u32, u32, vp, vp, u64, u32, vp, vp, vp, vp, u64,
                                                    vp, vp, vp]
        L.yume_pdq_amd_prefix_plane_bytes.restype = u64
        L.yume_pdq_amd_prefix_plane_bytes.argtypes = [u64]
        L.yume_pdq_amd_build_prefix_plane.restype = i
        L.yume_pdq_amd_build_prefix_plane.argtypes = [vp, u64, vp, vp]
        pp = C.POINTER(C.c_void_p)
        for name, res, args in (
                ("yume_pdq_amd_shards_create", i, [pp]),
                ("yume_pdq_amd_shards_destroy", None, [vp]),
                ("yume_pdq_amd_shards_add", i, [vp, vp, u64, u64, i, C.POINTER(C.c_uint32)]),
                ("yume_pdq_amd_shards_load_raw", i, [vp, C.c_char_p, u64, u64, C.POINTER(C.c_uint32)]),
                ("yume_pdq_amd_shards_replace", i, [vp, u32, vp, vp, u32, vp, vp]),
                ("yume_pdq_amd_shards_info", i, [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                                 C.POINTER(C.c_uint64)]),
                ("yume_pdq_amd_shards_match", i, [vp, vp, u32, u32, vp, u64, vp, u32, vp]),
                ("yume_pdq_amd_batcher_create", i, [vp, u32, i, u32, u32, u64, pp]),
                ("yume_pdq_amd_batcher_destroy", None, [vp]),
                ("yume_pdq_amd_batcher_buffers", i, [vp, C.POINTER(BatchBuffers)]),
                ("yume_pdq_amd_batcher_run", i, [vp, u32, i]),
                ("yume_pdq_amd_batcher_autotune", i, [vp, i]),
                ("yume_pdq_amd_batcher_tune_records", i, [vp, C.POINTER(TuneRecord), u32,
                                                          C.POINTER(C.c_uint32)]),
                ("yume_pdq_amd_batcher_choice", i, [vp, u32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
                ("yume_pdq_amd_tune_save", i, [C.c_char_p]),
                ("yume_pdq_amd_tune_load", i, [C.c_char_p]),
                ("yume_pdq_amd_tune_clear", i, []),
                ("yume_pdq_amd_server_create", i, [vp, C.POINTER(ServerConfig), pp]),
                ("yume_pdq_amd_server_destroy", None, [vp]),
                ("yume_pdq_amd_server_submit", i, [vp, vp, u32, C.POINTER(C.c_uint64)]),
                ("yume_pdq_amd_server_submit_pinned", i, [vp, vp, u32, C.POINTER(C.c_uint64)]),
                ("yume_pdq_amd_server_wait", i, [vp, u64, i, C.POINTER(ServerResult)]),
                ("yume_pdq_amd_server_release", i, [vp, u64]),
                ("yume_pdq_amd_server_stats", i, [vp, C.POINTER(ServerStats)])):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L.yume_pdq_amd_match_path.restype = i
        L.yume_pdq_amd_match_path.argtypes = [u32, u32, u32]
        for name, args in (("compress", [vp, i, u64, u32, vp, vp]),
                           ("dct2d", [vp, u32, vp, vp, vp]),
                           ("sum_of_gradients", [vp, u32, vp, vp]),
                           ("quantize", [vp, u32, vp, vp, vp]),
                           ("threshold", [vp, vp, u32, vp, vp]),
                           ("transform", [vp, u32, i, vp]),
                           ("luma", [vp, i, u64, f, f, f, vp, vp])):
            fn = getattr(L, "yume_pdq_amd_stage_" + name)
            fn.restype = i
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc < 0:
        raise YumePdqError(f"{what}: {lib().yume_pdq_amd_strerror(rc).decode()} ({rc})")
    return rc


def version():
    L = lib()
    return tuple(C.c_uint32.in_dll(L, n).value for n in
                 ("YUME_PDQ_VERSION_MAJOR", "YUME_PDQ_VERSION_MINOR", "YUME_PDQ_VERSION_PATCH"))


def dihedral_constant(name):
    return C.c_uint32.in_dll(lib(), "YUME_PDQ_DIHEDRAL_" + name).value


def device_info(device=0):
    name = C.create_string_buffer(256)
    cus = C.c_int(0)
    n = _check(lib().yume_pdq_amd_device_info(device, name, 256, C.byref(cus)), "device_info")
    return {"devices": n, "name": name.value.decode(), "compute_units": cus.value}


# ---------------------------------------------------------------------------
# Drop-in ABI (the reference's include/yume_pdq.h)
# ---------------------------------------------------------------------------
def hash_smart_kernel(image):
    """yume_pdq_hash_smart_kernel on one 512x512 f32 image (host arrays).

    Returns (hash[32] u8, threshold, quality, pdqf[16,16], buf1[128,128], tmp[128])."""
    a = np.ascontiguousarray(image, dtype=np.float32)
    if a.shape != (512, 512):
        raise ValueError("image must be 512x512")
    thr = np.zeros(1, np.float32)
    out = np.zeros(32, np.uint8)
    buf1 = np.zeros((128, 128), np.float32)
    tmp = np.zeros(128, np.float32)
    pdqf = np.zeros((16, 16), np.float32)
    p = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
    q = lib().yume_pdq_hash_smart_kernel(p(a), p(thr), p(out), p(buf1), p(tmp), p(pdqf))
    return out, float(thr[0]), float(q), pdqf, buf1, tmp


def visit_dihedrals_smart_kernel(threshold, pdqf, callback):
    """yume_pdq_visit_dihedrals_smart_kernel.  callback(code, threshold, quality,
    hash[32], pdqf[16,16]) -> bool (False stops).  Returns the ABI's bool."""
    out = np.zeros(32, np.uint8)
    pd = np.ascontiguousarray(pdqf, dtype=np.float32).copy()

    def tramp(_ctx, code, thr, quality):
        return bool(callback(code, thr, quality, out.copy(), pd.reshape(16, 16).copy()))

    cb = DihedralCallback(tramp)
    return bool(lib().yume_pdq_visit_dihedrals_smart_kernel(
        None, C.c_float(threshold), out.ctypes.data_as(C.c_void_p), pd.ctypes.data_as(C.c_void_p), cb))


# ---------------------------------------------------------------------------
# Batched device API over torch tensors (include/yume_pdq_amd.h)
# ---------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def _stream_ptr(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def hash_batch(planes, dihedral_mode=DIHEDRAL_AVX2, want_pdqf=False, stream=None, out=None, luma=None,
               dct=DCT_EXACT, layout=LAYOUT_AUTO, want_buf1=False):
    """Hash n 512x512 planes resident on the GPU: torch f32 or u8 luma
    (n, 512, 512), or u8 RGB8 / RGBA8 pixels (n, 512, 512, 3 or 4) converted to
    luma while read, with (R, G, B) coefficients `luma` (default BT.601).

    dct: DCT_EXACT (the reference's FP32 arithmetic, bit-identical) or a
    matrix-core precision (DCT_BF16, DCT_F16, DCT_BF16X3, DCT_F16X3); layout:
    LAYOUT_AUTO, LAYOUT_WAVE (one wavefront per image), LAYOUT_BLOCK4,
    LAYOUT_BLOCK8 or LAYOUT_BLOCK16 (four, eight or sixteen wavefronts per
    image; sixteen with exact arithmetic only).

    Returns (hashes (n,8,32) u8, quality (n,8) f32, threshold (n,8) f32
    [, pdqf (n,16,16)][, buf1 (n,128,128) compressed planes])."""
    torch = _torch()
    colour = planes.dim() >= 3 and planes.shape[-1] in (3, 4) and planes.shape[-3:-1] == (512, 512)
    if colour and planes.dim() == 3:
        planes = planes.unsqueeze(0)
    if not colour and planes.dim() == 2:
        planes = planes.unsqueeze(0)
    if not planes.is_cuda or (planes.shape[1:3] != (512, 512)):
        raise ValueError("planes must be a CUDA tensor of shape (n, 512, 512[, 3|4])")
    if colour:
        if planes.dtype != torch.uint8:
            raise ValueError("colour planes must be uint8")
        kind, elem = (INPUT_RGB8, 3) if planes.shape[-1] == 3 else (INPUT_RGBA8, 4)
        planes = planes.contiguous()
    elif planes.dtype == torch.float32:
        kind, elem = INPUT_F32, 1
    elif planes.dtype == torch.uint8:
        kind, elem = INPUT_U8, 1
    else:
        raise ValueError("planes must be float32 or uint8")
    if planes.dtype == torch.float32:
        elem = 4
    if not colour and (planes.stride(-1) != 1 or planes.stride(-2) != 512):
        planes = planes.contiguous()
    n = planes.shape[0]
    per = 512 * 512 * (planes.shape[-1] if colour else 1)  # elements per plane
    # a size-1 batch dimension may carry any stride (e.g. 0 from x[None])
    stride = planes.stride(0) if n > 1 else per
    if stride < per:
        planes = planes.contiguous()
        stride = per
    if colour:
        elem = 1  # stride counted in bytes of u8 elements
    dev = planes.device
    if out is None:
        hashes = torch.empty((n, 8, 32), dtype=torch.uint8, device=dev)
        quality = torch.empty((n, 8), dtype=torch.float32, device=dev)
        threshold = torch.empty((n, 8), dtype=torch.float32, device=dev)
        pdqf = torch.empty((n, 16, 16), dtype=torch.float32, device=dev) if want_pdqf else None
    else:
        hashes, quality, threshold, pdqf = out
    buf1 = torch.empty((n, 128, 128), dtype=torch.float32, device=dev) if want_buf1 else None
    pd = C.c_void_p(pdqf.data_ptr() if pdqf is not None else 0)
    if dct == DCT_EXACT and layout == LAYOUT_AUTO and not want_buf1:
        if luma is None:
            _check(lib().yume_pdq_amd_hash_batch(
                C.c_void_p(planes.data_ptr()), kind, stride * elem, n, dihedral_mode,
                C.c_void_p(hashes.data_ptr()), C.c_void_p(quality.data_ptr()),
                C.c_void_p(threshold.data_ptr()), pd, _stream_ptr(stream)), "hash_batch")
        else:
            coeff = (C.c_float * 3)(*[float(x) for x in luma])
            _check(lib().yume_pdq_amd_hash_batch_luma(
                C.c_void_p(planes.data_ptr()), kind, stride * elem, n, dihedral_mode, coeff,
                C.c_void_p(hashes.data_ptr()), C.c_void_p(quality.data_ptr()),
                C.c_void_p(threshold.data_ptr()), pd, _stream_ptr(stream)), "hash_batch_luma")
    else:
        coeff = (C.c_float * 3)(*[float(x) for x in luma]) if luma is not None else None
        opts = HashOptions(dihedral_mode, dct, layout, C.cast(coeff, C.POINTER(C.c_float)) if coeff else None)
        _check(lib().yume_pdq_amd_hash_batch_ex(
            C.c_void_p(planes.data_ptr()), kind, stride * elem, n, C.byref(opts),
            C.c_void_p(hashes.data_ptr()), C.c_void_p(quality.data_ptr()),
            C.c_void_p(threshold.data_ptr()), pd,
            C.c_void_p(buf1.data_ptr() if buf1 is not None else 0), _stream_ptr(stream)), "hash_batch_ex")
    res = (hashes, quality, threshold)
    if want_pdqf:
        res = res + (pdqf,)
    if want_buf1:
        res = res + (buf1,)
    return res


def dihedrals_batch(pdqf, thr_in, dihedral_mode=DIHEDRAL_AVX2, stream=None):
    """visit_dihedrals on device PDQFs (n,16,16) with thresholds (n,): returns
    (hashes (n,8,32), quality (n,8), threshold (n,8), pdqfs (n,8,16,16))."""
    torch = _torch()
    pdqf = pdqf.contiguous().to(torch.float32)
    thr_in = thr_in.contiguous().to(torch.float32)
    n = pdqf.shape[0]
    dev = pdqf.device
    hashes = torch.empty((n, 8, 32), dtype=torch.uint8, device=dev)
    quality = torch.empty((n, 8), dtype=torch.float32, device=dev)
    threshold = torch.empty((n, 8), dtype=torch.float32, device=dev)
    pds = torch.empty((n, 8, 16, 16), dtype=torch.float32, device=dev)
    _check(lib().yume_pdq_amd_dihedrals_batch(
        C.c_void_p(pdqf.data_ptr()), C.c_void_p(thr_in.data_ptr()), n, dihedral_mode,
        C.c_void_p(hashes.data_ptr()), C.c_void_p(quality.data_ptr()),
        C.c_void_p(threshold.data_ptr()), C.c_void_p(pds.data_ptr()), _stream_ptr(stream)),
        "dihedrals_batch")
    return hashes, quality, threshold, pds


HIT_DTYPE = np.dtype([("query", "<u4"), ("distance", "<u4"), ("index", "<u8")])


def build_prefix_plane(db, stream=None):
    """The database's prefix plane (a u8 device tensor) for match(..., plane=)."""
    torch = _torch()
    db = db.reshape(-1, 32)
    if not db.is_contiguous():
        raise ValueError("db must be contiguous")
    n = db.shape[0]
    plane = torch.empty(int(lib().yume_pdq_amd_prefix_plane_bytes(n)), dtype=torch.uint8, device=db.device)
    _check(lib().yume_pdq_amd_build_prefix_plane(C.c_void_p(db.data_ptr()), n, C.c_void_p(plane.data_ptr()),
                                                 _stream_ptr(stream)), "build_prefix_plane")
    return plane


def match(db, queries, threshold, capacity=1 << 16, index_base=0, stream=None, hits=None,
          count=None, append=False, sync=True, path=None, plane=None, noalloc=False):
    """All (query, index, distance) with popcount(query ^ row) <= threshold.

    db: CUDA u8 tensor (n_rows, 32); queries: CUDA u8 tensor (n_queries, 32);
    plane: the prefix plane of db (build_prefix_plane) or None; noalloc: run
    yume_pdq_amd_match_noalloc (the library allocates nothing; the caller
    still passes hits and count to avoid torch allocations).
    Returns (total_count, hits ndarray of HIT_DTYPE sorted by (index, query))
    when sync, else the device (hits, count) tensors."""
    torch = _torch()
    db = db.reshape(-1, 32)
    queries = queries.reshape(-1, 32)
    if not (db.is_cuda and queries.is_cuda) or db.dtype != torch.uint8 or queries.dtype != torch.uint8:
        raise ValueError("db and queries must be CUDA uint8 tensors of 32-byte rows")
    db = db.contiguous()
    queries = queries.contiguous()
    dev = db.device
    if hits is None:
        hits = torch.empty((max(capacity, 1), 16), dtype=torch.uint8, device=dev)
    if count is None:
        count = torch.zeros(1, dtype=torch.int64, device=dev)
    capacity = hits.shape[0]
    flags = (MATCH_APPEND if append else 0) | {None: 0, "valu": MATCH_VALU, "mfma": MATCH_MFMA}[path]
    if noalloc:
        _check(lib().yume_pdq_amd_match_noalloc(
            C.c_void_p(db.data_ptr()), C.c_void_p(plane.data_ptr() if plane is not None else 0), db.shape[0],
            C.c_void_p(queries.data_ptr()), queries.shape[0], int(threshold), C.c_void_p(hits.data_ptr()),
            capacity, C.c_void_p(count.data_ptr()), index_base, flags, _stream_ptr(stream)), "match_noalloc")
    elif plane is None:
        _check(lib().yume_pdq_amd_match(
            C.c_void_p(db.data_ptr()), db.shape[0], C.c_void_p(queries.data_ptr()), queries.shape[0],
            int(threshold), C.c_void_p(hits.data_ptr()), capacity, C.c_void_p(count.data_ptr()),
            index_base, flags, _stream_ptr(stream)), "match")
    else:
        _check(lib().yume_pdq_amd_match_with_plane(
            C.c_void_p(db.data_ptr()), C.c_void_p(plane.data_ptr()), db.shape[0],
            C.c_void_p(queries.data_ptr()), queries.shape[0], int(threshold), C.c_void_p(hits.data_ptr()),
            capacity, C.c_void_p(count.data_ptr()), index_base, flags, _stream_ptr(stream)), "match")
    if not sync:
        return hits, count
    if stream is not None:
        stream.synchronize()  # the reads below run on torch's current stream
    total = int(count.item())
    k = min(total, capacity)
    arr = hits[:k].cpu().numpy().copy().view(HIT_DTYPE).reshape(-1)
    arr = np.sort(arr, order=("index", "query"))
    return total, arr


class HashMatch:
    """Hash and match one batch shape repeatedly with one native call per batch
    (yume_pdq_amd_hash_match_batch): n u8 or f32 planes (n, 512, 512) resident
    on the GPU, their 8n dihedral hashes matched against db (n_rows, 32) --
    over its prefix plane when given -- at `threshold`.  Every buffer is
    allocated here once; run() enqueues the batch on `stream` and, unless
    sync is False, waits for it and returns the hit count, which the GPU
    stores into page-locked host memory.  Outputs: .hashes (n, 8, 32),
    .quality / .threshold (n, 8), .hits (capacity, 16) u8 records of
    HIT_DTYPE, .count (1,) int64 on the device."""

    def __init__(self, planes, db, threshold, plane=None, capacity=1 << 16, dihedral_mode=DIHEDRAL_AVX2,
                 stream=None):
        torch = _torch()
        if not planes.is_cuda or planes.dim() != 3 or tuple(planes.shape[1:]) != (512, 512):
            raise ValueError("planes must be a CUDA tensor of shape (n, 512, 512)")
        if planes.dtype == torch.uint8:
            kind, elem = INPUT_U8, 1
        elif planes.dtype == torch.float32:
            kind, elem = INPUT_F32, 4
        else:
            raise ValueError("planes must be uint8 or float32")
        self.planes = planes.contiguous()
        self.db = db.reshape(-1, 32).contiguous()
        if not self.db.is_cuda or self.db.dtype != torch.uint8:
            raise ValueError("db must be a CUDA uint8 tensor of 32-byte rows")
        self.plane = plane
        n, dev = self.planes.shape[0], self.planes.device
        self.hashes = torch.empty((n, 8, 32), dtype=torch.uint8, device=dev)
        self.quality = torch.empty((n, 8), dtype=torch.float32, device=dev)
        self.threshold = torch.empty((n, 8), dtype=torch.float32, device=dev)
        self.hits = torch.empty((max(capacity, 1), 16), dtype=torch.uint8, device=dev)
        self.count = torch.zeros(1, dtype=torch.int64, device=dev)
        self.h_count = torch.zeros(1, dtype=torch.int64).pin_memory()
        self.stream = stream if stream is not None else torch.cuda.current_stream(dev)
        self._args = (C.c_void_p(self.planes.data_ptr()), kind, 512 * 512 * elem, n, int(dihedral_mode),
                      C.c_void_p(self.db.data_ptr()), C.c_void_p(plane.data_ptr() if plane is not None else 0),
                      self.db.shape[0], int(threshold), C.c_void_p(self.hashes.data_ptr()),
                      C.c_void_p(self.quality.data_ptr()), C.c_void_p(self.threshold.data_ptr()),
                      C.c_void_p(self.hits.data_ptr()), self.hits.shape[0], C.c_void_p(self.count.data_ptr()),
                      C.c_void_p(self.h_count.data_ptr()), C.c_void_p(self.stream.cuda_stream))
        self._fn = lib().yume_pdq_amd_hash_match_batch

    def run(self, sync=True):
        rc = self._fn(*self._args)
        if rc < 0:
            _check(rc, "hash_match_batch")
        if not sync:
            return None
        self.stream.synchronize()
        return int(self.h_count[0])

    def results(self):
        """(total, hits sorted by (index, query)) of the last run, after it
        has finished."""
        total = int(self.h_count[0])
        k = min(total, self.hits.shape[0])
        arr = self.hits[:k].cpu().numpy().copy().view(HIT_DTYPE).reshape(-1)
        return total, np.sort(arr, order=("index", "query"))


def hash_match(planes, db, threshold, plane=None, capacity=1 << 16, dihedral_mode=DIHEDRAL_AVX2, stream=None):
    """One batch through yume_pdq_amd_hash_match_batch: returns (hashes (n,8,32),
    quality (n,8), threshold (n,8), total, hits sorted by (index, query))."""
    hm = HashMatch(planes, db, threshold, plane=plane, capacity=capacity, dihedral_mode=dihedral_mode,
                   stream=stream)
    hm.run()
    total, arr = hm.results()
    return hm.hashes, hm.quality, hm.threshold, total, arr


def set_match_record_cap(cap):
    """Lower the matrix-core matcher's candidate-record cap (1 .. 2**20; 0 =
    the default): calls past it still return the exact set, through the
    per-chunk recheck.  For tests and measurement."""
    _check(lib().yume_pdq_amd_match_set_record_cap(int(cap)), "match_set_record_cap")


def match_uses_matrix_cores(n_queries, threshold, path=None):
    """True when match() runs the matrix-core kernel for these arguments."""
    flags = {None: 0, "valu": MATCH_VALU, "mfma": MATCH_MFMA}[path]
    rc = lib().yume_pdq_amd_match_path(int(n_queries), int(threshold), flags)
    _check(rc, "match_path")
    return rc == MATCH_MFMA


def match_first(db, queries, threshold, stream=None):
    """(row, query) of the first hit in (row, query) order, or None."""
    torch = _torch()
    db = db.reshape(-1, 32).contiguous()
    queries = queries.reshape(-1, 32).contiguous()
    first = torch.empty(1, dtype=torch.int64, device=db.device)
    _check(lib().yume_pdq_amd_match_first(
        C.c_void_p(db.data_ptr()), db.shape[0], C.c_void_p(queries.data_ptr()), queries.shape[0],
        int(threshold), C.c_void_p(first.data_ptr()), _stream_ptr(stream)), "match_first")
    if stream is not None:
        stream.synchronize()  # the read below runs on torch's current stream
    key = int(first.item()) & 0xFFFFFFFFFFFFFFFF
    if key == 0xFFFFFFFFFFFFFFFF:
        return None
    nq = queries.shape[0]
    return key // nq, key % nq


# ---------------------------------------------------------------------------
# Staged API: one launch per method of the reference's Kernel trait
# (src/kernel/mod.rs:268-497), batched over n images of device tensors.
# ---------------------------------------------------------------------------
T_TRANSPOSE = 0
T_NEG_COLS_EVEN_AVX2 = 1
T_NEG_COLS_ODD_AVX2 = 2
T_NEG_COLS_EVEN = 3
T_NEG_COLS_ODD = 4
T_NEG_ROWS_ODD = 5
T_NEG_ROWS_EVEN = 6
T_NEG_OFF_DIAGONALS = 7

# BT.601 luma coefficients (src/kernel/constants.rs:23).
RGB8_TO_LUMA8_TABLE_ITU = (0.299, 0.587, 0.114)


def _vp(t):
    return C.c_void_p(t.data_ptr() if t is not None else 0)


def stage_compress(planes, stream=None):
    """Kernel::jarosz_compress on (n,512,512) f32/u8 planes -> (n,128,128) f32."""
    torch = _torch()
    if planes.dim() == 2:
        planes = planes.unsqueeze(0)
    planes = planes.contiguous()
    kind, elem = (INPUT_F32, 4) if planes.dtype == torch.float32 else (INPUT_U8, 1)
    if planes.dtype not in (torch.float32, torch.uint8) or planes.shape[-2:] != (512, 512):
        raise ValueError("planes must be (n, 512, 512) float32 or uint8")
    n = planes.shape[0]
    buf1 = torch.empty((n, 128, 128), dtype=torch.float32, device=planes.device)
    _check(lib().yume_pdq_amd_stage_compress(_vp(planes), kind, 512 * 512 * elem, n, _vp(buf1),
                                             _stream_ptr(stream)), "stage_compress")
    return buf1


def stage_dct2d(buf1, stream=None):
    """Kernel::dct2d on (n,128,128) -> (tmp (n,128), pdqf (n,16,16))."""
    torch = _torch()
    buf1 = buf1.reshape(-1, 128, 128).contiguous()
    n = buf1.shape[0]
    tmp = torch.empty((n, 128), dtype=torch.float32, device=buf1.device)
    pdqf = torch.empty((n, 16, 16), dtype=torch.float32, device=buf1.device)
    _check(lib().yume_pdq_amd_stage_dct2d(_vp(buf1), n, _vp(tmp), _vp(pdqf), _stream_ptr(stream)),
           "stage_dct2d")
    return tmp, pdqf


def stage_sum_of_gradients(pdqf, stream=None):
    torch = _torch()
    pdqf = pdqf.reshape(-1, 16, 16).contiguous()
    out = torch.empty(pdqf.shape[0], dtype=torch.float32, device=pdqf.device)
    _check(lib().yume_pdq_amd_stage_sum_of_gradients(_vp(pdqf), pdqf.shape[0], _vp(out),
                                                     _stream_ptr(stream)), "stage_sum_of_gradients")
    return out


def stage_quantize(pdqf, stream=None):
    """Kernel::quantize: (n,16,16) -> (threshold (n,), hash (n,32))."""
    torch = _torch()
    pdqf = pdqf.reshape(-1, 16, 16).contiguous()
    n = pdqf.shape[0]
    thr = torch.empty(n, dtype=torch.float32, device=pdqf.device)
    h = torch.empty((n, 32), dtype=torch.uint8, device=pdqf.device)
    _check(lib().yume_pdq_amd_stage_quantize(_vp(pdqf), n, _vp(thr), _vp(h), _stream_ptr(stream)),
           "stage_quantize")
    return thr, h


def stage_threshold(pdqf, thr, stream=None):
    """threshold_2d_f32 with one threshold per image: -> hash (n,32)."""
    torch = _torch()
    pdqf = pdqf.reshape(-1, 16, 16).contiguous()
    thr = thr.reshape(-1).contiguous().to(torch.float32)
    n = pdqf.shape[0]
    h = torch.empty((n, 32), dtype=torch.uint8, device=pdqf.device)
    _check(lib().yume_pdq_amd_stage_threshold(_vp(pdqf), _vp(thr), n, _vp(h), _stream_ptr(stream)),
           "stage_threshold")
    return h


def stage_transform(pdqf, op, stream=None):
    """In-place PDQF transform (T_* op) of a contiguous (n,16,16) f32 tensor."""
    if not pdqf.is_contiguous():
        raise ValueError("stage_transform works in place on a contiguous tensor")
    _check(lib().yume_pdq_amd_stage_transform(_vp(pdqf), pdqf.numel() // 256, int(op), _stream_ptr(stream)),
           "stage_transform")
    return pdqf


def stage_luma(pixels, coeff=RGB8_TO_LUMA8_TABLE_ITU, stream=None):
    """RGB8 / RGBA8 pixels (..., 3 or 4) u8 -> luma f32 (...)."""
    torch = _torch()
    ch = pixels.shape[-1]
    pixels = pixels.contiguous()
    out = torch.empty(pixels.shape[:-1], dtype=torch.float32, device=pixels.device)
    _check(lib().yume_pdq_amd_stage_luma(_vp(pixels), ch, out.numel(), float(coeff[0]), float(coeff[1]),
                                         float(coeff[2]), _vp(out), _stream_ptr(stream)), "stage_luma")
    return out


# ---------------------------------------------------------------------------
# Batching layer: shard sets and batch contexts (include/yume_pdq_amd.h)
# ---------------------------------------------------------------------------
class _DeviceArray:
    """A device allocation owned by the library, seen by torch through the
    CUDA array interface (no copy)."""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr,
                                         "data": (int(ptr), False), "version": 3, "strides": None}


def _alias(ptr, shape, dtype):
    torch = _torch()
    typestr = {torch.uint8: "|u1", torch.float32: "<f4", torch.int64: "<i8"}[dtype]
    return torch.as_tensor(_DeviceArray(ptr, shape, typestr), device="cuda")


def _hits_to_numpy(hits_u8, total, capacity):
    k = min(int(total), capacity)
    arr = hits_u8[:k].cpu().numpy().copy().view(HIT_DTYPE).reshape(-1)
    return np.sort(arr, order=("index", "query"))


class Shards:
    """Database shards resident in HBM (rows + prefix plane each)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib().yume_pdq_amd_shards_create(C.byref(h)), "shards_create")
        self._h = h
        self._next_base = 0
        self._keep = []  # device tensors referenced by shards

    def close(self):
        if self._h:
            lib().yume_pdq_amd_shards_destroy(self._h)
            self._h = None
            self._keep = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, rows, index_base=None):
        """rows: host numpy (n, 32) u8 (copied into HBM) or CUDA u8 tensor
        (referenced; kept alive by this object).  Returns the shard id."""
        torch = _torch()
        base = self._next_base if index_base is None else int(index_base)
        sid = C.c_uint32(0)
        if isinstance(rows, np.ndarray):
            r = np.ascontiguousarray(rows, np.uint8).reshape(-1, 32)
            _check(lib().yume_pdq_amd_shards_add(self._h, r.ctypes.data_as(C.c_void_p), r.shape[0], base, 0,
                                                 C.byref(sid)), "shards_add")
            n = r.shape[0]
        else:
            r = rows.reshape(-1, 32).contiguous()
            if r.dtype != torch.uint8 or not r.is_cuda:
                raise ValueError("device rows must be a CUDA uint8 tensor")
            torch.cuda.synchronize()
            _check(lib().yume_pdq_amd_shards_add(self._h, C.c_void_p(r.data_ptr()), r.shape[0], base, 1,
                                                 C.byref(sid)), "shards_add")
            self._keep.append(r)
            n = r.shape[0]
        self._next_base = max(self._next_base, base + n)
        return sid.value

    def load_raw(self, path, rows_per_shard, index_base=0):
        n = C.c_uint32(0)
        _check(lib().yume_pdq_amd_shards_load_raw(self._h, str(path).encode(), int(rows_per_shard),
                                                  int(index_base), C.byref(n)), "shards_load_raw")
        return n.value

    def replace(self, shard, local_rows, new_rows):
        """Rows local_rows (CUDA int64) of `shard` become new_rows (CUDA u8 (n, 32));
        returns the number of out-of-range indices skipped."""
        torch = _torch()
        idx = local_rows.reshape(-1).to(torch.int64).contiguous()
        new = new_rows.reshape(-1, 32).contiguous()
        bad = torch.zeros(1, dtype=torch.int64, device=new.device)
        _check(lib().yume_pdq_amd_shards_replace(self._h, int(shard), C.c_void_p(idx.data_ptr()),
                                                 C.c_void_p(new.data_ptr()), idx.numel(),
                                                 C.c_void_p(bad.data_ptr()), _stream_ptr(None)),
               "shards_replace")
        return int(bad.item())

    def info(self):
        n, rows, b = C.c_uint32(0), C.c_uint64(0), C.c_uint64(0)
        _check(lib().yume_pdq_amd_shards_info(self._h, C.byref(n), C.byref(rows), C.byref(b)), "shards_info")
        return {"shards": n.value, "rows": rows.value, "device_bytes": b.value}

    def match(self, queries, threshold, capacity=1 << 16, path=None):
        """Exhaustive match against every shard (one persistent launch):
        (total, hits sorted by (index, query)) with global indices."""
        torch = _torch()
        q = queries.reshape(-1, 32).contiguous()
        hits = torch.empty((max(capacity, 1), 16), dtype=torch.uint8, device=q.device)
        count = torch.zeros(1, dtype=torch.int64, device=q.device)
        flags = {None: 0, "valu": MATCH_VALU, "mfma": MATCH_MFMA}[path]
        _check(lib().yume_pdq_amd_shards_match(self._h, C.c_void_p(q.data_ptr()), q.shape[0], int(threshold),
                                               C.c_void_p(hits.data_ptr()), hits.shape[0],
                                               C.c_void_p(count.data_ptr()), flags, _stream_ptr(None)),
               "shards_match")
        total = int(count.item())
        return total, _hits_to_numpy(hits, total, hits.shape[0])


class Batcher:
    """Batch context: hash up to max_images planes and match their 8n hashes
    against every shard of `shards`, replayed as one HIP graph per batch."""

    def __init__(self, shards, max_images, input_kind=INPUT_U8, dihedral_mode=DIHEDRAL_AVX2, threshold=31,
                 hit_capacity=1 << 20):
        h = C.c_void_p()
        _check(lib().yume_pdq_amd_batcher_create(shards._h, int(max_images), int(input_kind), int(dihedral_mode),
                                                 int(threshold), int(hit_capacity), C.byref(h)), "batcher_create")
        self._h = h
        self._shards = shards  # keep the set alive
        b = BatchBuffers()
        _check(lib().yume_pdq_amd_batcher_buffers(h, C.byref(b)), "batcher_buffers")
        self.buffers = b
        torch = _torch()
        self.max_images = b.max_images
        elem = torch.float32 if input_kind == INPUT_F32 else torch.uint8
        self.input = _alias(b.d_input, (b.max_images, 512, 512), elem)
        self.hashes = _alias(b.d_hashes, (b.max_images, 8, 32), torch.uint8)
        self.quality = _alias(b.d_quality, (b.max_images, 8), torch.float32)
        self.threshold = _alias(b.d_threshold, (b.max_images, 8), torch.float32)
        self.hits = _alias(b.d_hits, (b.hit_capacity, 16), torch.uint8)
        self.count = _alias(b.d_count, (1,), torch.int64)
        self.stream = torch.cuda.ExternalStream(b.stream)

    def close(self):
        if self._h:
            self.stream.synchronize()
            lib().yume_pdq_amd_batcher_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, planes, use_graph=True, sync=True):
        """Copy n planes (CUDA tensor (n, 512, 512)) in and run one batch.
        With sync, returns (hashes (n,8,32), quality, threshold, total, hits)."""
        torch = _torch()
        n = planes.shape[0]
        if n > self.max_images:
            raise ValueError("batch larger than max_images")
        torch.cuda.current_stream().synchronize()  # planes written on torch's stream
        with torch.cuda.stream(self.stream):
            self.input[:n].copy_(planes)
        _check(lib().yume_pdq_amd_batcher_run(self._h, n, 1 if use_graph else 0), "batcher_run")
        if not sync:
            return None
        return self.results(n)

    def results(self, n):
        torch = _torch()
        with torch.cuda.stream(self.stream):
            h = self.hashes[:n].clone()
            q = self.quality[:n].clone()
            t = self.threshold[:n].clone()
            c = self.count.clone()
        self.stream.synchronize()
        total = int(c.item())
        return h, q, t, total, _hits_to_numpy(self.hits, total, self.hits.shape[0])

    def autotune(self, iterations=5):
        _check(lib().yume_pdq_amd_batcher_autotune(self._h, int(iterations)), "batcher_autotune")
        n = C.c_uint32(0)
        _check(lib().yume_pdq_amd_batcher_tune_records(self._h, None, 0, C.byref(n)), "tune_records")
        recs = (TuneRecord * max(n.value, 1))()
        _check(lib().yume_pdq_amd_batcher_tune_records(self._h, recs, n.value, C.byref(n)), "tune_records")
        return [{"n_queries": r.n_queries, "path": "mfma" if r.path == MATCH_MFMA else "valu",
                 "query_tiles": r.query_tiles, "ms": r.ms} for r in recs[:n.value]]

    def choice(self, n_queries):
        path, qt = C.c_uint32(0), C.c_uint32(0)
        _check(lib().yume_pdq_amd_batcher_choice(self._h, int(n_queries), C.byref(path), C.byref(qt)),
               "batcher_choice")
        return ("mfma" if path.value == MATCH_MFMA else "valu"), qt.value


def tune_save(path):
    """Write the process-wide matcher tuning cache (autotuner choices per GPU,
    threshold and shard-set shape) to a text file; returns the entries."""
    n = lib().yume_pdq_amd_tune_save(os.fsencode(path))
    if n < 0:
        raise YumePdqError(f"tune_save: {lib().yume_pdq_amd_strerror(n).decode()} ({n})")
    return n


def tune_load(path):
    """Merge a file written by tune_save into the tuning cache; returns the
    entries read (a malformed file loads nothing and raises)."""
    n = lib().yume_pdq_amd_tune_load(os.fsencode(path))
    if n < 0:
        raise YumePdqError(f"tune_load: {lib().yume_pdq_amd_strerror(n).decode()} ({n})")
    return n


def tune_clear():
    """Drop every cached tuning entry; returns how many there were."""
    return lib().yume_pdq_amd_tune_clear()


class Server:
    """Continuous-batching server over a shard set: submit() u8 planes from
    any thread, wait() for each request's hashes, qualities, thresholds and
    hits (queries numbered 8 * image + dihedral within the request).

    scan_chunk_rows > 0 selects the shared circular scan instead of batches:
    the shard set is matched chunk by chunk, round and round, against every
    active request's hashes; a request joins at the next chunk and is done
    after one full turn. Results are the same in both modes."""

    def __init__(self, shards, max_batch_images=64, hbm_budget_bytes=8 << 30, max_contexts=4, threshold=31,
                 dihedral_mode=DIHEDRAL_AVX2, hit_capacity=1 << 20, use_graph=True, busy_poll=True,
                 in_order_compute=True, autotune_iterations=0, scan_chunk_rows=0):
        cfg = ServerConfig(int(max_batch_images), int(hbm_budget_bytes), int(max_contexts), int(threshold),
                           int(dihedral_mode), int(hit_capacity), 1 if use_graph else 0, 1 if busy_poll else 0,
                           1 if in_order_compute else 0, int(autotune_iterations), int(scan_chunk_rows))
        h = C.c_void_p()
        _check(lib().yume_pdq_amd_server_create(shards._h, C.byref(cfg), C.byref(h)), "server_create")
        self._h = h
        self._shards = shards
        self.max_batch_images = int(max_batch_images)

    def close(self):
        if self._h:
            lib().yume_pdq_amd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, planes):
        p = np.ascontiguousarray(planes, np.uint8).reshape(-1, 512, 512)
        rid = C.c_uint64(0)
        _check(lib().yume_pdq_amd_server_submit(self._h, p.ctypes.data_as(C.c_void_p), p.shape[0], C.byref(rid)),
               "server_submit")
        return rid.value

    def submit_pinned(self, planes):
        """Zero-copy submit of a pinned CPU torch tensor (n, 512, 512) u8 that
        stays unchanged until the request's result has been waited for."""
        if not planes.is_pinned() or not planes.is_contiguous() or planes.dtype != _torch().uint8:
            raise ValueError("planes must be a contiguous pinned uint8 CPU tensor")
        n = planes.numel() // (512 * 512)
        rid = C.c_uint64(0)
        _check(lib().yume_pdq_amd_server_submit_pinned(self._h, C.c_void_p(planes.data_ptr()), n, C.byref(rid)),
               "server_submit_pinned")
        return rid.value

    def wait(self, rid, timeout_ms=-1):
        r = ServerResult()
        _check(lib().yume_pdq_amd_server_wait(self._h, int(rid), int(timeout_ms), C.byref(r)), "server_wait")
        try:
            n = r.n_images

            def arr(ptr, count, dtype):
                if count == 0:
                    return np.zeros(0, dtype)
                buf = (C.c_uint8 * (count * np.dtype(dtype).itemsize)).from_address(ptr)
                return np.frombuffer(buf, dtype=dtype).copy()

            out = {
                "status": r.status,
                "hashes": arr(r.hashes, n * 256, np.uint8).reshape(n, 8, 32),
                "quality": arr(r.quality, n * 8, np.float32).reshape(n, 8),
                "threshold": arr(r.threshold, n * 8, np.float32).reshape(n, 8),
                "hits": arr(r.hits, int(r.n_hits), HIT_DTYPE) if r.n_hits else np.zeros(0, HIT_DTYPE),
                "latency_ms": r.latency_ms,
                "batch": r.batch,
                "batch_images": r.batch_images,
            }
        finally:
            lib().yume_pdq_amd_server_release(self._h, int(rid))
        return out

    def stats(self):
        st = ServerStats()
        _check(lib().yume_pdq_amd_server_stats(self._h, C.byref(st)), "server_stats")
        return {"contexts": st.contexts, "device_bytes": st.device_bytes, "batches": st.batches, "images": st.images}


# ---------------------------------------------------------------------------
# Stream ingestion (the reference's `pipe` subcommand) and its output formats
# ---------------------------------------------------------------------------
PIPE_INPUT_FORMATS = {"luma8": INPUT_U8, "LUMA8": INPUT_U8, "rgb8": INPUT_RGB8, "RGB8": INPUT_RGB8,
                      "rgba8": INPUT_RGBA8, "RGBA8": INPUT_RGBA8}


def format_frames(fmt, hashes, quality, threshold, dihedrals=False, dihedral_pdqf=None):
    """Records of hashed frames in a pipe output format (host-only): hashes
    (n, 8, 32) u8, quality and threshold (n, 8) f32 as hash_batch returns them,
    dihedral_pdqf (n, 8, 16, 16) f32 for the diagnostic format.  Returns bytes."""
    h = np.ascontiguousarray(hashes, np.uint8).reshape(-1, 8, 32)
    q = np.ascontiguousarray(quality, np.float32).reshape(-1, 8)
    t = np.ascontiguousarray(threshold, np.float32).reshape(-1, 8)
    n = h.shape[0]
    p = None if dihedral_pdqf is None else np.ascontiguousarray(dihedral_pdqf, np.float32).reshape(n, 8, 256)
    f = fmt.encode()
    cap = int(lib().yume_pdq_amd_format_bound(f, n, int(bool(dihedrals))))
    if cap == 0 and n > 0:
        raise YumePdqError(f"invalid output format: {fmt!r}")
    buf = C.create_string_buffer(max(cap, 1))
    got = _check(lib().yume_pdq_amd_format_frames(
        f, n, int(bool(dihedrals)), h.ctypes.data, q.ctypes.data, t.ctypes.data,
        p.ctypes.data if p is not None else None, buf, cap), "format_frames")
    return buf.raw[:got]


def pipe(source, dest, fmt="bin", input_format="luma8", dihedrals=False, batch=64, stats=False,
         dihedral_mode=DIHEDRAL_AVX2):
    """Hash every whole 512x512 frame of `source` on the GPU and write the
    records to `dest` (paths or open file descriptors), as the reference's
    `yume-pdq pipe -f fmt [--dihedrals] --input-format ...`.  Returns the
    number of frames; raises YumePdqError (EEOF after writing every whole
    frame if the input ends inside a frame)."""
    if input_format not in PIPE_INPUT_FORMATS:
        raise ValueError(f"invalid input format: {input_format!r}")
    own = []
    try:
        if isinstance(source, int):
            fd_in = source
        else:
            fd_in = os.open(source, os.O_RDONLY)
            own.append(fd_in)
        if isinstance(dest, int):
            fd_out = dest
        else:
            fd_out = os.open(dest, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
            own.append(fd_out)
        cfg = PipeConfig(fd_in, fd_out, PIPE_INPUT_FORMATS[input_format], fmt.encode(), int(bool(dihedrals)),
                         dihedral_mode, batch, int(bool(stats)))
        frames = C.c_uint64(0)
        rc = lib().yume_pdq_amd_pipe_run(C.byref(cfg), C.byref(frames))
        if rc < 0:
            raise YumePdqError(f"pipe: {lib().yume_pdq_amd_strerror(rc).decode()} ({rc}) after "
                               f"{frames.value} frames")
        return frames.value
    finally:
        for fd in own:
            os.close(fd)


def pipe_trim():
    """Free the buffer sets finished pipe runs keep for the next run of the
    same shape; returns how many were freed."""
    return lib().yume_pdq_amd_pipe_trim()


def hash_ref_batch(planes, want_pdqf=False, stream=None):
    """The precise ReferenceKernel<f32> (Jarosz box filter resampling) on n
    512x512 f32 or u8 planes resident on the GPU: NORMAL orientation only.
    Returns (hashes (n,32) u8, quality (n,) f32, threshold (n,) f32[, pdqf (n,16,16)])."""
    torch = _torch()
    if planes.dim() == 2:
        planes = planes.unsqueeze(0)
    if not planes.is_cuda or planes.shape[1:] != (512, 512):
        raise ValueError("planes must be a CUDA tensor of shape (n, 512, 512)")
    if planes.dtype == torch.float32:
        kind, elem = INPUT_F32, 4
    elif planes.dtype == torch.uint8:
        kind, elem = INPUT_U8, 1
    else:
        raise ValueError("planes must be float32 or uint8")
    planes = planes.contiguous()
    n = planes.shape[0]
    dev = planes.device
    ws_bytes = int(lib().yume_pdq_amd_hash_ref_workspace_bytes(n))
    ws = torch.empty(max(ws_bytes // 4, 1), dtype=torch.float32, device=dev)
    hashes = torch.empty((n, 32), dtype=torch.uint8, device=dev)
    quality = torch.empty(n, dtype=torch.float32, device=dev)
    threshold = torch.empty(n, dtype=torch.float32, device=dev)
    pdqf = torch.empty((n, 16, 16), dtype=torch.float32, device=dev) if want_pdqf else None
    _check(lib().yume_pdq_amd_hash_ref_batch(
        C.c_void_p(planes.data_ptr()), kind, 512 * 512 * elem, n, C.c_void_p(ws.data_ptr()), ws_bytes,
        C.c_void_p(hashes.data_ptr()), C.c_void_p(quality.data_ptr()), C.c_void_p(threshold.data_ptr()),
        C.c_void_p(pdqf.data_ptr() if pdqf is not None else 0), _stream_ptr(stream)), "hash_ref_batch")
    if want_pdqf:
        return hashes, quality, threshold, pdqf
    return hashes, quality, threshold
