// C ABI of libyume_pdq.so: the drop-in entry points of the reference's
// include/yume_pdq.h and the batched device API of include/yume_pdq_amd.h.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/yume_pdq.h"
#include "../../include/yume_pdq_amd.h"
#include "pdq_device.h"
#include "pipe.h"

// ---------------------------------------------------------------------------
// Data symbols of the reference ABI.
// ---------------------------------------------------------------------------
// Declared outside any extern "C" block, as in the reference header; global
// variables are not name-mangled, so C and C++ callers see the same symbols.
const uint32_t YUME_PDQ_VERSION_MAJOR = 1;
const uint32_t YUME_PDQ_VERSION_MINOR = 1;
const uint32_t YUME_PDQ_VERSION_PATCH = 0;
const char YUME_PDQ_VERSION_STRING[] = "1.1.0";

extern "C" {

// Dihedrals::from_tuples packs (dx0, dx1, dy0, dy1) big-endian (src/lib.rs:96-133).
#define YPDQ_PACK(a, b, c, d) \
  ((uint32_t)(uint8_t)(a) << 24 | (uint32_t)(uint8_t)(b) << 16 | (uint32_t)(uint8_t)(c) << 8 | (uint32_t)(uint8_t)(d))
const struct YumePDQ_Dihedrals YUME_PDQ_DIHEDRAL_NORMAL = {YPDQ_PACK(1, 0, 0, 1)};
const struct YumePDQ_Dihedrals YUME_PDQ_DIHEDRAL_FLIPPED = {YPDQ_PACK(1, 0, 0, -1)};
const struct YumePDQ_Dihedrals YUME_PDQ_DIHEDRAL_FLOPPED = {YPDQ_PACK(-1, 0, 0, 1)};
const struct YumePDQ_Dihedrals YUME_PDQ_DIHEDRAL_ROTATED_180 = {YPDQ_PACK(-1, 0, 0, -1)};
const struct YumePDQ_Dihedrals YUME_PDQ_DIHEDRAL_ROTATED_90 = {YPDQ_PACK(0, 1, -1, 0)};
const struct YumePDQ_Dihedrals YUME_PDQ_DIHEDRAL_ROTATED_270 = {YPDQ_PACK(0, 1, 1, 0)};
const struct YumePDQ_Dihedrals YUME_PDQ_DIHEDRAL_ROTATED_90_FLOPPED = {YPDQ_PACK(0, -1, -1, 0)};
const struct YumePDQ_Dihedrals YUME_PDQ_DIHEDRAL_FLOPPED_ROTATED_270 = {YPDQ_PACK(0, -1, 1, 0)};
#undef YPDQ_PACK
}

namespace {

using ypdq::HashOutputs;

// Result block of one drop-in call.
struct DropInResult {
  uint8_t hashes[8 * 32];
  float quality[8];
  float threshold[8];
  float pdqf[256];
  float tmp[128];
  float buf1[128 * 128];
  float dihedral_pdqf[8 * 256];
};

// Per-thread state for the drop-in entry points (the reference clones its
// kernel per call, src/lib.rs:266, so calls are reentrant).  A hash call
// copies the image into page-locked h_img with memcpy, in two halves; a copy
// kernel per half on the call's stream pulls it into d_img (a DMA-engine copy followed by a kernel
// costs a cross-engine dependency of a few tens of microseconds); the hash
// kernel writes its results straight into page-locked h_res over the bus: no
// copy back, one synchronisation per call.  The visit call's 1 KiB PDQF is
// read by its kernel from h_pdqf directly.
struct DropInState {
  hipStream_t stream = nullptr;
  float* h_img = nullptr;         // 512 x 512 f32, page-locked, coherent
  float* d_img = nullptr;         // its device copy
  float* h_pdqf = nullptr;        // 256 PDQF + 1 threshold, page-locked, coherent
  DropInResult* h_res = nullptr;  // the kernels' outputs, page-locked, coherent
  float* dv_img = nullptr;        // device addresses of the page-locked three
  float* dv_pdqf = nullptr;
  DropInResult* dv_res = nullptr;
};

typedef float pull_f4 __attribute__((ext_vector_type(4)));

// dst[i] = src[i] for n 16-byte elements, src in page-locked host memory.
__global__ void __launch_bounds__(256) pull_plane(pull_f4* __restrict__ dst, const pull_f4* __restrict__ src,
                                                  uint32_t n) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    dst[i] = __builtin_nontemporal_load(src + i);
}

std::mutex g_pool_mu;
std::vector<DropInState*> g_pool;

[[noreturn]] void die(const char* what, hipError_t e) {
  fprintf(stderr, "libyume_pdq: %s failed: %s\n", what, hipGetErrorString(e));
  abort();
}

DropInState* state_new() {
  DropInState* s = new DropInState();
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess) die("hipStreamCreate", e);
  if ((e = hipMalloc(&s->d_img, 512 * 512 * sizeof(float))) != hipSuccess) die("hipMalloc", e);
  // coherent (fine-grained): the GPU does not cache these lines, so a kernel
  // reads the bytes just written by memcpy and its stores are in host memory
  // when the stream synchronisation returns
  if ((e = hipHostMalloc(&s->h_img, 512 * 512 * sizeof(float), hipHostMallocCoherent)) != hipSuccess)
    die("hipHostMalloc", e);
  if ((e = hipHostMalloc(&s->h_pdqf, 257 * sizeof(float), hipHostMallocCoherent)) != hipSuccess)
    die("hipHostMalloc", e);
  if ((e = hipHostMalloc(&s->h_res, sizeof(DropInResult), hipHostMallocCoherent)) != hipSuccess)
    die("hipHostMalloc", e);
  if ((e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->dv_img), s->h_img, 0)) != hipSuccess)
    die("hipHostGetDevicePointer", e);
  if ((e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->dv_pdqf), s->h_pdqf, 0)) != hipSuccess)
    die("hipHostGetDevicePointer", e);
  if ((e = hipHostGetDevicePointer(reinterpret_cast<void**>(&s->dv_res), s->h_res, 0)) != hipSuccess)
    die("hipHostGetDevicePointer", e);
  return s;
}

// Returns the state to a process-wide pool when its thread exits (no HIP calls
// in thread teardown); states are reused by later threads.
struct StateHolder {
  DropInState* s = nullptr;
  ~StateHolder() {
    if (s) {
      std::lock_guard<std::mutex> g(g_pool_mu);
      g_pool.push_back(s);
    }
  }
};

DropInState* state() {
  static thread_local StateHolder h;
  if (!h.s) {
    {
      std::lock_guard<std::mutex> g(g_pool_mu);
      if (!g_pool.empty()) {
        h.s = g_pool.back();
        g_pool.pop_back();
      }
    }
    if (!h.s) h.s = state_new();
  }
  return h.s;
}

HashOutputs outputs_of(DropInResult* r, bool with_buf1) {
  HashOutputs o;
  o.hashes = r->hashes;
  o.quality = r->quality;
  o.threshold = r->threshold;
  o.pdqf = r->pdqf;
  o.buf1 = with_buf1 ? r->buf1 : nullptr;
  o.tmp = with_buf1 ? r->tmp : nullptr;
  o.dihedral_pdqf = r->dihedral_pdqf;
  return o;
}

int from_hip(hipError_t e) {
  if (e == hipSuccess) return YUME_PDQ_AMD_OK;
  if (e == hipErrorOutOfMemory) return YUME_PDQ_AMD_ENOMEM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return YUME_PDQ_AMD_ENODEV;
  return YUME_PDQ_AMD_EHIP;
}

__global__ void fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }
// the hit count into page-locked host memory (its device address)
__global__ void store_count(const unsigned long long* d, unsigned long long* h) { *h = *d; }

}  // namespace

extern "C" {

float yume_pdq_hash_smart_kernel(const float* input, float* threshold, uint8_t* output, float* buf1, float* tmp,
                                 float* pdqf) {
  DropInState* s = state();
  hipError_t e;
  // in two halves: the device pulls the first while the host copies the
  // second (77.7 against 95.6 us per call in one piece, same box)
  constexpr int kHalves = 2;
  constexpr uint32_t kVec = 512 * 512 / 4 / kHalves;  // 16-byte elements per half
  for (int h = 0; h < kHalves; h++) {
    memcpy(s->h_img + h * kVec * 4, input + h * kVec * 4, kVec * sizeof(pull_f4));
    hipLaunchKernelGGL(pull_plane, dim3(64), dim3(256), 0, s->stream,
                       reinterpret_cast<pull_f4*>(s->d_img) + h * kVec,
                       reinterpret_cast<const pull_f4*>(s->dv_img) + h * kVec, kVec);
    if ((e = hipGetLastError()) != hipSuccess) die("copy kernel launch", e);
  }
  HashOutputs o = outputs_of(s->dv_res, true);
  o.dihedral_pdqf = nullptr;
  if ((e = ypdq::launch_hash(s->d_img, ypdq::kInputF32, 512 * 512 * sizeof(float), 1, ypdq::kDihedralAvx2, o,
                             s->stream)) != hipSuccess)
    die("hash kernel launch", e);
  if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) die("hipStreamSynchronize", e);
  const DropInResult* r = s->h_res;
  memcpy(output, r->hashes, 32);
  if (threshold) *threshold = r->threshold[0];
  memcpy(buf1, r->buf1, sizeof r->buf1);
  memcpy(tmp, r->tmp, sizeof r->tmp);
  memcpy(pdqf, r->pdqf, sizeof r->pdqf);
  return r->quality[0];
}

bool yume_pdq_visit_dihedrals_smart_kernel(void* ctx, float threshold, uint8_t* output, float* pdqf,
                                           DihedralCallback callback) {
  DropInState* s = state();
  hipError_t e;
  memcpy(s->h_pdqf, pdqf, 256 * sizeof(float));
  s->h_pdqf[256] = threshold;
  HashOutputs o = outputs_of(s->dv_res, false);
  o.pdqf = nullptr;
  if ((e = ypdq::launch_dihedrals(s->dv_pdqf, s->dv_pdqf + 256, 1, ypdq::kDihedralAvx2, o, s->stream)) != hipSuccess)
    die("dihedral kernel launch", e);
  if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) die("hipStreamSynchronize", e);
  const DropInResult* r = s->h_res;
  // orientations 1..7 are the reference's visiting order (src/lib.rs:389-418)
  static const uint32_t codes[8] = {
      0, YUME_PDQ_DIHEDRAL_FLOPPED.packed, YUME_PDQ_DIHEDRAL_ROTATED_180.packed, YUME_PDQ_DIHEDRAL_FLIPPED.packed,
      YUME_PDQ_DIHEDRAL_ROTATED_90_FLOPPED.packed, YUME_PDQ_DIHEDRAL_ROTATED_270.packed,
      YUME_PDQ_DIHEDRAL_FLOPPED_ROTATED_270.packed, YUME_PDQ_DIHEDRAL_ROTATED_90.packed};
  for (int d = 1; d < 8; d++) {
    memcpy(output, r->hashes + 32 * d, 32);
    memcpy(pdqf, r->dihedral_pdqf + 256 * d, 256 * sizeof(float));
    if (!callback(ctx, codes[d], threshold, r->quality[d])) return false;
  }
  return true;
}

const char* yume_pdq_amd_strerror(int code) {
  switch (code) {
    case YUME_PDQ_AMD_OK: return "ok";
    case YUME_PDQ_AMD_EINVAL: return "invalid argument";
    case YUME_PDQ_AMD_EHIP: return "HIP runtime error";
    case YUME_PDQ_AMD_ENOMEM: return "device out of memory";
    case YUME_PDQ_AMD_ENODEV: return "no usable GPU";
    case YUME_PDQ_AMD_EOVERFLOW: return "hit buffer overflow";
    case YUME_PDQ_AMD_ETIMEOUT: return "timed out";
    case YUME_PDQ_AMD_EIO: return "stream read or write failed";
    case YUME_PDQ_AMD_EEOF: return "input ended inside a frame";
    case YUME_PDQ_AMD_ESCRATCH:
      return "the call needs more candidate records than the no-allocation matcher's static scratch holds";
    default: return "unknown error";
  }
}

int yume_pdq_amd_device_info(int device, char* name, size_t name_len, int* compute_units) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return YUME_PDQ_AMD_ENODEV;
  if (device < 0 || device >= n) return YUME_PDQ_AMD_EINVAL;
  hipDeviceProp_t p;
  hipError_t e = hipGetDeviceProperties(&p, device);
  if (e != hipSuccess) return from_hip(e);
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (compute_units) *compute_units = p.multiProcessorCount;
  return n;
}

int yume_pdq_amd_hash_batch_luma(const void* d_planes, int input_kind, uint64_t plane_stride_bytes, uint32_t n_images,
                                 uint32_t dihedral_mode, const float* luma_coeff, uint8_t* d_hashes, float* d_quality,
                                 float* d_threshold, float* d_pdqf, void* stream) {
  if (n_images == 0) return YUME_PDQ_AMD_OK;
  if (!d_planes || !d_hashes || !d_quality || !d_threshold) return YUME_PDQ_AMD_EINVAL;
  const uint64_t plane_bytes = ypdq::input_plane_bytes(input_kind);
  if (plane_bytes == 0) return YUME_PDQ_AMD_EINVAL;
  if (dihedral_mode > YUME_PDQ_AMD_DIHEDRAL_AVX2) return YUME_PDQ_AMD_EINVAL;
  if (plane_stride_bytes < plane_bytes || plane_stride_bytes % 16 || (uintptr_t)d_planes % 16)
    return YUME_PDQ_AMD_EINVAL;
  HashOutputs o;
  o.hashes = d_hashes;
  o.quality = d_quality;
  o.threshold = d_threshold;
  o.pdqf = d_pdqf;
  o.buf1 = nullptr;
  o.tmp = nullptr;
  o.dihedral_pdqf = nullptr;
  return from_hip(ypdq::launch_hash(d_planes, input_kind, plane_stride_bytes, n_images, dihedral_mode, o,
                                    static_cast<hipStream_t>(stream), luma_coeff));
}

int yume_pdq_amd_hash_batch(const void* d_planes, int input_kind, uint64_t plane_stride_bytes, uint32_t n_images,
                            uint32_t dihedral_mode, uint8_t* d_hashes, float* d_quality, float* d_threshold,
                            float* d_pdqf, void* stream) {
  if (n_images == 0) return YUME_PDQ_AMD_OK;
  if (!d_planes || !d_hashes || !d_quality || !d_threshold) return YUME_PDQ_AMD_EINVAL;
  const uint64_t plane_bytes = ypdq::input_plane_bytes(input_kind);
  if (plane_bytes == 0) return YUME_PDQ_AMD_EINVAL;
  if (dihedral_mode > YUME_PDQ_AMD_DIHEDRAL_AVX2) return YUME_PDQ_AMD_EINVAL;
  if (plane_stride_bytes < plane_bytes || plane_stride_bytes % 16 || (uintptr_t)d_planes % 16)
    return YUME_PDQ_AMD_EINVAL;
  HashOutputs o;
  o.hashes = d_hashes;
  o.quality = d_quality;
  o.threshold = d_threshold;
  o.pdqf = d_pdqf;
  o.buf1 = nullptr;
  o.tmp = nullptr;
  o.dihedral_pdqf = nullptr;
  return from_hip(ypdq::launch_hash(d_planes, input_kind, plane_stride_bytes, n_images, dihedral_mode, o,
                                    static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_hash_batch_ex(const void* d_planes, int input_kind, uint64_t plane_stride_bytes, uint32_t n_images,
                               const YumePdqAmdHashOptions* opts, uint8_t* d_hashes, float* d_quality,
                               float* d_threshold, float* d_pdqf, float* d_buf1, void* stream) {
  const YumePdqAmdHashOptions def = {YUME_PDQ_AMD_DIHEDRAL_AVX2, YUME_PDQ_AMD_DCT_EXACT, YUME_PDQ_AMD_LAYOUT_AUTO,
                                     nullptr};
  const YumePdqAmdHashOptions& o = opts ? *opts : def;
  if (n_images == 0) return YUME_PDQ_AMD_OK;
  if (!d_planes || !d_hashes || !d_quality || !d_threshold) return YUME_PDQ_AMD_EINVAL;
  const uint64_t plane_bytes = ypdq::input_plane_bytes(input_kind);
  if (plane_bytes == 0) return YUME_PDQ_AMD_EINVAL;
  if (o.dihedral_mode > YUME_PDQ_AMD_DIHEDRAL_AVX2) return YUME_PDQ_AMD_EINVAL;
  if (o.dct_precision > YUME_PDQ_AMD_DCT_F16X3 || o.layout > YUME_PDQ_AMD_LAYOUT_BLOCK16) return YUME_PDQ_AMD_EINVAL;
  if ((o.layout == YUME_PDQ_AMD_LAYOUT_WAVE || o.layout == YUME_PDQ_AMD_LAYOUT_BLOCK16) &&
      o.dct_precision != YUME_PDQ_AMD_DCT_EXACT)
    return YUME_PDQ_AMD_EINVAL;
  if (plane_stride_bytes < plane_bytes || plane_stride_bytes % 16 || (uintptr_t)d_planes % 16)
    return YUME_PDQ_AMD_EINVAL;
  HashOutputs out;
  out.hashes = d_hashes;
  out.quality = d_quality;
  out.threshold = d_threshold;
  out.pdqf = d_pdqf;
  out.buf1 = d_buf1;
  out.tmp = nullptr;
  out.dihedral_pdqf = nullptr;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  int layout = (int)o.layout;
  if (layout == YUME_PDQ_AMD_LAYOUT_AUTO) {
    const int a = ypdq::hash_layout_for(n_images, (int)o.dct_precision);
    layout = a == 1 ? YUME_PDQ_AMD_LAYOUT_WAVE : a == 16 ? YUME_PDQ_AMD_LAYOUT_BLOCK16 : YUME_PDQ_AMD_LAYOUT_BLOCK8;
  }
  if (layout == YUME_PDQ_AMD_LAYOUT_WAVE)
    return from_hip(ypdq::launch_hash_wave(d_planes, input_kind, plane_stride_bytes, n_images, nullptr,
                                           o.dihedral_mode, out, o.luma_coeff, s));
  const int waves = layout == YUME_PDQ_AMD_LAYOUT_BLOCK16 ? 16 : layout == YUME_PDQ_AMD_LAYOUT_BLOCK8 ? 8 : 4;
  return from_hip(ypdq::launch_hash_block(d_planes, input_kind, plane_stride_bytes, n_images, nullptr,
                                          o.dihedral_mode, (int)o.dct_precision, waves, out, o.luma_coeff, s));
}

uint64_t yume_pdq_amd_hash_ref_workspace_bytes(uint32_t n_images) { return ypdq::hash_ref_workspace_bytes(n_images); }

int yume_pdq_amd_hash_ref_batch(const void* d_planes, int input_kind, uint64_t plane_stride_bytes, uint32_t n_images,
                                void* d_workspace, uint64_t workspace_bytes, uint8_t* d_hashes, float* d_quality,
                                float* d_threshold, float* d_pdqf, void* stream) {
  if (n_images == 0) return YUME_PDQ_AMD_OK;
  if (!d_planes || !d_workspace || !d_hashes || !d_quality || !d_threshold) return YUME_PDQ_AMD_EINVAL;
  if (input_kind != YUME_PDQ_AMD_INPUT_F32 && input_kind != YUME_PDQ_AMD_INPUT_U8) return YUME_PDQ_AMD_EINVAL;
  const uint64_t plane_bytes = ypdq::input_plane_bytes(input_kind);
  if (plane_stride_bytes < plane_bytes || plane_stride_bytes % 16 || (uintptr_t)d_planes % 16)
    return YUME_PDQ_AMD_EINVAL;
  if (workspace_bytes < ypdq::hash_ref_workspace_bytes(n_images) || (uintptr_t)d_workspace % 16)
    return YUME_PDQ_AMD_EINVAL;
  return from_hip(ypdq::launch_hash_ref(d_planes, input_kind, plane_stride_bytes, n_images, d_workspace, d_hashes,
                                        d_quality, d_threshold, d_pdqf, static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_dihedrals_batch(const float* d_pdqf, const float* d_thr_in, uint32_t n, uint32_t dihedral_mode,
                                 uint8_t* d_hashes, float* d_quality, float* d_threshold, float* d_dihedral_pdqf,
                                 void* stream) {
  if (n == 0) return YUME_PDQ_AMD_OK;
  if (!d_pdqf || !d_thr_in || !d_hashes || !d_quality || !d_threshold) return YUME_PDQ_AMD_EINVAL;
  if (dihedral_mode > YUME_PDQ_AMD_DIHEDRAL_AVX2) return YUME_PDQ_AMD_EINVAL;
  HashOutputs o;
  o.hashes = d_hashes;
  o.quality = d_quality;
  o.threshold = d_threshold;
  o.pdqf = nullptr;
  o.buf1 = nullptr;
  o.tmp = nullptr;
  o.dihedral_pdqf = d_dihedral_pdqf;
  return from_hip(ypdq::launch_dihedrals(d_pdqf, d_thr_in, n, dihedral_mode, o, static_cast<hipStream_t>(stream)));
}

static int match_impl(const uint8_t* d_db, const void* d_plane, uint64_t n_rows, const uint8_t* d_queries,
                      uint32_t n_queries, uint32_t threshold, YumePdqAmdHit* d_hits, uint64_t capacity,
                      unsigned long long* d_count, uint64_t index_base, uint32_t flags, void* stream,
                      bool noalloc = false) {
  if (!d_count || (capacity && !d_hits)) return YUME_PDQ_AMD_EINVAL;
  if ((uintptr_t)d_plane % 16) return YUME_PDQ_AMD_EINVAL;
  if ((n_rows && !d_db) || (n_queries && !d_queries)) return YUME_PDQ_AMD_EINVAL;
  if ((uintptr_t)d_db % 16 || (uintptr_t)d_queries % 4) return YUME_PDQ_AMD_EINVAL;
  const bool force_valu = flags & YUME_PDQ_AMD_MATCH_VALU, force_mfma = flags & YUME_PDQ_AMD_MATCH_MFMA;
  if (force_valu && force_mfma) return YUME_PDQ_AMD_EINVAL;
  if (force_mfma && threshold >= 64) return YUME_PDQ_AMD_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  static_assert(sizeof(YumePdqAmdHit) == sizeof(ypdq::MatchHit), "hit layout");
  ypdq::MatchHit* hits = reinterpret_cast<ypdq::MatchHit*>(d_hits);
  const bool reset = !(flags & YUME_PDQ_AMD_MATCH_APPEND);
  const bool mfma = force_mfma || (!force_valu && ypdq::match_mfma_applicable(n_queries, threshold));
  // the matrix-core matcher clears the counter in its own first kernel
  if (mfma && noalloc) {
    if (ypdq::match_mfma_scratch_records(n_rows, n_queries) > ypdq::match_scratch_records())
      return YUME_PDQ_AMD_ESCRATCH;
    return from_hip(ypdq::launch_match_mfma_noalloc(d_db, d_plane, n_rows, d_queries, n_queries, threshold, hits,
                                                    capacity, d_count, index_base, st, reset));
  }
  if (mfma)
    return from_hip(ypdq::launch_match_mfma(d_db, d_plane, n_rows, d_queries, n_queries, threshold, hits, capacity,
                                            d_count, index_base, st, reset));
  if (reset) {
    // a kernel, not hipMemsetAsync, so the sequence can be graph-captured
    hipLaunchKernelGGL(fill_u64, dim3(1), dim3(1), 0, st, d_count, 0ull);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return from_hip(e);
  }
  if (d_plane)
    return from_hip(ypdq::launch_match_plane(d_db, d_plane, n_rows, d_queries, n_queries, threshold, hits, capacity,
                                             d_count, index_base, st));
  return from_hip(
      ypdq::launch_match(d_db, n_rows, d_queries, n_queries, threshold, hits, capacity, d_count, index_base, st));
}

int yume_pdq_amd_match(const uint8_t* d_db, uint64_t n_rows, const uint8_t* d_queries, uint32_t n_queries,
                       uint32_t threshold, YumePdqAmdHit* d_hits, uint64_t capacity, unsigned long long* d_count,
                       uint64_t index_base, uint32_t flags, void* stream) {
  return match_impl(d_db, nullptr, n_rows, d_queries, n_queries, threshold, d_hits, capacity, d_count, index_base,
                    flags, stream);
}

int yume_pdq_amd_match_with_plane(const uint8_t* d_db, const void* d_plane, uint64_t n_rows, const uint8_t* d_queries,
                                  uint32_t n_queries, uint32_t threshold, YumePdqAmdHit* d_hits, uint64_t capacity,
                                  unsigned long long* d_count, uint64_t index_base, uint32_t flags, void* stream) {
  if (!d_plane && n_rows) return YUME_PDQ_AMD_EINVAL;
  return match_impl(d_db, d_plane, n_rows, d_queries, n_queries, threshold, d_hits, capacity, d_count, index_base,
                    flags, stream);
}

int yume_pdq_amd_match_noalloc(const uint8_t* d_db, const void* d_plane, uint64_t n_rows, const uint8_t* d_queries,
                               uint32_t n_queries, uint32_t threshold, YumePdqAmdHit* d_hits, uint64_t capacity,
                               unsigned long long* d_count, uint64_t index_base, uint32_t flags, void* stream) {
  return match_impl(d_db, d_plane, n_rows, d_queries, n_queries, threshold, d_hits, capacity, d_count, index_base,
                    flags, stream, true);
}

int yume_pdq_amd_match_set_record_cap(uint32_t cap) {
  return ypdq::set_match_record_cap(cap) ? YUME_PDQ_AMD_OK : YUME_PDQ_AMD_EINVAL;
}

int yume_pdq_amd_hash_match_batch(const void* d_planes, int input_kind, uint64_t plane_stride_bytes,
                                  uint32_t n_images, uint32_t dihedral_mode, const uint8_t* d_db,
                                  const void* d_plane, uint64_t n_rows, uint32_t threshold, uint8_t* d_hashes,
                                  float* d_quality, float* d_threshold, YumePdqAmdHit* d_hits, uint64_t capacity,
                                  unsigned long long* d_count, unsigned long long* h_count, void* stream) {
  // every argument first, so that an error enqueues nothing
  if (n_images > UINT32_MAX / 8) return YUME_PDQ_AMD_EINVAL;
  if (n_images && (!d_planes || !d_hashes || !d_quality || !d_threshold)) return YUME_PDQ_AMD_EINVAL;
  const uint64_t plane_bytes = ypdq::input_plane_bytes(input_kind);
  if (plane_bytes == 0 || dihedral_mode > YUME_PDQ_AMD_DIHEDRAL_AVX2) return YUME_PDQ_AMD_EINVAL;
  if (n_images && (plane_stride_bytes < plane_bytes || plane_stride_bytes % 16 || (uintptr_t)d_planes % 16))
    return YUME_PDQ_AMD_EINVAL;
  if (!d_count || (capacity && !d_hits) || (n_rows && !d_db)) return YUME_PDQ_AMD_EINVAL;
  if ((uintptr_t)d_db % 16 || (uintptr_t)d_plane % 16) return YUME_PDQ_AMD_EINVAL;
  unsigned long long* d_h_count = nullptr;
  if (h_count) {
    // page-locked host memory from hipHostMalloc only: the GPU stores into it
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, h_count) != hipSuccess || a.type != hipMemoryTypeHost ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&d_h_count), h_count, 0) != hipSuccess || !d_h_count) {
      (void)hipGetLastError();  // the failed query must not surface in a later call
      return YUME_PDQ_AMD_EINVAL;
    }
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  int rc = yume_pdq_amd_hash_batch(d_planes, input_kind, plane_stride_bytes, n_images, dihedral_mode, d_hashes,
                                   d_quality, d_threshold, nullptr, stream);
  if (rc == YUME_PDQ_AMD_OK)
    rc = match_impl(d_db, d_plane, n_rows, d_hashes, 8 * n_images, threshold, d_hits, capacity, d_count, 0, 0,
                    stream);
  if (rc == YUME_PDQ_AMD_OK && d_h_count) {
    hipLaunchKernelGGL(store_count, dim3(1), dim3(1), 0, st, d_count, d_h_count);
    rc = from_hip(hipGetLastError());
  }
  return rc;
}

uint64_t yume_pdq_amd_prefix_plane_bytes(uint64_t n_rows) { return ypdq::prefix_plane_bytes(n_rows); }

int yume_pdq_amd_build_prefix_plane(const uint8_t* d_db, uint64_t n_rows, void* d_plane, void* stream) {
  if (n_rows == 0) return YUME_PDQ_AMD_OK;
  if (!d_db || !d_plane || (uintptr_t)d_db % 16 || (uintptr_t)d_plane % 16) return YUME_PDQ_AMD_EINVAL;
  return from_hip(ypdq::launch_build_prefix_plane(d_db, n_rows, d_plane, static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_match_path(uint32_t n_queries, uint32_t threshold, uint32_t flags) {
  const bool force_valu = flags & YUME_PDQ_AMD_MATCH_VALU, force_mfma = flags & YUME_PDQ_AMD_MATCH_MFMA;
  if ((force_valu && force_mfma) || (force_mfma && threshold >= 64)) return YUME_PDQ_AMD_EINVAL;
  if (force_mfma || (!force_valu && ypdq::match_mfma_applicable(n_queries, threshold)))
    return YUME_PDQ_AMD_MATCH_MFMA;
  return YUME_PDQ_AMD_MATCH_VALU;
}

int yume_pdq_amd_match_first(const uint8_t* d_db, uint64_t n_rows, const uint8_t* d_queries, uint32_t n_queries,
                             uint32_t threshold, unsigned long long* d_first, void* stream) {
  if (!d_first || (n_rows && !d_db) || (n_queries && !d_queries)) return YUME_PDQ_AMD_EINVAL;
  if ((uintptr_t)d_db % 16 || (uintptr_t)d_queries % 4) return YUME_PDQ_AMD_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(1), 0, st, d_first, ~0ull);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return from_hip(e);
  return from_hip(ypdq::launch_match_first(d_db, n_rows, d_queries, n_queries, threshold, d_first, st));
}

int yume_pdq_amd_stage_compress(const void* d_planes, int input_kind, uint64_t plane_stride_bytes, uint32_t n,
                                float* d_buf1, void* stream) {
  if (n == 0) return YUME_PDQ_AMD_OK;
  if (!d_planes || !d_buf1) return YUME_PDQ_AMD_EINVAL;
  if (input_kind != YUME_PDQ_AMD_INPUT_F32 && input_kind != YUME_PDQ_AMD_INPUT_U8) return YUME_PDQ_AMD_EINVAL;
  const uint64_t plane_bytes = 512ull * 512 * (input_kind == YUME_PDQ_AMD_INPUT_F32 ? 4 : 1);
  if (plane_stride_bytes < plane_bytes || plane_stride_bytes % 4) return YUME_PDQ_AMD_EINVAL;
  return from_hip(ypdq::launch_stage_compress(d_planes, input_kind, plane_stride_bytes, n, d_buf1,
                                              static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_stage_dct2d(const float* d_buf1, uint32_t n, float* d_tmp, float* d_pdqf, void* stream) {
  if (n == 0) return YUME_PDQ_AMD_OK;
  if (!d_buf1 || !d_pdqf) return YUME_PDQ_AMD_EINVAL;
  return from_hip(ypdq::launch_stage_dct2d(d_buf1, n, d_tmp, d_pdqf, static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_stage_sum_of_gradients(const float* d_pdqf, uint32_t n, float* d_out, void* stream) {
  if (n == 0) return YUME_PDQ_AMD_OK;
  if (!d_pdqf || !d_out || (uintptr_t)d_pdqf % 16) return YUME_PDQ_AMD_EINVAL;
  return from_hip(ypdq::launch_stage_gradient(d_pdqf, n, d_out, static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_stage_quantize(const float* d_pdqf, uint32_t n, float* d_threshold, uint8_t* d_hash, void* stream) {
  if (n == 0) return YUME_PDQ_AMD_OK;
  if (!d_pdqf || !d_threshold || !d_hash) return YUME_PDQ_AMD_EINVAL;
  return from_hip(ypdq::launch_stage_quantize(d_pdqf, n, d_threshold, d_hash, static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_stage_threshold(const float* d_pdqf, const float* d_threshold, uint32_t n, uint8_t* d_hash,
                                 void* stream) {
  if (n == 0) return YUME_PDQ_AMD_OK;
  if (!d_pdqf || !d_threshold || !d_hash) return YUME_PDQ_AMD_EINVAL;
  return from_hip(ypdq::launch_stage_threshold(d_pdqf, d_threshold, n, d_hash, static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_stage_transform(float* d_pdqf, uint32_t n, int op, void* stream) {
  if (n == 0) return YUME_PDQ_AMD_OK;
  if (!d_pdqf || op < 0 || op > 7) return YUME_PDQ_AMD_EINVAL;
  return from_hip(ypdq::launch_stage_transform(d_pdqf, n, op, static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_stage_luma(const uint8_t* d_pixels, int channels, uint64_t n_pixels, float r_coeff, float g_coeff,
                            float b_coeff, float* d_out, void* stream) {
  if (n_pixels == 0) return YUME_PDQ_AMD_OK;
  if (!d_pixels || !d_out || (channels != 3 && channels != 4)) return YUME_PDQ_AMD_EINVAL;
  const float c[3] = {r_coeff, g_coeff, b_coeff};
  return from_hip(ypdq::launch_stage_luma(d_pixels, channels, n_pixels, c, d_out, static_cast<hipStream_t>(stream)));
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Stream ingestion and pipe output formats
// ---------------------------------------------------------------------------
int yume_pdq_amd_pipe_run(const YumePdqAmdPipeConfig* cfg, uint64_t* frames) {
  if (frames) *frames = 0;
  if (!cfg) return YUME_PDQ_AMD_EINVAL;
  ypdq::PipeOptions o;
  if (!ypdq::parse_output_format(cfg->output_format, &o.format)) return YUME_PDQ_AMD_EINVAL;
  if (cfg->input_kind != YUME_PDQ_AMD_INPUT_U8 && cfg->input_kind != YUME_PDQ_AMD_INPUT_RGB8 &&
      cfg->input_kind != YUME_PDQ_AMD_INPUT_RGBA8)
    return YUME_PDQ_AMD_EINVAL;
  if (cfg->dihedral_mode > YUME_PDQ_AMD_DIHEDRAL_AVX2 || cfg->input_fd < 0 || cfg->output_fd < 0)
    return YUME_PDQ_AMD_EINVAL;
  o.input_fd = cfg->input_fd;
  o.output_fd = cfg->output_fd;
  o.input_kind = cfg->input_kind;
  o.dihedrals = cfg->dihedrals != 0;
  o.dihedral_mode = cfg->dihedral_mode;
  o.batch = cfg->batch_frames ? cfg->batch_frames : 64;
  if (o.batch > 4096) return YUME_PDQ_AMD_EINVAL;
  o.stats = cfg->stats != 0;
  std::string err;
  const int rc = ypdq::run_pipe(o, frames, &err);
  if (rc == 0) return YUME_PDQ_AMD_OK;
  if (!err.empty()) fprintf(stderr, "yume_pdq_amd_pipe_run: %s\n", err.c_str());
  if (rc == ypdq::kPipeEio) return YUME_PDQ_AMD_EIO;
  if (rc == ypdq::kPipeEof) return YUME_PDQ_AMD_EEOF;
  if (rc == -1) return YUME_PDQ_AMD_EINVAL;
  return YUME_PDQ_AMD_EHIP;
}

int yume_pdq_amd_pipe_trim(void) { return (int)ypdq::trim_pipe_cache(); }

int64_t yume_pdq_amd_format_frames(const char* output_format, uint32_t n_frames, int dihedrals, const uint8_t* hashes,
                                   const float* quality, const float* threshold, const float* dihedral_pdqf, char* out,
                                   uint64_t out_capacity) {
  ypdq::OutputFormat f;
  if (!ypdq::parse_output_format(output_format, &f)) return YUME_PDQ_AMD_EINVAL;
  if (n_frames == 0) return 0;
  if (!hashes || !quality || !threshold || (!out && out_capacity)) return YUME_PDQ_AMD_EINVAL;
  if (f.type == ypdq::OutputFormat::kDiagnostic && !dihedral_pdqf) return YUME_PDQ_AMD_EINVAL;
  std::string text;
  ypdq::format_frames(f, n_frames, dihedrals != 0, hashes, quality, threshold, dihedral_pdqf, text);
  if (text.size() > out_capacity) return YUME_PDQ_AMD_EOVERFLOW;
  if (!text.empty()) memcpy(out, text.data(), text.size());
  return (int64_t)text.size();
}

uint64_t yume_pdq_amd_format_bound(const char* output_format, uint32_t n_frames, int dihedrals) {
  ypdq::OutputFormat f;
  if (!ypdq::parse_output_format(output_format, &f)) return 0;
  return ypdq::format_bound(f, n_frames, dihedrals != 0);
}
