// Fused PDQ-variant hash kernel for gfx950 (CDNA4): one wavefront per image.
//
// Computes, per 512x512 grayscale plane (f32 or u8), the 256-bit hash of the
// reference's optimised kernel (x86::Avx2F32Kernel, src/kernel/x86.rs) and the
// seven other dihedral hashes (src/lib.rs:358-420), with the quality and the
// threshold of each.  The arithmetic reproduces the reference's FP32 rounding
// sequence exactly: same FMAs, same two-accumulator alternation, same
// reduction trees, same iterative median search.  Nothing is approximated, so
// the hash bits equal the reference's for every finite input.
//
// Data flow of one wavefront (64 lanes, lane t owns input columns 8t..8t+7):
//   1. stream input rows 1..510 once, coalesced (2 KiB per f32 row, 512 B per
//      u8 row); per compressed row i (0..126) evaluate the 5x5 tent
//      (x86.rs:626-682) for output columns 2t and 2t+1 and immediately fold
//      them into DCT stage 1 (x86.rs:161-225): 16 frequencies x 2 columns x 2
//      alternating accumulators stay in registers; the 127x127 plane is never
//      materialised.
//   2. stage 2 (x86.rs:227-272) from LDS, 4 outputs per lane.
//   3. the 8 dihedral PDQFs by sign flips / transpose, their gradient sums
//      (x86.rs:840-865), the AVX2 median search (x86.rs:440-621) for the four
//      untransposed ones and threshold_2d_f32 (threshold.rs:37-119) for the
//      transposed four, and bit packing.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "pdq_hash_common.h"
#include "pdq_hash_tail.h"

namespace ypdq {
// ---------------------------------------------------------------------------
// Compress + stage 1 over one plane.
// ---------------------------------------------------------------------------
// Rows in flight per wave: D compressed-row steps of four input rows each.
// u8 rows cost two registers per lane, so the u8 path keeps four steps in
// flight; f32 rows cost eight and keep one.
template <typename T>
struct Prefetch;
template <>
struct Prefetch<float> {
  static constexpr int D = 1;
};
template <>
struct Prefetch<uint8_t> {
  static constexpr int D = 4;
};
template <>
struct Prefetch<Rgb8> {
  static constexpr int D = 1;
};
template <>
struct Prefetch<Rgba8> {
  static constexpr int D = 1;
};

// One compressed row i (ring slot S = i mod D): rows 4i+3 .. 4i+6 arrive from
// pre[S], each slot is refilled with the row D steps ahead as soon as it has
// been converted; the chains of row i complete and those of row i+1 start, and
// the two compressed values go into stage-1 accumulator P.
template <int P, int S, int D, typename T, typename Raw>
__device__ __forceinline__ void stream_step(const T* __restrict__ plane, int lane, int i, float (&acc)[16][2][2],
                                            Chains& cur, Chains& nxt, Raw (&pre)[D][4], float* __restrict__ buf1,
                                            const Luma& lc) {
  // the rows D steps ahead, loaded unconditionally and clamped into the
  // plane (past row 510 they are not used): a load under a branch makes the
  // compiler copy the ring's registers and wait for every load in flight
  const int r0 = 4 * (i + D) + 3;
  auto row = [&](int d) { return min(r0 + d, 511); };
  if constexpr (std::is_same<T, float>::value) {
    // f32 rows feed from their ring slots, which reload after their last use
    chains_feed<2>(cur, pre[S][0]);
    load_raw(plane, row(0), lane, pre[S][0]);
    chains_feed<3>(cur, pre[S][1]);
    load_raw(plane, row(1), lane, pre[S][1]);
    chains_feed<4>(cur, pre[S][2]);
    chains_feed<5>(cur, pre[S][3]);
    chains_start(nxt, pre[S][2], pre[S][3]);
    load_raw(plane, row(2), lane, pre[S][2]);
    load_raw(plane, row(3), lane, pre[S][3]);
  } else {
    Row8 x, x2, x3;
    to_f32(pre[S][0], x, lc);
    load_raw(plane, row(0), lane, pre[S][0]);
    chains_feed<2>(cur, x);
    to_f32(pre[S][1], x, lc);
    load_raw(plane, row(1), lane, pre[S][1]);
    chains_feed<3>(cur, x);
    to_f32(pre[S][2], x2, lc);
    load_raw(plane, row(2), lane, pre[S][2]);
    chains_feed<4>(cur, x2);
    to_f32(pre[S][3], x3, lc);
    load_raw(plane, row(3), lane, pre[S][3]);
    chains_feed<5>(cur, x3);
    chains_start(nxt, x2, x3);
  }
  float c0, c1;
  chains_finish(cur, c0, c1);
  if (buf1) {
    buf1[i * 128 + 2 * lane] = c0;
    buf1[i * 128 + 2 * lane + 1] = (lane == 63) ? 0.0f : c1;
  }
  stage1_acc<P>(acc, c0, c1, kDctT.v + i * 16);
}

template <typename T, int D>
__device__ __forceinline__ void stream_image(const T* __restrict__ plane, int lane, float (&acc)[16][2][2],
                                             float* __restrict__ buf1, const Luma& lc) {
  using Raw = typename RawRow<T>::type;
  static_assert(4 % D == 0, "ring depth must divide the unroll factor");
#pragma unroll
  for (int k = 0; k < 16; k++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++) acc[k][jj][0] = acc[k][jj][1] = 0.0f;

  // two chain sets that trade roles from step to step (the next row's
  // chains are the current ones of the step after): no copies
  Chains a, b;
  Raw pre[D][4];
  {
    Raw r1, r2;
    load_raw(plane, 1, lane, r1);
    load_raw(plane, 2, lane, r2);
#pragma unroll
    for (int s = 0; s < D; s++)
#pragma unroll
      for (int d = 0; d < 4; d++) load_raw(plane, 4 * s + 3 + d, lane, pre[s][d]);
    Row8 x1, x2;
    to_f32(r1, x1, lc);
    to_f32(r2, x2, lc);
    chains_start(a, x1, x2);
  }
  // The accumulator a compressed row feeds is fixed by i mod 4 (alt), so the
  // loop is unrolled by four to keep every accumulator and ring index a
  // constant.
#pragma unroll 1
  for (int i = 0; i < 124; i += 4) {
    stream_step<1, 0 % D, D>(plane, lane, i, acc, a, b, pre, buf1, lc);
    stream_step<0, 1 % D, D>(plane, lane, i + 1, acc, b, a, pre, buf1, lc);
    stream_step<0, 2 % D, D>(plane, lane, i + 2, acc, a, b, pre, buf1, lc);
    stream_step<1, 3 % D, D>(plane, lane, i + 3, acc, b, a, pre, buf1, lc);
  }
  stream_step<1, 124 % D, D>(plane, lane, 124, acc, a, b, pre, buf1, lc);
  stream_step<0, 125 % D, D>(plane, lane, 125, acc, b, a, pre, buf1, lc);
  stream_step<0, 126 % D, D>(plane, lane, 126, acc, a, b, pre, buf1, lc);
}

// Stage-1 rows into LDS: tmp[k][j] = nan0(acc0) + nan0(acc1) (x86.rs:207-224);
// column 127 is the zero padding column.
__device__ __forceinline__ void write_stage1(const float (&acc)[16][2][2], WaveLds& W, int lane) {
#pragma unroll
  for (int k = 0; k < 16; k++) {
#pragma unroll
    for (int jj = 0; jj < 2; jj++) {
      float a0 = acc[k][jj][0], a1 = acc[k][jj][1];
      a0 = __builtin_isnan(a0) ? 0.0f : a0;
      a1 = __builtin_isnan(a1) ? 0.0f : a1;
      const float v = a0 + a1;
      const int j = 2 * lane + jj;
      W.t[k * kTS + j] = (j == 127) ? 0.0f : v;
    }
  }
}

template <int P, typename T>
__device__ __forceinline__ void literal_step(const T* __restrict__ plane, int lane, int i, int k0,
                                             float (&acc)[8][2][2], float* __restrict__ buf1, const Luma& lc) {
  const float c0 = compress_one_literal(plane, i, 2 * lane, lc);
  // output column 127 does not exist: lane 63 evaluates column 126 again (in
  // bounds) and discards it
  float c1 = compress_one_literal(plane, i, lane == 63 ? 126 : 2 * lane + 1, lc);
  if (lane == 63) c1 = 0.0f;
  if (buf1) {
    buf1[i * 128 + 2 * lane] = c0;
    buf1[i * 128 + 2 * lane + 1] = c1;
  }
  const float* dk = kDctT.v + i * 16 + k0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    acc[k][0][P] = __builtin_fmaf(c0, dk[k], acc[k][0][P]);
    acc[k][1][P] = __builtin_fmaf(c1, dk[k], acc[k][1][P]);
  }
}

// Literal (all-taps) recomputation for planes that hold NaN/Inf; writes the
// stage-1 rows to W.t.  Rarely taken, kept out of line.
template <typename T>
__device__ __noinline__ void stream_image_literal(const T* __restrict__ plane, int lane, WaveLds& W,
                                                  float* __restrict__ buf1, Luma lc) {
  // two passes of eight frequencies each keep the accumulators small; the
  // compressed values are recomputed identically in the second pass
#pragma unroll 1
  for (int k0 = 0; k0 < 16; k0 += 8) {
    float acc[8][2][2];
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
      for (int jj = 0; jj < 2; jj++) acc[k][jj][0] = acc[k][jj][1] = 0.0f;
#pragma unroll 1
    for (int i = 0; i < 124; i += 4) {
      literal_step<1>(plane, lane, i, k0, acc, buf1, lc);
      literal_step<0>(plane, lane, i + 1, k0, acc, buf1, lc);
      literal_step<0>(plane, lane, i + 2, k0, acc, buf1, lc);
      literal_step<1>(plane, lane, i + 3, k0, acc, buf1, lc);
    }
    literal_step<1>(plane, lane, 124, k0, acc, buf1, lc);
    literal_step<0>(plane, lane, 125, k0, acc, buf1, lc);
    literal_step<0>(plane, lane, 126, k0, acc, buf1, lc);
#pragma unroll
    for (int k = 0; k < 8; k++) {
#pragma unroll
      for (int jj = 0; jj < 2; jj++) {
        float a0 = acc[k][jj][0], a1 = acc[k][jj][1];
        a0 = __builtin_isnan(a0) ? 0.0f : a0;
        a1 = __builtin_isnan(a1) ? 0.0f : a1;
        const float v = a0 + a1;
        const int j = 2 * lane + jj;
        W.t[(k0 + k) * kTS + j] = (j == 127) ? 0.0f : v;
      }
    }
  }
}

// Does the plane hold a value the fast form would treat differently?  Every
// position of rows 2..510 x cols 2..510 carries a non-zero tap of some output,
// so a NaN/Inf there makes a stage-1 accumulator non-finite; the remaining
// positions the reference reads (rows 1 and 511, cols 0, 1 and 511) are
// checked directly.
__device__ __forceinline__ bool plane_needs_literal(const float* __restrict__ plane, int lane,
                                                    const float (&acc)[16][2][2]) {
  bool bad = false;
#pragma unroll
  for (int k = 0; k < 16; k++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++) bad = bad || nonfinite(acc[k][jj][0]) || nonfinite(acc[k][jj][1]);
  Row8 r;
  load_row(plane, 1, lane, r);
#pragma unroll
  for (int m = 0; m < 8; m++) bad |= nonfinite(r.v[m]);
  load_row(plane, 511, lane, r);
#pragma unroll
  for (int m = 0; m < 8; m++) bad |= nonfinite(r.v[m]);
  for (int row = 2 + lane; row <= 510; row += 64) {
    const float* pr = plane + (size_t)row * 512;
    bad = bad || nonfinite(pr[0]) || nonfinite(pr[1]) || nonfinite(pr[511]);
  }
  return wave_any(bad);
}

// u8 and colour inputs are finite by construction.
__device__ __forceinline__ bool plane_needs_literal(const uint8_t* __restrict__, int, const float (&)[16][2][2]) {
  return false;
}
__device__ __forceinline__ bool plane_needs_literal(const Rgb8* __restrict__, int, const float (&)[16][2][2]) {
  return false;
}
__device__ __forceinline__ bool plane_needs_literal(const Rgba8* __restrict__, int, const float (&)[16][2][2]) {
  return false;
}

template <typename T, int D, typename Lds>
__device__ __forceinline__ void hash_one(const T* __restrict__ planes, uint64_t plane_stride, uint32_t img, int lane,
                                         int wv, uint32_t mode, const HashOutputs& out, Lds& L, const Luma& lc) {
  const T* plane = reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(planes) + (size_t)img * plane_stride);
  float* buf1 = out.buf1 ? out.buf1 + (size_t)img * 128 * 128 : nullptr;
  WaveLds& W = L.w[wv];
  {
    float acc[16][2][2];
    stream_image<T, D>(plane, lane, acc, buf1, lc);
    const bool literal = plane_needs_literal(plane, lane, acc);
    write_stage1(acc, W, lane);
    if (literal) stream_image_literal(plane, lane, W, buf1, lc);
  }
  wave_lds_sync();
  if (buf1) {
    buf1[127 * 128 + 2 * lane] = 0.0f;
    buf1[127 * 128 + 2 * lane + 1] = 0.0f;
    buf1[(2 * lane) * 128 + 127] = 0.0f;
    buf1[(2 * lane + 1) * 128 + 127] = 0.0f;
  }
  stage2_to_pdqf(L.d2, W, lane, out, img);
  dihedral_tail(W, lane, mode, out, img, nullptr);
}

// One wave per image, kW images per block; the grid covers n_images.  With
// d_n set, the image count is read from device memory (capped at n_images)
// and the waves past it return at once, so one captured launch sized for the
// largest batch serves batches of any size.  Blocks of 4 waves run 3 per CU
// (46.7 KB of LDS and 168 registers each); blocks of 16 (161.5 KB) run one
// per CU, four waves per SIMD within 128 registers, which u8 rows reach with
// one prefetch step in flight (D = 1).
template <typename T, int kW, int D>
__global__ void __launch_bounds__(64 * kW) __attribute__((amdgpu_waves_per_eu(kW >= 16 ? 4 : 3))) pdq_hash_kernel(
    const T* __restrict__ planes, uint64_t plane_stride, uint32_t n_images, const uint32_t* __restrict__ d_n,
    uint32_t mode, HashOutputs out, Luma lc) {
  const uint32_t n = d_n ? min(*d_n, n_images) : n_images;
  if (blockIdx.x * kW >= n) return;
  __shared__ BlockLdsOf<kW> L;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  // padded DCT rows for stage 2: d2[j][m] = D_flat[j*127 + m] (m = 127 reads
  // the next row's first coefficient, or the zero pad after row 15)
  for (int e = threadIdx.x; e < 16 * 128; e += blockDim.x) {
    const int j = e >> 7, m = e & 127;
    L.d2[j * kTS + m] = kDct[j * 127 + m];
  }
  __syncthreads();
  const uint32_t img = blockIdx.x * kW + wv;
  if (img >= n) return;
  hash_one<T, D>(planes, plane_stride, img, lane, wv, mode, out, L, lc);
}

// Dihedrals of given PDQFs (the visit entry point): pdqf [n][256], thr_in [n].
__global__ void __launch_bounds__(256) pdq_dihedral_kernel(const float* __restrict__ pdqf,
                                                           const float* __restrict__ thr_in, uint32_t n,
                                                           uint32_t mode, HashOutputs out) {
  __shared__ WaveLds Ws[kWavesPerBlock];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const uint32_t img = blockIdx.x * kWavesPerBlock + wv;
  if (img >= n) return;
  WaveLds& W = Ws[wv];
#pragma unroll
  for (int e = 0; e < 4; e++) W.p[lane + 64 * e] = pdqf[(size_t)img * 256 + lane + 64 * e];
  wave_lds_sync();
  HashOutputs o = out;
  o.pdqf = nullptr;
  dihedral_tail(W, lane, mode, o, img, thr_in + img);
}

hipError_t launch_dihedrals(const float* pdqf, const float* thr_in, uint32_t n, uint32_t mode,
                            const HashOutputs& out, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(pdq_dihedral_kernel, dim3((n + kWavesPerBlock - 1) / kWavesPerBlock),
                     dim3(64 * kWavesPerBlock), 0, stream, pdqf, thr_in, n, mode, out);
  return hipGetLastError();
}



template <typename T, int kW, int D>
static void launch_kind_w(const void* planes, uint64_t plane_stride, uint32_t n_max, const uint32_t* d_n,
                          uint32_t mode, const HashOutputs& out, const Luma& lc, hipStream_t stream) {
  const dim3 grid((n_max + kW - 1) / kW);
  hipLaunchKernelGGL((pdq_hash_kernel<T, kW, D>), grid, dim3(64 * kW), 0, stream, static_cast<const T*>(planes),
                     plane_stride, n_max, d_n, mode, out, lc);
}

// u8 batches of at least 16 images per CU take the 16-wave blocks.
static bool sixteen_wave_blocks(uint32_t n_max) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess)
    return false;
  return cus > 0 && n_max >= 16u * (uint32_t)cus;
}

template <typename T>
static void launch_kind(const void* planes, uint64_t plane_stride, uint32_t n_max, const uint32_t* d_n, uint32_t mode,
                        const HashOutputs& out, const Luma& lc, hipStream_t stream) {
  if constexpr (std::is_same<T, uint8_t>::value) {
    if (sixteen_wave_blocks(n_max)) {
      launch_kind_w<T, 16, 1>(planes, plane_stride, n_max, d_n, mode, out, lc, stream);
      return;
    }
  }
  launch_kind_w<T, kWavesPerBlock, Prefetch<T>::D>(planes, plane_stride, n_max, d_n, mode, out, lc, stream);
}

static hipError_t launch_any(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_max,
                             const uint32_t* d_n, uint32_t mode, const HashOutputs& out, const float* luma,
                             hipStream_t stream) {
  if (n_max == 0) return hipSuccess;
  Luma lc{0.299f, 0.587f, 0.114f};
  if (luma) lc = Luma{luma[0], luma[1], luma[2]};
  switch (input_kind) {
    case kInputF32: launch_kind<float>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kInputU8: launch_kind<uint8_t>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kInputRgb8: launch_kind<Rgb8>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kInputRgba8: launch_kind<Rgba8>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_hash_wave(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_max,
                            const uint32_t* d_n, uint32_t mode, const HashOutputs& out, const float* luma,
                            hipStream_t stream) {
  return launch_any(planes, input_kind, plane_stride, n_max, d_n, mode, out, luma, stream);
}

}  // namespace ypdq
