// Block-per-image form of the fused hash kernel for gfx950 (CDNA4): the
// wavefronts of one block share one 512x512 plane, so a launch of a few
// hundred images fills the chip and one image takes a fraction of the
// wave-per-image form's latency.
//
// Per image (kWaves wavefronts, 64 lanes each):
//   A. wave w compresses its band of rows of the 127x127 plane with the same
//      tent FMA chains as the wave-per-image form (x86.rs:626-682 order,
//      lane t owns output columns 2t and 2t+1) and stores them in LDS as a
//      128x128 f32 plane whose row and column 127 are zero.
//   B. DCT stage 1, T = D C (16 x 128).
//        exact:        thread (column j, group of frequencies) runs the
//                      reference's two alternating FMA chains over
//                      k2 = 0..126 (x86.rs:161-225) -- the chains, operands
//                      and order of the wave form, so the values are the
//                      same bits;
//        matrix cores: v_mfma_f32_16x16x32_{bf16,f16}, operands rounded to
//                      bf16 or f16 (round to nearest even) with f32
//                      accumulation, or the three products hi*hi + hi*lo +
//                      lo*hi of split operands x = hi + lo.
//   C. stage 2 (exact: x86.rs:227-272, one column of each quad per wave;
//      matrix cores: four 16x16x32 products on wave 0), then the shared tail
//      (dihedrals, gradients, median search, bits) on wave 0.
// Planes holding NaN/Inf take the literal 56-tap compression and the exact
// stage 1 in every form.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "pdq_hash_common.h"
#include "pdq_hash_tail.h"

namespace ypdq {

namespace {

struct ImageTail {
  WaveLds w;
  float d2[16 * kTS];  // DCT rows, stride kTS
};

struct ImageLds {
  union {
    float c[128 * 128];  // compressed plane, row-major, stride 128
    ImageTail tail;      // after stage 1
  } u;
  float dct_t[127 * 16];  // exact stage 1: D transposed, dct_t[k2 * 16 + k] = D[k][k2]
};

// Rows in flight per wave in the band loop: D compressed-row steps of four
// input rows each.
template <typename T>
struct BandPrefetch {
  static constexpr int D = 2;
};
template <>
struct BandPrefetch<uint8_t> {
  static constexpr int D = 4;
};

// ---------------------------------------------------------------------------
// A. compression of one band of rows into LDS
// ---------------------------------------------------------------------------
template <int S, int D, typename T, typename Raw>
__device__ __forceinline__ void band_step(const T* __restrict__ plane, int lane, int i, int i_end, Chains& cur,
                                          Chains& nxt, Raw (&pre)[D][4], float* __restrict__ c, const Luma& lc,
                                          bool& bad) {
  // The rows D steps ahead are loaded unconditionally, clamped into the
  // plane (past the band's end they are not used): a load under a branch
  // makes the compiler copy the ring's registers and wait for every load in
  // flight.
  const int r0 = 4 * (i + D) + 3;
  auto row = [&](int d) { return min(r0 + d, 511); };
  (void)i_end;
  if constexpr (std::is_same<T, float>::value) {
    // f32 rows feed from their ring slots, which reload after their last use
    // (a copy out of the slot would cost eight moves per row)
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
    // converted rows free their slots at once
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
  if (lane == 63) c1 = 0.0f;  // column 127 is the zero pad
  if constexpr (std::is_same<T, float>::value) bad = bad || nonfinite(c0) || nonfinite(c1);
  *reinterpret_cast<float2*>(&c[i * 128 + 2 * lane]) = make_float2(c0, c1);
}

// D steps; the two chain sets trade roles from step to step (the next row's
// chains are the current ones of the step after), so no step copies them.
// kAll: every step runs (no test, so the ring's registers stay fixed);
// otherwise the steps up to the band's end.
template <int S, int D, bool kAll, typename T, typename Raw>
__device__ __forceinline__ void band_steps(const T* __restrict__ plane, int lane, int i, int i_end, Chains& a,
                                           Chains& b, Raw (&pre)[D][4], float* __restrict__ c, const Luma& lc,
                                           bool& bad) {
  static_assert(D % 2 == 0, "an even number of steps returns the chain sets to their roles");
  if (kAll || i + S < i_end) {
    if constexpr (S % 2 == 0)
      band_step<S, D>(plane, lane, i + S, i_end, a, b, pre, c, lc, bad);
    else
      band_step<S, D>(plane, lane, i + S, i_end, b, a, pre, c, lc, bad);
  }
  if constexpr (S + 1 < D) band_steps<S + 1, D, kAll>(plane, lane, i, i_end, a, b, pre, c, lc, bad);
}

// Compressed rows [i0, i1) into c; true if one of them is not finite.
template <typename T>
__device__ __forceinline__ bool compress_band(const T* __restrict__ plane, int lane, int i0, int i1,
                                              float* __restrict__ c, const Luma& lc) {
  using Raw = typename RawRow<T>::type;
  constexpr int D = BandPrefetch<T>::D;
  Chains cur, nxt;
  Raw pre[D][4];
  {
    Raw r1, r2;
    load_raw(plane, 4 * i0 + 1, lane, r1);
    load_raw(plane, 4 * i0 + 2, lane, r2);
#pragma unroll
    for (int s = 0; s < D; s++)
#pragma unroll
      for (int d = 0; d < 4; d++) load_raw(plane, min(4 * (i0 + s) + 3 + d, 511), lane, pre[s][d]);
    Row8 x1, x2;
    to_f32(r1, x1, lc);
    to_f32(r2, x2, lc);
    chains_start(cur, x1, x2);
  }
  bool bad = false;
  int i = i0;
#pragma unroll 1
  for (; i + D <= i1; i += D) band_steps<0, D, true>(plane, lane, i, i1, cur, nxt, pre, c, lc, bad);
  if (i < i1) band_steps<0, D, false>(plane, lane, i, i1, cur, nxt, pre, c, lc, bad);
  return bad;
}

// Positions the reference reads that carry no non-zero tap of any output:
// rows 1 and 511, and columns 0, 1 and 511 of rows 2..510.
__device__ __forceinline__ bool edges_nonfinite(const float* __restrict__ plane, int t, int n_threads) {
  bool bad = false;
  for (int e = t; e < 512; e += n_threads)
    bad = bad || nonfinite(plane[512 + e]) || nonfinite(plane[511 * 512 + e]);
  for (int row = 2 + t; row <= 510; row += n_threads) {
    const float* pr = plane + (size_t)row * 512;
    bad = bad || nonfinite(pr[0]) || nonfinite(pr[1]) || nonfinite(pr[511]);
  }
  return bad;
}

// ---------------------------------------------------------------------------
// B. stage 1, exact: thread (column j, frequencies k0 .. k0+NK-1).
// ---------------------------------------------------------------------------
// The plane and the transposed DCT both sit in LDS, so every operand load is
// an LDS load (those complete in order) and the loads of the next chunk of
// rows are issued before the FMAs of the current one.
template <int NK>
struct DctSlice {
  float v[NK];
};

template <int NK>
__device__ __forceinline__ void load_slice(const float* __restrict__ p, DctSlice<NK>& s) {
  if constexpr (NK % 4 == 0) {
#pragma unroll
    for (int q = 0; q < NK / 4; q++) {
      const f32x4 v = reinterpret_cast<const f32x4*>(p)[q];
      s.v[4 * q] = v.x; s.v[4 * q + 1] = v.y; s.v[4 * q + 2] = v.z; s.v[4 * q + 3] = v.w;
    }
  } else {
    static_assert(NK == 2, "two, four or eight frequencies per thread");
    const float2 v = *reinterpret_cast<const float2*>(p);
    s.v[0] = v.x; s.v[1] = v.y;
  }
}

template <int NK, int R>
__device__ __forceinline__ void stage1_load_chunk(const float* __restrict__ c, const float* __restrict__ dt, int i0,
                                                  float (&cv)[R], DctSlice<NK> (&dv)[R], int n = R) {
#pragma unroll
  for (int r = 0; r < R; r++)
    if (r < n) {
      cv[r] = c[(i0 + r) * 128];
      load_slice<NK>(dt + (i0 + r) * 16, dv[r]);
    }
}

// Rows i0 .. i0 + n - 1 with i0 a multiple of 4, so row i0 + r takes
// accumulator alt(r).
template <int NK, int R>
__device__ __forceinline__ void stage1_fma_chunk(const float (&cv)[R], const DctSlice<NK> (&dv)[R],
                                                 float (&acc)[NK][2], int n = R) {
#pragma unroll
  for (int r = 0; r < R; r++)
    if (r < n) {
      const int p = alt(r);
#pragma unroll
      for (int k = 0; k < NK; k++) acc[k][p] = __builtin_fmaf(cv[r], dv[r].v[k], acc[k][p]);
    }
}

// acc[k][p] = the reference's two alternating chains over k2 = 0..126
// (x86.rs:161-225) for output (k0 + k, j); c is column j of the plane, dt
// the transposed DCT at frequency k0.  Chunks of eight rows, the loads of
// chunk b + 1 in flight while chunk b is accumulated.
template <int NK>
__device__ __forceinline__ void stage1_exact(const float* __restrict__ c, const float* __restrict__ dt,
                                             float (&acc)[NK][2]) {
  constexpr int R = 8;
#pragma unroll
  for (int k = 0; k < NK; k++) acc[k][0] = acc[k][1] = 0.0f;
  float cv0[R], cv1[R];
  DctSlice<NK> dv0[R], dv1[R];
  stage1_load_chunk<NK, R>(c, dt, 0, cv0, dv0);
#pragma unroll 1
  for (int i0 = 0; i0 < 112; i0 += 2 * R) {
    stage1_load_chunk<NK, R>(c, dt, i0 + R, cv1, dv1);
    stage1_fma_chunk<NK, R>(cv0, dv0, acc);
    stage1_load_chunk<NK, R>(c, dt, i0 + 2 * R, cv0, dv0);
    stage1_fma_chunk<NK, R>(cv1, dv1, acc);
  }
  // rows 112 .. 119 are in cv0; rows 120 .. 126 remain
  stage1_load_chunk<NK, R>(c, dt, 120, cv1, dv1, 7);
  stage1_fma_chunk<NK, R>(cv0, dv0, acc);
  stage1_fma_chunk<NK, R>(cv1, dv1, acc, 7);
}

// ---------------------------------------------------------------------------
// B'. stage 1 and 2 on matrix cores.
//
// v_mfma_f32_16x16x32_{bf16,f16}: lane l supplies row l%16 of A and column
// l%16 of B, eight consecutive k each (group g = l/16 holds k = 8g .. 8g+7 of
// the 32); it receives D[4g + r][l%16], r = 0..3.  A and B use the same k
// assignment, so the contraction is right whatever order the unit sums in.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4v __attribute__((ext_vector_type(4)));

// bf16 bits of a finite float, rounded to nearest even.
__device__ __forceinline__ uint32_t bf16_rne(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// Eight floats as an operand pair: hi = round(x), lo = round(x - hi).
template <int kPrec>
struct Operand;

template <>
struct Operand<kDctBf16> {
  bf16x8 hi, lo;
  __device__ __forceinline__ void set(const float (&x)[8], bool split) {
    u32x4 h, l;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const uint32_t a = bf16_rne(x[2 * e]), b = bf16_rne(x[2 * e + 1]);
      h[e] = a | (b << 16);
      if (split) {
        const uint32_t al = bf16_rne(x[2 * e] - __uint_as_float(a << 16));
        const uint32_t bl = bf16_rne(x[2 * e + 1] - __uint_as_float(b << 16));
        l[e] = al | (bl << 16);
      } else {
        l[e] = 0;
      }
    }
    hi = __builtin_bit_cast(bf16x8, h);
    lo = __builtin_bit_cast(bf16x8, l);
  }
  static __device__ __forceinline__ f32x4v mma(const bf16x8& a, const bf16x8& b, const f32x4v& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

template <>
struct Operand<kDctF16> {
  f16x8 hi, lo;
  __device__ __forceinline__ void set(const float (&x)[8], bool split) {
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const _Float16 h = (_Float16)x[e];
      hi[e] = h;
      lo[e] = split ? (_Float16)(x[e] - (float)h) : (_Float16)0.0f;
    }
  }
  static __device__ __forceinline__ f32x4v mma(const f16x8& a, const f16x8& b, const f32x4v& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

// Operand type and split flag of each matrix-core precision.
template <int kPrec>
struct Precision {
  static constexpr int kBase = (kPrec == kDctBf16 || kPrec == kDctBf16x3) ? kDctBf16 : kDctF16;
  static constexpr bool kSplit = kPrec == kDctBf16x3 || kPrec == kDctF16x3;
  using Op = Operand<kBase>;
};

// acc += a * b (hi + lo split: the two cross terms first, then hi * hi).
template <int kPrec>
__device__ __forceinline__ f32x4v mma_acc(const typename Precision<kPrec>::Op& a,
                                          const typename Precision<kPrec>::Op& b, f32x4v acc) {
  using Op = typename Precision<kPrec>::Op;
  if constexpr (Precision<kPrec>::kSplit) {
    acc = Op::mma(a.lo, b.hi, acc);
    acc = Op::mma(a.hi, b.lo, acc);
  }
  return Op::mma(a.hi, b.hi, acc);
}

// Row l%16 of the DCT matrix, k = 32s + 8g .. +7 (column 127 is zero).
template <int kPrec>
__device__ __forceinline__ void dct_operands(int lane, typename Precision<kPrec>::Op (&d)[4]) {
  const int m = lane & 15, g = lane >> 4;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    float x[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int k = 32 * s + 8 * g + e;
      x[e] = k < 127 ? kDct[m * 127 + k] : 0.0f;
    }
    d[s].set(x, Precision<kPrec>::kSplit);
  }
}

// T[:, 16 jt .. 16 jt + 15] = D C on matrix cores.
template <int kPrec>
__device__ __forceinline__ f32x4v stage1_mfma(const float* __restrict__ c, int lane, int jt,
                                              const typename Precision<kPrec>::Op (&d)[4]) {
  const int n = lane & 15, g = lane >> 4;
  f32x4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int s = 0; s < 4; s++) {
    float x[8];
#pragma unroll
    for (int e = 0; e < 8; e++) x[e] = c[(32 * s + 8 * g + e) * 128 + 16 * jt + n];
    typename Precision<kPrec>::Op b;
    b.set(x, Precision<kPrec>::kSplit);
    acc = mma_acc<kPrec>(d[s], b, acc);
  }
  return acc;
}

// PDQF = T D^T on matrix cores (one wave): A = T rows, B = the DCT operands.
template <int kPrec>
__device__ __forceinline__ void stage2_mfma(WaveLds& W, int lane, const typename Precision<kPrec>::Op (&d)[4]) {
  const int m = lane & 15, g = lane >> 4;
  f32x4v acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const f32x4* tp = reinterpret_cast<const f32x4*>(&W.t[m * kTS + 32 * s + 8 * g]);
    const f32x4 ta = tp[0], tb = tp[1];
    const float x[8] = {ta.x, ta.y, ta.z, ta.w, tb.x, tb.y, tb.z, tb.w};
    typename Precision<kPrec>::Op a;
    a.set(x, Precision<kPrec>::kSplit);
    acc = mma_acc<kPrec>(a, d[s], acc);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) W.p[(4 * g + r) * 16 + m] = acc[r];
  wave_lds_sync();
}

// ---------------------------------------------------------------------------
// The kernel: one block of kWaves wavefronts per image.
// ---------------------------------------------------------------------------
template <typename T, int kPrec, int kWaves>
__global__ void __launch_bounds__(64 * kWaves) pdq_hash_block_kernel(const T* __restrict__ planes,
                                                                     uint64_t plane_stride, uint32_t n_images,
                                                                     const uint32_t* __restrict__ d_n, uint32_t mode,
                                                                     HashOutputs out, Luma lc) {
  static_assert(kWaves == 4 || kWaves == 8 || kWaves == 16, "four, eight or sixteen waves per image");
  static_assert(kWaves <= 8 || kPrec == kDctExact, "sixteen waves: exact arithmetic only");
  constexpr int kBand = (127 + kWaves - 1) / kWaves;  // compressed rows per wave
  // exact stage 1 on at most eight waves: with two frequencies per thread
  // its LDS reads of the plane double
  constexpr int kS1Waves = kWaves > 8 ? 8 : kWaves;
  constexpr int kNK = 16 / (kS1Waves / 2);  // exact stage 1: frequencies per thread
  const uint32_t n = d_n ? min(*d_n, n_images) : n_images;
  const uint32_t img = blockIdx.x;
  if (img >= n) return;
  __shared__ ImageLds L;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const T* plane = reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(planes) + (size_t)img * plane_stride);
  float* c = L.u.c;

  // A. compression; the DCT table for stage 1 is copied into LDS after it,
  // so the first plane rows are requested at the kernel's start
  const int i0 = wv * kBand, i1 = min(i0 + kBand, 127);
  bool bad = compress_band(plane, lane, i0, i1, c, lc);
  if (wv == kWaves - 1) *reinterpret_cast<float2*>(&c[127 * 128 + 2 * lane]) = make_float2(0.0f, 0.0f);
  if (kPrec == kDctExact || std::is_same<T, float>::value)
    for (int e = t; e < 127 * 16; e += 64 * kWaves) L.dct_t[e] = kDctT.v[e];
  if constexpr (std::is_same<T, float>::value) bad = bad || edges_nonfinite(plane, t, 64 * kWaves);
  const bool literal = __syncthreads_or(bad ? 1 : 0) != 0;
  if constexpr (std::is_same<T, float>::value) {
    if (literal) {
      // 56 literal taps per output (NaN/Inf propagate through zero weights
      // as in the reference)
#pragma unroll 1
      for (int i = i0; i < i1; i++) {
        const float a = compress_one_literal(plane, i, 2 * lane, lc);
        const float b = lane == 63 ? 0.0f : compress_one_literal(plane, i, 2 * lane + 1, lc);
        *reinterpret_cast<float2*>(&c[i * 128 + 2 * lane]) = make_float2(a, b);
      }
      __syncthreads();
    }
  }
  if (out.buf1) {
    float* b1 = out.buf1 + (size_t)img * 128 * 128;
    for (int e = t; e < 128 * 128; e += 64 * kWaves) b1[e] = c[e];
  }

  WaveLds& W = L.u.tail.w;
  float* d2 = L.u.tail.d2;
  if (kPrec == kDctExact || literal) {
    // B. exact stage 1: thread (j = t % 128, frequencies kNK * (t / 128) ..)
    const int j = t & 127;
    const int k0 = kNK * __builtin_amdgcn_readfirstlane(t >> 7);
    const bool s1 = wv < kS1Waves;  // wave-uniform
    float acc[kNK][2];
    if (s1) stage1_exact<kNK>(c + j, L.dct_t + k0, acc);
    __syncthreads();  // every read of the plane is done: T and d2 overwrite it
    if (s1) {
#pragma unroll
      for (int k = 0; k < kNK; k++) {
        float a0 = acc[k][0], a1 = acc[k][1];
        a0 = __builtin_isnan(a0) ? 0.0f : a0;
        a1 = __builtin_isnan(a1) ? 0.0f : a1;
        W.t[(k0 + k) * kTS + j] = (j == 127) ? 0.0f : a0 + a1;
      }
    }
  } else if constexpr (kPrec != kDctExact) {
    // B'. stage 1 on matrix cores: wave w computes column tiles w, w + kWaves
    typename Precision<kPrec>::Op d[4];
    dct_operands<kPrec>(lane, d);
    constexpr int kTiles = 8 / kWaves;
    f32x4v acc[kTiles];
#pragma unroll
    for (int q = 0; q < kTiles; q++) acc[q] = stage1_mfma<kPrec>(c, lane, wv + kWaves * q, d);
    __syncthreads();
    const int nn = lane & 15, g = lane >> 4;
#pragma unroll
    for (int q = 0; q < kTiles; q++)
#pragma unroll
      for (int r = 0; r < 4; r++) W.t[(4 * g + r) * kTS + 16 * (wv + kWaves * q) + nn] = acc[q][r];
  }
  if (kPrec == kDctExact || literal) {
    // d2[j][m] = D[j][m] from the transposed copy in LDS; slot 127 holds the
    // reference's over-read D_flat[127 j + 127] = D[j + 1][0] (0 after row 15)
    for (int e = t; e < 16 * 128; e += 64 * kWaves) {
      const int jj = e & 15, m = e >> 4;
      d2[jj * kTS + m] = m < 127 ? L.dct_t[e] : (jj < 15 ? L.dct_t[jj + 1] : 0.0f);
    }
  }
  __syncthreads();

  // C. stage 2 and the tail
  if (kPrec == kDctExact || literal) {
    if (wv < 4) stage2_column(d2, W, lane, wv);
  } else if (wv == 0) {
    typename Precision<kPrec>::Op d[4];
    dct_operands<kPrec>(lane, d);
    stage2_mfma<kPrec>(W, lane, d);
  }
  __syncthreads();
  if (wv == 0) write_tmp(W, lane, out, img);
  // the eight orientations' PDQFs, written by every wave (two orientations
  // per element and thread with eight waves)
  {
    constexpr int kPer = 8 / (kWaves / 4);  // orientations per thread
    const int idx = t & 255, d0 = (t >> 8) * kPer;
    dihedral_pdqf_element(W, idx, d0, kPer, mode);
  }
  __syncthreads();
  if constexpr (kWaves >= 8) {
    // waves 0..3 search the four thresholds and write the hashes, wave 4
    // the qualities, the others the optional PDQF outputs
    if (wv < 4) {
      tail_rows_thresholds_bits(W, lane, wv, out, img, nullptr);
    } else if (wv == 4) {
      tail_quality(W, lane, out, img, true);
    } else {
      tail_copy_pdqfs(W, t - 5 * 64, 64 * (kWaves - 5), out, img);
    }
  } else {
    // the quality on wave 1 while wave 0 searches the thresholds
    if (wv == 1) tail_quality(W, lane, out, img, true);
    if (wv == 0) tail_thresholds_bits(W, lane, out, img, nullptr);
  }
}

template <typename T, int kPrec, int kWaves>
void launch_block_kind(const void* planes, uint64_t plane_stride, uint32_t n_max, const uint32_t* d_n, uint32_t mode,
                       const HashOutputs& out, const Luma& lc, hipStream_t stream) {
  hipLaunchKernelGGL((pdq_hash_block_kernel<T, kPrec, kWaves>), dim3(n_max), dim3(64 * kWaves), 0, stream,
                     static_cast<const T*>(planes), plane_stride, n_max, d_n, mode, out, lc);
}

template <typename T, int kWaves>
hipError_t launch_block_dct(int dct, const void* planes, uint64_t plane_stride, uint32_t n_max, const uint32_t* d_n,
                            uint32_t mode, const HashOutputs& out, const Luma& lc, hipStream_t stream) {
  switch (dct) {
    case kDctExact: launch_block_kind<T, kDctExact, kWaves>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kDctBf16: launch_block_kind<T, kDctBf16, kWaves>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kDctF16: launch_block_kind<T, kDctF16, kWaves>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kDctBf16x3: launch_block_kind<T, kDctBf16x3, kWaves>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kDctF16x3: launch_block_kind<T, kDctF16x3, kWaves>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int kWaves>
hipError_t launch_block_waves(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_max,
                              const uint32_t* d_n, uint32_t mode, int dct, const HashOutputs& out, const Luma& lc,
                              hipStream_t stream) {
  switch (input_kind) {
    case kInputF32: return launch_block_dct<float, kWaves>(dct, planes, plane_stride, n_max, d_n, mode, out, lc, stream);
    case kInputU8: return launch_block_dct<uint8_t, kWaves>(dct, planes, plane_stride, n_max, d_n, mode, out, lc, stream);
    case kInputRgb8: return launch_block_dct<Rgb8, kWaves>(dct, planes, plane_stride, n_max, d_n, mode, out, lc, stream);
    case kInputRgba8: return launch_block_dct<Rgba8, kWaves>(dct, planes, plane_stride, n_max, d_n, mode, out, lc, stream);
    default: return hipErrorInvalidValue;
  }
}

// Sixteen waves per image: exact arithmetic only.
hipError_t launch_block_waves16(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_max,
                                const uint32_t* d_n, uint32_t mode, int dct, const HashOutputs& out, const Luma& lc,
                                hipStream_t stream) {
  if (dct != kDctExact) return hipErrorInvalidValue;
  switch (input_kind) {
    case kInputF32: launch_block_kind<float, kDctExact, 16>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kInputU8: launch_block_kind<uint8_t, kDctExact, 16>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kInputRgb8: launch_block_kind<Rgb8, kDctExact, 16>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    case kInputRgba8: launch_block_kind<Rgba8, kDctExact, 16>(planes, plane_stride, n_max, d_n, mode, out, lc, stream); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess)
    return 0;
  return cus;
}

}  // namespace

int hash_layout_for(uint32_t n_images, int dct) {
  // Exact arithmetic: sixteen waves per image while the images fit one per
  // CU, then eight (18 us for one u8 image, 23 us for 128, 71 us for 1024,
  // against 60 / 71 / 78 us for a wave per image) until the batch is large
  // enough that a wave per image fills the GPU with less LDS per image (1536
  // u8 images: 91 us against 105 us; 4096: 193-208 against 261 us).
  // Matrix-core arithmetic exists only in the block forms of 4 and 8 waves.
  if (dct != kDctExact) return 8;
  if (n_images > 1024) return 1;
  const int cus = device_cus();
  return (cus > 0 && n_images <= (uint32_t)cus) ? 16 : 8;
}

hipError_t launch_hash(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_images, uint32_t mode,
                       const HashOutputs& out, hipStream_t stream, const float* luma) {
  const int layout = hash_layout_for(n_images, kDctExact);
  if (layout == 1)
    return launch_hash_wave(planes, input_kind, plane_stride, n_images, nullptr, mode, out, luma, stream);
  return launch_hash_block(planes, input_kind, plane_stride, n_images, nullptr, mode, kDctExact, layout, out, luma,
                           stream);
}

hipError_t launch_hash_device_count(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_max,
                                    const uint32_t* d_n, uint32_t mode, const HashOutputs& out, hipStream_t stream) {
  const int layout = hash_layout_for(n_max, kDctExact);
  if (layout == 1)
    return launch_hash_wave(planes, input_kind, plane_stride, n_max, d_n, mode, out, nullptr, stream);
  return launch_hash_block(planes, input_kind, plane_stride, n_max, d_n, mode, kDctExact, layout, out, nullptr,
                           stream);
}

hipError_t launch_hash_block(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_max,
                             const uint32_t* d_n, uint32_t mode, int dct, int waves, const HashOutputs& out,
                             const float* luma, hipStream_t stream) {
  if (n_max == 0) return hipSuccess;
  Luma lc{0.299f, 0.587f, 0.114f};
  if (luma) lc = Luma{luma[0], luma[1], luma[2]};
  if (waves == 16) return launch_block_waves16(planes, input_kind, plane_stride, n_max, d_n, mode, dct, out, lc, stream);
  if (waves == 8) return launch_block_waves<8>(planes, input_kind, plane_stride, n_max, d_n, mode, dct, out, lc, stream);
  if (waves == 4) return launch_block_waves<4>(planes, input_kind, plane_stride, n_max, d_n, mode, dct, out, lc, stream);
  return hipErrorInvalidValue;
}

}  // namespace ypdq
