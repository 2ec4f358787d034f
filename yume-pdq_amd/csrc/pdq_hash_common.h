// Device code shared by the fused and the staged hash kernels: constant
// tables, row loaders, the tent-compression FMA chains, DCT stage-1
// accumulation, and the reductions, median search and gradient sum of the
// reference's AVX2 kernel (src/kernel/x86.rs).  Every piece reproduces the
// reference's FP32 operation order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pdq_device.h"
#include "pdq_tables.h"

namespace ypdq {

constexpr float kTent[56] = {PDQ_TENT_X8_INIT};
static __constant__ float kDct[16 * 127 + 16] = {PDQ_DCT_F32_INIT};

constexpr float kDctRowMajor[16 * 127 + 16] = {PDQ_DCT_F32_INIT};

// DCT matrix transposed (kDctT[k2][k] = D[k][k2]) so one compressed row's 16
// coefficients are contiguous and the wave reads them with one scalar load.
struct DctTransposed {
  float v[127 * 16];
  constexpr DctTransposed() : v() {
    for (int k2 = 0; k2 < 127; k2++)
      for (int k = 0; k < 16; k++) v[k2 * 16 + k] = kDctRowMajor[k * 127 + k2];
  }
};
static __constant__ DctTransposed kDctT = DctTransposed();

// Tent weight w[di][l] of the reference's 7x8 table (build.rs:729-745).
__device__ __forceinline__ constexpr float tw(int di, int l) { return kTent[di * 8 + l]; }

// Which of the two stage-1/stage-2 accumulators takes term n: the reference's
// macro expansion alternates them as 1,0,0,1 (x86.rs:163-202, 241-265).
__device__ __forceinline__ constexpr int alt(int n) { return (n & 3) == 0 || (n & 3) == 3 ? 1 : 0; }

// ---------------------------------------------------------------------------
// Input row loaders: 8 consecutive columns per lane.
// ---------------------------------------------------------------------------
struct Row8 {
  float v[8];
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// Planes are read once: non-temporal loads keep them from displacing the
// L2/MALL lines other waves still need.
__device__ __forceinline__ void load_row(const float* __restrict__ plane, int r, int lane, Row8& o) {
  const f32x4* p = reinterpret_cast<const f32x4*>(plane + (size_t)r * 512 + 8 * lane);
  f32x4 a = __builtin_nontemporal_load(p);
  f32x4 b = __builtin_nontemporal_load(p + 1);
  o.v[0] = a.x; o.v[1] = a.y; o.v[2] = a.z; o.v[3] = a.w;
  o.v[4] = b.x; o.v[5] = b.y; o.v[6] = b.z; o.v[7] = b.w;
}

struct RowU8 {
  uint32_t lo, hi;
};

__device__ __forceinline__ void load_raw(const uint8_t* __restrict__ plane, int r, int lane, RowU8& o) {
  const u32x2* p = reinterpret_cast<const u32x2*>(plane + (size_t)r * 512 + 8 * lane);
  u32x2 a = __builtin_nontemporal_load(p);
  o.lo = a.x;
  o.hi = a.y;
}

__device__ __forceinline__ void load_raw(const float* __restrict__ plane, int r, int lane, Row8& o) {
  load_row(plane, r, lane, o);
}

__device__ __forceinline__ void to_f32(const Row8& in, Row8& o) { o = in; }

__device__ __forceinline__ void to_f32(const RowU8& in, Row8& o) {
  // u8 -> f32 is exact (v_cvt_f32_ubyte0..3)
  o.v[0] = (float)(in.lo & 0xff);
  o.v[1] = (float)((in.lo >> 8) & 0xff);
  o.v[2] = (float)((in.lo >> 16) & 0xff);
  o.v[3] = (float)(in.lo >> 24);
  o.v[4] = (float)(in.hi & 0xff);
  o.v[5] = (float)((in.hi >> 8) & 0xff);
  o.v[6] = (float)((in.hi >> 16) & 0xff);
  o.v[7] = (float)(in.hi >> 24);
}

// Colour inputs, converted to luma in the row loader: RGB8 by the Kernel
// trait's default cvt_rgb8_to_luma8f (mod.rs:303-322: R*r + G*g + B*b, left
// to right, no fused multiply-add), RGBA8 by the AVX2 kernel's
// cvt_rgba8_to_luma8f (x86.rs:686-753: r*R, then fma(g, G, .), then
// fma(b, B, .)).  Coefficients (R, G, B), BT.601 by default
// (constants.rs:23).
struct Luma {
  float r, g, b;
};
struct Rgb8 {
  uint8_t c[3];
};
struct Rgba8 {
  uint8_t c[4];
};

__device__ __forceinline__ float luma_rgb(uint32_t r, uint32_t g, uint32_t b, const Luma& lc) {
  const float pr = lc.r * (float)r;
  const float pg = lc.g * (float)g;
  const float pb = lc.b * (float)b;
  return (pr + pg) + pb;
}
__device__ __forceinline__ float luma_rgba(uint32_t r, uint32_t g, uint32_t b, const Luma& lc) {
  float v = (float)r * lc.r;
  v = __builtin_fmaf((float)g, lc.g, v);
  return __builtin_fmaf((float)b, lc.b, v);
}

struct RowRgb {  // 8 pixels x 3 bytes
  uint32_t w[6];
};
struct RowRgba {  // 8 pixels x 4 bytes
  uint32_t w[8];
};
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void load_raw(const Rgb8* __restrict__ plane, int r, int lane, RowRgb& o) {
  // 1536 bytes per row, 24 per lane: three 8-byte loads
  const u32x2* p = reinterpret_cast<const u32x2*>(reinterpret_cast<const uint8_t*>(plane) + (size_t)r * 1536 + 24 * lane);
  const u32x2 a = __builtin_nontemporal_load(p), b = __builtin_nontemporal_load(p + 1),
              c = __builtin_nontemporal_load(p + 2);
  o.w[0] = a.x; o.w[1] = a.y; o.w[2] = b.x; o.w[3] = b.y; o.w[4] = c.x; o.w[5] = c.y;
}

__device__ __forceinline__ void load_raw(const Rgba8* __restrict__ plane, int r, int lane, RowRgba& o) {
  const u32x4* p = reinterpret_cast<const u32x4*>(reinterpret_cast<const uint8_t*>(plane) + (size_t)r * 2048 + 32 * lane);
  const u32x4 a = __builtin_nontemporal_load(p), b = __builtin_nontemporal_load(p + 1);
  o.w[0] = a.x; o.w[1] = a.y; o.w[2] = a.z; o.w[3] = a.w;
  o.w[4] = b.x; o.w[5] = b.y; o.w[6] = b.z; o.w[7] = b.w;
}

__device__ __forceinline__ void to_f32(const Row8& in, Row8& o, const Luma&) { o = in; }
__device__ __forceinline__ void to_f32(const RowU8& in, Row8& o, const Luma&) { to_f32(in, o); }

__device__ __forceinline__ void to_f32(const RowRgb& in, Row8& o, const Luma& lc) {
#pragma unroll
  for (int k = 0; k < 8; k++) {
    uint32_t ch[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const int byte = 3 * k + c;
      ch[c] = (in.w[byte >> 2] >> (8 * (byte & 3))) & 0xffu;
    }
    o.v[k] = luma_rgb(ch[0], ch[1], ch[2], lc);
  }
}

__device__ __forceinline__ void to_f32(const RowRgba& in, Row8& o, const Luma& lc) {
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t w = in.w[k];
    o.v[k] = luma_rgba(w & 0xffu, (w >> 8) & 0xffu, (w >> 16) & 0xffu, lc);
  }
}

// One input element as f32 (the literal compression path).
__device__ __forceinline__ float px_f32(float v, const Luma&) { return v; }
__device__ __forceinline__ float px_f32(uint8_t v, const Luma&) { return (float)v; }
__device__ __forceinline__ float px_f32(const Rgb8& v, const Luma& lc) { return luma_rgb(v.c[0], v.c[1], v.c[2], lc); }
__device__ __forceinline__ float px_f32(const Rgba8& v, const Luma& lc) {
  return luma_rgba(v.c[0], v.c[1], v.c[2], lc);
}

template <typename T>
struct RawRow;
template <>
struct RawRow<float> {
  using type = Row8;
};
template <>
struct RawRow<uint8_t> {
  using type = RowU8;
};
template <>
struct RawRow<Rgb8> {
  using type = RowRgb;
};
template <>
struct RawRow<Rgba8> {
  using type = RowRgba;
};

// ---------------------------------------------------------------------------
// Tent compression, fast form.
//
// The reference accumulates, per output (i, j), eight lanes l = 0..7 over the
// seven rows di: lane[l] = fma(in[4i+1+di][4j+1+l], w[di][l], lane[l]), then
// ((l0+l1)+(l2+l3))+((l4+l5)+(l6+l7)).  w is zero outside di = 0..5 x l = 1..5
// and in row di = 0 except l = 4, 5 (denormal residues of the 128-bit table).
// For finite inputs a zero-weight FMA adds a signed zero to an accumulator
// that started at +0, which leaves its value unchanged; lanes 0, 6, 7 stay
// +0, and (+0 + a1) = a1, (a4+a5) + (+0 + +0) = a4+a5 up to the sign of an
// exact zero.  The result is therefore (a1 + (a2 + a3)) + (a4 + a5) over the
// 27 non-zero taps: bit for bit the reference's value (a zero can differ in
// sign only when the plane holds negative subnormals, and no signed zero
// reaches a hash bit, quality or threshold).  Planes holding NaN/Inf take
// the literal form below.
//
// Column c = 8t+m of lane t plays role l = c - (4j+1) for output j:
//   m=0: l3 of output 2t-1     m=4: l3 of 2t
//   m=1: l4 of 2t-1            m=5: l4 of 2t
//   m=2: l5 of 2t-1, l1 of 2t  m=6: l5 of 2t, l1 of 2t+1
//   m=3: l2 of 2t              m=7: l2 of 2t+1
// giving ten vertical FMA chains per lane; roles of columns m = 0..2 for
// output 2t+1 come from lane t+1.  The chains advance one input row at a
// time (di ascending, the reference's order), so each row is consumed as it
// arrives; rows 4i+5 and 4i+6 also start the chains of compressed row i+1.
// ---------------------------------------------------------------------------
constexpr int kChainCol[10] = {0, 1, 2, 2, 3, 4, 5, 6, 6, 7};
constexpr int kChainRole[10] = {3, 4, 5, 1, 2, 3, 4, 5, 1, 2};

struct Chains {
  float s[10];
};

// Feed input row `x` as tap row DI to every chain with a non-zero weight there.
template <int DI>
__device__ __forceinline__ void chains_feed(Chains& c, const Row8& x) {
#pragma unroll
  for (int n = 0; n < 10; n++) {
    const int l = kChainRole[n];
    if (DI >= 1 || l == 4 || l == 5) c.s[n] = __builtin_fmaf(x.v[kChainCol[n]], tw(DI, l), c.s[n]);
  }
}

// The chains of a compressed row from its first two input rows (tap rows 0
// and 1): each accumulator starts from +0 as the reference's do, so the
// first term is fma(x, w, +0); written as one expression per chain, with no
// zeroed register for the compiler to materialise.
__device__ __forceinline__ void chains_start(Chains& c, const Row8& x0, const Row8& x1) {
#pragma unroll
  for (int n = 0; n < 10; n++) {
    const int l = kChainRole[n];
    const float s0 = (l == 4 || l == 5) ? __builtin_fmaf(x0.v[kChainCol[n]], tw(0, l), 0.0f) : 0.0f;
    c.s[n] = __builtin_fmaf(x1.v[kChainCol[n]], tw(1, l), s0);
  }
}

__device__ __forceinline__ void chains_finish(const Chains& c, float& c0, float& c1) {
  const float n0 = shfl_down1(c.s[0]);
  const float n1 = shfl_down1(c.s[1]);
  const float n2h = shfl_down1(c.s[2]);
  c0 = (c.s[3] + (c.s[4] + c.s[5])) + (c.s[6] + c.s[7]);
  c1 = (c.s[8] + (c.s[9] + n0)) + (n1 + n2h);
}

// Literal evaluation of one compressed row for outputs 2t, 2t+1 (all 56 taps,
// the one-element rewind at (126,126,di=6), reads past a row's end), used when
// the plane holds a non-finite value.  Reads the plane from memory directly.
template <typename T>
__device__ __forceinline__ float compress_one_literal(const T* __restrict__ plane, int outi, int outj,
                                                      const Luma& lc = Luma{0.0f, 0.0f, 0.0f}) {
  const int in_i = 4 * (outi + 1) - 3;
  const int in_j = 4 * (outj + 1) - 3;
  float lane[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
  for (int di = 0; di < 7; di++) {
    long off = (long)(in_i + di) * 512 + in_j;
    float x[8];
    if (di == 6 && outi == 126 && outj == 126) {
#pragma unroll
      for (int k = 0; k < 7; k++) x[k] = px_f32(plane[off + k], lc);
      x[7] = px_f32(plane[off + 6], lc);
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) x[k] = px_f32(plane[off + k], lc);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) lane[k] = __builtin_fmaf(x[k], tw(di, k), lane[k]);
  }
  return ((lane[0] + lane[1]) + (lane[2] + lane[3])) + ((lane[4] + lane[5]) + (lane[6] + lane[7]));
}

// Stage-1 accumulate of one compressed row (16 frequencies x 2 columns).
template <int P>
__device__ __forceinline__ void stage1_acc(float (&acc)[16][2][2], float c0, float c1, const float* __restrict__ dk) {
#pragma unroll
  for (int k = 0; k < 16; k++) {
    const float d = dk[k];
    acc[k][0][P] = __builtin_fmaf(c0, d, acc[k][0][P]);
    acc[k][1][P] = __builtin_fmaf(c1, d, acc[k][1][P]);
  }
}


__device__ __forceinline__ bool nonfinite(float x) { return !__builtin_isfinite(x); }

// ---------------------------------------------------------------------------
// Per-wave LDS layout.
// ---------------------------------------------------------------------------
// Row stride of the stage-1 rows and of the DCT rows in LDS: 128 columns
// plus four pad floats, so the sixteen rows one ds_read_b128 of stage 2
// touches start in different banks.
constexpr int kTS = 132;

struct WaveLds {
  float t[16 * kTS];  // stage-1 rows (stride kTS); later the 8 dihedral PDQFs (8 x 256)
  float p[256];       // PDQF of the NORMAL orientation
  float q[8];         // per-dihedral quality
  float th[8];        // per-dihedral threshold
  float bt[8];        // per-dihedral bit threshold (last compared guess)
};

constexpr int kWavesPerBlock = 4;

template <int kW>
struct BlockLdsOf {
  float d2[16 * kTS];  // DCT rows padded to 128 (slot 127 = the reference's over-read), stride kTS
  WaveLds w[kW];
};
using BlockLds = BlockLdsOf<kWavesPerBlock>;

// 8-lane group reductions with the AVX2 tree shapes.  Lanes of a group are
// g*8 .. g*8+7.  The operand moves are DPP lane permutations within rows of
// 16 lanes: quad_perm [1,0,3,2] (lane ^ 1), quad_perm [2,3,0,1] (lane ^ 2),
// row_half_mirror (lane 7 - i of the group) and row_ror:12 (lane i takes
// lane i - 12 mod 16 = i + 4 for i < 12).  For
// addition the tree's value appears in every lane (IEEE addition commutes,
// and after two steps lanes 4..7 all hold the same upper half); max and min
// are read from the group's lane 0, which takes the operands of the xor tree.
template <int kCtrl>
__device__ __forceinline__ uint32_t dpp_u(uint32_t x) {
  return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(x), kCtrl, 0xf, 0xf, false));
}
template <int kCtrl>
__device__ __forceinline__ float dpp_f(float x) {
  return __uint_as_float(dpp_u<kCtrl>(__float_as_uint(x)));
}
constexpr int kDppXor1 = 0xb1;        // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4e;        // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror
constexpr int kDppPlus4 = 0x12c;      // row_ror:12

__device__ __forceinline__ float grp_sum8(float a) {
  const float s1 = a + dpp_f<kDppXor1>(a);
  const float s2 = s1 + dpp_f<kDppXor2>(s1);
  return s2 + dpp_f<kDppHalfMirror>(s2);
}

__device__ __forceinline__ float x86max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float x86min(float a, float b) { return a < b ? a : b; }

__device__ __forceinline__ float grp_max8(float a, int lane) {
  float p = x86max(a, dpp_f<kDppXor1>(a));  // lane 2m: max(a_2m, a_2m+1)
  float q = x86max(p, dpp_f<kDppXor2>(p));  // lanes 0,4: max(p0,p2), max(p4,p6)
  float r = x86max(q, dpp_f<kDppPlus4>(q));  // lane 0: max(q0, q4)
  return shfl_idx(r, lane & ~7);
}

__device__ __forceinline__ float grp_min8(float a, int lane) {
  float p = x86min(a, dpp_f<kDppXor1>(a));
  float q = x86min(p, dpp_f<kDppXor2>(p));
  float r = x86min(q, dpp_f<kDppPlus4>(q));
  return shfl_idx(r, lane & ~7);
}

__device__ __forceinline__ uint32_t grp_isum8(uint32_t a) {
  a += dpp_u<kDppXor1>(a);
  a += dpp_u<kDppXor2>(a);
  return a + dpp_u<kDppHalfMirror>(a);
}


// Serial sum of gradients of one 16x16 PDQF (x86.rs:840-865): a single f32
// accumulation, vertical differences first, then horizontal.
__device__ __forceinline__ float sum_of_gradients16(const float* __restrict__ P) {
  float g = 0.0f;
  float4 cur[4];
#pragma unroll
  for (int q4 = 0; q4 < 4; q4++) cur[q4] = reinterpret_cast<const float4*>(P)[q4];
#pragma unroll 1
  for (int i = 0; i < 15; i++) {
    float4 nxt[4];
#pragma unroll
    for (int q4 = 0; q4 < 4; q4++) nxt[q4] = reinterpret_cast<const float4*>(P + (i + 1) * 16)[q4];
#pragma unroll
    for (int q4 = 0; q4 < 4; q4++) {
      g += __builtin_fabsf(cur[q4].x - nxt[q4].x);
      g += __builtin_fabsf(cur[q4].y - nxt[q4].y);
      g += __builtin_fabsf(cur[q4].z - nxt[q4].z);
      g += __builtin_fabsf(cur[q4].w - nxt[q4].w);
    }
#pragma unroll
    for (int q4 = 0; q4 < 4; q4++) cur[q4] = nxt[q4];
  }
#pragma unroll 1
  for (int i = 0; i < 16; i++) {
    float r[16];
#pragma unroll
    for (int q4 = 0; q4 < 4; q4++) {
      float4 v = reinterpret_cast<const float4*>(P + i * 16)[q4];
      r[4 * q4] = v.x; r[4 * q4 + 1] = v.y; r[4 * q4 + 2] = v.z; r[4 * q4 + 3] = v.w;
    }
#pragma unroll
    for (int j = 0; j < 15; j++) g += __builtin_fabsf(r[j] - r[j + 1]);
  }
  return g;
}

// adjust_quality (x86.rs:769-773, QUALITY_ADJUST_DIVISOR = 180 at
// mod.rs:257): min(g / 180, 1), where Rust's f32::min ignores a NaN operand.
__device__ __forceinline__ float adjust_quality(float g) {
  const float s = g / 180.0f;
  return __builtin_isnan(s) ? 1.0f : (s < 1.0f ? s : 1.0f);
}

// The AVX2 median search (quantize_impl, x86.rs:440-621) over the PDQF P by
// the eight lanes of one group (AVX2 lane l = lane & 7 holds elements 8c + l,
// c = 0..31).  Every group of the wave runs its own search in lockstep; the
// loop ends when all groups have converged (or after 32 iterations, as the
// reference).  Returns the threshold and the guess of the last comparison
// (the bits the reference stores are x > that guess) in all lanes of the group.
__device__ __forceinline__ void median_search8(const float* __restrict__ P, int lane, float& threshold,
                                               float& bit_threshold) {
  const int l = lane & 7;
  float x[32];
#pragma unroll
  for (int c = 0; c < 32; c++) x[c] = P[8 * c + l];
  float mxl = -3.40282347e+38f, mnl = 3.40282347e+38f;  // f32::MIN, f32::MAX
#pragma unroll
  for (int c = 0; c < 32; c++) {
    mnl = x86min(mnl, x[c]);
    mxl = x86max(mxl, x[c]);
  }
  float mx = grp_max8(mxl, lane);
  float mn = grp_min8(mnl, lane);
  float guess = (mn + mx) * 0.5f;
  float last = guess;
  uint32_t num_over = 0, num_under = 0;
  bool done = false;
#pragma unroll 1
  for (int iter = 0; iter < 32; iter++) {
    float rg = 0.0f, rl = 0.0f;
    uint32_t gt = 0, lt = 0;
#pragma unroll
    for (int c = 0; c < 32; c++) {
      const float v = x[c];
      const bool mg = (v > guess) && (v <= mx);
      const bool ml = (v < guess) && (v >= mn);
      gt += mg ? 1u : 0u;
      lt += ml ? 1u : 0u;
      rg = rg + (mg ? v : 0.0f);
      rl = rl + (ml ? v : 0.0f);
    }
    const uint32_t gtc = grp_isum8(gt), ltc = grp_isum8(lt);
    const float sg = grp_sum8(rg), sl = grp_sum8(rl);
    if (!done) {
      last = guess;
      if (gtc + num_over > 128) {
        num_under += ltc;
        mn = guess;
        guess = sg / (float)gtc;
      } else if (ltc + num_under > 128) {
        num_over += gtc;
        mx = guess;
        guess = sl / (float)ltc;
      } else {
        done = true;
      }
    }
    if (wave_all(done)) break;
  }
  threshold = guess;
  bit_threshold = last;
}

// The value of lane l ^ 32 (v_permlane32_swap: the upper half of the first
// operand trades places with the lower half of the second).
__device__ __forceinline__ uint32_t half_swap(uint32_t x, int lane) {
  const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
  return lane < 32 ? r[1] : r[0];
}

// median_search8 for four PDQFs at once, the sums of each search split over
// the two halves of the wave: group s = (lane >> 3) & 3 of the lower half
// accumulates the elements above the guess, the same group of the upper half
// the elements below it, so each lane's sequential sum covers one side (half
// the instructions per element).  The upper half works on negated elements,
// y = -x, where x < guess and x >= mn read y > -guess and y <= -mn -- the
// same two comparisons as the lower half with other bounds.  IEEE negation
// is exact and round-to-nearest is symmetric, so the sum of negated elements
// in the reference's order is the negated sum, except that an exactly-zero
// sum is +0 both ways: the lower side's sum is recovered as +0 or as the
// negation.  Counts and sums cross halves with one lane swap each.
__device__ __forceinline__ void median_search8_split(const float* __restrict__ P, int lane, float& threshold,
                                                     float& bit_threshold) {
  const int l = lane & 7;
  const bool lower_side = lane >= 32;
  float x[32];
#pragma unroll
  for (int c = 0; c < 32; c++) x[c] = P[8 * c + l];
  float mxl = -3.40282347e+38f, mnl = 3.40282347e+38f;  // f32::MIN, f32::MAX
#pragma unroll
  for (int c = 0; c < 32; c++) {
    mnl = x86min(mnl, x[c]);
    mxl = x86max(mxl, x[c]);
  }
  float mx = grp_max8(mxl, lane);
  float mn = grp_min8(mnl, lane);
#pragma unroll
  for (int c = 0; c < 32; c++) x[c] = lower_side ? -x[c] : x[c];
  float guess = (mn + mx) * 0.5f;
  float last = guess;
  uint32_t num_over = 0, num_under = 0;
  bool done = false;
#pragma unroll 1
  for (int iter = 0; iter < 32; iter++) {
    // this half's side: (guess, mx] above, or [mn, guess) below (negated)
    const float lo_b = lower_side ? -guess : guess;
    const float hi_b = lower_side ? -mn : mx;
    float r = 0.0f;
    uint32_t n = 0;
#pragma unroll
    for (int c = 0; c < 32; c++) {
      const float v = x[c];
      const bool m = (v > lo_b) && (v <= hi_b);
      n += m ? 1u : 0u;
      r = r + (m ? v : 0.0f);
    }
    const uint32_t cnt = grp_isum8(n);
    float sum = grp_sum8(r);
    if (lower_side) sum = sum == 0.0f ? 0.0f : -sum;  // back to the original sign (+0 for zero)
    const uint32_t cnt_o = half_swap(cnt, lane);
    const float sum_o = __uint_as_float(half_swap(__float_as_uint(sum), lane));
    const uint32_t gtc = lower_side ? cnt_o : cnt, ltc = lower_side ? cnt : cnt_o;
    const float sg = lower_side ? sum_o : sum, sl = lower_side ? sum : sum_o;
    if (!done) {
      last = guess;
      if (gtc + num_over > 128) {
        num_under += ltc;
        mn = guess;
        guess = sg / (float)gtc;
      } else if (ltc + num_under > 128) {
        num_over += gtc;
        mx = guess;
        guess = sl / (float)ltc;
      } else {
        done = true;
      }
    }
    if (wave_all(done)) break;
  }
  threshold = guess;
  bit_threshold = last;
}

// The value of the same lane in the partner row of 16 lanes, rows 0 <-> 1
// and 2 <-> 3 (v_permlane16_swap: the odd rows of the first operand trade
// places with the even rows of the second).
__device__ __forceinline__ uint32_t row_swap(uint32_t x, int lane) {
  const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
  return (lane & 16) ? r[0] : r[1];
}
__device__ __forceinline__ float row_swapf(float x, int lane) {
  return __uint_as_float(row_swap(__float_as_uint(x), lane));
}
__device__ __forceinline__ float half_swapf(float x, int lane) {
  return __uint_as_float(half_swap(__float_as_uint(x), lane));
}

// median_search8 of one PDQF on a whole wave, for latency: each of the
// sixteen sequential sums (8 AVX lanes x the two sides of the guess, as in
// median_search8_split) is cut into four runs of 8 elements on four rows of
// the wave.  Every lane selects its run's elements and counts them at once;
// then the runs are summed in order, row 0, 1, 3, 2, each row starting from
// the previous row's partial sum (v_permlane16/32_swap move it), so every
// sum adds the same values in the same order as the reference's AVX lane.
// Lane layout: l = lane & 7 the AVX lane, bit 3 the side (set: below the
// guess, on negated elements), row = lane >> 4 holds run row ^ (row >> 1).
// The counts are integers and are summed in any order.  The guess and the
// bounds are wave-uniform, so the search's branches are scalar.
__device__ __forceinline__ void median_search_rows(const float* __restrict__ P, int lane, float& threshold,
                                                   float& bit_threshold) {
  const int l = lane & 7;
  const bool lower_side = (lane & 8) != 0;
  const int row = lane >> 4;
  const int run = row ^ (row >> 1);
  float mx, mn;
  {
    // the reference's running extremes over the 32 elements of AVX lane l
    float mxl = -3.40282347e+38f, mnl = 3.40282347e+38f;  // f32::MIN, f32::MAX
#pragma unroll
    for (int c = 0; c < 32; c++) {
      const float v = P[8 * c + l];
      mnl = x86min(mnl, v);
      mxl = x86max(mxl, v);
    }
    mx = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(grp_max8(mxl, lane))));
    mn = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(grp_min8(mnl, lane))));
  }
  float x[8];
#pragma unroll
  for (int e = 0; e < 8; e++) {
    const float v = P[8 * (8 * run + e) + l];
    x[e] = lower_side ? -v : v;
  }
  float guess = (mn + mx) * 0.5f;
  float last = guess;
  uint32_t num_over = 0, num_under = 0;
#pragma unroll 1
  for (int iter = 0; iter < 32; iter++) {
    // this lane's side: (guess, mx] above, or [mn, guess) below (negated)
    const float lo_b = lower_side ? -guess : guess;
    const float hi_b = lower_side ? -mn : mx;
    float s[8];
    uint32_t n = 0;
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const bool m = (x[e] > lo_b) && (x[e] <= hi_b);
      n += m ? 1u : 0u;
      s[e] = m ? x[e] : 0.0f;
    }
    // the sequential sum, run after run
    float acc = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; e++) acc = acc + s[e];
    acc = row_swapf(acc, lane);  // row 1 continues row 0
#pragma unroll
    for (int e = 0; e < 8; e++) acc = acc + s[e];
    acc = half_swapf(acc, lane);  // row 3 continues row 1
#pragma unroll
    for (int e = 0; e < 8; e++) acc = acc + s[e];
    acc = row_swapf(acc, lane);  // row 2 continues row 3
#pragma unroll
    for (int e = 0; e < 8; e++) acc = acc + s[e];
    float sum = grp_sum8(acc);  // the AVX tree; rows 2 hold the whole sums
    if (lower_side) sum = sum == 0.0f ? 0.0f : -sum;  // back to the original sign (+0 for zero)
    n = grp_isum8(n);
    n += row_swap(n, lane);
    n += half_swap(n, lane);
    const uint32_t gtc = __builtin_amdgcn_readlane(n, 0);
    const uint32_t ltc = __builtin_amdgcn_readlane(n, 8);
    const float sg = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(sum), 32));
    const float sl = __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(sum), 40));
    last = guess;
    if (gtc + num_over > 128) {
      num_under += ltc;
      mn = guess;
      guess = sg / (float)gtc;
    } else if (ltc + num_under > 128) {
      num_over += gtc;
      mx = guess;
      guess = sl / (float)ltc;
    } else {
      break;
    }
  }
  threshold = guess;
  bit_threshold = last;
}

// Byte b (0..31) of the hash of P at bit threshold t: bit k is
// P[8(31-b) + k] > t (element i at byte 31 - i/8, bit i%8; x86.rs:497-503,
// threshold.rs:45-115).
__device__ __forceinline__ uint32_t hash_byte(const float* __restrict__ P, int b, float t) {
  const int e0 = 8 * (31 - b);
  uint32_t byte = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) byte |= (P[e0 + k] > t ? 1u : 0u) << k;
  return byte;
}

}  // namespace ypdq
