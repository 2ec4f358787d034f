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

#include "pdq_device.h"
#include "pdq_tables.h"

namespace ypdq {

constexpr float kTent[56] = {PDQ_TENT_X8_INIT};
__constant__ float kDct[16 * 127 + 16] = {PDQ_DCT_F32_INIT};

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
__constant__ DctTransposed kDctT = DctTransposed();

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

__device__ __forceinline__ void chains_zero(Chains& c) {
#pragma unroll
  for (int n = 0; n < 10; n++) c.s[n] = 0.0f;
}

// Feed input row `x` as tap row DI to every chain with a non-zero weight there.
template <int DI>
__device__ __forceinline__ void chains_feed(Chains& c, const Row8& x) {
#pragma unroll
  for (int n = 0; n < 10; n++) {
    const int l = kChainRole[n];
    if (DI >= 1 || l == 4 || l == 5) c.s[n] = __builtin_fmaf(x.v[kChainCol[n]], tw(DI, l), c.s[n]);
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
__device__ __forceinline__ float compress_one_literal(const T* __restrict__ plane, int outi, int outj) {
  const int in_i = 4 * (outi + 1) - 3;
  const int in_j = 4 * (outj + 1) - 3;
  float lane[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
  for (int di = 0; di < 7; di++) {
    long off = (long)(in_i + di) * 512 + in_j;
    float x[8];
    if (di == 6 && outi == 126 && outj == 126) {
#pragma unroll
      for (int k = 0; k < 7; k++) x[k] = (float)plane[off + k];
      x[7] = (float)plane[off + 6];
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) x[k] = (float)plane[off + k];
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
struct WaveLds {
  float t[16 * 128];  // stage-1 rows; later the 8 dihedral PDQFs (8 x 256)
  float p[256];       // PDQF of the NORMAL orientation
  float q[8];         // per-dihedral quality
  float th[8];        // per-dihedral threshold
  float bt[8];        // per-dihedral bit threshold (last compared guess)
};

constexpr int kWavesPerBlock = 4;

struct BlockLds {
  float d2[16 * 128];  // DCT rows padded to 128 (slot 127 = the reference's over-read)
  WaveLds w[kWavesPerBlock];
};

// 8-lane group reductions with the AVX2 tree shapes.  Lanes of a group are
// g*8 .. g*8+7; the value of the tree is produced in every lane by xor
// shuffles, which for addition gives identical results in all lanes (IEEE
// addition commutes) and for max/min is read from the group's lane 0.
__device__ __forceinline__ float grp_sum8(float a) {
  float s1 = a + shfl_xor(a, 1);
  float s2 = s1 + shfl_xor(s1, 2);
  return s2 + shfl_xor(s2, 4);
}

__device__ __forceinline__ float x86max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float x86min(float a, float b) { return a < b ? a : b; }

__device__ __forceinline__ float grp_max8(float a, int lane) {
  float p = x86max(a, shfl_xor(a, 1));   // lane 2m: max(a_2m, a_2m+1)
  float q = x86max(p, shfl_xor(p, 2));   // lanes 0,4: max(p0,p2), max(p4,p6)
  float r = x86max(q, shfl_xor(q, 4));   // lane 0: max(q0, q4)
  return shfl_idx(r, lane & ~7);
}

__device__ __forceinline__ float grp_min8(float a, int lane) {
  float p = x86min(a, shfl_xor(a, 1));
  float q = x86min(p, shfl_xor(p, 2));
  float r = x86min(q, shfl_xor(q, 4));
  return shfl_idx(r, lane & ~7);
}

__device__ __forceinline__ uint32_t grp_isum8(uint32_t a) {
  a += shfl_xor_u(a, 1);
  a += shfl_xor_u(a, 2);
  return a + shfl_xor_u(a, 4);
}

// ---------------------------------------------------------------------------
// The tail: stage 2, dihedrals, gradients, quantisation, packing.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stage2_to_pdqf(const BlockLds& L, WaveLds& W, int lane, const HashOutputs& out,
                                               uint32_t img) {
  if (out.tmp) {
    // the reference leaves the k = 15 stage-1 row in its tmp buffer
    float* dst = out.tmp + (size_t)img * 128;
    dst[2 * lane] = W.t[15 * 128 + 2 * lane];
    dst[2 * lane + 1] = W.t[15 * 128 + 2 * lane + 1];
  }
  // Stage 2 (x86.rs:227-272): lane (k = lane & 15, jq = lane >> 4) computes
  // out[k][4jq + r], r = 0..3, one output at a time: 8 AVX2 lanes x 2
  // alternating accumulators over 16 chunks, then the horizontal tree.
  const int k = lane & 15, jq = lane >> 4;
#pragma unroll 1
  for (int r = 0; r < 4; r++) {
    const int j = 4 * jq + r;
    float h[2][8];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int l = 0; l < 8; l++) h[p][l] = 0.0f;
#pragma unroll
    for (int c = 0; c < 16; c++) {
      const f32x4* tp = reinterpret_cast<const f32x4*>(&W.t[k * 128 + 8 * c]);
      const f32x4* dp = reinterpret_cast<const f32x4*>(&L.d2[j * 128 + 8 * c]);
      const f32x4 ta = tp[0], tb = tp[1], da = dp[0], db = dp[1];
      const float tv[8] = {ta.x, ta.y, ta.z, ta.w, tb.x, tb.y, tb.z, tb.w};
      const float dv[8] = {da.x, da.y, da.z, da.w, db.x, db.y, db.z, db.w};
      const int p = alt(c);
#pragma unroll
      for (int l = 0; l < 8; l++) h[p][l] = __builtin_fmaf(tv[l], dv[l], h[p][l]);
    }
    float v[8];
#pragma unroll
    for (int l = 0; l < 8; l++) v[l] = h[0][l] + h[1][l];
    W.p[k * 16 + j] = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
  }
  wave_lds_sync();
}

// From the NORMAL PDQF in W.p: the eight dihedral PDQFs, their qualities,
// thresholds and hashes.  thr_normal < 0 (null) means "use NORMAL's own
// quantisation threshold" for ROTATED_90_FLOPPED; the visit entry point passes
// the caller's threshold instead (src/lib.rs:403-404).
__device__ __forceinline__ void dihedral_tail(WaveLds& W, int lane, uint32_t mode, const HashOutputs& out, uint32_t img,
                              const float* thr_normal) {
  // The eight dihedral PDQFs into W.t (d-major, 256 each):
  //   0 NORMAL P          1 FLOPPED  sc.P       2 ROTATED_180 sr.sc.P   3 FLIPPED sr.P
  //   4 ROT90_FLOPPED P^T 5 ROT270   sr.P^T     6 FLOPPED_ROT270 sr.sc.P^T  7 ROT90 sc.P^T
  // sr negates even rows (pdqf_negate_alt_rows::<true>, x86.rs:285-326); sc
  // negates even columns (pdqf_negate_alt_cols::<false>) -- all 16 in exact
  // mode, columns 0..7 only in the AVX2-compatible mode (x86.rs:330-376 loads
  // 8 floats per row).  Negation is exact, so the sequence of in-place flips
  // in visit_dihedrals reduces to these closed forms.
  {
    const int ncols = mode == kDihedralAvx2 ? 8 : 16;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int idx = lane + 64 * e;  // element (i, j) of the output orientation
      const int i = idx >> 4, j = idx & 15;
      const float a = W.p[idx];
      const float at = W.p[j * 16 + i];
      const bool nr = (i & 1) == 0;
      const bool nc = (j & 1) == 0 && j < ncols;
      W.t[0 * 256 + idx] = a;
      W.t[1 * 256 + idx] = nc ? a * -1.0f : a;
      W.t[2 * 256 + idx] = (nr != nc) ? a * -1.0f : a;
      W.t[3 * 256 + idx] = nr ? a * -1.0f : a;
      W.t[4 * 256 + idx] = at;
      W.t[5 * 256 + idx] = nr ? at * -1.0f : at;
      W.t[6 * 256 + idx] = (nr != nc) ? at * -1.0f : at;
      W.t[7 * 256 + idx] = nc ? at * -1.0f : at;
    }
  }
  wave_lds_sync();

  // Gradient sums (x86.rs:840-865): one sequential f32 accumulation per
  // dihedral, vertical then horizontal differences, lane d owns dihedral d.
  if (lane < 8) {
    const float* P = &W.t[lane * 256];
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
    // adjust_quality (x86.rs:769-773): min(g / 180, 1); Rust's min ignores NaN
    float s = g / 180.0f;
    W.q[lane] = __builtin_isnan(s) ? 1.0f : (s < 1.0f ? s : 1.0f);
  }

  // Median search (x86.rs:440-621) for dihedrals 0..3: group s = lane >> 3
  // (lanes 32..63 repeat groups 0..3), AVX2 lane l = lane & 7 holds elements
  // 8c + l, c = 0..31.
  {
    const int s = (lane >> 3) & 3, l = lane & 7;
    const float* P = &W.t[s * 256];
    float x[32];
#pragma unroll
    for (int c = 0; c < 32; c++) x[c] = P[8 * c + l];
    float mxl = -3.40282347e+38f, mnl = 3.40282347e+38f;
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
    if (lane < 32 && l == 0) {
      W.th[s] = guess;
      W.bt[s] = last;
    }
  }
  wave_lds_sync();
  if (lane < 4) {
    // threshold_2d_f32 with the thresholds visit_dihedrals passes (lib.rs:403-416)
    const float t = (lane == 0 && thr_normal) ? *thr_normal : W.th[lane];
    W.th[4 + lane] = t;
    W.bt[4 + lane] = t;
  }
  wave_lds_sync();

  // Bits: lane (d = lane >> 3, w = lane & 7) builds bytes 4w..4w+3 of hash d;
  // byte b holds elements 8(31-b) .. 8(31-b)+7 at bits 0..7 (x86.rs:497-503).
  {
    const int d = lane >> 3, w = lane & 7;
    const float* P = &W.t[d * 256];
    const float bt = W.bt[d];
    uint32_t word = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int b = 4 * w + q;
      const int e0 = 8 * (31 - b);
      uint32_t byte = 0;
#pragma unroll
      for (int k = 0; k < 8; k++) byte |= (P[e0 + k] > bt ? 1u : 0u) << k;
      word |= byte << (8 * q);
    }
    reinterpret_cast<uint32_t*>(out.hashes + (size_t)img * 256 + d * 32)[w] = word;
  }
  if (lane < 8) {
    out.quality[(size_t)img * 8 + lane] = W.q[lane];
    out.threshold[(size_t)img * 8 + lane] = W.th[lane];
  }
  if (out.pdqf) {
#pragma unroll
    for (int e = 0; e < 4; e++) out.pdqf[(size_t)img * 256 + lane + 64 * e] = W.p[lane + 64 * e];
  }
  if (out.dihedral_pdqf) {
    float* dst = out.dihedral_pdqf + (size_t)img * 8 * 256;
#pragma unroll
    for (int e = 0; e < 32; e++) dst[lane + 64 * e] = W.t[lane + 64 * e];
  }
  wave_lds_sync();
}

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

// One compressed row i (ring slot S = i mod D): rows 4i+3 .. 4i+6 arrive from
// pre[S], each slot is refilled with the row D steps ahead as soon as it has
// been converted; the chains of row i complete and those of row i+1 start, and
// the two compressed values go into stage-1 accumulator P.
template <int P, int S, int D, typename T, typename Raw>
__device__ __forceinline__ void stream_step(const T* __restrict__ plane, int lane, int i, float (&acc)[16][2][2],
                                            Chains& cur, Chains& nxt, Raw (&pre)[D][4], float* __restrict__ buf1) {
  const bool more = i + D < 127;
  const int r0 = 4 * (i + D) + 3;
  Row8 x;
  to_f32(pre[S][0], x);
  if (more) load_raw(plane, r0, lane, pre[S][0]);
  chains_feed<2>(cur, x);
  to_f32(pre[S][1], x);
  if (more) load_raw(plane, r0 + 1, lane, pre[S][1]);
  chains_feed<3>(cur, x);
  chains_zero(nxt);
  to_f32(pre[S][2], x);
  if (more) load_raw(plane, r0 + 2, lane, pre[S][2]);
  chains_feed<4>(cur, x);
  chains_feed<0>(nxt, x);
  to_f32(pre[S][3], x);
  if (more) load_raw(plane, r0 + 3, lane, pre[S][3]);
  chains_feed<5>(cur, x);
  chains_feed<1>(nxt, x);
  float c0, c1;
  chains_finish(cur, c0, c1);
  if (buf1) {
    buf1[i * 128 + 2 * lane] = c0;
    buf1[i * 128 + 2 * lane + 1] = (lane == 63) ? 0.0f : c1;
  }
  stage1_acc<P>(acc, c0, c1, kDctT.v + i * 16);
  cur = nxt;
}

template <typename T>
__device__ __forceinline__ void stream_image(const T* __restrict__ plane, int lane, float (&acc)[16][2][2],
                                             float* __restrict__ buf1) {
  using Raw = typename RawRow<T>::type;
  constexpr int D = Prefetch<T>::D;
  static_assert(4 % D == 0, "ring depth must divide the unroll factor");
#pragma unroll
  for (int k = 0; k < 16; k++)
#pragma unroll
    for (int jj = 0; jj < 2; jj++) acc[k][jj][0] = acc[k][jj][1] = 0.0f;

  Chains cur, nxt;
  chains_zero(cur);
  Raw pre[D][4];
  {
    Raw r1, r2;
    load_raw(plane, 1, lane, r1);
    load_raw(plane, 2, lane, r2);
#pragma unroll
    for (int s = 0; s < D; s++)
#pragma unroll
      for (int d = 0; d < 4; d++) load_raw(plane, 4 * s + 3 + d, lane, pre[s][d]);
    Row8 x;
    to_f32(r1, x);
    chains_feed<0>(cur, x);
    to_f32(r2, x);
    chains_feed<1>(cur, x);
  }
  // The accumulator a compressed row feeds is fixed by i mod 4 (alt), so the
  // loop is unrolled by four to keep every accumulator and ring index a
  // constant.
#pragma unroll 1
  for (int i = 0; i < 124; i += 4) {
    stream_step<1, 0 % D, D>(plane, lane, i, acc, cur, nxt, pre, buf1);
    stream_step<0, 1 % D, D>(plane, lane, i + 1, acc, cur, nxt, pre, buf1);
    stream_step<0, 2 % D, D>(plane, lane, i + 2, acc, cur, nxt, pre, buf1);
    stream_step<1, 3 % D, D>(plane, lane, i + 3, acc, cur, nxt, pre, buf1);
  }
  stream_step<1, 124 % D, D>(plane, lane, 124, acc, cur, nxt, pre, buf1);
  stream_step<0, 125 % D, D>(plane, lane, 125, acc, cur, nxt, pre, buf1);
  stream_step<0, 126 % D, D>(plane, lane, 126, acc, cur, nxt, pre, buf1);
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
      W.t[k * 128 + j] = (j == 127) ? 0.0f : v;
    }
  }
}

template <int P, typename T>
__device__ __forceinline__ void literal_step(const T* __restrict__ plane, int lane, int i, int k0,
                                             float (&acc)[8][2][2], float* __restrict__ buf1) {
  const float c0 = compress_one_literal(plane, i, 2 * lane);
  // output column 127 does not exist: lane 63 evaluates column 126 again (in
  // bounds) and discards it
  float c1 = compress_one_literal(plane, i, lane == 63 ? 126 : 2 * lane + 1);
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
                                                  float* __restrict__ buf1) {
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
      literal_step<1>(plane, lane, i, k0, acc, buf1);
      literal_step<0>(plane, lane, i + 1, k0, acc, buf1);
      literal_step<0>(plane, lane, i + 2, k0, acc, buf1);
      literal_step<1>(plane, lane, i + 3, k0, acc, buf1);
    }
    literal_step<1>(plane, lane, 124, k0, acc, buf1);
    literal_step<0>(plane, lane, 125, k0, acc, buf1);
    literal_step<0>(plane, lane, 126, k0, acc, buf1);
#pragma unroll
    for (int k = 0; k < 8; k++) {
#pragma unroll
      for (int jj = 0; jj < 2; jj++) {
        float a0 = acc[k][jj][0], a1 = acc[k][jj][1];
        a0 = __builtin_isnan(a0) ? 0.0f : a0;
        a1 = __builtin_isnan(a1) ? 0.0f : a1;
        const float v = a0 + a1;
        const int j = 2 * lane + jj;
        W.t[(k0 + k) * 128 + j] = (j == 127) ? 0.0f : v;
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

__device__ __forceinline__ bool plane_needs_literal(const uint8_t* __restrict__, int, const float (&)[16][2][2]) {
  return false;
}

template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) pdq_hash_kernel(const T* __restrict__ planes, uint64_t plane_stride,
                                                        uint32_t n_images, uint32_t mode, HashOutputs out) {
  __shared__ BlockLds L;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  // padded DCT rows for stage 2: d2[j][m] = D_flat[j*127 + m] (m = 127 reads
  // the next row's first coefficient, or the zero pad after row 15)
  for (int e = threadIdx.x; e < 16 * 128; e += blockDim.x) {
    const int j = e >> 7, m = e & 127;
    L.d2[e] = kDct[j * 127 + m];
  }
  __syncthreads();
  const uint32_t img = blockIdx.x * kWavesPerBlock + wv;
  if (img >= n_images) return;
  const T* plane = reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(planes) + (size_t)img * plane_stride);
  float* buf1 = out.buf1 ? out.buf1 + (size_t)img * 128 * 128 : nullptr;
  WaveLds& W = L.w[wv];
  {
    float acc[16][2][2];
    stream_image(plane, lane, acc, buf1);
    const bool literal = plane_needs_literal(plane, lane, acc);
    write_stage1(acc, W, lane);
    if (literal) stream_image_literal(plane, lane, W, buf1);
  }
  wave_lds_sync();
  if (buf1) {
    buf1[127 * 128 + 2 * lane] = 0.0f;
    buf1[127 * 128 + 2 * lane + 1] = 0.0f;
    buf1[(2 * lane) * 128 + 127] = 0.0f;
    buf1[(2 * lane + 1) * 128 + 127] = 0.0f;
  }
  stage2_to_pdqf(L, W, lane, out, img);
  dihedral_tail(W, lane, mode, out, img, nullptr);
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

template __global__ void pdq_hash_kernel<float>(const float* __restrict__, uint64_t, uint32_t, uint32_t, HashOutputs);
template __global__ void pdq_hash_kernel<uint8_t>(const uint8_t* __restrict__, uint64_t, uint32_t, uint32_t,
                                                  HashOutputs);

hipError_t launch_hash(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_images, uint32_t mode,
                       const HashOutputs& out, hipStream_t stream) {
  if (n_images == 0) return hipSuccess;
  const dim3 grid((n_images + kWavesPerBlock - 1) / kWavesPerBlock);
  const dim3 block(64 * kWavesPerBlock);
  if (input_kind == kInputF32)
    hipLaunchKernelGGL(pdq_hash_kernel<float>, grid, block, 0, stream, static_cast<const float*>(planes),
                       plane_stride, n_images, mode, out);
  else
    hipLaunchKernelGGL(pdq_hash_kernel<uint8_t>, grid, block, 0, stream, static_cast<const uint8_t*>(planes),
                       plane_stride, n_images, mode, out);
  return hipGetLastError();
}

}  // namespace ypdq
