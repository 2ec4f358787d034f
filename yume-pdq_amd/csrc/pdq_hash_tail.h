// The tail of the fused hash kernels, shared by the wave-per-image and the
// block-per-image forms: DCT stage 2 from the stage-1 rows in LDS, the eight
// dihedral PDQFs, their gradient sums, quantisation thresholds and hash bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pdq_hash_common.h"

namespace ypdq {
// ---------------------------------------------------------------------------
// The tail: stage 2, dihedrals, gradients, quantisation, packing.
// ---------------------------------------------------------------------------
// The reference leaves the k = 15 stage-1 row in its tmp buffer.
__device__ __forceinline__ void write_tmp(const WaveLds& W, int lane, const HashOutputs& out, uint32_t img) {
  if (out.tmp) {
    float* dst = out.tmp + (size_t)img * 128;
    dst[2 * lane] = W.t[15 * kTS + 2 * lane];
    dst[2 * lane + 1] = W.t[15 * kTS + 2 * lane + 1];
  }
}

// Stage 2 (x86.rs:227-272) for one column of each quad: lane (k = lane & 15,
// jq = lane >> 4) computes out[k][4jq + r] -- 8 AVX2 lanes x 2 alternating
// accumulators over 16 chunks, then the horizontal tree.  d2: the DCT rows
// padded to 128 (slot 127 = the reference's over-read).
__device__ __forceinline__ void stage2_column(const float* __restrict__ d2, WaveLds& W, int lane, int r) {
  const int k = lane & 15, jq = lane >> 4;
  const int j = 4 * jq + r;
  float h[2][8];
#pragma unroll
  for (int p = 0; p < 2; p++)
#pragma unroll
    for (int l = 0; l < 8; l++) h[p][l] = 0.0f;
#pragma unroll
  for (int c = 0; c < 16; c++) {
    const f32x4* tp = reinterpret_cast<const f32x4*>(&W.t[k * kTS + 8 * c]);
    const f32x4* dp = reinterpret_cast<const f32x4*>(&d2[j * kTS + 8 * c]);
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

// Stage 2 by one wave: all four columns of each quad.
__device__ __forceinline__ void stage2_to_pdqf(const float* __restrict__ d2, WaveLds& W, int lane,
                                               const HashOutputs& out, uint32_t img) {
  write_tmp(W, lane, out, img);
#pragma unroll 1
  for (int r = 0; r < 4; r++) stage2_column(d2, W, lane, r);
  wave_lds_sync();
}

// The eight dihedral PDQFs into W.t (d-major, 256 each), from the NORMAL
// PDQF in W.p, element idx of orientations d0 .. d0 + nd - 1:
//   0 NORMAL P          1 FLOPPED  sc.P       2 ROTATED_180 sr.sc.P   3 FLIPPED sr.P
//   4 ROT90_FLOPPED P^T 5 ROT270   sr.P^T     6 FLOPPED_ROT270 sr.sc.P^T  7 ROT90 sc.P^T
// sr negates even rows (pdqf_negate_alt_rows::<true>, x86.rs:285-326); sc
// negates even columns (pdqf_negate_alt_cols::<false>) -- all 16 in exact
// mode, columns 0..7 only in the AVX2-compatible mode (x86.rs:330-376 loads
// 8 floats per row).  Negation is exact, so the sequence of in-place flips
// in visit_dihedrals reduces to these closed forms.
__device__ __forceinline__ void dihedral_pdqf_element(WaveLds& W, int idx, int d0, int nd, uint32_t mode) {
  const int ncols = mode == kDihedralAvx2 ? 8 : 16;
  const int i = idx >> 4, j = idx & 15;
  const float a = W.p[idx];
  const float at = W.p[j * 16 + i];
  const bool nr = (i & 1) == 0;
  const bool nc = (j & 1) == 0 && j < ncols;
  for (int d = d0; d < d0 + nd; d++) {
    const float v = d < 4 ? a : at;
    const int kind = d & 3;  // 0: as is, 1: sc (FLOPPED / ROT90), 2: sr.sc, 3: sr
    // the transposed four negate (none, sr, sr.sc, sc) for d = 4, 5, 6, 7
    const int k = d < 4 ? kind : (kind == 1 ? 3 : kind == 3 ? 1 : kind);
    const bool neg = k == 1 ? nc : k == 2 ? (nr != nc) : k == 3 ? nr : false;
    W.t[d * 256 + idx] = neg ? v * -1.0f : v;
  }
}

// Quality of the eight orientations (gradient sums, x86.rs:840-865, one
// sequential accumulation per orientation, lane d owns orientation d) into
// W.q and, when write_out, the output.
__device__ __forceinline__ void tail_quality(WaveLds& W, int lane, const HashOutputs& out, uint32_t img,
                                             bool write_out) {
  if (lane < 8) {
    const float q = adjust_quality(sum_of_gradients16(&W.t[lane * 256]));
    W.q[lane] = q;
    if (write_out) out.quality[(size_t)img * 8 + lane] = q;
  }
}

// Thresholds and hashes of the eight orientations in W.t, and the outputs
// other than quality.  thr_normal null means "use NORMAL's own quantisation
// threshold" for ROTATED_90_FLOPPED; the visit entry point passes the
// caller's threshold instead (src/lib.rs:403-404).
__device__ __forceinline__ void tail_thresholds_bits(WaveLds& W, int lane, const HashOutputs& out, uint32_t img,
                                                     const float* thr_normal) {
  // Median search (x86.rs:440-621) for dihedrals 0..3: group s =
  // (lane >> 3) & 3, the two halves of the wave summing the two sides.
  {
    const int s = (lane >> 3) & 3;
    float thr, bthr;
    median_search8_split(&W.t[s * 256], lane, thr, bthr);
    if (lane < 32 && (lane & 7) == 0) {
      W.th[s] = thr;
      W.bt[s] = bthr;
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
    for (int q = 0; q < 4; q++) word |= hash_byte(P, 4 * w + q, bt) << (8 * q);
    reinterpret_cast<uint32_t*>(out.hashes + (size_t)img * 256 + d * 32)[w] = word;
  }
  if (lane < 8) out.threshold[(size_t)img * 8 + lane] = W.th[lane];
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

// The latency form of the thresholds and bits, on four waves: wave s
// (0..3) searches the threshold of orientation s on all its lanes
// (median_search_rows) and writes hash s at s's bit threshold and hash s + 4
// at s's threshold -- the caller's for s == 0 when thr_normal is given
// (src/lib.rs:403-416, as tail_thresholds_bits) -- one byte per lane.
__device__ __forceinline__ void tail_rows_thresholds_bits(WaveLds& W, int lane, int s, const HashOutputs& out,
                                                          uint32_t img, const float* thr_normal) {
  float thr, bthr;
  median_search_rows(&W.t[s * 256], lane, thr, bthr);
  const float t4 = (s == 0 && thr_normal) ? *thr_normal : thr;
  const int d = lane < 32 ? s : s + 4, b = lane & 31;
  out.hashes[(size_t)img * 256 + d * 32 + b] = (uint8_t)hash_byte(&W.t[d * 256], b, lane < 32 ? bthr : t4);
  if (lane == 0) {
    out.threshold[(size_t)img * 8 + s] = thr;
    out.threshold[(size_t)img * 8 + s + 4] = t4;
  }
}

// The optional PDQF outputs (NORMAL's, and all eight orientations'), thread
// t0 of n.
__device__ __forceinline__ void tail_copy_pdqfs(const WaveLds& W, int t0, int n, const HashOutputs& out,
                                                uint32_t img) {
  if (out.pdqf)
    for (int e = t0; e < 256; e += n) out.pdqf[(size_t)img * 256 + e] = W.p[e];
  if (out.dihedral_pdqf) {
    float* dst = out.dihedral_pdqf + (size_t)img * 8 * 256;
    for (int e = t0; e < 8 * 256; e += n) dst[e] = W.t[e];
  }
}

// The whole tail on one wave: from the NORMAL PDQF in W.p, the eight
// dihedral PDQFs, their qualities, thresholds and hashes.
__device__ __forceinline__ void dihedral_tail(WaveLds& W, int lane, uint32_t mode, const HashOutputs& out, uint32_t img,
                              const float* thr_normal) {
#pragma unroll
  for (int e = 0; e < 4; e++) dihedral_pdqf_element(W, lane + 64 * e, 0, 8, mode);
  wave_lds_sync();
  tail_quality(W, lane, out, img, true);
  tail_thresholds_bits(W, lane, out, img, thr_normal);
}

}  // namespace ypdq
