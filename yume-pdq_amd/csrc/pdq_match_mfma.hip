// Matrix-core Hamming matcher for large query batches (gfx950, fp4 MFMA).
//
// The reference's predicate is popcount(q ^ r) <= threshold over 256 bits
// (src/matching/mod.rs:232-257, 336-362).  Writing each bit as +1/-1, the
// dot product of two such vectors over the first 128 bits is
//     dot128 = 128 - 2 * d128,     d128 = popcount of the first 128 XOR bits,
// and d128 <= d256 always, so d256 <= t  implies  dot128 >= 128 - 2t.
// +-1 are exact fp4 (e2m1) values and |dot128| <= 128 is exact in the fp32
// accumulator, so one v_mfma_scale_f32_16x16x128_f8f6f4 computes d128 for a
// 16-query x 16-row tile exactly.  Tiles whose maximum clears the bound
// (for random rows with t = 31: about 1e-9 of pairs, plus the true hits) are
// re-checked pair by pair with the full 256-bit integer distance, so the
// result is the exact set the reference defines -- the matrix pass is only a
// filter that can never drop a pair.  Used for t < 64 (above that the prefix
// bound passes everything) and batches of >= 32 queries.
//
// Operand layout (measured on gfx950): lane l holds row/column (l & 15) and
// K quarter (l >> 4), i.e. 32 fp4 values = one 32-bit word of the 128-bit
// prefix; the output element C[m][n] is in lane (n + 16*(m >> 2)), register
// (m & 3).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <utility>

#include "pdq_device.h"

namespace ypdq {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaWaves = 4;
constexpr int kMfmaTilesPerWave = 64;  // 16-row tiles streamed per wave per block

// fp4 e2m1 codes: 0x2 = +1.0 (bit set), 0xA = -1.0 (bit clear), 0x0 = 0.
__device__ __forceinline__ uint32_t spread_byte(uint32_t b) {
  uint32_t x = b & 0xffu;
  x = (x | (x << 12)) & 0x000f000fu;
  x = (x | (x << 6)) & 0x03030303u;
  x = (x | (x << 3)) & 0x11111111u;
  return 0xaaaaaaaau ^ (x << 3);
}

struct Expanded {
  uint32_t d[4];
};

__device__ __forceinline__ void expand_word(uint32_t w, const uint32_t* __restrict__ lut, Expanded& e) {
  e.d[0] = lut[w & 0xffu];
  e.d[1] = lut[(w >> 8) & 0xffu];
  e.d[2] = lut[(w >> 16) & 0xffu];
  e.d[3] = lut[w >> 24];
}

__device__ __forceinline__ v4f mfma_prefix(const Expanded& a, const Expanded& b) {
  v8i av = {(int)a.d[0], (int)a.d[1], (int)a.d[2], (int)a.d[3], 0, 0, 0, 0};
  v8i bv = {(int)b.d[0], (int)b.d[1], (int)b.d[2], (int)b.d[3], 0, 0, 0, 0};
  v4f c = {0.0f, 0.0f, 0.0f, 0.0f};
  // cbsz = blgp = 4: both operands fp4 e2m1; scales 127 = 2^0
  return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(av, bv, c, 4, 4, 0, 127, 0, 127);
}

__device__ __forceinline__ uint32_t full_distance(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b) {
  uint32_t d = 0;
#pragma unroll
  for (int w = 0; w < 8; w++) d += (uint32_t)__builtin_popcount(a[w] ^ b[w]);
  return d;
}

// Exact check of query tile t's candidates in this lane's four outputs.
__device__ __forceinline__ void check_candidates(v4f cv, int t, uint64_t tbase, uint32_t qbase, int rn, int kq,
                                                 float bound, const uint32_t* __restrict__ queries,
                                                 const uint32_t* __restrict__ db, uint32_t n_queries, uint64_t n_rows,
                                                 uint32_t threshold, MatchHit* __restrict__ hits, uint64_t capacity,
                                                 unsigned long long* __restrict__ counter, uint64_t index_base) {
  const uint64_t row = tbase + rn;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float v = r == 0 ? cv.x : r == 1 ? cv.y : r == 2 ? cv.z : cv.w;
    if (v >= bound) {
      const uint32_t qi = qbase + 16 * t + 4 * kq + r;
      if (qi < n_queries && row < n_rows) {
        const uint32_t d = full_distance(queries + (size_t)qi * 8, db + row * 8);
        if (d <= threshold) {
          const unsigned long long slot = atomicAdd(counter, 1ull);
          if (slot < capacity) {
            MatchHit h;
            h.query = qi;
            h.distance = d;
            h.index = index_base + row;
            hits[slot] = h;
          }
        }
      }
    }
  }
}

// The same out of line, for the per-wave layout, whose rare path keeps the
// stream's query operands (inlined there it spills registers into the tile
// loop).  (The shared-rows streaming kernel checks nothing itself: a call
// anywhere in it, even one that never runs, cost its tile loop about 12 %,
// and inline checks 9 %; its second kernel checks with stage_candidates.)
__device__ __noinline__ void resolve_candidates(v4f cv, int t, uint64_t tbase, uint32_t qbase, int rn, int kq,
                                                float bound, const uint32_t* __restrict__ queries,
                                                const uint32_t* __restrict__ db, uint32_t n_queries, uint64_t n_rows,
                                                uint32_t threshold, MatchHit* __restrict__ hits, uint64_t capacity,
                                                unsigned long long* __restrict__ counter, uint64_t index_base) {
  check_candidates(cv, t, tbase, qbase, rn, kq, bound, queries, db, n_queries, n_rows, threshold, hits, capacity,
                   counter, index_base);
}

// The shared-rows layout's second kernel appends to its block's hit list in
// LDS (stage, n_stage) and takes its slots in the hit buffer once per block:
// thousands of appends to the one global counter from a short kernel
// serialise on it (measured: 49 of 68 microseconds for 4938 hits).  The list
// (64 KiB) holds 4096 hits: with one block per CU, about a million per
// launch.  Hits past it go straight to the buffer.
constexpr uint32_t kStageHits = 4096;

__device__ __forceinline__ void stage_candidates(v4f cv, int t, uint64_t tbase, uint32_t qbase, int rn, int kq,
                                                float bound, const uint32_t* __restrict__ queries,
                                                const uint32_t* __restrict__ db, uint32_t n_queries, uint64_t n_rows,
                                                uint32_t threshold, MatchHit* stage, uint32_t* n_stage,
                                                MatchHit* __restrict__ hits, uint64_t capacity,
                                                unsigned long long* __restrict__ counter, uint64_t index_base) {
  const uint64_t row = tbase + rn;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const float v = r == 0 ? cv.x : r == 1 ? cv.y : r == 2 ? cv.z : cv.w;
    if (v >= bound) {
      const uint32_t qi = qbase + 16 * t + 4 * kq + r;
      if (qi < n_queries && row < n_rows) {
        const uint32_t d = full_distance(queries + (size_t)qi * 8, db + row * 8);
        if (d <= threshold) {
          MatchHit h;
          h.query = qi;
          h.distance = d;
          h.index = index_base + row;
          const uint32_t s = atomicAdd(n_stage, 1u);
          if (s < kStageHits) {
            stage[s] = h;
          } else {
            const unsigned long long slot = atomicAdd(counter, 1ull);
            if (slot < capacity) hits[slot] = h;
          }
        }
      }
    }
  }
}

// Maximum over one 16-row tile of the dot products with every query tile of
// the wave, kept in four running maxima (query tile t feeds mx[t % 4]) so the
// max instructions of consecutive products do not wait on each other.  The
// products are integers in [-128, 128] and the bound is positive, so
// comparing float bit patterns as signed integers orders them exactly
// (negative floats are negative integers); integer max3 needs no NaN
// canonicalisation.  (Element-wise __float_as_int: a bit cast of a vector
// element expression is miscompiled to element 0 by this toolchain.)
constexpr int kMaxChains = 4;

template <int QT>
__device__ __forceinline__ void tile_max(const Expanded (&qa)[QT], const Expanded& rb, int (&mx)[kMaxChains]) {
#pragma unroll
  for (int t = 0; t < QT; t++) {
    const v4f c = mfma_prefix(qa[t], rb);
    const int c0 = __float_as_int(c.x), c1 = __float_as_int(c.y);
    const int c2 = __float_as_int(c.z), c3 = __float_as_int(c.w);
    int& m = mx[t % kMaxChains];
    m = max(max(m, c0), c1);
    m = max(max(m, c2), c3);
  }
}

// Packed candidate flags: the filter of the streaming loops.  The products
// of one row tile with two query tiles share one accumulator.  The first
// MFMA adds dot_a to C0 = 2^23 + K + 512 K, the second adds 512 dot_b (its
// query operand scaled by 2^9 through the MFMA's block scale), so each
// result is
//     2^23 + f_a + 512 f_b,   f = dot + K,   K = 256 - bound,
// an integer in [2^23, 2^24) -- as is every partial sum, in any order of
// accumulation -- so it is exact, and its low 23 bits are the two fields.
// dot is in [-128, 128], so f is in [128 - bound, 384 - bound], below 512
// (no carry between fields; bound = 128 - 2t is in [2, 128] for t < 64), and
// f >= 256 exactly when dot >= bound: bit 8 flags pair a, bit 17 pair b.
// One OR of the four results' bits then tests 8 pairs, where a maximum
// needs one instruction per 2 values.
constexpr uint32_t kFlagMask = (1u << 8) | (1u << 17);
constexpr int kFlagChains = 4;
constexpr int kScale1 = 127;      // 2^0
constexpr int kScale512 = 127 + 9;  // 2^9

__device__ __forceinline__ v4f flags_c0(float bound) {
  const float k = 256.0f - bound;
  const float c = 8388608.0f + k + 512.0f * k;  // exact: an integer below 2^24
  return v4f{c, c, c, c};
}

template <int kScaleA>
__device__ __forceinline__ v4f mfma_acc(const Expanded& a, const Expanded& b, const v4f& c) {
  v8i av = {(int)a.d[0], (int)a.d[1], (int)a.d[2], (int)a.d[3], 0, 0, 0, 0};
  v8i bv = {(int)b.d[0], (int)b.d[1], (int)b.d[2], (int)b.d[3], 0, 0, 0, 0};
  return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(av, bv, c, 4, 4, 0, kScaleA, 0, kScale1);
}

// The empty asm pins m here: without it the compiler gathers every fold of a
// group into one OR tree at the group's end, which keeps all the group's
// results live at once.
__device__ __forceinline__ void fold_flags(const v4f& v, uint32_t& m) {
  m = m | __float_as_uint(v.x) | __float_as_uint(v.y);
  m = m | __float_as_uint(v.z) | __float_as_uint(v.w);
  asm volatile("" : "+v"(m));
}

// The query pair of pair P of a row tile (of QT / 2 pairs).  Pair P's
// products go to result register P % 4 and are folded into flag chain P % 4;
// at QT = 16 the pairs are ordered so that chains 0 and 1 hold query tiles
// 0-7 and chains 2 and 3 tiles 8-15, so one vote per half tells which half
// of the query tiles has candidates.
template <int kHalf>
__device__ constexpr int query_pair(int p) {
  if constexpr (kHalf == 8) {
    constexpr int order[8] = {0, 1, 4, 5, 2, 3, 6, 7};
    return order[p % 8];
  }
  return p % kHalf;
}

// Pair P (row tile P / (QT/2), query tiles 2 query_pair(P) and +1): its first
// product issues at step P into result register P % 4, its second at step
// P + 1 (one product after, so the accumulation does not wait), and its bits
// are folded at step P + 4, before the register takes pair P + 4.  The four
// result registers are named, not an array, so the compiler keeps each in
// place.
template <int QT, int kTiles, int S>
__device__ __forceinline__ void group_flags_step(const Expanded (&qa)[QT], const Expanded (&rb)[kTiles],
                                                 const v4f& c0, uint32_t (&acc)[kFlagChains], v4f& r0, v4f& r1,
                                                 v4f& r2, v4f& r3) {
  constexpr int kHalf = QT / 2;
  constexpr int kN = kHalf * kTiles;  // pairs
  v4f& cur = (S % 4 == 0) ? r0 : (S % 4 == 1) ? r1 : (S % 4 == 2) ? r2 : r3;
  v4f& prev = (S % 4 == 1) ? r0 : (S % 4 == 2) ? r1 : (S % 4 == 3) ? r2 : r3;
  if constexpr (S >= 4) fold_flags(cur, acc[(S - 4) % kFlagChains]);
  if constexpr (S < kN) cur = mfma_acc<kScale1>(qa[2 * query_pair<kHalf>(S)], rb[S / kHalf], c0);
  if constexpr (S >= 1 && S - 1 < kN)
    prev = mfma_acc<kScale512>(qa[2 * query_pair<kHalf>(S - 1) + 1], rb[(S - 1) / kHalf], prev);
  // keep this order: hoisting the products runs out of registers
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (S + 1 < kN + 4) group_flags_step<QT, kTiles, S + 1>(qa, rb, c0, acc, r0, r1, r2, r3);
}

template <int QT, int kTiles>
__device__ __forceinline__ void group_flags(const Expanded (&qa)[QT], const Expanded (&rb)[kTiles], const v4f& c0,
                                            uint32_t (&acc)[kFlagChains]) {
  static_assert(QT % 2 == 0 && (QT / 2) * kTiles >= 4, "pairs of query tiles, four pairs in flight");
  v4f r0, r1, r2, r3;
  group_flags_step<QT, kTiles, 0>(qa, rb, c0, acc, r0, r1, r2, r3);
}

__device__ __forceinline__ bool flags_set(const uint32_t (&acc)[kFlagChains]) {
  return ((acc[0] | acc[1] | acc[2] | acc[3]) & kFlagMask) != 0;
}

// Flags of query tiles 0-7 (chains 0 and 1) and 8-15 (chains 2 and 3) at
// QT = 16.
__device__ __forceinline__ bool flags_set_low(const uint32_t (&acc)[kFlagChains]) {
  return ((acc[0] | acc[1]) & kFlagMask) != 0;
}
__device__ __forceinline__ bool flags_set_high(const uint32_t (&acc)[kFlagChains]) {
  return ((acc[2] | acc[3]) & kFlagMask) != 0;
}

// Three products per accumulator (sixteen query tiles, thresholds 28..63).
// The products of one row tile with query tiles a, b, c add dot_a (scale
// 2^0), 2^7 dot_b and 2^15 dot_c to
//     C0 = 2^23 + K_a + 2^8 K_b + 2^16 K_c,
//     K_a = 128 - bound,  K_b = 129 - bound/2,  K_c = 64 - bound/2,
// so, dot being even, the result is 2^23 + f_a + 2^8 f_b + 2^16 f_c with
// f_a = dot_a + K_a, f_b = dot_b/2 + K_b, f_c = dot_c/2 + K_c.  Every
// partial sum, in any order, is an integer in (2^22, 2^24): exact.
//   bit 7:  f_a >= 128, i.e. dot_a >= bound (f_a <= 254: no carry);
//   bit 15: f_b >= 128, i.e. dot_b >= bound - 2 (f_b in [1, 192]);
//   bit 22: f_c >= 64, i.e. dot_c >= bound (f_c <= 127).
// f_a or f_c below zero (dot < bound - 128: a prefix distance above 64 + t,
// about 6e-9 of random pairs at t = 31) borrows: a borrow from f_a shows
// in bit 7 (f_a + 256 >= 128) and lowers f_b by one, which the slack of
// K_b (a flag from dot_b >= bound - 2) absorbs; a borrow from f_c takes the
// result below 2^23, which changes the exponent's lowest bit, bit 23.  So a
// pair at the bound always sets one of the four bits, and the other causes
// are rare false candidates that the recheck discards.  tests/test_flag_fields.py
// checks this over every (dot_a, dot_b, dot_c) and threshold.
// Per row tile the sixteen query tiles form accumulators of (0,1,2),
// (3,4,5), (6,7), (8,9,10), (11,12,13), (14,15): twelve folds instead of
// sixteen; accumulators 0-2 hold the first half of the query tiles.
constexpr uint32_t kFlagMaskTri = (1u << 7) | (1u << 15) | (1u << 22) | (1u << 23);
constexpr uint32_t kTriMinThreshold = 28;  // below it, distant pairs borrow too often
constexpr int kScale128 = 127 + 7;         // 2^7
constexpr int kScale32k = 127 + 15;        // 2^15

__device__ __forceinline__ v4f flags_c0_tri(float bound) {
  const int b = (int)bound;
  const int c = 8388608 + (128 - b) + 256 * (129 - b / 2) + 65536 * (64 - b / 2);  // below 2^24: exact
  const float f = (float)c;
  return v4f{f, f, f, f};
}

// Accumulator m (0..5) of a row tile: its first query tile and its size.
__device__ constexpr int tri_first(int m) { return m < 3 ? 3 * m : 8 + 3 * (m - 3); }
__device__ constexpr int tri_size(int m) { return (m == 2 || m == 5) ? 2 : 3; }
// Flag chain of accumulator P: two per half of the query tiles.
__device__ constexpr int tri_chain(int p) { return 2 * ((p % 6) >= 3 ? 1 : 0) + (p & 1); }

// Accumulator P (row tile P / 6, accumulator P % 6): its first product issues
// at step P into result register P % 4, its second at step P + 1 and its
// third at step P + 2, and its bits are folded at step P + 4, before the
// register takes accumulator P + 4.
template <int kTiles, int S>
__device__ __forceinline__ void group_flags_tri_step(const Expanded (&qa)[16], const Expanded (&rb)[kTiles],
                                                     const v4f& c0, uint32_t (&acc)[kFlagChains], v4f& r0, v4f& r1,
                                                     v4f& r2, v4f& r3) {
  constexpr int kN = 6 * kTiles;  // accumulators
  v4f& cur = (S % 4 == 0) ? r0 : (S % 4 == 1) ? r1 : (S % 4 == 2) ? r2 : r3;
  v4f& p1 = (S % 4 == 1) ? r0 : (S % 4 == 2) ? r1 : (S % 4 == 3) ? r2 : r3;
  v4f& p2 = (S % 4 == 2) ? r0 : (S % 4 == 3) ? r1 : (S % 4 == 0) ? r2 : r3;
  if constexpr (S >= 4) fold_flags(cur, acc[tri_chain(S - 4)]);
  if constexpr (S < kN) cur = mfma_acc<kScale1>(qa[tri_first(S % 6)], rb[S / 6], c0);
  if constexpr (S >= 1 && S - 1 < kN)
    p1 = mfma_acc<kScale128>(qa[tri_first((S - 1) % 6) + 1], rb[(S - 1) / 6], p1);
  if constexpr (S >= 2 && S - 2 < kN) {
    if constexpr (tri_size((S - 2) % 6) == 3)
      p2 = mfma_acc<kScale32k>(qa[tri_first((S - 2) % 6) + 2], rb[(S - 2) / 6], p2);
  }
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (S + 1 < kN + 4) group_flags_tri_step<kTiles, S + 1>(qa, rb, c0, acc, r0, r1, r2, r3);
}

template <int kTiles>
__device__ __forceinline__ void group_flags_tri(const Expanded (&qa)[16], const Expanded (&rb)[kTiles], const v4f& c0,
                                                uint32_t (&acc)[kFlagChains]) {
  v4f r0, r1, r2, r3;
  group_flags_tri_step<kTiles, 0>(qa, rb, c0, acc, r0, r1, r2, r3);
}

// The rare path: recompute one tile's products and re-check every pair that
// clears the bound with the exact 256-bit distance.
template <int QT>
__device__ __forceinline__ void resolve_tile(const Expanded (&qa)[QT], const Expanded& rb, uint64_t tbase,
                                             uint32_t qbase, int rn, int kq, float bound,
                                             const uint32_t* __restrict__ queries, const uint32_t* __restrict__ db,
                                             uint32_t n_queries, uint64_t n_rows, uint32_t threshold,
                                             MatchHit* __restrict__ hits, uint64_t capacity,
                                             unsigned long long* __restrict__ counter, uint64_t index_base) {
  for (int t = 0; t < QT; t++) {
    const v4f c = mfma_prefix(qa[t], rb);
    resolve_candidates(c, t, tbase, qbase, rn, kq, bound, queries, db, n_queries, n_rows, threshold, hits, capacity,
                       counter, index_base);
  }
}

// Where a wave's row words come from.  Lane (rn, kq) needs word kq of row
// 16 * tile + rn.  Raw rows: db[r * 8 + kq] -- 16 of every 32 bytes,
// scattered over 4 cache lines per tile.  Prefix plane (the first 16 bytes of
// every row, packed): plane[r * 4 + kq], so a tile is 256 contiguous bytes and
// a launch reads half the bytes.
template <bool kPlane>
struct RowSource {
  const uint32_t* __restrict__ p;
  __device__ __forceinline__ RowSource(const uint32_t* db, const uint32_t* plane, uint64_t row0, int lane)
      : p(kPlane ? plane + (row0 + (lane & 15)) * 4 + (lane >> 4) : db + (row0 + (lane & 15)) * 8 + (lane >> 4)) {}
  __device__ __forceinline__ uint32_t operator()(uint32_t tile) const {
    return p[(size_t)tile * (kPlane ? 64 : 128)];
  }
};

// Stream the wave's tiles: loads run kAhead tiles ahead, each tile's word is
// expanded to fp4 codes just before its products, and the bound is tested
// once per kGroup tiles.  A group that reaches it (about once per
// million tiles on random data, plus the true hits) is only noted in a
// wave-uniform bit mask; its tiles are recomputed and checked pair by pair
// after the stream, so the loop itself holds no call and no branch.
template <int QT, bool kFull, bool kPlane>
__device__ __forceinline__ void stream_tiles(const Expanded (&qa)[QT], const uint32_t* __restrict__ lut,
                                             const RowSource<kPlane>& src, const uint32_t* __restrict__ db,
                                             uint64_t row0, uint64_t n_rows, uint32_t qbase, int rn, int kq,
                                             float bound, const uint32_t* __restrict__ queries, uint32_t n_queries,
                                             uint32_t threshold, MatchHit* __restrict__ hits, uint64_t capacity,
                                             unsigned long long* __restrict__ counter, uint64_t index_base) {
  constexpr int kAhead = 8;
  constexpr int kGroup = kFull ? 4 : 1;
  static_assert(kMfmaTilesPerWave % kAhead == 0 && kAhead % kGroup == 0, "tile loop shape");
  static_assert(kMfmaTilesPerWave / kGroup <= 64, "one mask bit per group");
  // rows of this wave that exist (the last wave of the database only)
  const uint32_t left = kFull ? 0u : (uint32_t)(n_rows - row0);
  const uint32_t tiles = kFull ? (uint32_t)kMfmaTilesPerWave : (left + 15) / 16;
  auto fetch = [&](uint32_t tile) -> uint32_t {
    if (kFull || kPlane) return src(tile);  // the plane is padded to whole tiles
    return (tile * 16 + rn < left) ? src(tile) : 0u;
  };
  uint32_t ring[kAhead];
#pragma unroll
  for (int a = 0; a < kAhead; a++) ring[a] = fetch(a < (int)tiles ? a : tiles - 1);
  int mx[kMaxChains];
  uint32_t acc[kFlagChains];
#pragma unroll
  for (int m = 0; m < kMaxChains; m++) mx[m] = INT32_MIN;
#pragma unroll
  for (int m = 0; m < kFlagChains; m++) acc[m] = 0u;
  const int ibound = __float_as_int(bound);
  const v4f c0 = flags_c0(bound);
  uint64_t pending = 0;  // groups whose products reached the bound
#pragma unroll 1
  for (uint32_t tile = 0; tile < tiles; tile += kAhead) {
#pragma unroll
    for (int a = 0; a < kAhead; a++) {
      const uint32_t t = tile + a;
      if (!kFull && t >= tiles) break;
      Expanded rb;
      expand_word(ring[a], lut, rb);
      if (!kFull && t * 16 + rn >= left) rb.d[0] = rb.d[1] = rb.d[2] = rb.d[3] = 0;
      // Always issue the load (the last kAhead tiles re-read tile `tiles - 1`
      // and drop it): a load under a branch makes the wait counters unknown
      // to the compiler, which then waits for every load in flight.
      ring[a] = fetch(t + kAhead < tiles ? t + kAhead : tiles - 1);
      if constexpr (QT >= 8) {
        const Expanded rbs[1] = {rb};
        group_flags<QT, 1>(qa, rbs, c0, acc);
      } else {
        tile_max<QT>(qa, rb, mx);
      }
      if (a % kGroup == kGroup - 1) {
        const int mall = max(max(mx[0], mx[1]), max(mx[2], mx[3]));
        pending |= (uint64_t)(wave_any(mall >= ibound || flags_set(acc)) ? 1u : 0u) << (t / kGroup);
#pragma unroll
        for (int m = 0; m < kMaxChains; m++) mx[m] = INT32_MIN;
#pragma unroll
        for (int m = 0; m < kFlagChains; m++) acc[m] = 0u;
      }
    }
  }
  while (pending) {
    const uint32_t g = (uint32_t)__builtin_ctzll(pending);
    pending &= pending - 1;
    for (uint32_t tg = g * kGroup; tg < (g + 1) * kGroup; tg++) {
      Expanded eg;
      expand_word(fetch(tg), lut, eg);
      if (!kFull && tg * 16 + rn >= left) eg.d[0] = eg.d[1] = eg.d[2] = eg.d[3] = 0;
      resolve_tile<QT>(qa, eg, row0 + (uint64_t)tg * 16, qbase, rn, kq, bound, queries, db, n_queries, n_rows,
                       threshold, hits, capacity, counter, index_base);
    }
  }
}

// One work item: row chunk `chunk` of a database (kMfmaWaves waves x
// kMfmaTilesPerWave tiles x 16 rows; each wave, `wave` = 0 .. kMfmaWaves - 1,
// streams its own rows) against query group `group` (QT x 16 queries).
template <int QT, bool kPlane>
__device__ __forceinline__ void mfma_item(const uint32_t* __restrict__ lut, const uint32_t* __restrict__ db,
                                          const uint32_t* __restrict__ plane, uint64_t n_rows, uint64_t chunk,
                                          uint32_t group, const uint32_t* __restrict__ queries, uint32_t n_queries,
                                          uint32_t threshold, MatchHit* __restrict__ hits, uint64_t capacity,
                                          unsigned long long* __restrict__ counter, uint64_t index_base,
                                          int wave) {
  const int lane = threadIdx.x & 63;
  const int rn = lane & 15;  // operand row (query m / db row n) of this lane
  const int kq = lane >> 4;  // K quarter = word index of the prefix
  const float bound = (float)(128 - 2 * (int)threshold);

  const uint64_t wave_rows = (uint64_t)kMfmaTilesPerWave * 16;
  const uint64_t row0 = (chunk * kMfmaWaves + wave) * wave_rows;
  if (row0 >= n_rows) return;

  // query operands: tile t, query qbase + 16t + rn, word kq
  const uint32_t qbase = group * (uint32_t)(QT * 16);
  // All QT words are loaded before any is used (loads clamped in bounds rather
  // than branched around, so they overlap instead of waiting one by one).
  uint32_t qw[QT];
#pragma unroll
  for (int t = 0; t < QT; t++) {
    const uint32_t qi = qbase + 16 * t + rn;
    qw[t] = queries[(size_t)(qi < n_queries ? qi : n_queries - 1) * 8 + kq];
  }
  Expanded qa[QT];
#pragma unroll
  for (int t = 0; t < QT; t++) {
    expand_word(qw[t], lut, qa[t]);
    if (qbase + 16 * t + rn >= n_queries)
      qa[t].d[0] = qa[t].d[1] = qa[t].d[2] = qa[t].d[3] = 0;  // zero rows never reach the bound
  }

  const RowSource<kPlane> src(db, plane, row0, lane);
  // Only the wave that reaches the end of the database needs bounds checks.
  if (row0 + wave_rows <= n_rows)
    stream_tiles<QT, true, kPlane>(qa, lut, src, db, row0, n_rows, qbase, rn, kq, bound, queries, n_queries,
                                   threshold, hits, capacity, counter, index_base);
  else
    stream_tiles<QT, false, kPlane>(qa, lut, src, db, row0, n_rows, qbase, rn, kq, bound, queries, n_queries,
                                    threshold, hits, capacity, counter, index_base);
}

// Shared-rows work item, for batches of kMfmaWaves or more query groups: the
// block's waves stream the same kSharedTiles 16-row tiles (rows
// sub * kSharedRows ...), wave w against query group kMfmaWaves * sgroup + w.
// For every group of kGroupTiles tiles, wave w loads and expands tiles
// kGroupTiles * g + kMfmaWaves * s + w (s < kStagePerWave) once, into LDS
// (double-buffered, one barrier per group), and every wave reads the group's
// expanded tiles with one 16-byte LDS read each.  Each wave would otherwise
// load and expand every tile itself; shared, a product costs a quarter of the
// loads and LUT reads, which is what more query tiles per wave would buy
// without their registers.  A wave whose group lies past the last query still
// loads and expands its tiles for the others.  The rare path is per wave, as
// in mfma_item, reading its tiles' rows directly.
constexpr int kSharedTiles = 64;
constexpr uint64_t kSharedRows = (uint64_t)kSharedTiles * 16;
constexpr int kStagePerWave = 1;                          // tiles a wave stages per group
constexpr int kGroupTiles = kMfmaWaves * kStagePerWave;  // tiles per barrier and per vote
static_assert(kSharedTiles % kGroupTiles == 0 && kSharedTiles / kGroupTiles <= 32,
              "one bit of a 32-bit record mask per group of tiles");
// At QT = 16 a record's mask holds query tiles 0-7's groups in its low bits
// and tiles 8-15's from bit kHalfShift.
constexpr int kHalfShift = 16;
static_assert(kSharedTiles / kGroupTiles <= kHalfShift, "two halves in a record mask");
static_assert((uint64_t)kMfmaWaves * kMfmaTilesPerWave * 16 % kSharedRows == 0, "sub-chunks of a chunk");
constexpr uint64_t kSubsPerChunk = (uint64_t)kMfmaWaves * kMfmaTilesPerWave * 16 / kSharedRows;

struct SharedRowsLds {
  u32x4 ex[2][kGroupTiles][64];  // two groups of expanded tiles, lane-major
};

// The shared-rows stream of one sub-chunk.  kFull (all kSharedTiles tiles
// exist, every sub-chunk but a database's last) has a loop with fixed bounds
// and no branch: a branch in the loop makes the compiler wait for every load
// in flight at each group.
template <int QT, bool kFull, bool kPlane, bool kTri>
__device__ __forceinline__ void stream_shared(const Expanded (&qa)[QT], const uint32_t* __restrict__ lut,
                                              SharedRowsLds& S, const RowSource<kPlane>& src,
                                              const uint32_t* __restrict__ db, uint64_t row0, uint64_t n_rows,
                                              uint32_t qbase, int wave, int lane, float bound,
                                              uint4* __restrict__ recs, uint32_t* __restrict__ rec_count,
                                              uint32_t rec_cap, uint64_t rec_sub_base) {
  const int rn = lane & 15, kq = lane >> 4;
  const uint64_t left = n_rows - row0;
  const uint32_t tiles = kFull ? (uint32_t)kSharedTiles : (uint32_t)((left + 15) / 16);
  const uint32_t ngroups =
      kFull ? (uint32_t)(kSharedTiles / kGroupTiles) : (tiles + kGroupTiles - 1) / kGroupTiles;
  // This wave's word of its s-th tile of group g, loaded unconditionally: the
  // tile is clamped into the sub-chunk (the plane is padded to whole tiles)
  // and a raw row past the end is clamped to the last row; stage() zeroes
  // both.
  auto word = [&](uint32_t g, int s) -> uint32_t {
    const uint32_t tt = min(kGroupTiles * g + kMfmaWaves * s + (uint32_t)wave, tiles - 1);
    if (kPlane || kFull) return src(tt);
    const uint64_t r = min((uint64_t)tt * 16 + rn, left - 1);
    return db[(row0 + r) * 8 + kq];
  };
  // staging in two halves: the table reads (expand_word), then the store,
  // with the group's products between them
  auto store = [&](uint32_t g, int s, const Expanded& e, int buf) {
    const uint32_t tt = kGroupTiles * g + kMfmaWaves * s + (uint32_t)wave;
    // zero operands (not the expansion of a zero word, which is all -1),
    // selected without a branch
    const bool have = kFull ? tt < tiles : (tt < tiles && (uint64_t)tt * 16 + rn < left);
    const uint32_t keep = have ? 0xffffffffu : 0u;
    u32x4 v;
    v.x = e.d[0] & keep;
    v.y = e.d[1] & keep;
    v.z = e.d[2] & keep;
    v.w = e.d[3] & keep;
    S.ex[buf][kMfmaWaves * s + wave][lane] = v;
  };
  auto stage = [&](uint32_t g, int s, uint32_t w, int buf) {
    Expanded e;
    expand_word(w, lut, e);
    store(g, s, e, buf);
  };
  // ring[a] holds this wave's words of group g0 + a + 1, loaded kRing groups
  // ahead; with the inner loop unrolled the slots are fixed registers
  constexpr int kRing = 4;
  static_assert((kSharedTiles / kGroupTiles) % kRing == 0, "whole rounds of the ring");
  uint32_t ring[kRing][kStagePerWave];
#pragma unroll
  for (int a = 0; a < kRing; a++)
#pragma unroll
    for (int st = 0; st < kStagePerWave; st++) ring[a][st] = word(a + 1 < (int)ngroups ? a + 1 : ngroups - 1, st);
#pragma unroll
  for (int st = 0; st < kStagePerWave; st++) stage(0, st, word(0, st), 0);
  __syncthreads();
  uint32_t acc[kFlagChains];
#pragma unroll
  for (int m = 0; m < kFlagChains; m++) acc[m] = 0u;
  static_assert(!kTri || QT == 16, "three products per accumulator at sixteen query tiles");
  const v4f c0 = kTri ? flags_c0_tri(bound) : flags_c0(bound);
  uint64_t pending = 0;  // groups of tiles whose products reached the bound
#pragma unroll 1
  for (uint32_t g0 = 0; g0 < ngroups; g0 += kRing) {
#pragma unroll
    for (int a = 0; a < kRing; a++) {
      const uint32_t g = g0 + a;
      if (!kFull && g >= ngroups) break;  // the same for the whole block
      // The group's expanded tiles, read together so their latencies
      // overlap, and before the staging below: LDS operations complete in
      // order, so the products then wait for these reads only, not for the
      // table reads and the write of the next group.
      u32x4 vt[kGroupTiles];
#pragma unroll
      for (int j = 0; j < kGroupTiles; j++) vt[j] = S.ex[g & 1][j][lane];
      __builtin_amdgcn_sched_barrier(0);
      // The next group's table reads are issued now and its store follows
      // the products, so neither waits for the other.  The buffer written
      // here was last read before the previous barrier; after the last group
      // this stages rows nobody reads.
      Expanded next[kStagePerWave];
#pragma unroll
      for (int st = 0; st < kStagePerWave; st++) expand_word(ring[a][st], lut, next[st]);
      __builtin_amdgcn_sched_barrier(0);
      Expanded rb[kGroupTiles];
#pragma unroll
      for (int j = 0; j < kGroupTiles; j++) {
        rb[j].d[0] = vt[j].x;
        rb[j].d[1] = vt[j].y;
        rb[j].d[2] = vt[j].z;
        rb[j].d[3] = vt[j].w;
      }
      if constexpr (kTri)
        group_flags_tri<kGroupTiles>(qa, rb, c0, acc);
      else
        group_flags<QT, kGroupTiles>(qa, rb, c0, acc);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int st = 0; st < kStagePerWave; st++) {
        if (kFull || g + 1 < ngroups) store(g + 1, st, next[st], (g + 1) & 1);
        ring[a][st] = word(g + 1 + kRing < ngroups ? g + 1 + kRing : ngroups - 1, st);
      }
      // the vote is a wave-uniform mask, so the pending bits stay in scalar
      // registers; at QT = 16 one vote per half of the query tiles, the
      // second half's bits kHalfShift above the first's
      if constexpr (kTri) {
        const uint64_t lo = __ballot(((acc[0] | acc[1]) & kFlagMaskTri) != 0);
        const uint64_t hi = __ballot(((acc[2] | acc[3]) & kFlagMaskTri) != 0);
        pending |= (uint64_t)(lo != 0) << g | (uint64_t)(hi != 0) << (g + kHalfShift);
      } else if constexpr (QT == 16) {
        const uint64_t lo = __ballot(flags_set_low(acc)), hi = __ballot(flags_set_high(acc));
        pending |= (uint64_t)(lo != 0) << g | (uint64_t)(hi != 0) << (g + kHalfShift);
      } else {
        const uint64_t vote = __ballot(flags_set(acc));
        pending |= (uint64_t)(vote != 0) << g;
      }
#pragma unroll
      for (int m = 0; m < kFlagChains; m++) acc[m] = 0u;
      __syncthreads();
    }
  }
  // The rare path is a record for the resolve kernel (sub-chunk number, plus
  // rec_sub_base for a shard of a set; this wave's first query; the pending
  // groups): checking pairs here, even in code that never runs, slows the
  // loop by a tenth.
  if (pending && lane == 0) {
    const uint32_t slot = atomicAdd(rec_count, 1u);
    const uint64_t sub = rec_sub_base + row0 / kSharedRows;
    if (slot < rec_cap) recs[slot] = make_uint4((uint32_t)sub, (uint32_t)(sub >> 32), qbase, (uint32_t)pending);
  }
}

template <int QT, bool kPlane, bool kTri>
__device__ __forceinline__ void mfma_item_shared(const uint32_t* __restrict__ lut, SharedRowsLds& S,
                                                 const uint32_t* __restrict__ db, const uint32_t* __restrict__ plane,
                                                 uint64_t n_rows, uint64_t sub, uint32_t sgroup,
                                                 const uint32_t* __restrict__ queries, uint32_t n_queries,
                                                 uint32_t threshold, uint4* __restrict__ recs,
                                                 uint32_t* __restrict__ rec_count, uint32_t rec_cap,
                                                 uint64_t rec_sub_base) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: the tests on it are scalar
  const int rn = lane & 15, kq = lane >> 4;
  const float bound = (float)(128 - 2 * (int)threshold);
  const uint64_t row0 = sub * kSharedRows;
  if (row0 >= n_rows) return;  // the same for every wave of the block

  // query operands; a wave whose group lies past the last query gets zero
  // operands (its products never reach the bound) and still stages its tiles
  const uint32_t qbase = (sgroup * kMfmaWaves + wave) * (uint32_t)(QT * 16);
  uint32_t qw[QT];
#pragma unroll
  for (int t = 0; t < QT; t++) {
    const uint32_t qi = qbase + 16 * t + rn;
    qw[t] = queries[(size_t)(qi < n_queries ? qi : n_queries - 1) * 8 + kq];
  }
  Expanded qa[QT];
#pragma unroll
  for (int t = 0; t < QT; t++) {
    expand_word(qw[t], lut, qa[t]);
    if (qbase + 16 * t + rn >= n_queries) qa[t].d[0] = qa[t].d[1] = qa[t].d[2] = qa[t].d[3] = 0;
  }

  const RowSource<kPlane> src(db, plane, row0, lane);
  if (row0 + kSharedRows <= n_rows)
    stream_shared<QT, true, kPlane, kTri>(qa, lut, S, src, db, row0, n_rows, qbase, wave, lane, bound, recs,
                                          rec_count, rec_cap, rec_sub_base);
  else
    stream_shared<QT, false, kPlane, kTri>(qa, lut, S, src, db, row0, n_rows, qbase, wave, lane, bound, recs,
                                           rec_count, rec_cap, rec_sub_base);
}

__device__ __forceinline__ void init_lut(uint32_t* lut) {
  lut[threadIdx.x] = spread_byte(threadIdx.x);  // blockDim.x == 256
  __syncthreads();
}

// Work item -> (row chunk, query group).  Consecutive workgroups are dealt
// round-robin to the 8 XCDs, each with its own L2; items are numbered so that
// all query groups of a row chunk run on the same XCD (row chunk = xcd +
// 8 * (slot / groups)) and read its rows from that XCD's L2 after the first
// one brought them in from HBM.
__device__ __forceinline__ void item_coords(uint64_t item, uint32_t groups, uint64_t& chunk, uint32_t& group) {
  const uint64_t xcd = item & 7, slot = item >> 3;
  chunk = xcd + 8 * (slot / groups);
  group = (uint32_t)(slot % groups);
}

// One launch, one block per work item: row chunks x query groups (the
// per-wave layout, which checks its candidates itself), or, with kShared
// (4+ query groups), 1024-row sub-chunks x groups of kMfmaWaves query groups,
// whose candidates become records for pdq_match_resolve_kernel.  Block 0
// first stores zero to clear_rec and clear_hits where they are given: the
// record count of the stream's next call, and (shared rows, which append
// no hits here) this call's hit counter.
template <int QT, bool kPlane, bool kShared, bool kTri = false>
__global__ void __launch_bounds__(64 * kMfmaWaves) __attribute__((amdgpu_waves_per_eu(QT >= 16 ? 3 : 4))) pdq_match_mfma_kernel(
    const uint32_t* __restrict__ db, const uint32_t* __restrict__ plane, uint64_t n_rows,
    const uint32_t* __restrict__ queries, uint32_t n_queries, uint32_t threshold, MatchHit* __restrict__ hits,
    uint64_t capacity, unsigned long long* __restrict__ counter, uint64_t index_base, uint32_t groups,
    uint4* __restrict__ recs, uint32_t* __restrict__ rec_count, uint32_t rec_cap, uint32_t* __restrict__ clear_rec,
    unsigned long long* __restrict__ clear_hits) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (clear_rec) *clear_rec = 0u;
    if (clear_hits) *clear_hits = 0ull;
  }
  __shared__ uint32_t lut[256];
  init_lut(lut);
  uint64_t chunk;
  uint32_t group;
  item_coords(blockIdx.x, groups, chunk, group);
  if constexpr (kShared) {
    __shared__ SharedRowsLds S;
    mfma_item_shared<QT, kPlane, kTri>(lut, S, db, plane, n_rows, chunk, group, queries, n_queries, threshold,
                                       recs, rec_count, rec_cap, 0);
  } else {
    mfma_item<QT, kPlane>(lut, db, plane, n_rows, chunk, group, queries, n_queries, threshold, hits, capacity,
                          counter, index_base, (int)(threadIdx.x >> 6));
  }
}

// The second kernel of the shared-rows layout, in blocks of kResolveWaves
// waves.  Each wave takes records in turn: for every pending group of the
// record's sub-chunk it recomputes the group's products, and for each query
// tile with a pair at the bound checks every such pair with the exact
// 256-bit distance.  The block's hits gather in LDS and take their slots in
// the hit buffer with one atomic.  If the records overflowed their area (more
// than rec_cap: dense hits), every work item is matched again in the
// per-wave layout instead (each half of the block an item), which checks its
// own candidates (the first kernel appended no hits), so the set stays exact.
constexpr int kResolveWaves = 8;
// Two blocks per CU (16 waves; the kernel fits 128 VGPRs at its at most 8
// query tiles per wave): a
// record's check is a chain of dependent loads, so waves in flight set its
// rate.  Measured against one block per CU: 19 against 22.5 us at the
// bench's sparse hits, 578 against 871 us at 271k hits per batch.
constexpr int kResolveBlocksPerCu = 2;

uint32_t match_resolve_blocks(int cus) { return (uint32_t)std::max(1, cus) * kResolveBlocksPerCu; }
static_assert(kResolveWaves % kMfmaWaves == 0, "whole per-wave items per block");

struct ResolveLds {
  uint32_t lut[256];
  MatchHit stage[kStageHits];
  uint32_t n_stage;
  unsigned long long stage_base;
};

__device__ __forceinline__ void resolve_init(ResolveLds& L) {
  if (threadIdx.x < 256) L.lut[threadIdx.x] = spread_byte(threadIdx.x);
  if (threadIdx.x == 0) L.n_stage = 0;
  __syncthreads();
}

// The block's gathered hits into the buffer: one run of slots.  Every thread
// of the block calls it.
__device__ __forceinline__ void resolve_flush(ResolveLds& L, MatchHit* __restrict__ hits, uint64_t capacity,
                                              unsigned long long* __restrict__ counter) {
  __syncthreads();
  const uint32_t n = min(L.n_stage, kStageHits);
  if (n == 0) return;  // the same for the whole block
  if (threadIdx.x == 0) L.stage_base = atomicAdd(counter, (unsigned long long)n);
  __syncthreads();
  const unsigned long long base = L.stage_base;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
    if (base + i < capacity) hits[base + i] = L.stage[i];
}

// The loads of one record part that do not depend on its products: the
// query words of its tiles and the tile words of its first pending group.  A
// wave issues them for its next part before it checks the current one, so a
// part's check waits only on its own exact-distance loads.
template <int QT>
struct PartLoads {
  uint32_t qw[QT];
  uint32_t w[kGroupTiles];
};

// Tiles of the sub-chunk at row0 (16 rows each; the last sub-chunk's may be
// partial).
__device__ __forceinline__ uint32_t sub_tiles(uint64_t n_rows, uint64_t row0) {
  const uint64_t left = n_rows - row0;
  return left >= kSharedRows ? (uint32_t)kSharedTiles : (uint32_t)((left + 15) / 16);
}

// Group g's tile words (lane rn, word kq of each tile), clamped into the
// sub-chunk's tiles (and, for raw rows, its rows).
template <bool kPlane>
__device__ __forceinline__ void load_group(uint32_t (&w)[kGroupTiles], const uint32_t* __restrict__ db,
                                           const uint32_t* __restrict__ plane, uint64_t n_rows, uint64_t row0,
                                           uint32_t g) {
  const int lane = threadIdx.x & 63;
  const int rn = lane & 15, kq = lane >> 4;
  const uint64_t left = n_rows - row0;
  const uint32_t tiles = sub_tiles(n_rows, row0);
  const RowSource<kPlane> src(db, plane, row0, lane);
#pragma unroll
  for (int j = 0; j < kGroupTiles; j++) {
    const uint32_t tt = min(kGroupTiles * g + (uint32_t)j, tiles - 1);
    if (kPlane) {
      w[j] = src(tt);
    } else {
      const uint64_t rr = min((uint64_t)tt * 16 + rn, left - 1);
      w[j] = db[(row0 + rr) * 8 + kq];
    }
  }
}

template <int QT, bool kPlane>
__device__ __forceinline__ void part_load(PartLoads<QT>& P, const uint32_t* __restrict__ db,
                                          const uint32_t* __restrict__ plane, uint64_t n_rows, uint64_t row0,
                                          uint32_t qbase, uint32_t pending, const uint32_t* __restrict__ queries,
                                          uint32_t n_queries) {
  if (pending == 0) return;  // nothing to check (the same for the whole wave)
  const int lane = threadIdx.x & 63;
  const int rn = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int t = 0; t < QT; t++) {
    const uint32_t qi = qbase + 16 * t + rn;
    P.qw[t] = queries[(size_t)(qi < n_queries ? qi : n_queries - 1) * 8 + kq];
  }
  load_group<kPlane>(P.w, db, plane, n_rows, row0, (uint32_t)__builtin_ctz(pending));
}

// One pending group g of a record part: its tiles' products with the part's
// query tiles qa, and the pairs of the query tiles that reach the bound
// checked exactly.  w holds the group's tile words.
template <int QT>
__device__ __forceinline__ void check_group(ResolveLds& L, const Expanded (&qa)[QT], uint32_t g,
                                            const uint32_t (&words)[kGroupTiles], uint32_t tiles, uint64_t left,
                                            uint64_t row0, uint32_t qbase, float bound,
                                            const uint32_t* __restrict__ queries, const uint32_t* __restrict__ db,
                                            uint32_t n_queries, uint64_t n_rows, uint32_t threshold,
                                            MatchHit* __restrict__ hits, uint64_t capacity,
                                            unsigned long long* __restrict__ counter, uint64_t index_base) {
  const int lane = threadIdx.x & 63;
  const int rn = lane & 15, kq = lane >> 4;
  const int ibound = __float_as_int(bound);
  // tile j's word is w[0] in iteration j (the words move down one place per
  // tile: a rolled loop cannot index registers, and a select by j becomes an
  // indexed copy in scratch memory)
  uint32_t w[kGroupTiles];
#pragma unroll
  for (int k = 0; k < kGroupTiles; k++) w[k] = words[k];
#pragma unroll 1
  for (int j = 0; j < kGroupTiles; j++) {
    const uint32_t tt = kGroupTiles * g + (uint32_t)j;
    if (tt >= tiles) break;  // the same for the whole wave
    const uint32_t wj = w[0];
#pragma unroll
    for (int k = 0; k + 1 < kGroupTiles; k++) w[k] = w[k + 1];
    Expanded eg;
    expand_word(wj, L.lut, eg);
    if ((uint64_t)tt * 16 + rn >= left) eg.d[0] = eg.d[1] = eg.d[2] = eg.d[3] = 0;
    // which query tiles have a pair at the bound (a wave-uniform mask); the
    // products go in pairs, so that one product's reads overlap the next's
    // matrix-core latency instead of each waiting on the last
    uint32_t tmask = 0;
#pragma unroll
    for (int t = 0; t < QT; t += 2) {
      const v4f c0 = mfma_prefix(qa[t], eg);
      const v4f c1 = t + 1 < QT ? mfma_prefix(qa[t + 1 < QT ? t + 1 : t], eg) : c0;
      const int m0 = max(max(__float_as_int(c0.x), __float_as_int(c0.y)), max(__float_as_int(c0.z), __float_as_int(c0.w)));
      const int m1 = max(max(__float_as_int(c1.x), __float_as_int(c1.y)), max(__float_as_int(c1.z), __float_as_int(c1.w)));
      tmask |= (wave_any(m0 >= ibound) ? 1u : 0u) << t;
      if (t + 1 < QT) tmask |= (wave_any(m1 >= ibound) ? 1u : 0u) << (t + 1);
    }
    // only those tiles' products are recomputed and checked, by one inlined
    // check in a loop over the mask: QT inlined copies keep several hundred
    // registers live, and an out-of-line check saves and restores registers
    // through scratch memory at every call
    while (tmask) {
      const int t = __builtin_ctz(tmask);
      tmask &= tmask - 1;
      Expanded a = qa[0];  // qa[t], selected (registers cannot be indexed)
#pragma unroll
      for (int k = 1; k < QT; k++)
        if (t == k) a = qa[k];
      stage_candidates(mfma_prefix(a, eg), t, row0 + (uint64_t)tt * 16, qbase, rn, kq, bound, queries, db,
                       n_queries, n_rows, threshold, L.stage, &L.n_stage, hits, capacity, counter, index_base);
    }
  }
}

// One record part, checked by one wave: sub-chunk rows row0 ... of a
// database of n_rows rows (global index index_base + row), queries qbase ...
// (QT x 16), pending groups of kGroupTiles tiles, with P holding its
// part_load.  Groups after the first load their tile words here.  `next` is
// called once P's words are taken (it issues the wave's next loads): a wait
// for P then never covers those newer loads.
template <int QT, bool kPlane, class Next>
__device__ __forceinline__ void resolve_record(ResolveLds& L, const PartLoads<QT>& P,
                                               const uint32_t* __restrict__ db,
                                               const uint32_t* __restrict__ plane, uint64_t n_rows, uint64_t row0,
                                               uint32_t qbase, uint32_t pending, const uint32_t* __restrict__ queries,
                                               uint32_t n_queries, uint32_t threshold, MatchHit* __restrict__ hits,
                                               uint64_t capacity, unsigned long long* __restrict__ counter,
                                               uint64_t index_base, Next&& next) {
  if (pending == 0) {  // a half of a record with no candidates
    next();
    return;
  }
  const int rn = threadIdx.x & 15;
  const float bound = (float)(128 - 2 * (int)threshold);
  const uint64_t left = n_rows - row0;
  const uint32_t tiles = sub_tiles(n_rows, row0);
  Expanded qa[QT];
#pragma unroll
  for (int t = 0; t < QT; t++) {
    expand_word(P.qw[t], L.lut, qa[t]);
    if (qbase + 16 * t + rn >= n_queries) qa[t].d[0] = qa[t].d[1] = qa[t].d[2] = qa[t].d[3] = 0;
  }
  // The first group's words come with P, outside the loop over the other
  // groups: where the two meet, a wait for the loop's own loads would also
  // wait for the loads `next` has issued.
  uint32_t w0[kGroupTiles];
#pragma unroll
  for (int j = 0; j < kGroupTiles; j++) w0[j] = P.w[j];
  next();
  check_group<QT>(L, qa, (uint32_t)__builtin_ctz(pending), w0, tiles, left, row0, qbase, bound, queries, db,
                  n_queries, n_rows, threshold, hits, capacity, counter, index_base);
  pending &= pending - 1;
  while (pending) {  // rare: a part with candidates in more than one group
    const uint32_t g = (uint32_t)__builtin_ctz(pending);
    pending &= pending - 1;
    uint32_t w[kGroupTiles];
    load_group<kPlane>(w, db, plane, n_rows, row0, g);
    check_group<QT>(L, qa, g, w, tiles, left, row0, qbase, bound, queries, db, n_queries, n_rows, threshold, hits,
                    capacity, counter, index_base);
  }
}

// Query tiles per wave in the recheck: at most eight.  A record of a 16-tile
// stream is checked as two halves (query tiles 0-7 and 8-15, each with its
// own pending groups), one after the other by the wave that holds the
// record: sixteen tiles' operands need more registers than two blocks per CU
// leave, and a half without candidates costs nothing.
template <int QT>
constexpr int kResolveQt = QT > 8 ? 8 : QT;

// The pending groups of part `part` of a record (of kParts parts).
template <int kParts>
__device__ __forceinline__ uint32_t part_groups(uint32_t mask, uint32_t part) {
  static_assert(kParts == 1 || kParts == 2, "whole records or halves of a 16-tile record");
  if constexpr (kParts == 1) return mask;
  else return (mask >> (kHalfShift * part)) & ((1u << kHalfShift) - 1);
}

template <int QT, bool kPlane>
__global__ void __launch_bounds__(64 * kResolveWaves) pdq_match_resolve_kernel(
    const uint32_t* __restrict__ db, const uint32_t* __restrict__ plane, uint64_t n_rows,
    const uint32_t* __restrict__ queries, uint32_t n_queries, uint32_t threshold, MatchHit* __restrict__ hits,
    uint64_t capacity, unsigned long long* __restrict__ counter, uint64_t index_base,
    const uint4* __restrict__ recs, const uint32_t* __restrict__ rec_count, uint32_t rec_cap, uint64_t chunks,
    uint32_t groups) {
  __shared__ ResolveLds L;
  resolve_init(L);
  const int wave = (int)(threadIdx.x >> 6);
  const uint32_t n_recs = *rec_count;
  if (n_recs > rec_cap) {
    // the per-wave layout, at most eight query tiles per wave
    constexpr int kRq = kResolveQt<QT>;
    const uint32_t rgroups = (n_queries + kRq * 16 - 1) / (kRq * 16);
    constexpr int kHalves = kResolveWaves / kMfmaWaves;
    const uint64_t items = chunks * rgroups;
    for (uint64_t item = (uint64_t)blockIdx.x * kHalves + wave / kMfmaWaves; item < items;
         item += (uint64_t)gridDim.x * kHalves) {
      uint64_t chunk;
      uint32_t group;
      item_coords(item, rgroups, chunk, group);
      mfma_item<kRq, kPlane>(L.lut, db, plane, n_rows, chunk, group, queries, n_queries, threshold, hits, capacity,
                             counter, index_base, wave % kMfmaWaves);
    }
    (void)groups;
    return;
  }
  // The wave's records are w, w + S, w + 2S, ... (w its number in the
  // grid, S the grid's waves); its item j is part j % kParts of the
  // (j / kParts)-th of them.  Whole records per wave, rather than parts
  // dealt out one by one, bound the parts with candidates that one wave
  // checks (a record's candidates are usually in one part).  The record of
  // the item after next is fetched, and the next item's loads are issued,
  // while this one is checked.
  constexpr int kRq = kResolveQt<QT>, kParts = QT / kRq;
  const uint32_t w0 = blockIdx.x * kResolveWaves + wave;
  const uint32_t S = gridDim.x * kResolveWaves;
  auto rec_of = [&](uint32_t j) { return w0 + (j / kParts) * S; };
  auto fetch = [&](uint32_t j) {
    const uint32_t i = rec_of(j);
    return i < n_recs ? recs[i] : make_uint4(0, 0, 0, 0);
  };
  auto row0_of = [](const uint4& x) { return ((uint64_t)x.y << 32 | x.x) * kSharedRows; };
  // the pending groups of item j (none past the last record)
  auto pending_of = [&](const uint4& x, uint32_t j) {
    return rec_of(j) < n_recs ? part_groups<kParts>(x.w, j % kParts) : 0u;
  };
  auto qbase_of = [](const uint4& x, uint32_t j) { return x.z + (j % kParts) * (kRq * 16); };
  uint32_t j = 0;
  uint4 rec = fetch(0);
  uint4 rec_next = fetch(1);
  PartLoads<kRq> P;
  part_load<kRq, kPlane>(P, db, plane, n_rows, row0_of(rec), qbase_of(rec, 0), pending_of(rec, 0), queries,
                         n_queries);
  // The first part's words arrive here (vmcnt 0; expcnt and lgkmcnt left
  // alone), as every later part's do at the end of the previous iteration:
  // the compiler then sees them complete on every path into the loop and
  // does not make the loop's waits for them cover the next part's loads.
  __builtin_amdgcn_s_waitcnt(0x0F70);
  for (; rec_of(j) < n_recs; j++) {
    uint4 rec_after;
    PartLoads<kRq> Pn;
    resolve_record<kRq, kPlane>(L, P, db, plane, n_rows, row0_of(rec), qbase_of(rec, j), pending_of(rec, j),
                                queries, n_queries, threshold, hits, capacity, counter, index_base, [&] {
                                  rec_after = fetch(j + 2);
                                  part_load<kRq, kPlane>(Pn, db, plane, n_rows, row0_of(rec_next),
                                                         qbase_of(rec_next, j + 1), pending_of(rec_next, j + 1),
                                                         queries, n_queries);
                                });
    rec = rec_next;
    rec_next = rec_after;
    P = Pn;
  }
  resolve_flush(L, hits, capacity, counter);
}

// The shard holding MFMA row chunk `chunk` (views in chunk order).
__device__ __forceinline__ uint32_t shard_of_chunk(const ShardView* __restrict__ shards, uint32_t n_shards,
                                                   uint64_t chunk) {
  uint32_t sh = 0;
  while (sh + 1 < n_shards && shards[sh + 1].chunk_begin <= chunk) sh++;
  return sh;
}

// The shared-rows layout over a set of shards, persistent: a fixed grid (a
// multiple of 8 blocks) walks (global sub-chunk, group of kMfmaWaves query
// groups) items, the sub-chunks of shard s numbered from its first chunk x
// kSubsPerChunk.  The query count is read from device memory (`d_nq`, capped
// at nq_cap), so one captured launch serves batches of any size.  Records
// carry the global sub-chunk number.
template <int QT, bool kPlane, bool kTri = false>
__global__ void __launch_bounds__(64 * kMfmaWaves) __attribute__((amdgpu_waves_per_eu(QT >= 16 ? 3 : 4))) pdq_match_mfma_shards_shared_kernel(
    const ShardView* __restrict__ shards, uint32_t n_shards, uint64_t total_chunks,
    const uint32_t* __restrict__ queries, const uint32_t* __restrict__ d_nq, uint32_t nq_cap, uint32_t threshold,
    uint4* __restrict__ recs, uint32_t* __restrict__ rec_count, uint32_t rec_cap) {
  const uint32_t nq = d_nq ? min(*d_nq, nq_cap) : nq_cap;
  if (nq == 0) return;
  __shared__ uint32_t lut[256];
  __shared__ SharedRowsLds S;
  init_lut(lut);
  const uint32_t groups = (nq + QT * 16 - 1) / (QT * 16);
  const uint32_t sgroups = (groups + kMfmaWaves - 1) / kMfmaWaves;
  const uint64_t items = total_chunks * kSubsPerChunk * sgroups;
  for (uint64_t item = blockIdx.x; item < items; item += gridDim.x) {
    uint64_t gsub;
    uint32_t sgroup;
    item_coords(item, sgroups, gsub, sgroup);
    const ShardView& v = shards[shard_of_chunk(shards, n_shards, gsub / kSubsPerChunk)];
    const uint64_t sub_base = v.chunk_begin * kSubsPerChunk;
    mfma_item_shared<QT, kPlane, kTri>(lut, S, reinterpret_cast<const uint32_t*>(v.rows),
                                       reinterpret_cast<const uint32_t*>(v.plane), v.n_rows, gsub - sub_base, sgroup,
                                       queries, nq, threshold, recs, rec_count, rec_cap, sub_base);
  }
}

// Its recheck kernel: as pdq_match_resolve_kernel, with each record's shard
// found from its global sub-chunk, and the per-wave rerun after an overflow
// walking the items of every shard.
template <int QT, bool kPlane>
__global__ void __launch_bounds__(64 * kResolveWaves) pdq_match_resolve_shards_kernel(
    const ShardView* __restrict__ shards, uint32_t n_shards, uint64_t total_chunks,
    const uint32_t* __restrict__ queries, const uint32_t* __restrict__ d_nq, uint32_t nq_cap, uint32_t threshold,
    MatchHit* __restrict__ hits, uint64_t capacity, unsigned long long* __restrict__ counter,
    const uint4* __restrict__ recs, const uint32_t* __restrict__ rec_count, uint32_t rec_cap) {
  const uint32_t nq = d_nq ? min(*d_nq, nq_cap) : nq_cap;
  if (nq == 0) return;
  __shared__ ResolveLds L;
  resolve_init(L);
  const int wave = (int)(threadIdx.x >> 6);
  const uint32_t n_recs = *rec_count;
  if (n_recs > rec_cap) {
    constexpr int kHalves = kResolveWaves / kMfmaWaves;
    constexpr int kRq = kResolveQt<QT>;
    const uint32_t groups = (nq + kRq * 16 - 1) / (kRq * 16);
    const uint64_t items = total_chunks * groups;
    for (uint64_t item = (uint64_t)blockIdx.x * kHalves + wave / kMfmaWaves; item < items;
         item += (uint64_t)gridDim.x * kHalves) {
      uint64_t chunk;
      uint32_t group;
      item_coords(item, groups, chunk, group);
      const ShardView& v = shards[shard_of_chunk(shards, n_shards, chunk)];
      mfma_item<kRq, kPlane>(L.lut, reinterpret_cast<const uint32_t*>(v.rows),
                            reinterpret_cast<const uint32_t*>(v.plane), v.n_rows, chunk - v.chunk_begin, group, queries,
                            nq, threshold, hits, capacity, counter, v.index_base, wave % kMfmaWaves);
    }
    return;
  }
  // whole records per wave, as in pdq_match_resolve_kernel
  constexpr int kRq = kResolveQt<QT>, kParts = QT / kRq;
  const uint32_t stride = gridDim.x * kResolveWaves;
  uint32_t i = blockIdx.x * kResolveWaves + wave;
  uint4 rec = i < n_recs ? recs[i] : make_uint4(0, 0, 0, 0);
  for (; i < n_recs; i += stride) {
    const uint4 next = i + stride < n_recs ? recs[i + stride] : make_uint4(0, 0, 0, 0);
    const uint64_t gsub = (uint64_t)rec.y << 32 | rec.x;
    const ShardView& v = shards[shard_of_chunk(shards, n_shards, gsub / kSubsPerChunk)];
    const uint64_t row0 = (gsub - v.chunk_begin * kSubsPerChunk) * kSharedRows;
    const uint32_t* rows = reinterpret_cast<const uint32_t*>(v.rows);
    const uint32_t* vplane = reinterpret_cast<const uint32_t*>(v.plane);
#pragma unroll 1
    for (int part = 0; part < kParts; part++) {
      const uint32_t qbase = rec.z + part * (kRq * 16);
      const uint32_t pending = part_groups<kParts>(rec.w, part);
      PartLoads<kRq> P;
      part_load<kRq, kPlane>(P, rows, vplane, v.n_rows, row0, qbase, pending, queries, nq);
      resolve_record<kRq, kPlane>(L, P, rows, vplane, v.n_rows, row0, qbase, pending, queries, nq, threshold, hits,
                                  capacity, counter, v.index_base, [] {});
    }
    rec = next;
  }
  resolve_flush(L, hits, capacity, counter);
}

// Persistent form over a set of database shards: a fixed grid (a multiple of
// 8 blocks) walks the work items of every shard, blockIdx.x + k * gridDim.x,
// which keeps each item on the XCD its number selects.  The query count is
// read from device memory (`d_nq`, capped at nq_cap), so one captured launch
// serves batches of any size; batches of fewer than min_nq queries are left
// to the vector-ALU kernel and this one returns at once.
template <int QT, bool kPlane>
__global__ void __launch_bounds__(64 * kMfmaWaves) __attribute__((amdgpu_waves_per_eu(QT >= 16 ? 3 : 4))) pdq_match_mfma_shards_kernel(
    const ShardView* __restrict__ shards, uint32_t n_shards, uint64_t total_chunks,
    const uint32_t* __restrict__ queries, const uint32_t* __restrict__ d_nq, uint32_t nq_cap, uint32_t min_nq,
    uint32_t threshold, MatchHit* __restrict__ hits, uint64_t capacity, unsigned long long* __restrict__ counter) {
  const uint32_t nq = d_nq ? min(*d_nq, nq_cap) : nq_cap;
  if (nq == 0 || nq < min_nq) return;
  __shared__ uint32_t lut[256];
  init_lut(lut);
  const uint32_t groups = (nq + QT * 16 - 1) / (QT * 16);
  const uint64_t items = total_chunks * groups;
  for (uint64_t item = blockIdx.x; item < items; item += gridDim.x) {
    uint64_t chunk;
    uint32_t group;
    item_coords(item, groups, chunk, group);
    uint32_t sh = 0;
    while (sh + 1 < n_shards && shards[sh + 1].chunk_begin <= chunk) sh++;
    const ShardView& v = shards[sh];
    mfma_item<QT, kPlane>(lut, reinterpret_cast<const uint32_t*>(v.rows), reinterpret_cast<const uint32_t*>(v.plane),
                          v.n_rows, chunk - v.chunk_begin, group, queries, nq, threshold, hits, capacity, counter,
                          v.index_base, (int)(threadIdx.x >> 6));
  }
}

// Record areas of the shared-rows layout: one per (device, stream), made on
// the stream's first such launch and kept for the life of the process, with
// a mutex so that the count reset, the filter and the resolve kernel of one
// call are enqueued together even when threads share a stream.  At most
// kMaxStreamAreas streams of a device get an area; calls on further streams,
// and on a stream being captured into a graph (which cannot allocate), run
// the per-wave layout, which needs none.
constexpr uint32_t kRecordCap = 1u << 20;  // 16 MiB; more records (dense hits) rerun every item per wave
constexpr size_t kMaxStreamAreas = 16;     // per device

// The record cap passed to the kernels: kRecordCap, or less when set by
// set_match_record_cap (tests reach the overflow path at small sizes with
// it); the environment's YUME_PDQ_AMD_MATCH_RECORD_CAP (1 .. kRecordCap) is
// its initial value, read once.
static uint32_t env_record_cap() {
  const char* v = getenv("YUME_PDQ_AMD_MATCH_RECORD_CAP");
  if (!v || !*v) return kRecordCap;
  char* end = nullptr;
  const unsigned long long c = strtoull(v, &end, 10);
  return (end && *end == 0 && c >= 1 && c <= kRecordCap) ? (uint32_t)c : kRecordCap;
}
static std::atomic<uint32_t> g_record_cap{env_record_cap()};
static uint32_t record_cap() { return g_record_cap.load(std::memory_order_relaxed); }

bool set_match_record_cap(uint32_t cap) {
  if (cap > kRecordCap) return false;
  g_record_cap.store(cap == 0 ? kRecordCap : cap, std::memory_order_relaxed);
  return true;
}

// A stream's record area: the records, then two record-count slots
// (kCountSlotStride words apart).  Calls on the stream take the slots in
// turn, and each call's first kernel clears the slot of the next call, so a
// call needs no launch of its own to clear its count while the previous
// call's kernels were all enqueued.  slot and slot_clear are host state.
constexpr uint32_t kCountSlotStride = 32;
static_assert(2 * kCountSlotStride * sizeof(uint32_t) <= 256, "both count slots in the area's 256-byte tail");
struct RecordArea {
  uint4* recs = nullptr;
  uint32_t* count = nullptr;
  int cus = 0;  // compute units of the area's device
  uint32_t slot = 0;        // the count slot of the stream's next call
  bool slot_clear = false;  // that slot holds zero once the enqueued work has run
};

static std::mutex g_record_mu;

static hipError_t record_area(hipStream_t stream, RecordArea** out, bool* usable) {
  *usable = false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(stream, &cs);
  if (e != hipSuccess) return e;
  if (cs != hipStreamCaptureStatusNone) return hipSuccess;
  int dev = 0;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  // keyed by device and stream; hipStreamPerThread names a different
  // stream in each host thread, so its areas are also keyed by the thread
  using Key = std::tuple<int, hipStream_t, std::thread::id>;
  static std::map<Key, RecordArea> areas;  // guarded by g_record_mu
  static std::map<int, size_t> per_device;  // areas made per device
  const Key key{dev, stream, stream == hipStreamPerThread ? std::this_thread::get_id() : std::thread::id()};
  auto it = areas.find(key);
  if (it == areas.end()) {
    if (per_device[dev] >= kMaxStreamAreas) return hipSuccess;  // not usable: the per-wave layout
    void* p = nullptr;
    if ((e = hipMalloc(&p, (size_t)kRecordCap * sizeof(uint4) + 256)) != hipSuccess) return e;
    RecordArea ra;
    if ((e = hipDeviceGetAttribute(&ra.cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) {
      (void)hipFree(p);
      return e;
    }
    ra.recs = static_cast<uint4*>(p);
    ra.count = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(p) + (size_t)kRecordCap * sizeof(uint4));
    it = areas.emplace(key, ra).first;
    per_device[dev]++;
  }
  *out = &it->second;
  *usable = true;
  return hipSuccess;
}

// The record area of the no-allocation entry point: kScratchRecords records,
// then the count, in the code object's own device memory.  Each GPU the
// library runs on holds one copy (16 MiB + 256 B) until the process ends.
// Host code never names the array in a launch: its device address, per
// device, comes from hipGetSymbolAddress (scratch_area below).
constexpr uint32_t kScratchRecords = 1u << 20;
static __device__ __align__(256) uint4 scratch[kScratchRecords + 16];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxScratchDevices];  // guarded by g_scratch_mu
static int g_scratch_cus[kMaxScratchDevices];

uint32_t match_scratch_records() { return kScratchRecords; }

// The current device's scratch area, its address looked up on the device's
// first call (with that device current) and kept.
static hipError_t scratch_area(RecordArea* out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    int cus = 0;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    g_scratch_cus[dev] = cus;
    g_scratch_ptr[dev] = p;
  }
  out->recs = static_cast<uint4*>(g_scratch_ptr[dev]);
  out->count = reinterpret_cast<uint32_t*>(out->recs + kScratchRecords);
  out->cus = g_scratch_cus[dev];
  return hipSuccess;
}

// Clears the candidate-record count and, when asked, the hit counter.
__global__ void reset_counts_kernel(uint32_t* rec_count, unsigned long long* counter) {
  if (rec_count) *rec_count = 0u;
  if (counter) *counter = 0ull;
}

template <int QT>
static hipError_t launch_qt(const void* db, const void* plane, uint64_t n_rows, const void* queries,
                            uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                            unsigned long long* counter, uint64_t index_base, hipStream_t stream, bool reset,
                            const RecordArea* fixed) {
  const uint32_t groups = (n_queries + QT * 16 - 1) / (QT * 16);
  const uint32_t* d = static_cast<const uint32_t*>(db);
  const uint32_t* p = static_cast<const uint32_t*>(plane);
  const uint32_t* q = static_cast<const uint32_t*>(queries);
  const dim3 block(64 * kMfmaWaves);
  // the per-wave layout's items: 4096-row chunks x query groups
  const uint64_t rows_per_chunk = (uint64_t)kMfmaWaves * kMfmaTilesPerWave * 16;
  const uint64_t chunks8 = ((n_rows + rows_per_chunk - 1) / rows_per_chunk + 7) / 8 * 8;  // whole XCD rounds
  if (groups >= (uint32_t)kMfmaWaves) {
    // the caller's fixed area (the no-allocation entry point's scratch,
    // sized for this call's worst case), else this stream's own area
    std::lock_guard<std::mutex> g(g_record_mu);
    RecordArea ra;
    RecordArea* area = nullptr;  // the stream's own area, whose count slots alternate
    bool usable = false;
    if (fixed) {
      ra = *fixed;
      usable = true;
    } else {
      hipError_t e = record_area(stream, &area, &usable);
      if (e != hipSuccess) return e;
      if (usable) ra = *area;
    }
    if (usable) {
      const uint32_t cap = fixed ? kScratchRecords : record_cap();
      const uint64_t subs8 = ((n_rows + kSharedRows - 1) / kSharedRows + 7) / 8 * 8;
      const uint32_t sgroups = (groups + kMfmaWaves - 1) / kMfmaWaves;
      const uint64_t blocks = subs8 * sgroups;
      if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
      // this call's count slot; the stream's area has the streaming kernel
      // clear the next call's slot, and the hit counter when this slot is
      // already clear
      uint32_t* rc = ra.count;
      uint32_t* rc_next = nullptr;
      unsigned long long* hits_clear = nullptr;
      if (area) {
        rc = area->count + kCountSlotStride * area->slot;
        rc_next = area->count + kCountSlotStride * (area->slot ^ 1u);
      }
      if (area && area->slot_clear)
        hits_clear = reset ? counter : nullptr;
      else
        hipLaunchKernelGGL(reset_counts_kernel, dim3(1), dim3(1), 0, stream, rc, reset ? counter : nullptr);
      // three products per accumulator where they apply
      const bool tri = QT == 16 && threshold >= kTriMinThreshold;
      if (plane && tri)
        hipLaunchKernelGGL((pdq_match_mfma_kernel<QT, true, true, QT == 16>), dim3((unsigned)blocks), block, 0,
                           stream, d, p, n_rows, q, n_queries, threshold, hits, capacity, counter, index_base,
                           sgroups, ra.recs, rc, cap, rc_next, hits_clear);
      else if (plane)
        hipLaunchKernelGGL((pdq_match_mfma_kernel<QT, true, true>), dim3((unsigned)blocks), block, 0, stream, d, p,
                           n_rows, q, n_queries, threshold, hits, capacity, counter, index_base, sgroups, ra.recs,
                           rc, cap, rc_next, hits_clear);
      else if (tri)
        hipLaunchKernelGGL((pdq_match_mfma_kernel<QT, false, true, QT == 16>), dim3((unsigned)blocks), block, 0,
                           stream, d, p, n_rows, q, n_queries, threshold, hits, capacity, counter, index_base,
                           sgroups, ra.recs, rc, cap, rc_next, hits_clear);
      else
        hipLaunchKernelGGL((pdq_match_mfma_kernel<QT, false, true>), dim3((unsigned)blocks), block, 0, stream, d, p,
                           n_rows, q, n_queries, threshold, hits, capacity, counter, index_base, sgroups, ra.recs,
                           rc, cap, rc_next, hits_clear);
      // the resident blocks of the recheck kernel; the per-wave rerun after
      // an overflow walks all items with them
      const uint64_t rgrid = std::min<uint64_t>(match_resolve_blocks(ra.cus), chunks8 * groups);
      const dim3 rblock(64 * kResolveWaves);
      if (plane)
        hipLaunchKernelGGL((pdq_match_resolve_kernel<QT, true>), dim3((unsigned)rgrid), rblock, 0, stream, d, p,
                           n_rows, q, n_queries, threshold, hits, capacity, counter, index_base, ra.recs, rc, cap,
                           chunks8, groups);
      else
        hipLaunchKernelGGL((pdq_match_resolve_kernel<QT, false>), dim3((unsigned)rgrid), rblock, 0, stream, d, p,
                           n_rows, q, n_queries, threshold, hits, capacity, counter, index_base, ra.recs, rc, cap,
                           chunks8, groups);
      const hipError_t e = hipGetLastError();
      if (area) {
        // the next call's slot is clear only if this call's kernels were
        // all enqueued; otherwise the next call clears it with a launch
        if (e == hipSuccess) area->slot ^= 1u;
        area->slot_clear = e == hipSuccess;
      }
      return e;
    }
  }
  const uint64_t blocks = chunks8 * groups;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  if (reset) hipLaunchKernelGGL(reset_counts_kernel, dim3(1), dim3(1), 0, stream, nullptr, counter);
  if (plane)
    hipLaunchKernelGGL((pdq_match_mfma_kernel<QT, true, false>), dim3((unsigned)blocks), block, 0, stream, d, p,
                       n_rows, q, n_queries, threshold, hits, capacity, counter, index_base, groups, nullptr, nullptr,
                       0u, nullptr, nullptr);
  else
    hipLaunchKernelGGL((pdq_match_mfma_kernel<QT, false, false>), dim3((unsigned)blocks), block, 0, stream, d, p,
                       n_rows, q, n_queries, threshold, hits, capacity, counter, index_base, groups, nullptr, nullptr,
                       0u, nullptr, nullptr);
  return hipGetLastError();
}

// Prefix plane: plane[r * 4 + k] = word k of row r (the row's first 16
// bytes), zero for the padding rows up to a whole 16-row tile.
__global__ void build_prefix_plane_kernel(const uint4* __restrict__ db, uint64_t n_rows, uint4* __restrict__ plane,
                                          uint64_t n_padded) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_padded) return;
  plane[r] = r < n_rows ? db[2 * r] : make_uint4(0, 0, 0, 0);
}

uint64_t prefix_plane_bytes(uint64_t n_rows) { return (n_rows + 15) / 16 * 256; }

hipError_t launch_build_prefix_plane(const void* db, uint64_t n_rows, void* plane, hipStream_t stream) {
  const uint64_t n_padded = (n_rows + 15) / 16 * 16;
  if (n_padded == 0) return hipSuccess;
  const uint64_t blocks = (n_padded + 255) / 256;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(build_prefix_plane_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                     static_cast<const uint4*>(db), n_rows, static_cast<uint4*>(plane), n_padded);
  return hipGetLastError();
}

uint64_t mfma_chunks(uint64_t n_rows) {
  const uint64_t rows_per_block = (uint64_t)kMfmaWaves * kMfmaTilesPerWave * 16;
  return ((n_rows + rows_per_block - 1) / rows_per_block + 7) / 8 * 8;
}

template <int QT>
static hipError_t launch_shards_qt(const ShardView* shards, uint32_t n_shards, uint64_t total_chunks, bool plane,
                                   const void* queries, const uint32_t* d_nq, uint32_t nq_cap, uint32_t min_nq,
                                   uint32_t threshold, MatchHit* hits, uint64_t capacity, unsigned long long* counter,
                                   uint32_t grid, hipStream_t stream) {
  if (plane)
    hipLaunchKernelGGL((pdq_match_mfma_shards_kernel<QT, true>), dim3(grid), dim3(64 * kMfmaWaves), 0, stream, shards,
                       n_shards, total_chunks, static_cast<const uint32_t*>(queries), d_nq, nq_cap, min_nq, threshold,
                       hits, capacity, counter);
  else
    hipLaunchKernelGGL((pdq_match_mfma_shards_kernel<QT, false>), dim3(grid), dim3(64 * kMfmaWaves), 0, stream,
                       shards, n_shards, total_chunks, static_cast<const uint32_t*>(queries), d_nq, nq_cap, min_nq,
                       threshold, hits, capacity, counter);
  return hipGetLastError();
}

hipError_t launch_match_mfma_shards(int qt, const ShardView* shards, uint32_t n_shards, uint64_t total_chunks,
                                    bool plane, const void* queries, const uint32_t* d_nq, uint32_t nq_cap,
                                    uint32_t min_nq, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                                    unsigned long long* counter, uint32_t grid, hipStream_t stream) {
  if (threshold >= 64 || grid == 0 || grid % 8) return hipErrorInvalidValue;
  switch (qt) {
    case 2: return launch_shards_qt<2>(shards, n_shards, total_chunks, plane, queries, d_nq, nq_cap, min_nq, threshold,
                                       hits, capacity, counter, grid, stream);
    case 4: return launch_shards_qt<4>(shards, n_shards, total_chunks, plane, queries, d_nq, nq_cap, min_nq, threshold,
                                       hits, capacity, counter, grid, stream);
    case 8: return launch_shards_qt<8>(shards, n_shards, total_chunks, plane, queries, d_nq, nq_cap, min_nq, threshold,
                                       hits, capacity, counter, grid, stream);
    case 16: return launch_shards_qt<16>(shards, n_shards, total_chunks, plane, queries, d_nq, nq_cap, min_nq,
                                         threshold, hits, capacity, counter, grid, stream);
    default: return hipErrorInvalidValue;
  }
}

uint64_t match_record_area_bytes() { return (uint64_t)kRecordCap * sizeof(uint4) + 256; }

template <int QT>
static hipError_t launch_shards_shared_qt(const ShardView* shards, uint32_t n_shards, uint64_t total_chunks,
                                          bool plane, const void* queries, const uint32_t* d_nq, uint32_t nq_cap,
                                          uint32_t threshold, MatchHit* hits, uint64_t capacity,
                                          unsigned long long* counter, void* area, uint32_t grid, uint32_t rgrid,
                                          hipStream_t stream) {
  uint4* recs = static_cast<uint4*>(area);
  uint32_t* count = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(area) + (size_t)kRecordCap * sizeof(uint4));
  const uint32_t* q = static_cast<const uint32_t*>(queries);
  const dim3 block(64 * kMfmaWaves), rblock(64 * kResolveWaves);
  const uint32_t cap = record_cap();
  hipLaunchKernelGGL(reset_counts_kernel, dim3(1), dim3(1), 0, stream, count, nullptr);
  const bool tri = QT == 16 && threshold >= kTriMinThreshold;  // three products per accumulator
  if (plane) {
    if (tri)
      hipLaunchKernelGGL((pdq_match_mfma_shards_shared_kernel<QT, true, QT == 16>), dim3(grid), block, 0, stream,
                         shards, n_shards, total_chunks, q, d_nq, nq_cap, threshold, recs, count, cap);
    else
      hipLaunchKernelGGL((pdq_match_mfma_shards_shared_kernel<QT, true>), dim3(grid), block, 0, stream, shards,
                         n_shards, total_chunks, q, d_nq, nq_cap, threshold, recs, count, cap);
    hipLaunchKernelGGL((pdq_match_resolve_shards_kernel<QT, true>), dim3(rgrid), rblock, 0, stream, shards, n_shards,
                       total_chunks, q, d_nq, nq_cap, threshold, hits, capacity, counter, recs, count, cap);
  } else {
    if (tri)
      hipLaunchKernelGGL((pdq_match_mfma_shards_shared_kernel<QT, false, QT == 16>), dim3(grid), block, 0, stream,
                         shards, n_shards, total_chunks, q, d_nq, nq_cap, threshold, recs, count, cap);
    else
      hipLaunchKernelGGL((pdq_match_mfma_shards_shared_kernel<QT, false>), dim3(grid), block, 0, stream, shards,
                         n_shards, total_chunks, q, d_nq, nq_cap, threshold, recs, count, cap);
    hipLaunchKernelGGL((pdq_match_resolve_shards_kernel<QT, false>), dim3(rgrid), rblock, 0, stream, shards, n_shards,
                       total_chunks, q, d_nq, nq_cap, threshold, hits, capacity, counter, recs, count, cap);
  }
  return hipGetLastError();
}

hipError_t launch_match_mfma_shards_shared(int qt, const ShardView* shards, uint32_t n_shards, uint64_t total_chunks,
                                           bool plane, const void* queries, const uint32_t* d_nq, uint32_t nq_cap,
                                           uint32_t threshold, MatchHit* hits, uint64_t capacity,
                                           unsigned long long* counter, void* record_area, uint32_t grid,
                                           uint32_t rgrid, hipStream_t stream) {
  if (threshold >= 64 || grid == 0 || grid % 8 || rgrid == 0 || !record_area) return hipErrorInvalidValue;
  switch (qt) {
    case 2: return launch_shards_shared_qt<2>(shards, n_shards, total_chunks, plane, queries, d_nq, nq_cap, threshold,
                                              hits, capacity, counter, record_area, grid, rgrid, stream);
    case 4: return launch_shards_shared_qt<4>(shards, n_shards, total_chunks, plane, queries, d_nq, nq_cap, threshold,
                                              hits, capacity, counter, record_area, grid, rgrid, stream);
    case 8: return launch_shards_shared_qt<8>(shards, n_shards, total_chunks, plane, queries, d_nq, nq_cap, threshold,
                                              hits, capacity, counter, record_area, grid, rgrid, stream);
    case 16: return launch_shards_shared_qt<16>(shards, n_shards, total_chunks, plane, queries, d_nq, nq_cap,
                                                threshold, hits, capacity, counter, record_area, grid, rgrid, stream);
    default: return hipErrorInvalidValue;
  }
}

bool match_mfma_applicable(uint32_t n_queries, uint32_t threshold) { return threshold < 64 && n_queries >= 32; }

// Query tiles per wave that launch_match_mfma_qt runs for n_queries.
static int qt_for(int max_qt, uint32_t n_queries) {
  const uint32_t tiles = (n_queries + 15) / 16;
  if (tiles <= 2 || max_qt == 2) return 2;
  if (tiles <= 4 || max_qt == 4) return 4;
  // sixteen tiles per wave only where the shared-rows layout runs (4 query
  // groups or more): the per-wave layout at sixteen needs more registers
  // than it has
  if (tiles < 16 * (uint32_t)kMfmaWaves || max_qt == 8) return 8;
  return 16;
}

static hipError_t launch_match_mfma_qt(int max_qt, const void* db, const void* plane, uint64_t n_rows,
                                       const void* queries, uint32_t n_queries, uint32_t threshold, MatchHit* hits,
                                       uint64_t capacity, unsigned long long* counter, uint64_t index_base,
                                       hipStream_t stream, bool reset_counter, const RecordArea* fixed) {
  if (threshold >= 64) return hipErrorInvalidValue;
  if (n_rows == 0 || n_queries == 0) {
    if (!reset_counter) return hipSuccess;
    hipLaunchKernelGGL(reset_counts_kernel, dim3(1), dim3(1), 0, stream, nullptr, counter);
    return hipGetLastError();
  }
  switch (qt_for(max_qt, n_queries)) {
    case 2:
      return launch_qt<2>(db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base,
                          stream, reset_counter, fixed);
    case 4:
      return launch_qt<4>(db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base,
                          stream, reset_counter, fixed);
    case 8:
      return launch_qt<8>(db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base,
                          stream, reset_counter, fixed);
    default:
      return launch_qt<16>(db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base,
                           stream, reset_counter, fixed);
  }
}

#ifndef YPDQ_DEFAULT_QT
#define YPDQ_DEFAULT_QT 16
#endif
constexpr int kDefaultQt = YPDQ_DEFAULT_QT;

hipError_t launch_match_mfma(const void* db, const void* plane, uint64_t n_rows, const void* queries,
                             uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                             unsigned long long* counter, uint64_t index_base, hipStream_t stream,
                             bool reset_counter) {
  // up to sixteen 16-query tiles per wave: sixteen from 64 tiles (the
  // shared-rows layout) measured 3.5 % faster than eight per call on MI355X
  // (10M rows x 1024 queries); the per-wave layout stops at eight
  return launch_match_mfma_qt(kDefaultQt, db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter,
                              index_base, stream, reset_counter, nullptr);
}

uint64_t match_mfma_scratch_records(uint64_t n_rows, uint32_t n_queries) {
  const int qt = qt_for(kDefaultQt, n_queries);
  const uint32_t groups = (n_queries + qt * 16 - 1) / (qt * 16);
  if (n_rows == 0 || groups < (uint32_t)kMfmaWaves) return 0;  // the per-wave layout keeps no records
  const uint64_t subs8 = ((n_rows + kSharedRows - 1) / kSharedRows + 7) / 8 * 8;
  const uint64_t sgroups = (groups + kMfmaWaves - 1) / kMfmaWaves;
  return subs8 * sgroups * kMfmaWaves;  // at most one record per wave of each block
}

hipError_t launch_match_mfma_noalloc(const void* db, const void* plane, uint64_t n_rows, const void* queries,
                                     uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                                     unsigned long long* counter, uint64_t index_base, hipStream_t stream,
                                     bool reset_counter) {
  if (match_mfma_scratch_records(n_rows, n_queries) > kScratchRecords) return hipErrorInvalidValue;
  RecordArea ra;
  hipError_t e = scratch_area(&ra);
  if (e != hipSuccess) return e;
  return launch_match_mfma_qt(kDefaultQt, db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter,
                              index_base, stream, reset_counter, &ra);
}

}  // namespace ypdq
