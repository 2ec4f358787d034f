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

#include "pdq_device.h"

namespace ypdq {

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

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
__device__ __noinline__ void resolve_candidates(const float* cv, int t, uint64_t tbase, uint32_t qbase, int rn, int kq,
                                                float bound, const uint32_t* __restrict__ queries,
                                                const uint32_t* __restrict__ db, uint32_t n_queries, uint64_t n_rows,
                                                uint32_t threshold, MatchHit* __restrict__ hits, uint64_t capacity,
                                                unsigned long long* __restrict__ counter, uint64_t index_base) {
  const uint64_t row = tbase + rn;
  for (int r = 0; r < 4; r++) {
    if (cv[r] >= bound) {
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

// One 16-row tile against every query tile of the wave.  The dot products are
// reduced to a running maximum as they come out of the matrix core; in the
// rare case the maximum clears the bound, the tile's products are recomputed
// and every surviving pair is re-checked with the exact 256-bit distance.
template <int QT>
__device__ __forceinline__ void tile_check(const Expanded (&qa)[QT], const Expanded& rb, uint64_t tbase, uint32_t qbase,
                                           int rn, int kq, float bound, const uint32_t* __restrict__ queries,
                                           const uint32_t* __restrict__ db, uint32_t n_queries, uint64_t n_rows,
                                           uint32_t threshold, MatchHit* __restrict__ hits, uint64_t capacity,
                                           unsigned long long* __restrict__ counter, uint64_t index_base) {
  float mx = -1.0e30f;
#pragma unroll
  for (int t = 0; t < QT; t++) {
    const v4f c = mfma_prefix(qa[t], rb);
    mx = __builtin_fmaxf(mx, __builtin_fmaxf(__builtin_fmaxf(c[0], c[1]), __builtin_fmaxf(c[2], c[3])));
  }
  if (__builtin_expect(wave_any(mx >= bound), 0)) {
    for (int t = 0; t < QT; t++) {
      const v4f c = mfma_prefix(qa[t], rb);
      float cv[4] = {c[0], c[1], c[2], c[3]};
      resolve_candidates(cv, t, tbase, qbase, rn, kq, bound, queries, db, n_queries, n_rows, threshold, hits,
                         capacity, counter, index_base);
    }
  }
}

// Grid: x = row chunks of kMfmaWaves * kMfmaTilesPerWave * 16 rows, y = query
// groups of QT * 16 queries.  Each wave holds its query group's operands in
// registers and streams its own 64 row tiles.
template <int QT>
__global__ void __launch_bounds__(64 * kMfmaWaves) pdq_match_mfma_kernel(
    const uint32_t* __restrict__ db, uint64_t n_rows, const uint32_t* __restrict__ queries, uint32_t n_queries,
    uint32_t threshold, MatchHit* __restrict__ hits, uint64_t capacity, unsigned long long* __restrict__ counter,
    uint64_t index_base) {
  __shared__ uint32_t lut[256];
  lut[threadIdx.x] = spread_byte(threadIdx.x);  // blockDim.x == 256
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rn = lane & 15;  // operand row (query m / db row n) of this lane
  const int kq = lane >> 4;  // K quarter = word index of the prefix
  const float bound = (float)(128 - 2 * (int)threshold);

  // query operands: tile t, query qbase + 16t + rn, word kq
  const uint32_t qbase = blockIdx.y * (uint32_t)(QT * 16);
  Expanded qa[QT];
#pragma unroll
  for (int t = 0; t < QT; t++) {
    const uint32_t qi = qbase + 16 * t + rn;
    if (qi < n_queries) {
      expand_word(queries[(size_t)qi * 8 + kq], lut, qa[t]);
    } else {
      qa[t].d[0] = qa[t].d[1] = qa[t].d[2] = qa[t].d[3] = 0;  // zero rows never reach the bound
    }
  }

  const uint64_t wave_rows = (uint64_t)kMfmaTilesPerWave * 16;
  const uint64_t row0 = ((uint64_t)blockIdx.x * kMfmaWaves + wave) * wave_rows;
  if (row0 >= n_rows) return;

  // four tiles in flight per wave: one 32-bit word per lane per tile
  constexpr int kAhead = 4;
  uint32_t ring[kAhead];
#pragma unroll
  for (int a = 0; a < kAhead; a++) {
    const uint64_t r = row0 + 16 * a + rn;
    ring[a] = r < n_rows ? db[r * 8 + kq] : 0u;
  }
#pragma unroll 1
  for (int tile = 0; tile < kMfmaTilesPerWave && row0 + (uint64_t)tile * 16 < n_rows; tile += kAhead) {
#pragma unroll
    for (int a = 0; a < kAhead; a++) {
      const uint64_t tbase = row0 + (uint64_t)(tile + a) * 16;
      const uint32_t w = ring[a];
      {
        const uint64_t r = tbase + 16 * kAhead + rn;
        ring[a] = (tile + a + kAhead < kMfmaTilesPerWave && r < n_rows) ? db[r * 8 + kq] : 0u;
      }
      if (tbase >= n_rows) break;
      Expanded rb;
      expand_word(w, lut, rb);
      if (tbase + rn >= n_rows) rb.d[0] = rb.d[1] = rb.d[2] = rb.d[3] = 0;
      tile_check<QT>(qa, rb, tbase, qbase, rn, kq, bound, queries, db, n_queries, n_rows, threshold, hits, capacity,
                     counter, index_base);
    }
  }
}

template <int QT>
static hipError_t launch_qt(const void* db, uint64_t n_rows, const void* queries, uint32_t n_queries,
                            uint32_t threshold, MatchHit* hits, uint64_t capacity, unsigned long long* counter,
                            uint64_t index_base, hipStream_t stream) {
  const uint64_t rows_per_block = (uint64_t)kMfmaWaves * kMfmaTilesPerWave * 16;
  const uint64_t gx = (n_rows + rows_per_block - 1) / rows_per_block;
  const uint32_t gy = (n_queries + QT * 16 - 1) / (QT * 16);
  if (gx > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((pdq_match_mfma_kernel<QT>), dim3((unsigned)gx, gy), dim3(64 * kMfmaWaves), 0, stream,
                     static_cast<const uint32_t*>(db), n_rows, static_cast<const uint32_t*>(queries), n_queries,
                     threshold, hits, capacity, counter, index_base);
  return hipGetLastError();
}

bool match_mfma_applicable(uint32_t n_queries, uint32_t threshold) { return threshold < 64 && n_queries >= 32; }

hipError_t launch_match_mfma(const void* db, uint64_t n_rows, const void* queries, uint32_t n_queries,
                             uint32_t threshold, MatchHit* hits, uint64_t capacity, unsigned long long* counter,
                             uint64_t index_base, hipStream_t stream) {
  if (n_rows == 0 || n_queries == 0) return hipSuccess;
  if (threshold >= 64) return hipErrorInvalidValue;
  const uint32_t tiles = (n_queries + 15) / 16;
  if (tiles <= 2) return launch_qt<2>(db, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base, stream);
  if (tiles <= 4) return launch_qt<4>(db, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base, stream);
  if (tiles <= 8) return launch_qt<8>(db, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base, stream);
  return launch_qt<16>(db, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base, stream);
}

}  // namespace ypdq
