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
__device__ __noinline__ void resolve_candidates(v4f cv, int t, uint64_t tbase, uint32_t qbase, int rn, int kq,
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
#pragma unroll
  for (int m = 0; m < kMaxChains; m++) mx[m] = INT32_MIN;
  const int ibound = __float_as_int(bound);
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
      tile_max<QT>(qa, rb, mx);
      if (a % kGroup == kGroup - 1) {
        const int mall = max(max(mx[0], mx[1]), max(mx[2], mx[3]));
        pending |= (uint64_t)(wave_any(mall >= ibound) ? 1u : 0u) << (t / kGroup);
#pragma unroll
        for (int m = 0; m < kMaxChains; m++) mx[m] = INT32_MIN;
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
// kMfmaTilesPerWave tiles x 16 rows; each wave streams its own rows) against
// query group `group` (QT x 16 queries).
template <int QT, bool kPlane>
__device__ __forceinline__ void mfma_item(const uint32_t* __restrict__ lut, const uint32_t* __restrict__ db,
                                          const uint32_t* __restrict__ plane, uint64_t n_rows, uint64_t chunk,
                                          uint32_t group, const uint32_t* __restrict__ queries, uint32_t n_queries,
                                          uint32_t threshold, MatchHit* __restrict__ hits, uint64_t capacity,
                                          unsigned long long* __restrict__ counter, uint64_t index_base) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
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

// One launch, one block per work item (row chunks x query groups).
template <int QT, bool kPlane>
__global__ void __launch_bounds__(64 * kMfmaWaves) __attribute__((amdgpu_waves_per_eu(QT >= 16 ? 3 : 5))) pdq_match_mfma_kernel(
    const uint32_t* __restrict__ db, const uint32_t* __restrict__ plane, uint64_t n_rows,
    const uint32_t* __restrict__ queries, uint32_t n_queries, uint32_t threshold, MatchHit* __restrict__ hits,
    uint64_t capacity, unsigned long long* __restrict__ counter, uint64_t index_base, uint32_t groups) {
  __shared__ uint32_t lut[256];
  init_lut(lut);
  uint64_t chunk;
  uint32_t group;
  item_coords(blockIdx.x, groups, chunk, group);
  mfma_item<QT, kPlane>(lut, db, plane, n_rows, chunk, group, queries, n_queries, threshold, hits, capacity, counter,
                        index_base);
}

// Persistent form over a set of database shards: a fixed grid (a multiple of
// 8 blocks) walks the work items of every shard, blockIdx.x + k * gridDim.x,
// which keeps each item on the XCD its number selects.  The query count is
// read from device memory (`d_nq`, capped at nq_cap), so one captured launch
// serves batches of any size; batches of fewer than min_nq queries are left
// to the vector-ALU kernel and this one returns at once.
template <int QT, bool kPlane>
__global__ void __launch_bounds__(64 * kMfmaWaves) __attribute__((amdgpu_waves_per_eu(QT >= 16 ? 3 : 5))) pdq_match_mfma_shards_kernel(
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
                          v.index_base);
  }
}

template <int QT>
static hipError_t launch_qt(const void* db, const void* plane, uint64_t n_rows, const void* queries,
                            uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                            unsigned long long* counter, uint64_t index_base, hipStream_t stream) {
  const uint64_t rows_per_block = (uint64_t)kMfmaWaves * kMfmaTilesPerWave * 16;
  const uint64_t chunks = (n_rows + rows_per_block - 1) / rows_per_block;
  const uint64_t chunks8 = (chunks + 7) / 8 * 8;  // whole rounds over the 8 XCDs
  const uint32_t groups = (n_queries + QT * 16 - 1) / (QT * 16);
  const uint64_t blocks = chunks8 * groups;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  if (plane)
    hipLaunchKernelGGL((pdq_match_mfma_kernel<QT, true>), dim3((unsigned)blocks), dim3(64 * kMfmaWaves), 0, stream,
                       static_cast<const uint32_t*>(db), static_cast<const uint32_t*>(plane), n_rows,
                       static_cast<const uint32_t*>(queries), n_queries, threshold, hits, capacity, counter,
                       index_base, groups);
  else
    hipLaunchKernelGGL((pdq_match_mfma_kernel<QT, false>), dim3((unsigned)blocks), dim3(64 * kMfmaWaves), 0, stream,
                       static_cast<const uint32_t*>(db), static_cast<const uint32_t*>(plane), n_rows,
                       static_cast<const uint32_t*>(queries), n_queries, threshold, hits, capacity, counter,
                       index_base, groups);
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

bool match_mfma_applicable(uint32_t n_queries, uint32_t threshold) { return threshold < 64 && n_queries >= 32; }

hipError_t launch_match_mfma_qt(int max_qt, const void* db, const void* plane, uint64_t n_rows, const void* queries,
                                uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                                unsigned long long* counter, uint64_t index_base, hipStream_t stream) {
  if (n_rows == 0 || n_queries == 0) return hipSuccess;
  if (threshold >= 64) return hipErrorInvalidValue;
  const uint32_t tiles = (n_queries + 15) / 16;
  if (tiles <= 2 || max_qt == 2)
    return launch_qt<2>(db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base, stream);
  if (tiles <= 4 || max_qt == 4)
    return launch_qt<4>(db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base, stream);
  if (tiles <= 8 || max_qt == 8)
    return launch_qt<8>(db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base, stream);
  return launch_qt<16>(db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base, stream);
}

#ifndef YPDQ_DEFAULT_QT
#define YPDQ_DEFAULT_QT 8
#endif
constexpr int kDefaultQt = YPDQ_DEFAULT_QT;

hipError_t launch_match_mfma(const void* db, const void* plane, uint64_t n_rows, const void* queries,
                             uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                             unsigned long long* counter, uint64_t index_base, hipStream_t stream) {
  // eight 16-query tiles per wave measured fastest on MI355X (10M rows x
  // 1024 / 8192 queries); larger tiles cost occupancy
  return launch_match_mfma_qt(kDefaultQt, db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter,
                              index_base, stream);
}

}  // namespace ypdq
