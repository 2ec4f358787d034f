// Exhaustive Hamming matcher for gfx950: every (query, row) pair with
// popcount(row ^ query) <= threshold over 256-bit PDQ hashes.
//
// Semantics follow the reference's matchers: CpuMatcher::scan/find
// (src/matching/mod.rs:232-257, 336-362) and VulkanMatcher
// (src/matching/vulkan.rs:111-146): inclusive threshold, distance = sum of
// popcounts over the 32-byte row.  The reference's matchers report only
// (row, needle); this one reports (query, index, distance) triples and also a
// first-hit mode equal to the lexicographically smallest (row, needle), which
// is what VulkanMatcher::find returns (atomicMin of needle + row * N).
//
// Layout: the database is resident in HBM as 32-byte rows (the hash bytes in
// the order the hash kernel writes them).  Each lane streams whole rows with
// two 16-byte loads; queries are wave-uniform and come through the scalar
// cache into SGPRs.  Per (row, query) the first 128 bits are compared first:
// their distance is a lower bound of the full distance, so a row whose
// 128-bit distance already exceeds the threshold is rejected exactly.  For
// random rows that happens for all but ~1e-9 of pairs, so the second half is
// almost never touched by the ALU (it still arrives with the same cache line).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pdq_device.h"

namespace ypdq {


// popcount(x) + acc: one v_bcnt_u32_b32
__device__ __forceinline__ uint32_t popc(uint32_t x, uint32_t acc) { return (uint32_t)__builtin_popcount(x) + acc; }

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Database rows are streamed once per scan: non-temporal loads.
__device__ __forceinline__ uint4 load_nt(const uint4* p) {
  const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}

struct Row256 {
  uint4 lo, hi;
};

__device__ __forceinline__ uint32_t dist128(const uint4& r, const uint32_t* q) {
  uint32_t d = popc(r.x ^ q[0], 0u);
  d = popc(r.y ^ q[1], d);
  d = popc(r.z ^ q[2], d);
  return popc(r.w ^ q[3], d);
}

template <bool kNT>
__device__ __forceinline__ uint4 load_row_half(const uint4* p) {
  if (kNT) return load_nt(p);
  return *p;
}

// kRows rows per lane per iteration, kBlock threads per block, grid-stride
// over blocks of kBlock * kRows rows.
template <bool kFirst, int kRows, bool kNT, int kBlock>
__global__ void __launch_bounds__(kBlock) pdq_match_kernel(const uint4* __restrict__ db, uint64_t n_rows,
                                                           const uint32_t* __restrict__ queries, uint32_t n_queries,
                                                           uint32_t threshold, MatchHit* __restrict__ hits,
                                                           uint64_t capacity, unsigned long long* __restrict__ counter,
                                                           uint64_t index_base) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock * kRows;
  for (uint64_t base = ((uint64_t)blockIdx.x * kBlock * kRows) + threadIdx.x; base < n_rows; base += stride) {
    Row256 row[kRows];
    bool valid[kRows];
#pragma unroll
    for (int u = 0; u < kRows; u++) {
      const uint64_t r = base + (uint64_t)u * kBlock;
      valid[u] = r < n_rows;
      if (valid[u]) {
        row[u].lo = load_row_half<kNT>(db + 2 * r);
        row[u].hi = load_row_half<kNT>(db + 2 * r + 1);
      } else {
        row[u].lo = make_uint4(0, 0, 0, 0);
        row[u].hi = make_uint4(0, 0, 0, 0);
      }
    }
    for (uint32_t qi = 0; qi < n_queries; qi++) {
      const uint32_t* q = queries + 8 * qi;
#pragma unroll
      for (int u = 0; u < kRows; u++) {
        const uint32_t d0 = dist128(row[u].lo, q);
        if (__builtin_expect(d0 <= threshold && valid[u], 0)) {
          const uint32_t d = d0 + dist128(row[u].hi, q + 4);
          if (d <= threshold) {
            const uint64_t r = base + (uint64_t)u * kBlock;
            if (kFirst) {
              atomicMin(counter, (unsigned long long)(r * n_queries + qi));
            } else {
              const unsigned long long slot = atomicAdd(counter, 1ull);
              if (slot < capacity) {
                MatchHit h;
                h.query = qi;
                h.distance = d;
                h.index = index_base + r;
                hits[slot] = h;
              }
            }
          }
        }
      }
    }
  }
}

// The same scan over the prefix plane (the first 16 bytes of every row,
// packed): a lane reads 16 bytes per row, and the second half of a row only
// when its prefix distance is within the threshold (for random rows, about
// once per 1e9 pairs).
template <bool kFirst, int kRows, int kBlock>
__global__ void __launch_bounds__(kBlock) pdq_match_plane_kernel(const uint4* __restrict__ db,
                                                                 const uint4* __restrict__ plane, uint64_t n_rows,
                                                                 const uint32_t* __restrict__ queries,
                                                                 uint32_t n_queries, uint32_t threshold,
                                                                 MatchHit* __restrict__ hits, uint64_t capacity,
                                                                 unsigned long long* __restrict__ counter,
                                                                 uint64_t index_base) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock * kRows;
  for (uint64_t base = ((uint64_t)blockIdx.x * kBlock * kRows) + threadIdx.x; base < n_rows; base += stride) {
    uint4 pre[kRows];
    bool valid[kRows];
#pragma unroll
    for (int u = 0; u < kRows; u++) {
      const uint64_t r = base + (uint64_t)u * kBlock;
      valid[u] = r < n_rows;
      pre[u] = valid[u] ? plane[r] : make_uint4(0, 0, 0, 0);
    }
    for (uint32_t qi = 0; qi < n_queries; qi++) {
      const uint32_t* q = queries + 8 * qi;
#pragma unroll
      for (int u = 0; u < kRows; u++) {
        const uint32_t d0 = dist128(pre[u], q);
        if (__builtin_expect(d0 <= threshold && valid[u], 0)) {
          const uint64_t r = base + (uint64_t)u * kBlock;
          const uint32_t d = d0 + dist128(db[2 * r + 1], q + 4);
          if (d <= threshold) {
            if (kFirst) {
              atomicMin(counter, (unsigned long long)(r * n_queries + qi));
            } else {
              const unsigned long long slot = atomicAdd(counter, 1ull);
              if (slot < capacity) {
                MatchHit h;
                h.query = qi;
                h.distance = d;
                h.index = index_base + r;
                hits[slot] = h;
              }
            }
          }
        }
      }
    }
  }
}

// One pass of kRows rows per lane (row base + u * kBlock) of one shard, over
// the prefix plane or the raw rows.
template <int kRows, int kBlock, bool kPlane>
__device__ __forceinline__ void valu_rows(const uint4* __restrict__ db, const uint4* __restrict__ plane,
                                          uint64_t n_rows, uint64_t base, const uint32_t* __restrict__ queries,
                                          uint32_t n_queries, uint32_t threshold, MatchHit* __restrict__ hits,
                                          uint64_t capacity, unsigned long long* __restrict__ counter,
                                          uint64_t index_base) {
  uint4 pre[kRows];
  bool valid[kRows];
#pragma unroll
  for (int u = 0; u < kRows; u++) {
    const uint64_t r = base + (uint64_t)u * kBlock;
    valid[u] = r < n_rows;
    pre[u] = valid[u] ? (kPlane ? plane[r] : db[2 * r]) : make_uint4(0, 0, 0, 0);
  }
  for (uint32_t qi = 0; qi < n_queries; qi++) {
    const uint32_t* q = queries + 8 * qi;
#pragma unroll
    for (int u = 0; u < kRows; u++) {
      const uint32_t d0 = dist128(pre[u], q);
      if (__builtin_expect(d0 <= threshold && valid[u], 0)) {
        const uint64_t r = base + (uint64_t)u * kBlock;
        const uint32_t d = d0 + dist128(db[2 * r + 1], q + 4);
        if (d <= threshold) {
          const unsigned long long slot = atomicAdd(counter, 1ull);
          if (slot < capacity) {
            MatchHit h;
            h.query = qi;
            h.distance = d;
            h.index = index_base + r;
            hits[slot] = h;
          }
        }
      }
    }
  }
}

// Persistent form over a set of database shards (work item = kBlock * kRows
// rows of one shard), the query count read from device memory; batches of
// max_nq or more queries are left to the matrix-core kernel.
template <int kRows, int kBlock, bool kPlane>
__global__ void __launch_bounds__(kBlock) pdq_match_shards_kernel(const ShardView* __restrict__ shards,
                                                                  uint32_t n_shards, uint64_t total_rblocks,
                                                                  const uint32_t* __restrict__ queries,
                                                                  const uint32_t* __restrict__ d_nq, uint32_t nq_cap,
                                                                  uint32_t max_nq, uint32_t threshold,
                                                                  MatchHit* __restrict__ hits, uint64_t capacity,
                                                                  unsigned long long* __restrict__ counter) {
  const uint32_t nq = d_nq ? min(*d_nq, nq_cap) : nq_cap;
  if (nq == 0 || nq >= max_nq) return;
  for (uint64_t item = blockIdx.x; item < total_rblocks; item += gridDim.x) {
    uint32_t sh = 0;
    while (sh + 1 < n_shards && shards[sh + 1].rblock_begin <= item) sh++;
    const ShardView& v = shards[sh];
    const uint64_t base = (item - v.rblock_begin) * (uint64_t)(kBlock * kRows) + threadIdx.x;
    valu_rows<kRows, kBlock, kPlane>(v.rows, v.plane, v.n_rows, base, queries, nq, threshold, hits, capacity,
                                     counter, v.index_base);
  }
}

uint64_t valu_rblocks(const MatchTuning& t, uint64_t n_rows) {
  const uint64_t per = (uint64_t)t.block * t.rows_per_lane;
  return (n_rows + per - 1) / per;
}

hipError_t launch_match_valu_shards(const MatchTuning& t, const ShardView* shards, uint32_t n_shards,
                                    uint64_t total_rblocks, bool plane, const void* queries, const uint32_t* d_nq,
                                    uint32_t nq_cap, uint32_t max_nq, uint32_t threshold, MatchHit* hits,
                                    uint64_t capacity, unsigned long long* counter, uint32_t grid, hipStream_t stream) {
  if (grid == 0) return hipErrorInvalidValue;
#define YPDQ_SHARDS(R, B)                                                                                           \
  if (t.rows_per_lane == R && t.block == B) {                                                                       \
    if (plane)                                                                                                      \
      hipLaunchKernelGGL((pdq_match_shards_kernel<R, B, true>), dim3(grid), dim3(B), 0, stream, shards, n_shards,   \
                         total_rblocks, static_cast<const uint32_t*>(queries), d_nq, nq_cap, max_nq, threshold,      \
                         hits, capacity, counter);                                                                  \
    else                                                                                                            \
      hipLaunchKernelGGL((pdq_match_shards_kernel<R, B, false>), dim3(grid), dim3(B), 0, stream, shards, n_shards,  \
                         total_rblocks, static_cast<const uint32_t*>(queries), d_nq, nq_cap, max_nq, threshold,      \
                         hits, capacity, counter);                                                                  \
    return hipGetLastError();                                                                                       \
  }
  YPDQ_SHARDS(1, 256) YPDQ_SHARDS(2, 256) YPDQ_SHARDS(4, 256) YPDQ_SHARDS(2, 512) YPDQ_SHARDS(4, 512)
#undef YPDQ_SHARDS
  return hipErrorInvalidValue;
}

static int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    hipDeviceProp_t p;
    cus = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) ? p.multiProcessorCount
                                                                                             : 256;
  }
  return cus;
}

template <bool kFirst, int kRows, bool kNT, int kBlock>
static hipError_t launch_cfg(const MatchTuning& t, const void* db, uint64_t n_rows, const void* queries,
                             uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                             unsigned long long* counter, uint64_t index_base, hipStream_t stream) {
  const uint64_t per_block = (uint64_t)kBlock * kRows;
  uint64_t need = (n_rows + per_block - 1) / per_block;
  const uint64_t cap = (uint64_t)device_cus() * (t.blocks_per_cu > 0 ? t.blocks_per_cu : 8);
  if (need > cap) need = cap;
  if (need == 0) need = 1;
  hipLaunchKernelGGL((pdq_match_kernel<kFirst, kRows, kNT, kBlock>), dim3((unsigned)need), dim3(kBlock), 0, stream,
                     static_cast<const uint4*>(db), n_rows, static_cast<const uint32_t*>(queries), n_queries,
                     threshold, hits, capacity, counter, index_base);
  return hipGetLastError();
}

template <bool kFirst>
static hipError_t dispatch(const MatchTuning& t, const void* db, uint64_t n_rows, const void* queries,
                           uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                           unsigned long long* counter, uint64_t index_base, hipStream_t stream) {
#define YPDQ_CASE(R, NT, B)                                                                                  \
  if (t.rows_per_lane == R && t.nontemporal == NT && t.block == B)                                         \
    return launch_cfg<kFirst, R, NT, B>(t, db, n_rows, queries, n_queries, threshold, hits, capacity, counter, \
                                        index_base, stream);
  YPDQ_CASE(1, true, 256) YPDQ_CASE(2, true, 256) YPDQ_CASE(4, true, 256) YPDQ_CASE(8, true, 256)
  YPDQ_CASE(1, false, 256) YPDQ_CASE(2, false, 256) YPDQ_CASE(4, false, 256) YPDQ_CASE(8, false, 256)
  YPDQ_CASE(2, true, 512) YPDQ_CASE(4, true, 512) YPDQ_CASE(2, false, 512) YPDQ_CASE(4, false, 512)
  YPDQ_CASE(2, true, 1024) YPDQ_CASE(4, true, 1024)
#undef YPDQ_CASE
  return hipErrorInvalidValue;
}

template <bool kFirst, int kRows, int kBlock>
static hipError_t launch_plane_cfg(const MatchTuning& t, const void* db, const void* plane, uint64_t n_rows,
                                   const void* queries, uint32_t n_queries, uint32_t threshold, MatchHit* hits,
                                   uint64_t capacity, unsigned long long* counter, uint64_t index_base,
                                   hipStream_t stream) {
  const uint64_t per_block = (uint64_t)kBlock * kRows;
  uint64_t need = (n_rows + per_block - 1) / per_block;
  const uint64_t cap = (uint64_t)device_cus() * (t.blocks_per_cu > 0 ? t.blocks_per_cu : 8);
  if (need > cap) need = cap;
  if (need == 0) need = 1;
  hipLaunchKernelGGL((pdq_match_plane_kernel<kFirst, kRows, kBlock>), dim3((unsigned)need), dim3(kBlock), 0, stream,
                     static_cast<const uint4*>(db), static_cast<const uint4*>(plane), n_rows,
                     static_cast<const uint32_t*>(queries), n_queries, threshold, hits, capacity, counter, index_base);
  return hipGetLastError();
}

template <bool kFirst>
static hipError_t dispatch_plane(const MatchTuning& t, const void* db, const void* plane, uint64_t n_rows,
                                 const void* queries, uint32_t n_queries, uint32_t threshold, MatchHit* hits,
                                 uint64_t capacity, unsigned long long* counter, uint64_t index_base,
                                 hipStream_t stream) {
#define YPDQ_CASE(R, B)                                                                                         \
  if (t.rows_per_lane == R && t.block == B)                                                                     \
    return launch_plane_cfg<kFirst, R, B>(t, db, plane, n_rows, queries, n_queries, threshold, hits, capacity, \
                                          counter, index_base, stream);
  YPDQ_CASE(1, 256) YPDQ_CASE(2, 256) YPDQ_CASE(4, 256) YPDQ_CASE(8, 256)
  YPDQ_CASE(2, 512) YPDQ_CASE(4, 512) YPDQ_CASE(8, 512) YPDQ_CASE(4, 1024)
#undef YPDQ_CASE
  return hipErrorInvalidValue;
}

MatchTuning default_plane_tuning(uint64_t n_rows, uint32_t n_queries) {
  (void)n_rows;
  (void)n_queries;
  MatchTuning t;
  // measured on MI355X, 10M rows x 8 queries: 512-thread blocks, two rows per
  // lane, 32 resident blocks per CU (25.5 us, 6.3 TB/s of plane bytes)
  t.rows_per_lane = 2;
  t.block = 512;
  t.blocks_per_cu = 32;
  t.nontemporal = false;
  return t;
}

hipError_t launch_match_plane_tuned(const MatchTuning& t, const void* db, const void* plane, uint64_t n_rows,
                                    const void* queries, uint32_t n_queries, uint32_t threshold, MatchHit* hits,
                                    uint64_t capacity, unsigned long long* counter, uint64_t index_base,
                                    hipStream_t stream) {
  if (n_rows == 0 || n_queries == 0) return hipSuccess;
  return dispatch_plane<false>(t, db, plane, n_rows, queries, n_queries, threshold, hits, capacity, counter,
                               index_base, stream);
}

hipError_t launch_match_plane(const void* db, const void* plane, uint64_t n_rows, const void* queries,
                              uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                              unsigned long long* counter, uint64_t index_base, hipStream_t stream) {
  return launch_match_plane_tuned(default_plane_tuning(n_rows, n_queries), db, plane, n_rows, queries, n_queries,
                                  threshold, hits, capacity, counter, index_base, stream);
}

MatchTuning default_match_tuning(uint64_t n_rows, uint32_t n_queries) {
  (void)n_rows;
  (void)n_queries;
  MatchTuning t;
  // measured on MI355X, 10M rows x 8 queries: 512-thread blocks, two rows per
  // lane, 32 resident blocks per CU, ordinary (cached) loads
  t.rows_per_lane = 2;
  t.block = 512;
  t.blocks_per_cu = 32;
  t.nontemporal = false;
  return t;
}

hipError_t launch_match_tuned(const MatchTuning& t, const void* db, uint64_t n_rows, const void* queries,
                              uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                              unsigned long long* counter, uint64_t index_base, hipStream_t stream) {
  if (n_rows == 0 || n_queries == 0) return hipSuccess;
  return dispatch<false>(t, db, n_rows, queries, n_queries, threshold, hits, capacity, counter, index_base, stream);
}

hipError_t launch_match(const void* db, uint64_t n_rows, const void* queries, uint32_t n_queries, uint32_t threshold,
                        MatchHit* hits, uint64_t capacity, unsigned long long* counter, uint64_t index_base,
                        hipStream_t stream) {
  return launch_match_tuned(default_match_tuning(n_rows, n_queries), db, n_rows, queries, n_queries, threshold, hits,
                            capacity, counter, index_base, stream);
}

hipError_t launch_match_first(const void* db, uint64_t n_rows, const void* queries, uint32_t n_queries,
                              uint32_t threshold, unsigned long long* first_key, hipStream_t stream) {
  if (n_rows == 0 || n_queries == 0) return hipSuccess;
  return dispatch<true>(default_match_tuning(n_rows, n_queries), db, n_rows, queries, n_queries, threshold, nullptr,
                        0, first_key, 0, stream);
}

}  // namespace ypdq
