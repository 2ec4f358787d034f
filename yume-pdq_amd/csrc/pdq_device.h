// Device-side helpers and the kernel/host interface shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ypdq {

constexpr uint32_t kDihedralExact = 0;  // sign algebra over all 16 columns
constexpr uint32_t kDihedralAvx2 = 1;   // the AVX2 kernel's column extent (x86.rs:330-376)
constexpr int kInputF32 = 0;
constexpr int kInputU8 = 1;
constexpr int kInputRgb8 = 2;   // 3 bytes per pixel, converted to luma while loading
constexpr int kInputRgba8 = 3;  // 4 bytes per pixel
// bytes of one 512 x 512 plane of an input kind (0 for an unknown kind)
inline uint64_t input_plane_bytes(int kind) {
  return 512ull * 512 * (kind == kInputF32 ? 4 : kind == kInputU8 ? 1 : kind == kInputRgb8 ? 3 : kind == kInputRgba8 ? 4 : 0);
}

// Per-image outputs of the hash kernels; every array is indexed by image.
struct HashOutputs {
  uint8_t* hashes;   // [n][8][32] NORMAL, FLOPPED, ROTATED_180, FLIPPED,
                     // ROTATED_90_FLOPPED, ROTATED_270, FLOPPED_ROTATED_270, ROTATED_90
  float* quality;    // [n][8]
  float* threshold;  // [n][8]
  float* pdqf;       // [n][256] NORMAL PDQF, or null
  float* buf1;       // [n][128*128] compressed plane (drop-in only), or null
  float* tmp;        // [n][128] last stage-1 row (drop-in only), or null
  float* dihedral_pdqf;  // [n][8][256] every orientation's PDQF, or null
};

struct MatchHit {
  uint32_t query;
  uint32_t distance;
  uint64_t index;
};

__device__ __forceinline__ float shfl_down1(float x) { return __shfl_down(x, 1, 64); }
__device__ __forceinline__ float shfl_xor(float x, int m) { return __shfl_xor(x, m, 64); }
__device__ __forceinline__ uint32_t shfl_xor_u(uint32_t x, int m) {
  return static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), m, 64));
}
__device__ __forceinline__ float shfl_idx(float x, int src) { return __shfl(x, src, 64); }

// Make LDS writes of this wavefront visible to its other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bool wave_any(bool b) { return __any(b ? 1 : 0) != 0; }
__device__ __forceinline__ bool wave_all(bool b) { return __all(b ? 1 : 0) != 0; }

// Fused hash of n_images planes, exact arithmetic, the layout chosen by
// hash_layout_for.  luma: (R, G, B) coefficients of the colour inputs, null
// for BT.601.
hipError_t launch_hash(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_images, uint32_t mode,
                       const HashOutputs& out, hipStream_t stream, const float* luma = nullptr);
// Grid for n_max images, image count read from d_n (capped at n_max).
hipError_t launch_hash_device_count(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_max,
                                    const uint32_t* d_n, uint32_t mode, const HashOutputs& out, hipStream_t stream);
// DCT arithmetic of the block-per-image hash kernel: the reference's exact
// FP32 chains, or matrix cores with bf16 / f16 operands (x3: hi + lo split
// operands, three products) and f32 accumulation.
enum DctPrecision { kDctExact = 0, kDctBf16 = 1, kDctF16 = 2, kDctBf16x3 = 3, kDctF16x3 = 4 };
// Wave-per-image form (d_n: device image count capped at n_max, or null).
hipError_t launch_hash_wave(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_max,
                            const uint32_t* d_n, uint32_t mode, const HashOutputs& out, const float* luma,
                            hipStream_t stream);
// Block-per-image form: `waves` (4 or 8) wavefronts per image; d_n as in
// launch_hash_device_count (or null); luma null for BT.601.
hipError_t launch_hash_block(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_max,
                             const uint32_t* d_n, uint32_t mode, int dct, int waves, const HashOutputs& out,
                             const float* luma, hipStream_t stream);
// Layout the automatic choice uses for n images: 1 = wave per image,
// 8 = block of eight waves per image.
int hash_layout_for(uint32_t n_images, int dct);
// The precise ReferenceKernel<f32> (f32 or u8 planes): hash, quality,
// threshold (the median) and PDQF of the NORMAL orientation per image;
// `work` holds hash_ref_workspace_bytes(n) bytes of scratch.
uint64_t hash_ref_workspace_bytes(uint32_t n_images);
hipError_t launch_hash_ref(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_images, void* work,
                           uint8_t* hashes, float* quality, float* threshold, float* pdqf, hipStream_t stream);
hipError_t launch_dihedrals(const float* pdqf, const float* thr_in, uint32_t n, uint32_t mode,
                            const HashOutputs& out, hipStream_t stream);
// Shape of a matcher launch: rows per lane per iteration, threads per block,
// resident blocks per CU for the grid-stride loop, non-temporal row loads.
struct MatchTuning {
  int rows_per_lane;
  int block;
  int blocks_per_cu;
  bool nontemporal;
};
MatchTuning default_match_tuning(uint64_t n_rows, uint32_t n_queries);
hipError_t launch_match_tuned(const MatchTuning& t, const void* db, uint64_t n_rows, const void* queries,
                              uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                              unsigned long long* counter, uint64_t index_base, hipStream_t stream);
hipError_t launch_match(const void* db, uint64_t n_rows, const void* queries, uint32_t n_queries, uint32_t threshold,
                        MatchHit* hits, uint64_t capacity, unsigned long long* counter, uint64_t index_base,
                        hipStream_t stream);
// A database shard as the persistent multi-shard matchers see it.  Work
// items are numbered over all shards: MFMA row chunks from chunk_begin (a
// multiple of 8), vector-ALU row blocks from rblock_begin.
struct ShardView {
  const uint4* rows;   // n_rows x 32 bytes
  const uint4* plane;  // prefix plane (may be null for the raw-row kernels)
  uint64_t n_rows;
  uint64_t index_base;  // global index of row 0
  uint64_t chunk_begin;
  uint64_t rblock_begin;
};
uint64_t mfma_chunks(uint64_t n_rows);  // MFMA work-item chunks of one shard (multiple of 8)
hipError_t launch_match_mfma_shards(int qt, const ShardView* shards, uint32_t n_shards, uint64_t total_chunks,
                                    bool plane, const void* queries, const uint32_t* d_nq, uint32_t nq_cap,
                                    uint32_t min_nq, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                                    unsigned long long* counter, uint32_t grid, hipStream_t stream);
// The shared-rows layout over a set of shards (4 or more query groups pay
// off): record-count reset, persistent streaming kernel, recheck kernel.
// record_area: match_record_area_bytes() of device memory used by no other
// launch in flight (a batch context's own).
uint64_t match_record_area_bytes();
// Grid of the shared-rows recheck kernel on a GPU of `cus` compute units.
uint32_t match_resolve_blocks(int cus);
hipError_t launch_match_mfma_shards_shared(int qt, const ShardView* shards, uint32_t n_shards, uint64_t total_chunks,
                                           bool plane, const void* queries, const uint32_t* d_nq, uint32_t nq_cap,
                                           uint32_t threshold, MatchHit* hits, uint64_t capacity,
                                           unsigned long long* counter, void* record_area, uint32_t grid,
                                           uint32_t rgrid, hipStream_t stream);
hipError_t launch_match_valu_shards(const MatchTuning& t, const ShardView* shards, uint32_t n_shards,
                                    uint64_t total_rblocks, bool plane, const void* queries, const uint32_t* d_nq,
                                    uint32_t nq_cap, uint32_t max_nq, uint32_t threshold, MatchHit* hits,
                                    uint64_t capacity, unsigned long long* counter, uint32_t grid, hipStream_t stream);
uint64_t valu_rblocks(const MatchTuning& t, uint64_t n_rows);  // VALU work items of one shard
// Vector-ALU scan over the prefix plane (16 bytes per row streamed).
MatchTuning default_plane_tuning(uint64_t n_rows, uint32_t n_queries);
hipError_t launch_match_plane_tuned(const MatchTuning& t, const void* db, const void* plane, uint64_t n_rows,
                                    const void* queries, uint32_t n_queries, uint32_t threshold, MatchHit* hits,
                                    uint64_t capacity, unsigned long long* counter, uint64_t index_base,
                                    hipStream_t stream);
hipError_t launch_match_plane(const void* db, const void* plane, uint64_t n_rows, const void* queries,
                              uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                              unsigned long long* counter, uint64_t index_base, hipStream_t stream);
// Matrix-core prefix filter + exact re-check (threshold < 64, >= 32 queries).
bool match_mfma_applicable(uint32_t n_queries, uint32_t threshold);
// plane: the database's prefix plane (launch_build_prefix_plane), or null to
// read the raw rows.  reset_counter: zero *counter first, in the one-thread
// kernel that clears the candidate-record count (one launch fewer than a
// separate reset).
hipError_t launch_match_mfma(const void* db, const void* plane, uint64_t n_rows, const void* queries,
                             uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                             unsigned long long* counter, uint64_t index_base, hipStream_t stream,
                             bool reset_counter = false);
// The same matcher with the shared-rows layout's candidate records in the
// library's static per-GPU scratch area instead of a per-stream allocation:
// no allocation at all.  match_mfma_scratch_records gives the records a call
// may write at most (0 when its layout keeps none); a call whose worst case
// exceeds match_scratch_records() fails with hipErrorInvalidValue before
// anything is enqueued.  Calls on one GPU share the area, so calls on
// different streams must not overlap.
uint32_t match_scratch_records();
uint64_t match_mfma_scratch_records(uint64_t n_rows, uint32_t n_queries);
hipError_t launch_match_mfma_noalloc(const void* db, const void* plane, uint64_t n_rows, const void* queries,
                                     uint32_t n_queries, uint32_t threshold, MatchHit* hits, uint64_t capacity,
                                     unsigned long long* counter, uint64_t index_base, hipStream_t stream,
                                     bool reset_counter);
// Lowers the shared-rows layout's record cap (0 = the default, 1 .. 2^20);
// false for a larger value.
bool set_match_record_cap(uint32_t cap);
uint64_t prefix_plane_bytes(uint64_t n_rows);
hipError_t launch_build_prefix_plane(const void* db, uint64_t n_rows, void* plane, hipStream_t stream);
hipError_t launch_match_first(const void* db, uint64_t n_rows, const void* queries, uint32_t n_queries,
                              uint32_t threshold, unsigned long long* first_key, hipStream_t stream);

// Staged kernels (one per Kernel trait method).
hipError_t launch_stage_compress(const void* planes, int input_kind, uint64_t stride, uint32_t n, float* buf1,
                                 hipStream_t s);
hipError_t launch_stage_dct2d(const float* buf1, uint32_t n, float* tmp, float* pdqf, hipStream_t s);
hipError_t launch_stage_gradient(const float* pdqf, uint32_t n, float* out, hipStream_t s);
hipError_t launch_stage_quantize(const float* pdqf, uint32_t n, float* thr, uint8_t* hash, hipStream_t s);
hipError_t launch_stage_threshold(const float* pdqf, const float* thr, uint32_t n, uint8_t* hash, hipStream_t s);
hipError_t launch_stage_transform(float* pdqf, uint32_t n, int op, hipStream_t s);
hipError_t launch_stage_luma(const uint8_t* px, int channels, uint64_t n, const float coeff[3], float* out,
                             hipStream_t s);

}  // namespace ypdq
