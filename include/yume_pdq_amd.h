/*
 * yume_pdq_amd.h -- batched, device-resident C ABI of the MI355X (gfx950)
 * implementation of yume-pdq's hash and match paths (libyume_pdq.so).
 *
 * These entry points are additive: the reference has no batched or device
 * API (its FFI is the per-image pair in yume_pdq.h).  Each one names the
 * reference function(s) whose behaviour it batches.  All pointers prefixed
 * d_ are device pointers (HBM); `stream` is a hipStream_t or NULL for the
 * default stream.  Functions return YUME_PDQ_AMD_OK (0) or a negative error
 * code and never abort.  Work is enqueued asynchronously on `stream`.
 */
#ifndef YUME_PDQ_AMD_H
#define YUME_PDQ_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  YUME_PDQ_AMD_OK = 0,
  YUME_PDQ_AMD_EINVAL = -1,    /* bad argument (null pointer, size, mode) */
  YUME_PDQ_AMD_EHIP = -2,      /* a HIP runtime call failed */
  YUME_PDQ_AMD_ENOMEM = -3,    /* device allocation failed */
  YUME_PDQ_AMD_ENODEV = -4,    /* no usable GPU */
  YUME_PDQ_AMD_EOVERFLOW = -5, /* more hits than the buffer holds (count still exact) */
  YUME_PDQ_AMD_ETIMEOUT = -6,  /* a wait ran out of time */
  YUME_PDQ_AMD_EIO = -7,       /* a read or write of the stream failed */
  YUME_PDQ_AMD_EEOF = -8,      /* the input stream ended inside a frame */
  YUME_PDQ_AMD_ESCRATCH = -9   /* more scratch than yume_pdq_amd_match_noalloc holds */
};

/* Input planes: f32 luma (0..255), u8 luma, or 8-bit colour converted to luma
 * while the plane is read -- RGB8 as the Kernel trait's cvt_rgb8_to_luma8f
 * (src/kernel/mod.rs:303-322), RGBA8 as the AVX2 kernel's
 * cvt_rgba8_to_luma8f (src/kernel/x86.rs:686-753), BT.601 coefficients
 * (src/kernel/constants.rs:23) unless given. */
enum {
  YUME_PDQ_AMD_INPUT_F32 = 0,
  YUME_PDQ_AMD_INPUT_U8 = 1,
  YUME_PDQ_AMD_INPUT_RGB8 = 2,
  YUME_PDQ_AMD_INPUT_RGBA8 = 3
};

/* Dihedral sign algebra.  AVX2 reproduces the reference's x86 smart kernel,
 * whose column negation reaches columns 0..7 only (src/kernel/x86.rs:330-376,
 * visited by src/lib.rs:391-418): its FLOPPED, ROTATED_180,
 * FLOPPED_ROTATED_270 and ROTATED_90 hashes are the reference's, not the true
 * hashes of those orientations.  AVX2 is the default wherever a default
 * exists (NULL options, the C++ ServerConfig, the Python binding).  EXACT
 * negates every second column (the AVX-512 and scalar kernels,
 * src/kernel/x86.rs:1026-1074, src/kernel/mod.rs:368-382): the hashes of the
 * truly flipped and rotated compressed plane, an explicit opt-in.  Both cost
 * the same. */
enum { YUME_PDQ_AMD_DIHEDRAL_EXACT = 0, YUME_PDQ_AMD_DIHEDRAL_AVX2 = 1 };

/* Order of the 8 hashes per image in every batched output. */
enum {
  YUME_PDQ_AMD_D_NORMAL = 0,
  YUME_PDQ_AMD_D_FLOPPED = 1,
  YUME_PDQ_AMD_D_ROTATED_180 = 2,
  YUME_PDQ_AMD_D_FLIPPED = 3,
  YUME_PDQ_AMD_D_ROTATED_90_FLOPPED = 4,
  YUME_PDQ_AMD_D_ROTATED_270 = 5,
  YUME_PDQ_AMD_D_FLOPPED_ROTATED_270 = 6,
  YUME_PDQ_AMD_D_ROTATED_90 = 7
};

/* One match: distance(query, row) <= threshold.  index is the database row
 * (plus the index_base given to the call). */
typedef struct YumePdqAmdHit {
  uint32_t query;
  uint32_t distance;
  uint64_t index;
} YumePdqAmdHit;

/* flags for yume_pdq_amd_match */
#define YUME_PDQ_AMD_MATCH_APPEND 1u /* keep *d_count: append after earlier hits */
#define YUME_PDQ_AMD_MATCH_VALU 2u   /* force the vector-ALU kernel */
#define YUME_PDQ_AMD_MATCH_MFMA 4u   /* force the matrix-core kernel (threshold < 64) */

const char *yume_pdq_amd_strerror(int code);

/* Number of visible GPUs and the name / CU count of `device`. */
int yume_pdq_amd_device_info(int device, char *name, size_t name_len, int *compute_units);

/*
 * Batched hash_get_threshold (src/lib.rs:427-455) + visit_dihedrals
 * (src/lib.rs:358-420), AVX2 arithmetic (src/kernel/x86.rs:100-682).
 *
 * d_planes: n_images planes of 512x512 (f32 0..255, u8, RGB8 or RGBA8 pixels),
 *           plane i at byte offset i * plane_stride_bytes (>= one plane,
 *           a multiple of 16; d_planes 16-byte aligned).
 * d_hashes: n_images x 8 x 32 bytes, orientations in YUME_PDQ_AMD_D_* order.
 * d_quality, d_threshold: n_images x 8 floats (per-orientation values; the
 *           threshold of a transposed orientation is the one it was
 *           re-thresholded with, as visit_dihedrals passes it).
 * d_pdqf:   n_images x 256 floats of NORMAL PDQF, or NULL.
 */
int yume_pdq_amd_hash_batch(const void *d_planes, int input_kind, uint64_t plane_stride_bytes,
                            uint32_t n_images, uint32_t dihedral_mode, uint8_t *d_hashes,
                            float *d_quality, float *d_threshold, float *d_pdqf, void *stream);

/* yume_pdq_amd_hash_batch with the (R, G, B) luma coefficients of colour
 * inputs given (luma_coeff[3]; NULL = BT.601).  Fuses cvt_*_to_luma8f into
 * the hash: the result equals converting each row first and hashing the f32
 * luma plane. */
int yume_pdq_amd_hash_batch_luma(const void *d_planes, int input_kind, uint64_t plane_stride_bytes,
                                 uint32_t n_images, uint32_t dihedral_mode, const float *luma_coeff,
                                 uint8_t *d_hashes, float *d_quality, float *d_threshold, float *d_pdqf,
                                 void *stream);

/* DCT arithmetic of yume_pdq_amd_hash_batch_ex.  EXACT is the reference's
 * FP32 operation order (bit-identical hashes).  The others run both DCT
 * stages on matrix cores with operands rounded to bf16 / f16 and f32
 * accumulation; the X3 forms split each operand into hi + lo and sum three
 * products.  Their hashes may differ from the reference's in bits whose
 * coefficient lies near the median; quality and threshold stay FP32
 * computations over the approximate coefficients.  Planes holding NaN/Inf
 * always take the exact arithmetic. */
enum {
  YUME_PDQ_AMD_DCT_EXACT = 0,
  YUME_PDQ_AMD_DCT_BF16 = 1,
  YUME_PDQ_AMD_DCT_F16 = 2,
  YUME_PDQ_AMD_DCT_BF16X3 = 3,
  YUME_PDQ_AMD_DCT_F16X3 = 4
};
/* Kernel layout: one wavefront per image, or 4 / 8 / 16 wavefronts sharing
 * an image (lower latency per image, fills the GPU with fewer images; 16:
 * exact arithmetic only).  AUTO picks by batch size.  The exact results do
 * not depend on the layout. */
enum {
  YUME_PDQ_AMD_LAYOUT_AUTO = 0,
  YUME_PDQ_AMD_LAYOUT_WAVE = 1,
  YUME_PDQ_AMD_LAYOUT_BLOCK4 = 2,
  YUME_PDQ_AMD_LAYOUT_BLOCK8 = 3,
  YUME_PDQ_AMD_LAYOUT_BLOCK16 = 4
};
typedef struct YumePdqAmdHashOptions {
  uint32_t dihedral_mode;   /* YUME_PDQ_AMD_DIHEDRAL_* */
  uint32_t dct_precision;   /* YUME_PDQ_AMD_DCT_* (not EXACT: a BLOCK layout) */
  uint32_t layout;          /* YUME_PDQ_AMD_LAYOUT_* */
  const float *luma_coeff;  /* (R, G, B) for colour inputs, NULL = BT.601 */
} YumePdqAmdHashOptions;

/* yume_pdq_amd_hash_batch with options (NULL = the reference's AVX2
 * dihedrals, the exact DCT, the automatic layout, BT.601).  d_buf1: n_images x 128 x 128 floats of each
 * compressed plane (jarosz_compress output with row and column 127 zero), or
 * NULL. */
int yume_pdq_amd_hash_batch_ex(const void *d_planes, int input_kind, uint64_t plane_stride_bytes,
                               uint32_t n_images, const YumePdqAmdHashOptions *opts, uint8_t *d_hashes,
                               float *d_quality, float *d_threshold, float *d_pdqf, float *d_buf1,
                               void *stream);

/*
 * The precise kernel, hash_get_threshold with ReferenceKernel<f32>
 * (src/kernel/mod.rs:691-821, src/kernel/ref.rs): two passes of a 3-wide box
 * filter along rows and columns, sampling at floor((2i + 1) 512 / 254), the
 * sequential DCT, Torben's median.  The NORMAL orientation only.
 * input_kind: YUME_PDQ_AMD_INPUT_F32 or _U8.  d_workspace: at least
 * yume_pdq_amd_hash_ref_workspace_bytes(n_images) bytes of device memory
 * (2 MiB per image).  d_hashes n x 32, d_quality and d_threshold n floats,
 * d_pdqf n x 256 floats or NULL.
 */
uint64_t yume_pdq_amd_hash_ref_workspace_bytes(uint32_t n_images);
int yume_pdq_amd_hash_ref_batch(const void *d_planes, int input_kind, uint64_t plane_stride_bytes,
                                uint32_t n_images, void *d_workspace, uint64_t workspace_bytes,
                                uint8_t *d_hashes, float *d_quality, float *d_threshold, float *d_pdqf,
                                void *stream);

/*
 * Batched visit_dihedrals (src/lib.rs:358-420) from given PDQFs.
 * d_pdqf: n x 256; d_thr_in: n thresholds (as returned with each PDQF).
 * Outputs as yume_pdq_amd_hash_batch (entry 0 = the NORMAL orientation);
 * d_dihedral_pdqf: n x 8 x 256 floats of each orientation's PDQF, or NULL.
 */
int yume_pdq_amd_dihedrals_batch(const float *d_pdqf, const float *d_thr_in, uint32_t n,
                                 uint32_t dihedral_mode, uint8_t *d_hashes, float *d_quality,
                                 float *d_threshold, float *d_dihedral_pdqf, void *stream);

/*
 * Exhaustive match (PDQMatcher::find with every hit visited,
 * src/matching/mod.rs:336-362, 365-435; VulkanMatcher src/matching/vulkan.rs).
 * d_db: n_rows x 32 bytes; d_queries: n_queries x 32 bytes (hash byte order).
 * Every (query, row) with popcount(query ^ row) <= threshold is written to
 * d_hits (in no particular order; sort by (index, query) for the reference's
 * visiting order) up to `capacity`; *d_count receives the exact number of
 * hits (reset first unless flags has YUME_PDQ_AMD_MATCH_APPEND).  The call
 * does not wait; read *d_count after synchronising `stream`.  Batches of 32+
 * queries at threshold < 64 run on the matrix cores (an exact fp4 +-1 dot
 * product over the first 128 bits as a filter, then the full 256-bit integer
 * distance for every surviving pair); smaller batches on the vector ALUs.
 * Both give the same set; flags MATCH_VALU / MATCH_MFMA force one.
 */
int yume_pdq_amd_match(const uint8_t *d_db, uint64_t n_rows, const uint8_t *d_queries,
                       uint32_t n_queries, uint32_t threshold, YumePdqAmdHit *d_hits,
                       uint64_t capacity, unsigned long long *d_count, uint64_t index_base,
                       uint32_t flags, void *stream);

/*
 * Prefix plane: the first 16 bytes of every database row, packed (row r at
 * byte 16 r, zero-padded to a whole number of 16-row tiles), built once per
 * database (and again after rows change).  Every pair is decided on these 128
 * bits except the rare ones whose prefix distance is within the threshold,
 * which read the rest of their row; so both matchers stream 16 bytes per row
 * instead of 32, in whole cache lines.
 * No reference counterpart (a device-side index of the database; the
 * reference's VulkanMatcher likewise keeps its own device copy,
 * src/matching/vulkan.rs:40-64).
 */
uint64_t yume_pdq_amd_prefix_plane_bytes(uint64_t n_rows);
int yume_pdq_amd_build_prefix_plane(const uint8_t *d_db, uint64_t n_rows, void *d_plane, void *stream);
/* yume_pdq_amd_match using the prefix plane of d_db (same results; the plane
 * must have been built from the current rows). */
int yume_pdq_amd_match_with_plane(const uint8_t *d_db, const void *d_plane, uint64_t n_rows,
                                  const uint8_t *d_queries, uint32_t n_queries, uint32_t threshold,
                                  YumePdqAmdHit *d_hits, uint64_t capacity, unsigned long long *d_count,
                                  uint64_t index_base, uint32_t flags, void *stream);

/*
 * yume_pdq_amd_match_with_plane (d_plane given) or yume_pdq_amd_match
 * (d_plane NULL) that allocates nothing, for callers that cannot call
 * hipMalloc: code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation.  No reference counterpart.  Same arguments, flags
 * and results.  Where the matrix-core matcher keeps candidate records (batches
 * of 385+ queries), they go to one static scratch area per GPU in the
 * library's own device memory (2^20 records of 16 bytes + 256 bytes, held
 * from the library's load until the process ends) instead of an area
 * allocated per stream.  Limit: a call's worst case -- one record per
 * query group of g queries and 1024 database rows (g = 256 from 1009
 * queries up, 128 below), i.e.
 * ceil(ceil(n_rows / 1024) / 8) * 8 * ceil(n_queries / (4 g)) * 4 records --
 * must fit the area (at 1024 queries: up to 268M rows; at 10M rows: up to
 * 26,624 queries); a larger call returns YUME_PDQ_AMD_ESCRATCH and enqueues
 * nothing.  Every call on one GPU shares the area, which keeps the last
 * call's records (the call clears the count it starts from): calls on
 * different streams of one GPU that overlap in time race and must be
 * serialised by the caller.
 */
int yume_pdq_amd_match_noalloc(const uint8_t *d_db, const void *d_plane, uint64_t n_rows,
                               const uint8_t *d_queries, uint32_t n_queries, uint32_t threshold,
                               YumePdqAmdHit *d_hits, uint64_t capacity, unsigned long long *d_count,
                               uint64_t index_base, uint32_t flags, void *stream);

/*
 * Lowers the number of candidate records the matrix-core matcher's
 * shared-rows layout keeps per call (cap 1 .. 2^20; 0 restores 2^20, the
 * initial value unless the environment's YUME_PDQ_AMD_MATCH_RECORD_CAP gives
 * another).  Past the cap a call still returns the exact set: it rechecks
 * every row chunk instead of its records.  For tests and measurement; takes
 * effect at the next call (batch contexts: at their next graph capture).
 * EINVAL for a cap above 2^20.  No reference counterpart.
 */
int yume_pdq_amd_match_set_record_cap(uint32_t cap);

/*
 * Hash and match in one call: yume_pdq_amd_hash_batch of n_images planes, then
 * the match of their 8 n_images dihedral hashes (query 8 i + d: image i,
 * orientation d) against the database -- over d_plane when given (as
 * yume_pdq_amd_match_with_plane), else over the raw rows -- with the hit
 * counter reset first, all enqueued back to back on `stream` after every
 * argument has been checked.  When h_count is not NULL it must be page-locked
 * host memory from hipHostMalloc: the GPU stores the hit count there after the
 * matcher, so once `stream` is synchronised the caller reads it without a
 * copy.  The call does not wait.
 * No reference counterpart: the batched hash + match of the new device ABI
 * (one host call per batch where the reference hashes per image,
 * src/lib.rs:251-283, and matches per 2048-row batch,
 * src/matching/mod.rs:232-257).
 */
int yume_pdq_amd_hash_match_batch(const void *d_planes, int input_kind, uint64_t plane_stride_bytes,
                                  uint32_t n_images, uint32_t dihedral_mode, const uint8_t *d_db,
                                  const void *d_plane, uint64_t n_rows, uint32_t threshold,
                                  uint8_t *d_hashes, float *d_quality, float *d_threshold,
                                  YumePdqAmdHit *d_hits, uint64_t capacity, unsigned long long *d_count,
                                  unsigned long long *h_count, void *stream);

/* Which kernel yume_pdq_amd_match would run for these arguments:
 * YUME_PDQ_AMD_MATCH_VALU or YUME_PDQ_AMD_MATCH_MFMA (or EINVAL for
 * contradictory flags).  No reference counterpart (path introspection for
 * schedulers and measurement). */
int yume_pdq_amd_match_path(uint32_t n_queries, uint32_t threshold, uint32_t flags);

/*
 * First hit in (row, query) order -- what VulkanMatcher::find reports
 * (atomicMin(needle + row * N), src/matching/vulkan.rs:133-136) and where
 * CpuMatcher::find stops when its callback returns Some on the first call.
 * *d_first receives row * n_queries + query, or UINT64_MAX for no hit.
 */
int yume_pdq_amd_match_first(const uint8_t *d_db, uint64_t n_rows, const uint8_t *d_queries,
                             uint32_t n_queries, uint32_t threshold, unsigned long long *d_first,
                             void *stream);

/*
 * Staged entry points: one GPU launch per method of the reference's Kernel
 * trait (src/kernel/mod.rs:268-497), with the AVX2 smart kernel's arithmetic.
 * n images per call; every buffer is device memory laid out image after image.
 */
/* Kernel::jarosz_compress (x86.rs:626-682): planes -> n x 128 x 128 (row and
 * column 127 zero). */
int yume_pdq_amd_stage_compress(const void *d_planes, int input_kind, uint64_t plane_stride_bytes,
                                uint32_t n, float *d_buf1, void *stream);
/* Kernel::dct2d (x86.rs:100-281): n x 128 x 128 -> n x 16 x 16; d_tmp (n x 128,
 * nullable) receives the last stage-1 row as the reference leaves it. */
int yume_pdq_amd_stage_dct2d(const float *d_buf1, uint32_t n, float *d_tmp, float *d_pdqf, void *stream);
/* Kernel::sum_of_gradients (x86.rs:840-865): n PDQFs -> n floats. */
int yume_pdq_amd_stage_sum_of_gradients(const float *d_pdqf, uint32_t n, float *d_out, void *stream);
/* Kernel::quantize (x86.rs:440-621): n PDQFs -> n thresholds, n x 32 hash bytes. */
int yume_pdq_amd_stage_quantize(const float *d_pdqf, uint32_t n, float *d_threshold, uint8_t *d_hash,
                                void *stream);
/* threshold_2d_f32 (threshold.rs:37-119): n PDQFs + n thresholds -> n x 32. */
int yume_pdq_amd_stage_threshold(const float *d_pdqf, const float *d_threshold, uint32_t n,
                                 uint8_t *d_hash, void *stream);

/* In-place PDQF transforms (Kernel::pdqf_t, pdqf_negate_alt_cols/rows,
 * pdqf_negate_off_diagonals; mod.rs:345-405, x86.rs:285-436). */
enum {
  YUME_PDQ_AMD_T_TRANSPOSE = 0,
  YUME_PDQ_AMD_T_NEG_COLS_EVEN_AVX2 = 1, /* negate_alt_cols::<false>, AVX2 extent (cols 0..7) */
  YUME_PDQ_AMD_T_NEG_COLS_ODD_AVX2 = 2,  /* negate_alt_cols::<true>,  AVX2 extent */
  YUME_PDQ_AMD_T_NEG_COLS_EVEN = 3,      /* negate_alt_cols::<false>, all columns */
  YUME_PDQ_AMD_T_NEG_COLS_ODD = 4,       /* negate_alt_cols::<true>,  all columns */
  YUME_PDQ_AMD_T_NEG_ROWS_ODD = 5,       /* negate_alt_rows::<false> (rows 1,3,..) */
  YUME_PDQ_AMD_T_NEG_ROWS_EVEN = 6,      /* negate_alt_rows::<true>  (rows 0,2,..) */
  YUME_PDQ_AMD_T_NEG_OFF_DIAGONALS = 7
};
int yume_pdq_amd_stage_transform(float *d_pdqf, uint32_t n, int op, void *stream);

/* Kernel::cvt_rgb8_to_luma8f / cvt_rgba8_to_luma8f (mod.rs:303-340,
 * x86.rs:686-753): n_pixels of 3 or 4 bytes -> n_pixels floats, coefficients
 * (r, g, b) (BT.601: 0.299, 0.587, 0.114, constants.rs:23). */
int yume_pdq_amd_stage_luma(const uint8_t *d_pixels, int channels, uint64_t n_pixels, float r_coeff,
                            float g_coeff, float b_coeff, float *d_out, void *stream);

/*
 * Batching layer.  No reference counterpart: the reference hashes one image
 * per call and matches one batch of needles against one database
 * (src/matching/mod.rs, src/matching/vulkan.rs); these entry points batch
 * both over many images and many database shards.
 *
 * Shard set: database shards of any size resident in HBM, each with its
 * prefix plane and the global index of its first row.  Rows given on the host
 * are copied in (owned); rows given on the device are referenced (they must
 * outlive the set, 16-byte aligned).  Do not add shards while batches that use
 * the set are running.
 */
typedef struct YumePdqAmdShards YumePdqAmdShards;
int yume_pdq_amd_shards_create(YumePdqAmdShards **out);
void yume_pdq_amd_shards_destroy(YumePdqAmdShards *shards);
int yume_pdq_amd_shards_add(YumePdqAmdShards *shards, const uint8_t *rows, uint64_t n_rows,
                            uint64_t index_base, int rows_on_device, uint32_t *shard_id);
/* A raw file of N x 32-byte rows (the hash byte order), split into shards of
 * rows_per_shard rows; global indices start at index_base. */
int yume_pdq_amd_shards_load_raw(YumePdqAmdShards *shards, const char *path, uint64_t rows_per_shard,
                                 uint64_t index_base, uint32_t *n_added);
/* In place, on `stream`: row d_local_rows[i] of shard `shard` becomes
 * d_new_rows[i] (32 bytes each, prefix plane updated too); indices must be
 * distinct; out-of-range indices are skipped and counted into *d_bad. */
int yume_pdq_amd_shards_replace(YumePdqAmdShards *shards, uint32_t shard, const uint64_t *d_local_rows,
                                const uint8_t *d_new_rows, uint32_t n, unsigned long long *d_bad,
                                void *stream);
int yume_pdq_amd_shards_info(const YumePdqAmdShards *shards, uint32_t *n_shards, uint64_t *total_rows,
                             uint64_t *device_bytes);
/* yume_pdq_amd_match over every shard of the set in one persistent launch;
 * hit indices are global (shard index_base + row). */
int yume_pdq_amd_shards_match(const YumePdqAmdShards *shards, const uint8_t *d_queries,
                              uint32_t n_queries, uint32_t threshold, YumePdqAmdHit *d_hits,
                              uint64_t capacity, unsigned long long *d_count, uint32_t flags,
                              void *stream);

/*
 * Batch context: device buffers for up to max_images planes and their
 * outputs, and one HIP graph per matcher choice holding the hash kernel, the
 * hit-counter reset and a persistent matcher over every shard.  Both kernels
 * read the batch size from device memory, so a batch of any size replays the
 * same graph.  run() enqueues on the context's own stream (it first waits
 * for the context's previous batch); read the outputs after synchronising
 * that stream.  Contexts over one shard set may run concurrently on their
 * streams.
 */
typedef struct YumePdqAmdBatcher YumePdqAmdBatcher;
typedef struct YumePdqAmdBatchBuffers {
  void *d_input;              /* max_images planes, input_stride_bytes apart */
  uint64_t input_stride_bytes;
  uint8_t *d_hashes;          /* max_images x 8 x 32 */
  float *d_quality;           /* max_images x 8 */
  float *d_threshold;         /* max_images x 8 */
  YumePdqAmdHit *d_hits;      /* hit_capacity records */
  uint64_t hit_capacity;
  unsigned long long *d_count;
  void *stream;               /* hipStream_t of the context */
  uint32_t max_images;
  uint64_t device_bytes;      /* HBM held by the context */
} YumePdqAmdBatchBuffers;
typedef struct YumePdqAmdTuneRecord {
  uint32_t n_queries;
  uint32_t path;              /* YUME_PDQ_AMD_MATCH_VALU or _MFMA */
  uint32_t query_tiles;       /* 16-query tiles per wave (matrix cores) */
  float ms;
} YumePdqAmdTuneRecord;
int yume_pdq_amd_batcher_create(YumePdqAmdShards *shards, uint32_t max_images, int input_kind,
                                uint32_t dihedral_mode, uint32_t threshold, uint64_t hit_capacity,
                                YumePdqAmdBatcher **out);
void yume_pdq_amd_batcher_destroy(YumePdqAmdBatcher *b);
int yume_pdq_amd_batcher_buffers(const YumePdqAmdBatcher *b, YumePdqAmdBatchBuffers *out);
/* Hash input planes [0, n_images) and match their 8 n_images hashes against
 * every shard (use_graph = 0 enqueues the same kernels without a graph). */
int yume_pdq_amd_batcher_run(YumePdqAmdBatcher *b, uint32_t n_images, int use_graph);
/* Time the vector-ALU matcher and the matrix-core matcher at 2/4/8/16 query
 * tiles per wave for query counts 8, 16, 32, ... 8 max_images on the set's
 * shards, and keep the fastest per bucket in the process-wide tuning cache:
 * every batcher and server context on a shard set of the same shape (its
 * shards' row counts, in order), GPU and threshold uses them from then on. */
int yume_pdq_amd_batcher_autotune(YumePdqAmdBatcher *b, int iterations);
/* The tuning cache as a text file, to carry choices to another process:
 * save returns the entries written, load the entries read and merged (a
 * malformed file loads nothing: EINVAL), clear the entries dropped.  No
 * counterpart in the reference. */
int yume_pdq_amd_tune_save(const char *path);
int yume_pdq_amd_tune_load(const char *path);
int yume_pdq_amd_tune_clear(void);
int yume_pdq_amd_batcher_tune_records(const YumePdqAmdBatcher *b, YumePdqAmdTuneRecord *out,
                                      uint32_t cap, uint32_t *n);
int yume_pdq_amd_batcher_choice(const YumePdqAmdBatcher *b, uint32_t n_queries, uint32_t *path,
                                uint32_t *query_tiles);

/*
 * Serving loop (continuous batching).  No reference counterpart.  Requests of
 * 1..max_batch_images u8 planes are copied into pinned staging by submit()
 * and queued; a worker thread admits queued requests in arrival order into
 * a free batch context (as many contexts as hbm_budget_bytes holds, at most
 * max_contexts), uploading each request's planes as it is admitted, replays
 * the context's hash -> match graph, and hands each request its own hashes,
 * qualities, thresholds and hits.  With in_order_compute (the default) one
 * context admits at a time and its batch starts when the previous batch's
 * kernels have finished or when the next request would not fit, so uploads
 * overlap the batches in flight; otherwise every free context takes what is
 * queued at once.  A batch's outputs come back through page-locked memory
 * written by a kernel at its end.  A request's results do not depend on which
 * other requests shared its batch.  Hit queries are 8 * (image in the
 * request) + dihedral; indices are global.
 */
typedef struct YumePdqAmdServer YumePdqAmdServer;
typedef struct YumePdqAmdServerConfig {
  uint32_t max_batch_images;
  uint64_t hbm_budget_bytes;
  uint32_t max_contexts;
  uint32_t threshold;
  uint32_t dihedral_mode;
  uint64_t hit_capacity;      /* per batch (a larger batch result is re-matched) */
  int use_graph;
  int busy_poll;              /* worker spins (yielding) while batches are in flight */
  int in_order_compute;       /* batches' kernels run one batch after another (uploads overlap) */
  uint32_t autotune_iterations; /* > 0: autotune at start unless choices for the shard set's
                                   shape are cached (yume_pdq_amd_batcher_autotune, _tune_load) */
  uint64_t scan_chunk_rows;     /* > 0: shared circular scan instead of batches: the shards are
                                   matched in chunks of this many rows, round and round, against
                                   every active request's hashes; a request joins at the next chunk
                                   and leaves after one full turn (up to max_batch_images x
                                   max_contexts images active) */
} YumePdqAmdServerConfig;
typedef struct YumePdqAmdResult {
  int status;                 /* YUME_PDQ_AMD_OK or an error of the batch */
  uint32_t n_images;
  const uint8_t *hashes;      /* n_images x 8 x 32 */
  const float *quality;       /* n_images x 8 */
  const float *threshold;     /* n_images x 8 */
  const YumePdqAmdHit *hits;  /* sorted by (index, query) */
  uint64_t n_hits;
  double latency_ms;          /* submit -> results ready */
  uint64_t batch;             /* sequence number of the serving batch */
  uint32_t batch_images;      /* images in that batch */
} YumePdqAmdResult;
typedef struct YumePdqAmdServerStats {
  uint32_t contexts;
  uint64_t device_bytes;
  uint64_t batches;
  uint64_t images;
} YumePdqAmdServerStats;
int yume_pdq_amd_server_create(YumePdqAmdShards *shards, const YumePdqAmdServerConfig *cfg,
                               YumePdqAmdServer **out);
void yume_pdq_amd_server_destroy(YumePdqAmdServer *server);
int yume_pdq_amd_server_submit(YumePdqAmdServer *server, const uint8_t *planes, uint32_t n_images,
                               uint64_t *request_id);
/* Zero-copy submit: `planes` is page-locked host memory (hipHostMalloc)
 * that the caller leaves unchanged until the request's result is ready. */
int yume_pdq_amd_server_submit_pinned(YumePdqAmdServer *server, const uint8_t *planes,
                                      uint32_t n_images, uint64_t *request_id);
/* Result pointers stay valid until yume_pdq_amd_server_release(id). */
int yume_pdq_amd_server_wait(YumePdqAmdServer *server, uint64_t request_id, int timeout_ms,
                             YumePdqAmdResult *out);
int yume_pdq_amd_server_release(YumePdqAmdServer *server, uint64_t request_id);
int yume_pdq_amd_server_stats(const YumePdqAmdServer *server, YumePdqAmdServerStats *out);

/*
 * Stream ingestion (the reference's `pipe` subcommand, src/main.rs:1091-1304).
 *
 * Reads raw 512x512 frames from input_fd until the end of the stream, hashes
 * them on the GPU in batches of batch_frames, and writes one record per frame
 * (or eight with dihedrals: the original, then visit_dihedrals' order) to
 * output_fd in `output_format`, the reference's -f spelling:
 *   raw | q+raw | hex | HEX | q+hex | q+HEX | bin | q+bin | diagnostic,
 *   optionally followed by +nl | +lf | +crlf | +cr | +nul (a separator after
 *   every record).
 * input_kind: YUME_PDQ_AMD_INPUT_U8 (luma8), _RGB8 or _RGBA8 (BT.601 luma).
 * dihedral_mode: YUME_PDQ_AMD_DIHEDRAL_AVX2 reproduces the x86 reference
 * binary.  Returns OK at a clean end of input, YUME_PDQ_AMD_EEOF after
 * writing every whole frame when the input ends inside a frame,
 * YUME_PDQ_AMD_EIO on a read/write error, EINVAL for a bad format.
 * *frames: frames written (may be NULL).  The page-locked ring, the device
 * buffers and the streams of a finished run are kept for the next run of the
 * same batch_frames, input_kind and diagnostic-or-not (at most two such sets;
 * yume_pdq_amd_pipe_trim frees them).
 */
typedef struct YumePdqAmdPipeConfig {
  int input_fd;
  int output_fd;
  int input_kind;
  const char *output_format;
  int dihedrals;
  uint32_t dihedral_mode;
  uint32_t batch_frames; /* 0 = 64, at most 4096 */
  int stats;             /* progress lines on stderr about once a second */
} YumePdqAmdPipeConfig;
int yume_pdq_amd_pipe_run(const YumePdqAmdPipeConfig *cfg, uint64_t *frames);

/* Frees the buffer sets finished pipe runs keep (a set for batches of 256
 * luma8 frames holds 192 MiB of page-locked and 128 MiB of device memory);
 * runs in progress keep theirs.  Returns the number of sets freed.  No
 * counterpart in the reference, whose pipe allocates per process. */
int yume_pdq_amd_pipe_trim(void);

/* Host-only formatting of hashed frames in a pipe output format (no GPU
 * call).  hashes n x 8 x 32, quality and threshold n x 8 as
 * yume_pdq_amd_hash_batch returns them; dihedral_pdqf n x 8 x 256 for the
 * diagnostic format (else NULL).  Writes at most out_capacity bytes and
 * returns the bytes written, or a negative error (EINVAL: bad format or
 * missing PDQFs; EOVERFLOW: out_capacity too small).
 * yume_pdq_amd_format_bound gives a sufficient capacity (0: bad format). */
int64_t yume_pdq_amd_format_frames(const char *output_format, uint32_t n_frames, int dihedrals,
                                   const uint8_t *hashes, const float *quality, const float *threshold,
                                   const float *dihedral_pdqf, char *out, uint64_t out_capacity);
uint64_t yume_pdq_amd_format_bound(const char *output_format, uint32_t n_frames, int dihedrals);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* YUME_PDQ_AMD_H */
