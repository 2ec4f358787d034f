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
  YUME_PDQ_AMD_EOVERFLOW = -5  /* more hits than the buffer holds (count still exact) */
};

enum { YUME_PDQ_AMD_INPUT_F32 = 0, YUME_PDQ_AMD_INPUT_U8 = 1 };

/* Dihedral sign algebra: EXACT negates every second column (the AVX-512
 * and scalar kernels, src/kernel/x86.rs:1026-1074, src/kernel/mod.rs:368-382);
 * AVX2 reproduces the reference's AVX2 kernel, whose column negation reaches
 * columns 0..7 only (src/kernel/x86.rs:330-376). */
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
 * d_planes: n_images planes of 512x512, f32 (values 0..255) or u8, plane i at
 *           byte offset i * plane_stride_bytes (>= 512*512*elem, 16-byte
 *           aligned).
 * d_hashes: n_images x 8 x 32 bytes, orientations in YUME_PDQ_AMD_D_* order.
 * d_quality, d_threshold: n_images x 8 floats (per-orientation values; the
 *           threshold of a transposed orientation is the one it was
 *           re-thresholded with, as visit_dihedrals passes it).
 * d_pdqf:   n_images x 256 floats of NORMAL PDQF, or NULL.
 */
int yume_pdq_amd_hash_batch(const void *d_planes, int input_kind, uint64_t plane_stride_bytes,
                            uint32_t n_images, uint32_t dihedral_mode, uint8_t *d_hashes,
                            float *d_quality, float *d_threshold, float *d_pdqf, void *stream);

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

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* YUME_PDQ_AMD_H */
