/* CPU oracle for yume-pdq's hash and match paths -- TEST INFRASTRUCTURE ONLY.
 * Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
 * the checker; never linked into the product. */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Dihedrals::* packed codes, lib.rs:103-133 (big-endian dx0,dx1,dy0,dy1). */
#define ORACLE_DIHEDRAL_NORMAL 0x01000001u
#define ORACLE_DIHEDRAL_FLIPPED 0x010000FFu
#define ORACLE_DIHEDRAL_FLOPPED 0xFF000001u
#define ORACLE_DIHEDRAL_ROTATED_180 0xFF0000FFu
#define ORACLE_DIHEDRAL_ROTATED_90 0x0001FF00u
#define ORACLE_DIHEDRAL_ROTATED_270 0x00010100u
#define ORACLE_DIHEDRAL_ROTATED_90_FLOPPED 0x00FFFF00u
#define ORACLE_DIHEDRAL_FLOPPED_ROTATED_270 0x00FF0100u

typedef struct OracleHit {
  uint32_t query;
  uint32_t distance;
  uint64_t index;
} OracleHit;

void oracle_tables(float *dct_f32, double *dct_f64, float *tent_x8);
void oracle_compress_avx2(const float *in, float *buf);
void oracle_dct2d_avx2(const float *buf, float *tmp, float *out);
float oracle_sum_of_gradients(const float *p);
float oracle_adjust_quality(float g);
void oracle_quantize_avx2(const float *p, float *threshold, uint8_t *out);
void oracle_threshold_2d(const float *p, uint8_t *out, float threshold);
float oracle_torben_median(const float *m);
void oracle_quantize_torben(const float *p, float *threshold, uint8_t *out);
void oracle_compress_scalar(const float *in, float *buf);
void oracle_dct2d_scalar(const float *buf, float *out);
float oracle_hash_ref32(const float *in, float *threshold, uint8_t *out, float *pdqf);
void oracle_pdqf_f64(const float *in, double *pdqf);
double oracle_hash_ref64(const float *in, double *threshold, uint8_t *out, double *pdqf);
void oracle_readme_fm_image(float *out);
float oracle_avx2_hash(const float *in, float *threshold, uint8_t *out, float *buf1, float *tmp,
                       float *pdqf);
int oracle_cpu_has_vpopcntdq(void);
uint64_t oracle_cpu_match_count(const uint8_t *db, uint64_t n_rows, const uint8_t *queries,
                                uint32_t n_queries, uint32_t thr, int use_avx512);
uint64_t oracle_cpu_baseline(const float *images, uint32_t n_pool, uint32_t n_images, const uint8_t *db,
                             uint64_t n_rows, uint32_t thr, uint32_t n_threads, int use_avx512,
                             uint8_t *hashes_out, double *t_hash, double *t_match);
float oracle_hash_get_threshold(const float *in, float *threshold, uint8_t *out, float *buf1,
                                float *tmp, float *pdqf);
void oracle_visit_dihedrals(const float *pdqf_in, float threshold, int mode, uint32_t *codes,
                            float *thresholds, float *qualities, uint8_t *hashes, float *pdqfs);
void oracle_hash_dihedrals(const float *in, int mode, uint8_t *hashes, float *qualities,
                           float *thresholds, float *pdqf);
uint32_t oracle_hamming(const uint8_t *a, const uint8_t *b);
uint64_t oracle_match_all(const uint8_t *db, uint64_t n_rows, const uint8_t *queries,
                          uint32_t n_queries, uint32_t threshold, OracleHit *hits, uint64_t cap);
/* oracle_match_all on n_threads threads (UINT64_MAX if a thread ran out of
 * memory). */
uint64_t oracle_match_all_mt(const uint8_t *db, uint64_t n_rows, const uint8_t *queries, uint32_t n_queries,
                             uint32_t threshold, OracleHit *hits, uint64_t cap, uint32_t n_threads);
int oracle_match_first(const uint8_t *db, uint64_t n_rows, const uint8_t *queries,
                       uint32_t n_queries, uint32_t threshold, uint64_t *row_out,
                       uint32_t *query_out);

#ifdef __cplusplus
}
#endif
