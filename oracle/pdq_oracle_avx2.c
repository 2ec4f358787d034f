/*
 * Intrinsic-for-intrinsic restatement of the reference's AVX2 kernel and of
 * CpuMatcher, plus a threaded CPU baseline driver.
 *
 * TEST INFRASTRUCTURE ONLY.  Two uses:
 *   1. cross-check of the lane-explicit scalar restatement in pdq_oracle.c
 *      (a transposed lane or a wrong accumulator alternation there would
 *      differ from this file, which follows the intrinsic sequence directly);
 *   2. the CPU baseline timed by bench.py (reference CPU kernel + matcher on
 *      the host cores).
 * Built with -mavx2 -mfma (and AVX-512 VPOPCNTDQ in a separately-targeted
 * function, used only when the host CPU reports it).
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "pdq_oracle.h"

static float t_dct[16 * 127 + 16];
static float t_tent[56];
static int t_init = 0;

static void init(void) {
  if (t_init) return;
  oracle_tables(t_dct, NULL, t_tent);
  t_init = 1;
}

/* jarosz_compress_avx2, x86.rs:626-682 */
static void avx2_compress(const float *in, float *out) {
  const __m256i shiftl1 = _mm256_set_epi32(7, 7, 6, 5, 4, 3, 2, 1);
  float ob[8] __attribute__((aligned(32)));
  for (int outi = 0; outi < 127; outi++) {
    int in_i = 4 * (outi + 1) - 3;
    for (int outj = 0; outj < 127; outj++) {
      int in_j = 4 * (outj + 1) - 3;
      __m256 sum = _mm256_setzero_ps();
      for (int di = 0; di < 7; di++) {
        long off = (long)(in_i + di) * 512 + in_j;
        int last = (di == 6 && outi == 126 && outj == 126);
        if (last) off -= 1;
        __m256 b = _mm256_loadu_ps(in + off);
        if (last) b = _mm256_permutevar8x32_ps(b, shiftl1);
        __m256 w = _mm256_loadu_ps(t_tent + di * 8);
        sum = _mm256_fmadd_ps(b, w, sum);
      }
      sum = _mm256_hadd_ps(sum, _mm256_setzero_ps());
      sum = _mm256_hadd_ps(sum, _mm256_setzero_ps());
      _mm256_store_ps(ob, sum);
      out[outi * 128 + outj] = ob[0] + ob[4];
    }
    out[outi * 128 + 127] = 0.0f;
  }
  memset(out + 127 * 128, 0, 128 * sizeof(float));
}

static float hsum_avx2(__m256 a) {
  const __m256i s1 = _mm256_set_epi32(0, 7, 6, 5, 4, 3, 2, 1);
  const __m256i s2 = _mm256_set_epi32(1, 0, 7, 6, 5, 4, 3, 2);
  const __m256i s4 = _mm256_set_epi32(3, 2, 1, 0, 7, 6, 5, 4);
  float ob[8] __attribute__((aligned(32)));
  __m256 p = _mm256_add_ps(a, _mm256_permutevar8x32_ps(a, s1));
  __m256 e = _mm256_add_ps(p, _mm256_permutevar8x32_ps(p, s2));
  __m256 f = _mm256_add_ps(e, _mm256_permutevar8x32_ps(e, s4));
  _mm256_store_ps(ob, f);
  return ob[0];
}

static float hmax_avx2(__m256 a) {
  const __m256i s1 = _mm256_set_epi32(0, 7, 6, 5, 4, 3, 2, 1);
  const __m256i s2 = _mm256_set_epi32(1, 0, 7, 6, 5, 4, 3, 2);
  const __m256i s4 = _mm256_set_epi32(3, 2, 1, 0, 7, 6, 5, 4);
  float ob[8] __attribute__((aligned(32)));
  __m256 p = _mm256_max_ps(a, _mm256_permutevar8x32_ps(a, s1));
  __m256 e = _mm256_max_ps(p, _mm256_permutevar8x32_ps(p, s2));
  __m256 f = _mm256_max_ps(e, _mm256_permutevar8x32_ps(e, s4));
  _mm256_store_ps(ob, f);
  return ob[0];
}

static float hmin_avx2(__m256 a) {
  const __m256i s1 = _mm256_set_epi32(0, 7, 6, 5, 4, 3, 2, 1);
  const __m256i s2 = _mm256_set_epi32(1, 0, 7, 6, 5, 4, 3, 2);
  const __m256i s4 = _mm256_set_epi32(3, 2, 1, 0, 7, 6, 5, 4);
  float ob[8] __attribute__((aligned(32)));
  __m256 p = _mm256_min_ps(a, _mm256_permutevar8x32_ps(a, s1));
  __m256 e = _mm256_min_ps(p, _mm256_permutevar8x32_ps(p, s2));
  __m256 f = _mm256_min_ps(e, _mm256_permutevar8x32_ps(e, s4));
  _mm256_store_ps(ob, f);
  return ob[0];
}

/* Avx2F32Kernel::dct2d_impl, x86.rs:100-281, following its macro expansion. */
static void avx2_dct2d(const float *buf, float *tmp, float *out) {
  for (int k = 0; k < 16; k++) {
    for (int jb = 0; jb < 128; jb += 8) {
      __m256 sumk[2] = {_mm256_setzero_ps(), _mm256_setzero_ps()};
      int t = 1;
      for (int kb = 0; kb < 128; kb += 8) {
        int k2 = kb;
#define DO1()                                                                         \
  do {                                                                                \
    __m256 row = _mm256_loadu_ps(buf + k2 * 128 + jb);                                \
    __m256 d = _mm256_set1_ps(t_dct[k * 127 + k2]);                                   \
    sumk[t] = _mm256_fmadd_ps(row, d, sumk[t]);                                       \
  } while (0)
#define DO2()      \
  do {             \
    DO1();         \
    t = !t;        \
    k2 += 1;       \
    DO1();         \
  } while (0)
#define DO4()      \
  do {             \
    DO2();         \
    k2 += 1;       \
    DO2();         \
  } while (0)
        DO4();
        k2 += 1;
        DO4();
#undef DO1
#undef DO2
#undef DO4
      }
      __m256 z = _mm256_setzero_ps();
      __m256 m0 = _mm256_cmp_ps(sumk[0], z, _CMP_ORD_S);
      __m256 a0 = _mm256_blendv_ps(z, sumk[0], m0);
      __m256 m1 = _mm256_cmp_ps(sumk[1], z, _CMP_ORD_S);
      __m256 a1 = _mm256_blendv_ps(z, sumk[1], m1);
      _mm256_storeu_ps(tmp + jb, _mm256_add_ps(a0, a1));
    }
    for (int j = 0; j < 16; j++) {
      __m256 h[2] = {_mm256_setzero_ps(), _mm256_setzero_ps()};
      int t = 1;
      int j2 = 0;
      const float *dp = t_dct + j * 127;
      /* do_loop!(128) -> 16 chunks, toggling inside each do_loop!(16) */
      for (int pair = 0; pair < 8; pair++) {
        h[t] = _mm256_fmadd_ps(_mm256_loadu_ps(tmp + j2), _mm256_loadu_ps(dp + j2), h[t]);
        t = !t;
        j2 += 8;
        h[t] = _mm256_fmadd_ps(_mm256_loadu_ps(tmp + j2), _mm256_loadu_ps(dp + j2), h[t]);
        j2 += 8;
      }
      out[k * 16 + j] = hsum_avx2(_mm256_add_ps(h[0], h[1]));
    }
  }
}

/* quantize_impl, x86.rs:440-621 */
static void avx2_quantize(const float *in, float *threshold, uint8_t *out) {
  __m256 maxv = _mm256_set1_ps(-3.40282347e+38f), minv = _mm256_set1_ps(3.40282347e+38f);
  for (int i = 0; i < 256; i += 8) {
    __m256 r = _mm256_loadu_ps(in + i);
    minv = _mm256_min_ps(minv, r);
    maxv = _mm256_max_ps(maxv, r);
  }
  float mx = hmax_avx2(maxv), mn = hmin_avx2(minv);
  float guess = (mn + mx) * 0.5f;
  unsigned num_over = 0, num_under = 0;
  for (int iter = 0; iter < 32; iter++) {
    __m256 g = _mm256_set1_ps(guess);
    __m256 rl = _mm256_setzero_ps(), rg = _mm256_setzero_ps();
    __m256 vmin = _mm256_set1_ps(mn), vmax = _mm256_set1_ps(mx);
    unsigned gt_count = 0, lt_count = 0;
    for (int c = 0; c < 32; c++) {
      __m256 row = _mm256_loadu_ps(in + 8 * c);
      __m256 cgt = _mm256_cmp_ps(row, g, _CMP_GT_OQ);
      out[31 - c] = (uint8_t)_mm256_movemask_ps(cgt);
      __m256 clt = _mm256_cmp_ps(row, g, _CMP_LT_OQ);
      __m256 cmax = _mm256_cmp_ps(row, vmax, _CMP_LE_OQ);
      __m256 cmin = _mm256_cmp_ps(row, vmin, _CMP_GE_OQ);
      __m256 mg = _mm256_and_ps(cgt, cmax), ml = _mm256_and_ps(clt, cmin);
      gt_count += (unsigned)__builtin_popcount(_mm256_movemask_ps(mg));
      lt_count += (unsigned)__builtin_popcount(_mm256_movemask_ps(ml));
      rg = _mm256_add_ps(rg, _mm256_and_ps(mg, row));
      rl = _mm256_add_ps(rl, _mm256_and_ps(ml, row));
    }
    if (gt_count + num_over > 128) {
      num_under += lt_count;
      mn = guess;
      guess = hsum_avx2(rg) / (float)gt_count;
    } else if (lt_count + num_under > 128) {
      num_over += gt_count;
      mx = guess;
      guess = hsum_avx2(rl) / (float)lt_count;
    } else {
      break;
    }
  }
  *threshold = guess;
}

/* hash_get_threshold (lib.rs:427-455) with the intrinsic kernel */
float oracle_avx2_hash(const float *in, float *threshold, uint8_t *out, float *buf1, float *tmp,
                       float *pdqf) {
  init();
  avx2_compress(in, buf1);
  avx2_dct2d(buf1, tmp, pdqf);
  float g = oracle_sum_of_gradients(pdqf);
  float q = oracle_adjust_quality(g);
  avx2_quantize(pdqf, threshold, out);
  return q;
}

/* ---------------------------------------------------------------------------
 * CpuMatcher (matching/mod.rs:117-436): scalar u64 popcount, and the
 * AVX-512 VPOPCNTDQ form with pre-complemented, pivoted needles.
 * ------------------------------------------------------------------------- */
static uint64_t scalar_scan_count(const uint64_t *rows, uint64_t n, const uint64_t *nd,
                                  uint32_t nq, uint32_t thr) {
  uint64_t hits = 0;
  for (uint64_t r = 0; r < n; r++) {
    const uint64_t *q = rows + 4 * r;
    for (uint32_t j = 0; j < nq; j++) {
      const uint64_t *b = nd + 4 * j;
      uint64_t d = (uint64_t)__builtin_popcountll(q[0] ^ b[0]) + __builtin_popcountll(q[1] ^ b[1]) +
                   __builtin_popcountll(q[2] ^ b[2]) + __builtin_popcountll(q[3] ^ b[3]);
      hits += d <= thr;
    }
  }
  return hits;
}

__attribute__((target("avx512f,avx512vpopcntdq"))) static uint64_t avx512_scan_count(
    const uint64_t *rows, uint64_t n, const uint64_t *nd, uint32_t nq, uint32_t thr) {
  /* needles in blocks of 8, complemented and pivoted (mod.rs:180-203) */
  uint32_t nb = (nq + 7) / 8;
  __m512i comp[8][4];
  const __m512i ones = _mm512_set1_epi64(-1);
  for (uint32_t b = 0; b < nb && b < 8; b++)
    for (int reg = 0; reg < 4; reg++) {
      uint64_t tmp[8];
      for (int i = 0; i < 8; i++) {
        uint32_t j = b * 8 + i;
        tmp[i] = j < nq ? nd[4 * j + reg] : ~0ull; /* padding never matches */
      }
      comp[b][reg] = _mm512_xor_si512(ones, _mm512_loadu_si512(tmp));
    }
  const __m512i th = _mm512_set1_epi64(256 - (long long)thr);
  uint64_t hits = 0;
  for (uint64_t r = 0; r < n; r++) {
    const uint64_t *q = rows + 4 * r;
    __m512i q0 = _mm512_set1_epi64((long long)q[0]), q1 = _mm512_set1_epi64((long long)q[1]);
    __m512i q2 = _mm512_set1_epi64((long long)q[2]), q3 = _mm512_set1_epi64((long long)q[3]);
    for (uint32_t b = 0; b < nb; b++) {
      __m512i c0 = _mm512_popcnt_epi64(_mm512_xor_si512(comp[b][0], q0));
      __m512i c1 = _mm512_popcnt_epi64(_mm512_xor_si512(comp[b][1], q1));
      __m512i c2 = _mm512_popcnt_epi64(_mm512_xor_si512(comp[b][2], q2));
      __m512i c3 = _mm512_popcnt_epi64(_mm512_xor_si512(comp[b][3], q3));
      __m512i s = _mm512_add_epi64(_mm512_add_epi64(c0, c1), _mm512_add_epi64(c2, c3));
      __mmask8 m = _mm512_cmpge_epi64_mask(s, th);
      hits += (uint64_t)__builtin_popcount((unsigned)m);
    }
  }
  return hits;
}

int oracle_cpu_has_vpopcntdq(void) {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512vpopcntdq") && __builtin_cpu_supports("avx512f");
}

/* Count hits of n_queries needles against n_rows DB rows (CpuMatcher::scan
 * semantics, counted).  use_avx512 selects the VPOPCNTDQ path when the host
 * supports it and n_queries <= 64 (the reference's limit, mod.rs:109-117). */
uint64_t oracle_cpu_match_count(const uint8_t *db, uint64_t n_rows, const uint8_t *queries,
                                uint32_t n_queries, uint32_t thr, int use_avx512) {
  const uint64_t *rows = (const uint64_t *)db;
  const uint64_t *nd = (const uint64_t *)queries;
  if (use_avx512 && oracle_cpu_has_vpopcntdq() && n_queries <= 64)
    return avx512_scan_count(rows, n_rows, nd, n_queries, thr);
  return scalar_scan_count(rows, n_rows, nd, n_queries, thr);
}

/* ---------------------------------------------------------------------------
 * Threaded CPU baseline: hash n_images (f32 planes, cycling through a pool of
 * n_pool distinct planes) then match the 8
 * dihedral hashes of each image against the DB, with n_threads std-thread
 * style workers (pthreads).  Returns elapsed seconds for each phase.
 * ------------------------------------------------------------------------- */
typedef struct {
  const float *images; /* pool of n_pool planes; image i is plane i % n_pool */
  uint32_t n_pool;
  uint32_t first, last;
  uint8_t *hashes; /* 8 x 32 per image */
} HashJob;

static void *hash_worker(void *arg) {
  HashJob *j = (HashJob *)arg;
  float *buf1 = (float *)aligned_alloc(64, 128 * 128 * sizeof(float));
  float tmp[128], pdqf[256], thr;
  memset(buf1, 0, 128 * 128 * sizeof(float));
  for (uint32_t i = j->first; i < j->last; i++) {
    uint8_t *h = j->hashes + (size_t)i * 256;
    oracle_avx2_hash(j->images + (size_t)(i % j->n_pool) * 512 * 512, &thr, h, buf1, tmp, pdqf);
    uint32_t codes[7];
    float th[7], q[7], pd[7 * 256];
    oracle_visit_dihedrals(pdqf, thr, 1, codes, th, q, h + 32, pd);
  }
  free(buf1);
  return NULL;
}

typedef struct {
  const uint8_t *db;
  uint64_t first, last;
  const uint8_t *queries;
  uint32_t nq, thr;
  int use_avx512;
  uint64_t hits;
} MatchJob;

static void *match_worker(void *arg) {
  MatchJob *j = (MatchJob *)arg;
  j->hits = 0;
  for (uint32_t q0 = 0; q0 < j->nq; q0 += 64) {
    uint32_t nq = j->nq - q0 < 64 ? j->nq - q0 : 64;
    j->hits += oracle_cpu_match_count(j->db + j->first * 32, j->last - j->first,
                                      j->queries + (size_t)q0 * 32, nq, j->thr, j->use_avx512);
  }
  return NULL;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

uint64_t oracle_cpu_baseline(const float *images, uint32_t n_pool, uint32_t n_images, const uint8_t *db,
                             uint64_t n_rows, uint32_t thr, uint32_t n_threads, int use_avx512,
                             uint8_t *hashes_out, double *t_hash, double *t_match) {
  init();
  if (n_threads < 1) n_threads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
  HashJob *hj = (HashJob *)malloc(sizeof(HashJob) * n_threads);
  double t0 = now_s();
  for (uint32_t t = 0; t < n_threads; t++) {
    hj[t].images = images;
    hj[t].n_pool = n_pool ? n_pool : 1;
    hj[t].first = (uint32_t)((uint64_t)n_images * t / n_threads);
    hj[t].last = (uint32_t)((uint64_t)n_images * (t + 1) / n_threads);
    hj[t].hashes = hashes_out;
    pthread_create(&th[t], NULL, hash_worker, &hj[t]);
  }
  for (uint32_t t = 0; t < n_threads; t++) pthread_join(th[t], NULL);
  double t1 = now_s();
  MatchJob *mj = (MatchJob *)malloc(sizeof(MatchJob) * n_threads);
  for (uint32_t t = 0; t < n_threads; t++) {
    mj[t].db = db;
    mj[t].first = n_rows * t / n_threads;
    mj[t].last = n_rows * (t + 1) / n_threads;
    mj[t].queries = hashes_out;
    mj[t].nq = n_images * 8;
    mj[t].thr = thr;
    mj[t].use_avx512 = use_avx512;
    pthread_create(&th[t], NULL, match_worker, &mj[t]);
  }
  uint64_t hits = 0;
  for (uint32_t t = 0; t < n_threads; t++) {
    pthread_join(th[t], NULL);
    hits += mj[t].hits;
  }
  double t2 = now_s();
  *t_hash = t1 - t0;
  *t_match = t2 - t1;
  free(th);
  free(hj);
  free(mj);
  return hits;
}

/* ---------------------------------------------------------------------------
 * Brute-force exact match on n_threads threads: oracle_match_all's triples
 * (every (query, row) with popcount(query ^ row) <= threshold, inclusive;
 * row-major order, queries ascending within a row), computed over contiguous
 * row ranges, one per thread, and concatenated in range order.  The checker
 * for database sizes where the single-threaded loop takes minutes.  The
 * first two words' distance is a lower bound of the full one, so rows past
 * the threshold there are skipped -- the same predicate, evaluated exactly.
 * Returns the total count; writes at most cap triples.
 * ------------------------------------------------------------------------- */
typedef struct {
  const uint8_t *db;
  uint64_t first, last;
  const uint64_t *q;
  uint32_t nq, thr;
  OracleHit *buf;
  uint64_t n, cap;
  int failed;
} AllJob;

static void *all_worker(void *arg) {
  AllJob *j = (AllJob *)arg;
  j->n = 0;
  for (uint64_t r = j->first; r < j->last; r++) {
    uint64_t x[4];
    memcpy(x, j->db + r * 32, 32);
    for (uint32_t k = 0; k < j->nq; k++) {
      const uint64_t *y = j->q + 4 * (size_t)k;
      uint32_t d = (uint32_t)(__builtin_popcountll(x[0] ^ y[0]) + __builtin_popcountll(x[1] ^ y[1]));
      if (d > j->thr) continue;
      d += (uint32_t)(__builtin_popcountll(x[2] ^ y[2]) + __builtin_popcountll(x[3] ^ y[3]));
      if (d > j->thr) continue;
      if (j->n == j->cap) {
        uint64_t cap = j->cap ? 2 * j->cap : 4096;
        OracleHit *nb = (OracleHit *)realloc(j->buf, cap * sizeof(OracleHit));
        if (!nb) {
          j->failed = 1;
          return NULL;
        }
        j->buf = nb;
        j->cap = cap;
      }
      j->buf[j->n].query = k;
      j->buf[j->n].distance = d;
      j->buf[j->n].index = r;
      j->n++;
    }
  }
  return NULL;
}

uint64_t oracle_match_all_mt(const uint8_t *db, uint64_t n_rows, const uint8_t *queries, uint32_t n_queries,
                             uint32_t threshold, OracleHit *hits, uint64_t cap, uint32_t n_threads) {
  if (n_threads < 1) n_threads = 1;
  uint64_t *q = (uint64_t *)malloc((size_t)n_queries * 32 + 32);
  memcpy(q, queries, (size_t)n_queries * 32);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
  AllJob *jobs = (AllJob *)calloc(n_threads, sizeof(AllJob));
  for (uint32_t t = 0; t < n_threads; t++) {
    jobs[t].db = db;
    jobs[t].first = n_rows * t / n_threads;
    jobs[t].last = n_rows * (t + 1) / n_threads;
    jobs[t].q = q;
    jobs[t].nq = n_queries;
    jobs[t].thr = threshold;
    pthread_create(&th[t], NULL, all_worker, &jobs[t]);
  }
  uint64_t total = 0;
  int failed = 0;
  for (uint32_t t = 0; t < n_threads; t++) {
    pthread_join(th[t], NULL);
    failed |= jobs[t].failed;
  }
  for (uint32_t t = 0; t < n_threads; t++) {
    for (uint64_t i = 0; i < jobs[t].n; i++, total++)
      if (total < cap) hits[total] = jobs[t].buf[i];
    free(jobs[t].buf);
  }
  free(jobs);
  free(th);
  free(q);
  return failed ? UINT64_MAX : total;
}
