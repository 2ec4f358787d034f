/*
 * CPU oracle for yume-pdq's hashing and matching hot paths.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference's algorithms in
 * plain C so that the HIP product can be checked against it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the product
 * never does.
 *
 * Floating-point discipline: compiled with -ffp-contract=off (Rust never
 * contracts), FMAs are written as fmaf() exactly where the reference uses
 * _mm256_fmadd_ps, and every AVX2 8-lane accumulator is kept as an explicit
 * 8-element array so the rounding sequence is the reference's, lane by lane.
 *
 * Citations are path:line into the reference crate (eternal-flame-AD/yume-pdq
 * v1.1.0).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pdq_tables.h"
#include "pdq_oracle.h"

static float g_dct[16 * 127 + 16];
static double g_dct64[16 * 127 + 16];
static float g_tent[7 * 8];
static int g_init = 0;

static float bits_f32(uint32_t b) {
  float f;
  memcpy(&f, &b, 4);
  return f;
}

static double bits_f64(uint64_t b) {
  double f;
  memcpy(&f, &b, 8);
  return f;
}

static void init_tables(void) {
  if (g_init) return;
  for (int i = 0; i < 16 * 127 + 16; i++) {
    g_dct[i] = bits_f32(ORACLE_DCT_F32_BITS[i]);
    g_dct64[i] = bits_f64(ORACLE_DCT_F64_BITS[i]);
  }
  for (int i = 0; i < 56; i++) g_tent[i] = bits_f32(ORACLE_TENT_X8_BITS[i]);
  g_init = 1;
}

void oracle_tables(float *dct_f32, double *dct_f64, float *tent_x8) {
  init_tables();
  if (dct_f32) memcpy(dct_f32, g_dct, sizeof g_dct);
  if (dct_f64) memcpy(dct_f64, g_dct64, sizeof g_dct64);
  if (tent_x8) memcpy(tent_x8, g_tent, sizeof g_tent);
}

/* The AVX2 horizontal reductions (x86.rs:131-161): lane i is combined with
 * lane i+1, then i+2, then i+4 (permutevar8x32 by shift1/2/4), and lane 0 of
 * the result is read.  For addition that is ((a0+a1)+(a2+a3))+((a4+a5)+(a6+a7)). */
static float hsum8(const float a[8]) {
  return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

/* x86 MAXPS/MINPS return the second operand unless the first compares
 * strictly greater/less (NaN and signed-zero behaviour included). */
static float x86_max(float a, float b) { return a > b ? a : b; }
static float x86_min(float a, float b) { return a < b ? a : b; }

static float hmax8(const float a[8]) {
  float p[8], e[8];
  for (int i = 0; i < 8; i++) p[i] = x86_max(a[i], a[(i + 1) & 7]);
  for (int i = 0; i < 8; i++) e[i] = x86_max(p[i], p[(i + 2) & 7]);
  return x86_max(e[0], e[4]);
}

static float hmin8(const float a[8]) {
  float p[8], e[8];
  for (int i = 0; i < 8; i++) p[i] = x86_min(a[i], a[(i + 1) & 7]);
  for (int i = 0; i < 8; i++) e[i] = x86_min(p[i], p[(i + 2) & 7]);
  return x86_min(e[0], e[4]);
}

/* ---------------------------------------------------------------------------
 * Tent compression 512x512 -> 127x127, AVX2 form: jarosz_compress_avx2,
 * x86.rs:626-682.  Output row/col 127 of the 128x128 buffer are zeroed (the
 * reference requires the caller to keep them zero, lib.rs:296-297).
 * ------------------------------------------------------------------------- */
void oracle_compress_avx2(const float *in, float *buf) {
  init_tables();
  for (int outi = 0; outi < 127; outi++) {
    int in_i = 4 * (outi + 1) - 3; /* CONVOLUTION_OFFSET - TENT_FILTER_COLUMN_OFFSET */
    for (int outj = 0; outj < 127; outj++) {
      int in_j = 4 * (outj + 1) - 3;
      float lane[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int di = 0; di < 7; di++) {
        long off = (long)(in_i + di) * 512 + in_j;
        float x[8];
        if (di == 6 && outi == 126 && outj == 126) {
          /* x86.rs:644-669: rewind by one, then permute lanes by
           * shiftl1 = (1,2,3,4,5,6,7,7) so lane k holds in[off+k], lane 7
           * repeats in[off+6]. */
          const float *p = in + off - 1;
          for (int k = 0; k < 7; k++) x[k] = p[k + 1];
          x[7] = p[7];
        } else {
          for (int k = 0; k < 8; k++) x[k] = in[off + k];
        }
        for (int k = 0; k < 8; k++) lane[k] = fmaf(x[k], g_tent[di * 8 + k], lane[k]);
      }
      /* two hadd_ps against zero, then lane0 + lane4 (x86.rs:673-676) */
      buf[outi * 128 + outj] = ((lane[0] + lane[1]) + (lane[2] + lane[3])) +
                               ((lane[4] + lane[5]) + (lane[6] + lane[7]));
    }
    buf[outi * 128 + 127] = 0.0f;
  }
  for (int j = 0; j < 128; j++) buf[127 * 128 + j] = 0.0f;
}

/* ---------------------------------------------------------------------------
 * Two-sided DCT, AVX2 form: Avx2F32Kernel::dct2d_impl, x86.rs:100-281.
 * Stage 1 keeps two accumulators per 8-column block; the macro expansion at
 * x86.rs:163-202 alternates them as (1,0,0,1) over k2 mod 4, starting from
 * accumulator 1.  k2 runs to 127 inclusive and reads one element past each
 * DCT row (the next row's first element, or the zero pad), multiplied by the
 * zero row 127 of the buffer.  NaN lanes are zeroed before the pair is added
 * (x86.rs:207-224).  Stage 2 (x86.rs:227-272) uses the same (1,0,0,1)
 * alternation over 8-wide chunks, then the horizontal sum tree.
 * ------------------------------------------------------------------------- */
static const int kAltPattern[4] = {1, 0, 0, 1};

void oracle_dct2d_avx2(const float *buf, float *tmp, float *out) {
  init_tables();
  for (int k = 0; k < 16; k++) {
    for (int jb = 0; jb < 128; jb += 8) {
      float s[2][8];
      memset(s, 0, sizeof s);
      for (int k2 = 0; k2 < 128; k2++) {
        float d = g_dct[k * 127 + k2];
        int t = kAltPattern[k2 & 3];
        for (int l = 0; l < 8; l++) s[t][l] = fmaf(buf[k2 * 128 + jb + l], d, s[t][l]);
      }
      for (int l = 0; l < 8; l++) {
        float a = isnan(s[0][l]) ? 0.0f : s[0][l];
        float b = isnan(s[1][l]) ? 0.0f : s[1][l];
        tmp[jb + l] = a + b;
      }
    }
    for (int j = 0; j < 16; j++) {
      float h[2][8];
      memset(h, 0, sizeof h);
      for (int c = 0; c < 16; c++) {
        int t = kAltPattern[c & 3];
        for (int l = 0; l < 8; l++)
          h[t][l] = fmaf(tmp[8 * c + l], g_dct[j * 127 + 8 * c + l], h[t][l]);
      }
      float v[8];
      for (int l = 0; l < 8; l++) v[l] = h[0][l] + h[1][l];
      out[k * 16 + j] = hsum8(v);
    }
  }
}

/* Sum of gradients, x86.rs:840-865 (identical to mod.rs:617-642): one
 * sequential f32 accumulation, vertical differences first. */
float oracle_sum_of_gradients(const float *p) {
  float g = 0.0f;
  for (int i = 0; i < 15; i++)
    for (int j = 0; j < 16; j++) g += fabsf(p[i * 16 + j] - p[(i + 1) * 16 + j]);
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < 15; j++) g += fabsf(p[i * 16 + j] - p[i * 16 + j + 1]);
  return g;
}

/* adjust_quality, x86.rs:769-773 / mod.rs:611-615 with QUALITY_ADJUST_DIVISOR
 * = 180 (mod.rs:257).  Rust's f32::min returns the non-NaN operand. */
float oracle_adjust_quality(float g) {
  float s = g / 180.0f;
  if (isnan(s)) return 1.0f;
  return s < 1.0f ? s : 1.0f;
}

/* ---------------------------------------------------------------------------
 * Median quantisation, AVX2 form: quantize_impl, x86.rs:440-621.
 * Iterative search whose next guess is the mean of the residual side.  Bits
 * are x > guess (ordered compare); element i goes to byte 31 - i/8, bit i%8.
 * ------------------------------------------------------------------------- */
void oracle_quantize_avx2(const float *p, float *threshold, uint8_t *out) {
  float mxv[8], mnv[8];
  for (int l = 0; l < 8; l++) {
    mxv[l] = -3.40282347e+38f; /* f32::MIN */
    mnv[l] = 3.40282347e+38f;  /* f32::MAX */
  }
  for (int c = 0; c < 32; c++)
    for (int l = 0; l < 8; l++) {
      mnv[l] = x86_min(mnv[l], p[8 * c + l]);
      mxv[l] = x86_max(mxv[l], p[8 * c + l]);
    }
  float mx = hmax8(mxv);
  float mn = hmin8(mnv);
  float guess = (mn + mx) * 0.5f;
  uint32_t num_over = 0, num_under = 0;
  for (int iter = 0; iter < 32; iter++) {
    float rg[8] = {0, 0, 0, 0, 0, 0, 0, 0}, rl[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t gt_count = 0, lt_count = 0;
    for (int c = 0; c < 32; c++) {
      uint8_t byte = 0;
      for (int l = 0; l < 8; l++) {
        float x = p[8 * c + l];
        int gt = x > guess, lt = x < guess;
        int mrg = gt && (x <= mx), mrl = lt && (x >= mn);
        byte |= (uint8_t)(gt << l);
        gt_count += (uint32_t)mrg;
        lt_count += (uint32_t)mrl;
        rg[l] = rg[l] + (mrg ? x : 0.0f);
        rl[l] = rl[l] + (mrl ? x : 0.0f);
      }
      out[31 - c] = byte;
    }
    if (gt_count + num_over > 128) {
      num_under += lt_count;
      mn = guess;
      guess = hsum8(rg) / (float)gt_count;
    } else if (lt_count + num_under > 128) {
      num_over += gt_count;
      mx = guess;
      guess = hsum8(rl) / (float)lt_count;
    } else {
      break;
    }
  }
  *threshold = guess;
}

/* threshold_2d_f32, threshold.rs:37-119: bit i is p[i] > threshold, packed
 * into byte 31 - i/8 at bit i%8. */
void oracle_threshold_2d(const float *p, uint8_t *out, float threshold) {
  for (int b = 0; b < 32; b++) {
    uint8_t byte = 0;
    for (int l = 0; l < 8; l++) byte |= (uint8_t)((p[8 * b + l] > threshold) << l);
    out[31 - b] = byte;
  }
}

/* Scalar DefaultKernel pieces (mod.rs:543-666), used for cross-checks. */
float oracle_torben_median(const float *m) {
  /* mod.rs:212-253 (pdqhash's Torben median, half = 128) */
  float mn = m[0], mx = m[0];
  for (int i = 1; i < 256; i++) {
    mn = fminf(mn, m[i]);
    mx = fmaxf(mx, m[i]);
  }
  const int half = (256 + 1) / 2;
  for (;;) {
    float guess = (mn + mx) / 2.0f;
    int less = 0, greater = 0, equal = 0;
    float maxlt = mn, mingt = mx;
    for (int i = 0; i < 256; i++) {
      float v = m[i];
      if (v < guess) {
        less++;
        if (v > maxlt) maxlt = v;
      } else if (v > guess) {
        greater++;
        if (v < mingt) mingt = v;
      } else {
        equal++;
      }
    }
    if (less <= half && greater <= half) {
      if (less >= half) return maxlt;
      if (less + equal >= half) return guess;
      return mingt;
    } else if (less > greater) {
      mx = maxlt;
    } else {
      mn = mingt;
    }
  }
}

void oracle_quantize_torben(const float *p, float *threshold, uint8_t *out) {
  /* mod.rs:595-608 */
  float med = oracle_torben_median(p);
  *threshold = med;
  memset(out, 0, 32);
  for (int i = 0; i < 256; i++)
    if (p[i] > med) out[31 - i / 8] += (uint8_t)(1u << (i % 8));
}

void oracle_compress_scalar(const float *in, float *buf) {
  /* mod.rs:644-666: sequential 49-term sum in (di, dj) order */
  init_tables();
  for (int outi = 0; outi < 127; outi++) {
    int in_i = 4 * (outi + 1) - 3;
    for (int outj = 0; outj < 127; outj++) {
      int in_j = 4 * (outj + 1) - 3;
      float s = 0.0f;
      for (int di = 0; di < 7; di++)
        for (int dj = 0; dj < 7; dj++) s += g_tent[di * 8 + dj] * in[(in_i + di) * 512 + in_j + dj];
      buf[outi * 128 + outj] = s;
    }
    buf[outi * 128 + 127] = 0.0f;
  }
  for (int j = 0; j < 128; j++) buf[127 * 128 + j] = 0.0f;
}

void oracle_dct2d_scalar(const float *buf, float *out) {
  /* mod.rs:543-574: four interleaved accumulators over 127 terms */
  init_tables();
  float tmp[127];
  for (int k = 0; k < 16; k++) {
    for (int j = 0; j < 127; j++) {
      float s[4] = {0, 0, 0, 0};
      for (int k2 = 0; k2 < 127; k2++) s[k2 & 3] += g_dct[k * 127 + k2] * buf[k2 * 128 + j];
      tmp[j] = s[0] + s[1] + s[2] + s[3];
    }
    for (int j = 0; j < 16; j++) {
      float s[4] = {0, 0, 0, 0};
      for (int m = 0; m < 127; m++) s[m & 3] += tmp[m] * g_dct[j * 127 + m];
      out[k * 16 + j] = s[0] + s[1] + s[2] + s[3];
    }
  }
}

/* ---------------------------------------------------------------------------
 * Reference (precise) kernel: Jarosz box filter x2 + centre sampling,
 * mod.rs:790-820 with ref.rs:7-194, DCT mod.rs:711-735 (f32) / 844-869 (f64).
 * ------------------------------------------------------------------------- */
#define DEFINE_BOX(T, NAME)                                                              \
  static void NAME(const T *invec, long start, T *outvec, long len, long stride,        \
                   long full) {                                                           \
    long half = (full + 2) / 2, p1 = half - 1, p2 = full - half + 1;                      \
    long oi_off = p1 * stride, li_off = p2 * stride;                                      \
    T sum = 0, cws = 0;                                                                   \
    long p1_end = oi_off + start;                                                         \
    for (long ri = start; ri < p1_end; ri += stride) {                                    \
      sum += invec[ri];                                                                   \
      cws += 1;                                                                           \
    }                                                                                     \
    long p2_end = full * stride + start;                                                  \
    for (long ri = p1_end; ri < p2_end; ri += stride) {                                   \
      long oi = ri - oi_off;                                                              \
      sum += invec[ri];                                                                   \
      cws += 1;                                                                           \
      outvec[oi] = sum / cws;                                                             \
    }                                                                                     \
    long p3_end = len * stride + start;                                                   \
    for (long ri = p2_end; ri < p3_end; ri += stride) {                                   \
      long oi = ri - oi_off, li = oi - li_off;                                            \
      sum += invec[ri];                                                                   \
      sum -= invec[li];                                                                   \
      outvec[oi] = sum / cws;                                                             \
    }                                                                                     \
    long p4_start = (len - half + 1) * stride + start;                                    \
    for (long oi = p4_start; oi < p3_end; oi += stride) {                                 \
      long li = oi - li_off;                                                              \
      sum -= invec[li];                                                                   \
      cws += (T)-1;                                                                       \
      outvec[oi] = sum / cws;                                                             \
    }                                                                                     \
  }

DEFINE_BOX(float, box1d_f32)
DEFINE_BOX(double, box1d_f64)

static void jarosz_f32(float *b) {
  float *t = (float *)calloc(512 * 512, sizeof(float));
  for (int rep = 0; rep < 2; rep++) {
    for (int i = 0; i < 512; i++) box1d_f32(b, (long)i * 512, t, 512, 1, 3);
    for (int j = 0; j < 512; j++) box1d_f32(t, j, b, 512, 512, 3);
  }
  free(t);
}

static void jarosz_f64(double *b) {
  double *t = (double *)calloc(512 * 512, sizeof(double));
  for (int rep = 0; rep < 2; rep++) {
    for (int i = 0; i < 512; i++) box1d_f64(b, (long)i * 512, t, 512, 1, 3);
    for (int j = 0; j < 512; j++) box1d_f64(t, j, b, 512, 512, 3);
  }
  free(t);
}

float oracle_hash_ref32(const float *in, float *threshold, uint8_t *out, float *pdqf) {
  init_tables();
  float *b = (float *)malloc(512 * 512 * sizeof(float));
  memcpy(b, in, 512 * 512 * sizeof(float));
  jarosz_f32(b);
  static float c[127 * 127];
  for (int oi = 0; oi < 127; oi++)
    for (int oj = 0; oj < 127; oj++)
      c[oi * 127 + oj] = b[((oi * 2 + 1) * 512) / 254 * 512 + ((oj * 2 + 1) * 512) / 254];
  free(b);
  float tmp[127];
  for (int i = 0; i < 16; i++) {
    for (int j = 0; j < 127; j++) {
      float s = 0.0f;
      for (int k = 0; k < 127; k++) s += g_dct[i * 127 + k] * c[k * 127 + j];
      tmp[j] = s;
    }
    for (int j = 0; j < 16; j++) {
      float s = 0.0f;
      for (int k = 0; k < 127; k++) s += tmp[k] * g_dct[j * 127 + k];
      pdqf[i * 16 + j] = s;
    }
  }
  float g = oracle_sum_of_gradients(pdqf);
  float q = oracle_adjust_quality(g);
  float med = oracle_torben_median(pdqf);
  *threshold = med;
  /* mod.rs:764-788 bit order */
  for (int i = 0; i < 32; i++) {
    uint8_t byte = 0;
    for (int j = 0; j < 8; j++)
      if (pdqf[i * 8 + j] > med) byte |= (uint8_t)(1u << j);
    out[31 - i] = byte;
  }
  return q;
}

/* FP64 oracle of the optimised variant: same 127 sampling centres and tent
 * weights (exact k/81 in double), DCT in double with the f64 table.  Used to
 * state per-coefficient error of the FP32 paths. */
void oracle_pdqf_f64(const float *in, double *pdqf) {
  init_tables();
  static const double a[7] = {0, 1, 2, 3, 2, 1, 0};
  static double c[127 * 127];
  for (int oi = 0; oi < 127; oi++) {
    int in_i = 4 * (oi + 1) - 3;
    for (int oj = 0; oj < 127; oj++) {
      int in_j = 4 * (oj + 1) - 3;
      double s = 0.0;
      for (int di = 0; di < 7; di++)
        for (int dj = 0; dj < 7; dj++)
          s += a[di] * a[dj] / 81.0 * (double)in[(in_i + di) * 512 + in_j + dj];
      c[oi * 127 + oj] = s;
    }
  }
  double tmp[127];
  for (int i = 0; i < 16; i++) {
    for (int j = 0; j < 127; j++) {
      double s = 0.0;
      for (int k = 0; k < 127; k++) s += g_dct64[i * 127 + k] * c[k * 127 + j];
      tmp[j] = s;
    }
    for (int j = 0; j < 16; j++) {
      double s = 0.0;
      for (int k = 0; k < 127; k++) s += tmp[k] * g_dct64[j * 127 + k];
      pdqf[i * 16 + j] = s;
    }
  }
}


/* ReferenceKernel<f64>: mod.rs:844-869 with ref.rs in double, f64 DCT table. */
double oracle_hash_ref64(const float *in, double *threshold, uint8_t *out, double *pdqf) {
  init_tables();
  double *b = (double *)malloc(512 * 512 * sizeof(double));
  for (int i = 0; i < 512 * 512; i++) b[i] = (double)in[i];
  jarosz_f64(b);
  static double c[127 * 127];
  for (int oi = 0; oi < 127; oi++)
    for (int oj = 0; oj < 127; oj++)
      c[oi * 127 + oj] = b[((oi * 2 + 1) * 512) / 254 * 512 + ((oj * 2 + 1) * 512) / 254];
  free(b);
  double tmp[127];
  for (int i = 0; i < 16; i++) {
    for (int j = 0; j < 127; j++) {
      double s = 0.0;
      for (int k = 0; k < 127; k++) s += g_dct64[i * 127 + k] * c[k * 127 + j];
      tmp[j] = s;
    }
    for (int j = 0; j < 16; j++) {
      double s = 0.0;
      for (int k = 0; k < 127; k++) s += tmp[k] * g_dct64[j * 127 + k];
      pdqf[i * 16 + j] = s;
    }
  }
  double g = 0.0;
  for (int i = 0; i < 15; i++)
    for (int j = 0; j < 16; j++) g += fabs(pdqf[i * 16 + j] - pdqf[(i + 1) * 16 + j]);
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < 15; j++) g += fabs(pdqf[i * 16 + j] - pdqf[i * 16 + j + 1]);
  /* sort-based lower median (exact rank 128 of 256, as Torben returns) */
  double s[256];
  memcpy(s, pdqf, sizeof s);
  for (int i = 1; i < 256; i++) {
    double v = s[i];
    int j = i - 1;
    while (j >= 0 && s[j] > v) { s[j + 1] = s[j]; j--; }
    s[j + 1] = v;
  }
  double med = s[127];
  *threshold = med;
  for (int i = 0; i < 32; i++) {
    uint8_t byte = 0;
    for (int j = 0; j < 8; j++)
      if (pdqf[i * 8 + j] > med) byte |= (uint8_t)(1u << j);
    out[31 - i] = byte;
  }
  double q = g / 180.0;
  return q < 1.0 ? q : 1.0;
}

/* Fixture: the README doctest FM-stripe image (README.md:313-322), computed
 * with the same f32 operation order (frequency accumulates over all 512x512
 * elements; sinf as f32::sin lowers to libm sinf). */
void oracle_readme_fm_image(float *out) {
  float frequency = 4.0f;
  float modulation = frequency / 512.0f;
  const float pi = 3.14159274101257324f; /* core::f32::consts::PI */
  for (int i = 0; i < 512; i++)
    for (int j = 0; j < 512; j++) {
      frequency += modulation;
      float arg = (float)j * 2.0f;
      arg = arg * pi;
      arg = arg * frequency;
      arg = arg / 512.0f;
      float value = sinf(arg);
      out[i * 512 + j] = value * 127.0f + 128.0f;
    }
}

/* ---------------------------------------------------------------------------
 * Pipeline: hash_get_threshold, lib.rs:427-455, with the AVX2 kernel.
 * ------------------------------------------------------------------------- */
float oracle_hash_get_threshold(const float *in, float *threshold, uint8_t *out, float *buf1,
                                float *tmp, float *pdqf) {
  oracle_compress_avx2(in, buf1);
  oracle_dct2d_avx2(buf1, tmp, pdqf);
  float g = oracle_sum_of_gradients(pdqf);
  float q = oracle_adjust_quality(g);
  oracle_quantize_avx2(pdqf, threshold, out);
  return q;
}

/* Sign algebra on the 16x16 PDQF. */
static void negate_rows_even(float *p) {
  /* pdqf_negate_alt_rows::<true>, AVX2 form x86.rs:285-326: rows 0,2,..,14 */
  for (int i = 0; i < 16; i += 2)
    for (int j = 0; j < 16; j++) p[i * 16 + j] = p[i * 16 + j] * -1.0f;
}

static void negate_cols_even(float *p, int ncols) {
  /* pdqf_negate_alt_cols::<false>: coefficient -1 at even columns.  The AVX2
   * form (x86.rs:330-376) loads 8 floats per row and advances 16, so it
   * reaches columns 0..7 only (ncols = 8); the AVX-512 form (x86.rs:1026-
   * 1074) and the scalar default (mod.rs:368-382) cover all 16. */
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < ncols; j += 2) p[i * 16 + j] = p[i * 16 + j] * -1.0f;
}

static void transpose16(float *p) {
  /* pdqf_t, mod.rs:345-361 */
  for (int i = 0; i < 16; i++)
    for (int j = i + 1; j < 16; j++) {
      float t = p[i * 16 + j];
      p[i * 16 + j] = p[j * 16 + i];
      p[j * 16 + i] = t;
    }
}

/* visit_dihedrals, lib.rs:358-420.  mode 0 = exact sign algebra (full
 * column negation), mode 1 = the AVX2 kernel's column extent.  Emits the 7
 * non-identity dihedrals in the reference's order; for each: the packed
 * dihedral code, the per-dihedral threshold, quality, hash and PDQF. */
void oracle_visit_dihedrals(const float *pdqf_in, float threshold, int mode, uint32_t *codes,
                            float *thresholds, float *qualities, uint8_t *hashes,
                            float *pdqfs) {
  float p[256];
  memcpy(p, pdqf_in, sizeof p);
  int nc = mode == 1 ? 8 : 16;
  float t_cols = threshold, t_rows = threshold, t_diag = threshold;
  int n = 0;
#define EMIT(code, thr)                                          \
  do {                                                           \
    codes[n] = (code);                                           \
    thresholds[n] = (thr);                                       \
    qualities[n] = oracle_adjust_quality(oracle_sum_of_gradients(p)); \
    memcpy(pdqfs + n * 256, p, sizeof p);                        \
    n++;                                                         \
  } while (0)
  negate_cols_even(p, nc);
  oracle_quantize_avx2(p, &t_cols, hashes + n * 32);
  EMIT(ORACLE_DIHEDRAL_FLOPPED, t_cols);
  negate_rows_even(p);
  oracle_quantize_avx2(p, &t_diag, hashes + n * 32);
  EMIT(ORACLE_DIHEDRAL_ROTATED_180, t_diag);
  negate_cols_even(p, nc);
  oracle_quantize_avx2(p, &t_rows, hashes + n * 32);
  EMIT(ORACLE_DIHEDRAL_FLIPPED, t_rows);
  negate_rows_even(p);
  transpose16(p);
  oracle_threshold_2d(p, hashes + n * 32, threshold);
  EMIT(ORACLE_DIHEDRAL_ROTATED_90_FLOPPED, threshold);
  negate_rows_even(p);
  oracle_threshold_2d(p, hashes + n * 32, t_cols);
  EMIT(ORACLE_DIHEDRAL_ROTATED_270, t_cols);
  negate_cols_even(p, nc);
  oracle_threshold_2d(p, hashes + n * 32, t_diag);
  EMIT(ORACLE_DIHEDRAL_FLOPPED_ROTATED_270, t_diag);
  negate_rows_even(p);
  oracle_threshold_2d(p, hashes + n * 32, t_rows);
  EMIT(ORACLE_DIHEDRAL_ROTATED_90, t_rows);
#undef EMIT
}

/* Full per-image result in the batch layout: 8 dihedrals, NORMAL first and
 * then the reference's visiting order. */
void oracle_hash_dihedrals(const float *in, int mode, uint8_t *hashes /*8x32*/,
                           float *qualities /*8*/, float *thresholds /*8*/,
                           float *pdqf /*256, NORMAL*/) {
  static float buf1[128 * 128];
  float tmp[128];
  float p[256];
  float thr;
  float q = oracle_hash_get_threshold(in, &thr, hashes, buf1, tmp, p);
  qualities[0] = q;
  thresholds[0] = thr;
  if (pdqf) memcpy(pdqf, p, sizeof p);
  uint32_t codes[7];
  float pd[7 * 256];
  oracle_visit_dihedrals(p, thr, mode, codes, thresholds + 1, qualities + 1, hashes + 32, pd);
}

/* ---------------------------------------------------------------------------
 * Matching (matching/mod.rs:232-257, 336-362): distance = sum of popcounts of
 * the XOR of the four u64 words; a hit is distance <= threshold (inclusive).
 * ------------------------------------------------------------------------- */
static inline uint32_t hamming256(const uint64_t *a, const uint64_t *b) {
  return (uint32_t)(__builtin_popcountll(a[0] ^ b[0]) + __builtin_popcountll(a[1] ^ b[1]) +
                    __builtin_popcountll(a[2] ^ b[2]) + __builtin_popcountll(a[3] ^ b[3]));
}

uint32_t oracle_hamming(const uint8_t *a, const uint8_t *b) {
  uint64_t x[4], y[4];
  memcpy(x, a, 32);
  memcpy(y, b, 32);
  return hamming256(x, y);
}

/* All (query, index, distance) triples with distance <= threshold, in
 * row-major order (db index, then query ascending), which is the visiting
 * order of CpuMatcher::find.  Returns the total count; writes at most cap. */
uint64_t oracle_match_all(const uint8_t *db, uint64_t n_rows, const uint8_t *queries,
                          uint32_t n_queries, uint32_t threshold, OracleHit *hits, uint64_t cap) {
  uint64_t count = 0;
  uint64_t *q = (uint64_t *)malloc((size_t)n_queries * 32);
  memcpy(q, queries, (size_t)n_queries * 32);
  for (uint64_t r = 0; r < n_rows; r++) {
    uint64_t row[4];
    memcpy(row, db + r * 32, 32);
    for (uint32_t j = 0; j < n_queries; j++) {
      uint32_t d = hamming256(row, q + 4 * j);
      if (d <= threshold) {
        if (count < cap) {
          hits[count].query = j;
          hits[count].distance = d;
          hits[count].index = r;
        }
        count++;
      }
    }
  }
  free(q);
  return count;
}

/* First hit in (row, needle) order -- what CpuMatcher::find stops at when the
 * callback returns Some on the first call, and what VulkanMatcher::find
 * reports through atomicMin(needle + row * N) (vulkan.rs:133-136). */
int oracle_match_first(const uint8_t *db, uint64_t n_rows, const uint8_t *queries,
                       uint32_t n_queries, uint32_t threshold, uint64_t *row_out,
                       uint32_t *query_out) {
  for (uint64_t r = 0; r < n_rows; r++)
    for (uint32_t j = 0; j < n_queries; j++)
      if (oracle_hamming(db + r * 32, queries + (size_t)j * 32) <= threshold) {
        *row_out = r;
        *query_out = j;
        return 1;
      }
  return 0;
}
