// GPU test of the C++ host API (yume-pdq_amd/host/yume_pdq.hpp): the Kernel
// trait mirror, the free functions hash / hash_get_threshold / hash_float /
// visit_dihedrals, and both PDQMatcher mirrors, each checked bit for bit
// against the CPU oracle (oracle/liboracle.so, the checker) and against the
// batched C ABI.  Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../oracle/pdq_oracle.h"
#include "yume_pdq.hpp"

namespace {

int g_failures = 0;

void expect(bool ok, const char* what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  std::fflush(stdout);
  if (!ok) g_failures++;
}

bool same_bits(const float* a, const float* b, size_t n) { return std::memcmp(a, b, n * sizeof(float)) == 0; }

// Smooth synthetic 512x512 image with texture, values in 0..255.
void make_image(uint32_t seed, float* img) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  const float fx = 2.0f + 10.0f * u(rng), fy = 2.0f + 10.0f * u(rng), ph = 6.0f * u(rng);
  for (int y = 0; y < 512; y++)
    for (int x = 0; x < 512; x++) {
      const float v = 127.5f + 60.0f * std::sin(fx * x / 512.0f * 6.2831853f + ph) * std::cos(fy * y / 512.0f * 3.14159f) +
                      40.0f * (u(rng) - 0.5f) + 0.1f * (x - y);
      img[y * 512 + x] = std::rint(std::fmin(255.0f, std::fmax(0.0f, v)));
    }
}

void test_kernel_methods(yume_pdq::GpuKernel& k, const float* img) {
  using namespace yume_pdq;
  static float buf1[128][128], obuf1[128 * 128];
  float tmp[128], otmp[128], pdqf[16][16], opdqf[256], thr = 0, othr = 0;
  uint8_t hash[16][2], ohash[32];

  k.jarosz_compress(*reinterpret_cast<const Image*>(img), buf1);
  oracle_compress_avx2(img, obuf1);
  expect(same_bits(&buf1[0][0], obuf1, 128 * 128), "jarosz_compress == oracle_compress_avx2");

  k.dct2d(buf1, tmp, pdqf);
  oracle_dct2d_avx2(obuf1, otmp, opdqf);
  expect(same_bits(&pdqf[0][0], opdqf, 256), "dct2d pdqf == oracle_dct2d_avx2");
  expect(same_bits(tmp, otmp, 128), "dct2d tmp row == oracle");

  const float g = k.sum_of_gradients(pdqf);
  const float og = oracle_sum_of_gradients(opdqf);
  expect(std::memcmp(&g, &og, 4) == 0, "sum_of_gradients == oracle");
  const float q = GpuKernel::adjust_quality(g), oq = oracle_adjust_quality(og);
  expect(std::memcmp(&q, &oq, 4) == 0, "adjust_quality == oracle");

  k.quantize(pdqf, thr, hash);
  oracle_quantize_avx2(opdqf, &othr, ohash);
  expect(std::memcmp(&thr, &othr, 4) == 0 && std::memcmp(hash, ohash, 32) == 0, "quantize == oracle_quantize_avx2");

  // re-threshold at a shifted value
  const float t2 = thr * 0.5f;
  k.threshold_2d(pdqf, hash, t2);
  oracle_threshold_2d(opdqf, ohash, t2);
  expect(std::memcmp(hash, ohash, 32) == 0, "threshold_2d == oracle_threshold_2d");

  // transforms vs a host restatement of the scalar trait defaults
  float p[16][16], e[16][16];
  std::memcpy(p, pdqf, sizeof p);
  std::memcpy(e, pdqf, sizeof e);
  k.pdqf_t(p);
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < 16; j++) e[i][j] = pdqf[j][i];
  expect(same_bits(&p[0][0], &e[0][0], 256), "pdqf_t");
  k.pdqf_negate_off_diagonals(p);
  for (int i = 0; i < 16; i++)
    for (int j = 0; j < 16; j++)
      if ((i + j) & 1) e[i][j] = -e[i][j];
  expect(same_bits(&p[0][0], &e[0][0], 256), "pdqf_negate_off_diagonals");
  k.pdqf_negate_alt_rows<true>(p);
  for (int i = 0; i < 16; i += 2)
    for (int j = 0; j < 16; j++) e[i][j] = -e[i][j];
  expect(same_bits(&p[0][0], &e[0][0], 256), "pdqf_negate_alt_rows<true> (rows 0,2,..)");
  k.pdqf_negate_alt_cols<true>(p);
  for (int i = 0; i < 16; i++)
    for (int j = 1; j < 8; j += 2) e[i][j] = -e[i][j];
  expect(same_bits(&p[0][0], &e[0][0], 256), "pdqf_negate_alt_cols<true> (AVX2 extent)");
}

void test_luma(yume_pdq::GpuKernel& k) {
  std::mt19937 rng(99);
  uint8_t rgb[512][3], rgba[512][4];
  for (auto& px : rgb)
    for (auto& c : px) c = (uint8_t)rng();
  for (auto& px : rgba)
    for (auto& c : px) c = (uint8_t)rng();
  float out[512];
  const float* c = yume_pdq::RGB8_TO_LUMA8_TABLE_ITU;
  k.cvt_rgb8_to_luma8f(rgb, out);
  bool ok = true;
  for (int i = 0; i < 512; i++) {
    // trait default (mod.rs:303-322): R*r + G*g + B*b, left to right, unfused
    volatile float a = c[0] * (float)rgb[i][0];
    volatile float b = c[1] * (float)rgb[i][1];
    volatile float d = c[2] * (float)rgb[i][2];
    const float e = (a + b) + d;
    ok &= std::memcmp(&e, &out[i], 4) == 0;
  }
  expect(ok, "cvt_rgb8_to_luma8f");
  k.cvt_rgba8_to_luma8f(rgba, out);
  ok = true;
  for (int i = 0; i < 512; i++) {
    // AVX2 override (x86.rs:686-753): r*R, fma(g, G, .), fma(b, B, .)
    float e = (float)rgba[i][0] * c[0];
    e = std::fmaf((float)rgba[i][1], c[1], e);
    e = std::fmaf((float)rgba[i][2], c[2], e);
    ok &= std::memcmp(&e, &out[i], 4) == 0;
  }
  expect(ok, "cvt_rgba8_to_luma8f");
}

void test_free_functions(yume_pdq::GpuKernel& k, const float* img) {
  using namespace yume_pdq;
  static float buf1[128][128], obuf1[128 * 128];
  float tmp[128], otmp[128], pdqf[16][16], opdqf[256], thr = 0, othr = 0;
  uint8_t hash[16][2], ohash[32];

  const float q = hash_get_threshold(k, *reinterpret_cast<const Image*>(img), thr, hash, buf1, tmp, pdqf);
  const float oq = oracle_hash_get_threshold(img, &othr, ohash, obuf1, otmp, opdqf);
  expect(std::memcmp(&q, &oq, 4) == 0 && std::memcmp(&thr, &othr, 4) == 0 && std::memcmp(hash, ohash, 32) == 0 &&
             same_bits(&pdqf[0][0], opdqf, 256),
         "hash_get_threshold == oracle");

  // the fused batched kernel agrees with the staged composition
  float* d_img = nullptr;
  uint8_t* d_h = nullptr;
  float *d_q = nullptr, *d_t = nullptr, *d_p = nullptr;
  yume_pdq::detail::check_hip(hipMalloc(&d_img, 512 * 512 * 4), "hipMalloc");
  yume_pdq::detail::check_hip(hipMalloc(&d_h, 8 * 32), "hipMalloc");
  yume_pdq::detail::check_hip(hipMalloc(&d_q, 8 * 4), "hipMalloc");
  yume_pdq::detail::check_hip(hipMalloc(&d_t, 8 * 4), "hipMalloc");
  yume_pdq::detail::check_hip(hipMalloc(&d_p, 256 * 4), "hipMalloc");
  yume_pdq::detail::check_hip(hipMemcpy(d_img, img, 512 * 512 * 4, hipMemcpyHostToDevice), "hipMemcpy");
  yume_pdq::detail::check(yume_pdq_amd_hash_batch(d_img, YUME_PDQ_AMD_INPUT_F32, 512 * 512 * 4, 1,
                                                  YUME_PDQ_AMD_DIHEDRAL_AVX2, d_h, d_q, d_t, d_p, nullptr),
                          "hash_batch");
  uint8_t fh[8][32];
  float fq[8], ft[8], fp[256];
  yume_pdq::detail::check_hip(hipDeviceSynchronize(), "sync");
  yume_pdq::detail::check_hip(hipMemcpy(fh, d_h, sizeof fh, hipMemcpyDeviceToHost), "hipMemcpy");
  yume_pdq::detail::check_hip(hipMemcpy(fq, d_q, sizeof fq, hipMemcpyDeviceToHost), "hipMemcpy");
  yume_pdq::detail::check_hip(hipMemcpy(ft, d_t, sizeof ft, hipMemcpyDeviceToHost), "hipMemcpy");
  yume_pdq::detail::check_hip(hipMemcpy(fp, d_p, sizeof fp, hipMemcpyDeviceToHost), "hipMemcpy");
  expect(std::memcmp(fh[0], hash, 32) == 0 && std::memcmp(&fq[0], &q, 4) == 0 && std::memcmp(&ft[0], &thr, 4) == 0 &&
             same_bits(fp, &pdqf[0][0], 256),
         "fused hash_batch NORMAL == staged composition");

  float hf[16][16];
  const float qf = hash_float(k, *reinterpret_cast<const Image*>(img), hf, buf1, tmp);
  expect(std::memcmp(&qf, &q, 4) == 0 && same_bits(&hf[0][0], opdqf, 256), "hash_float == oracle PDQF");

  // visit_dihedrals through the trait methods vs the oracle and the fused kernel
  uint32_t codes[8];
  float othrs[8], oqs[8], opdqfs[8 * 256];
  uint8_t ohashes[8 * 32];
  oracle_visit_dihedrals(opdqf, othr, 1, codes, othrs, oqs, ohashes, opdqfs);
  int n = 0;
  bool ok = true;
  float vp[16][16];
  std::memcpy(vp, pdqf, sizeof vp);
  const bool completed = visit_dihedrals(k, vp, hash, thr, [&](Dihedrals d, float t, float qq, PDQHashF& pp, PDQHash& hh) {
    ok &= d.packed == codes[n];
    ok &= std::memcmp(&t, &othrs[n], 4) == 0 && std::memcmp(&qq, &oqs[n], 4) == 0;
    ok &= std::memcmp(hh, ohashes + 32 * n, 32) == 0 && same_bits(&pp[0][0], opdqfs + 256 * n, 256);
    ok &= std::memcmp(hh, fh[n + 1], 32) == 0 && std::memcmp(&qq, &fq[n + 1], 4) == 0 &&
          std::memcmp(&t, &ft[n + 1], 4) == 0;
    n++;
    return true;
  });
  expect(completed && n == 7 && ok, "visit_dihedrals == oracle and fused kernel (7 dihedrals)");
  std::memcpy(vp, pdqf, sizeof vp);
  n = 0;
  const bool stopped = !visit_dihedrals(k, vp, hash, thr, [&](Dihedrals, float, float, PDQHashF&, PDQHash&) {
    return ++n < 3;
  });
  expect(stopped && n == 3, "visit_dihedrals stops when the callback fails");

  // drop-in C ABI on the same image
  float dthr = 0;
  uint8_t dh[32];
  const float dq = yume_pdq_hash_smart_kernel(img, &dthr, dh, &buf1[0][0], tmp, &pdqf[0][0]);
  expect(std::memcmp(&dq, &oq, 4) == 0 && std::memcmp(dh, ohash, 32) == 0 && std::memcmp(&dthr, &othr, 4) == 0,
         "yume_pdq_hash_smart_kernel == oracle");
  (void)hipFree(d_img);
  (void)hipFree(d_h);
  (void)hipFree(d_q);
  (void)hipFree(d_t);
  (void)hipFree(d_p);
}

void test_matchers() {
  using namespace yume_pdq;
  std::mt19937_64 rng(7);
  const size_t n_rows = 20011, n_needles = 37;
  std::vector<uint8_t> db(n_rows * 32), nd(n_needles * 32);
  for (auto& b : db) b = (uint8_t)rng();
  for (auto& b : nd) b = (uint8_t)rng();
  // plant near-duplicates of needles at distances 0..40
  for (int p = 0; p < 300; p++) {
    const size_t r = rng() % n_rows, q = rng() % n_needles;
    std::memcpy(&db[r * 32], &nd[q * 32], 32);
    const int flips = (int)(rng() % 41);
    for (int f = 0; f < flips; f++) db[r * 32 + rng() % 32] ^= (uint8_t)(1u << (rng() % 8));
  }
  std::vector<OracleHit> exp(n_rows * n_needles);
  const uint64_t n_exp = oracle_match_all(db.data(), n_rows, nd.data(), (uint32_t)n_needles, 31, exp.data(), exp.size());
  exp.resize(n_exp);

  auto rows = reinterpret_cast<const Row*>(db.data());
  auto needles = reinterpret_cast<const Row*>(nd.data());
  auto same = [&](const std::vector<Hit>& got) {
    if (got.size() != exp.size()) return false;
    for (size_t i = 0; i < got.size(); i++)
      if (got[i].query != exp[i].query || got[i].index != exp[i].index || got[i].distance != exp[i].distance)
        return false;
    return true;
  };

  GpuMatcher m(rows, n_rows, 31);
  expect(n_exp > 0 && same(m.find_all(needles, n_needles)), "GpuMatcher::find_all == oracle (row, needle) order");
  expect(m.scan(needles, n_needles), "GpuMatcher::scan");
  auto first = m.find<std::pair<size_t, size_t>>(needles, n_needles, [](size_t r, size_t q) {
    return std::optional<std::pair<size_t, size_t>>(std::make_pair(r, q));
  });
  uint64_t orow = 0;
  uint32_t oq = 0;
  const int ofound = oracle_match_first(db.data(), n_rows, nd.data(), (uint32_t)n_needles, 31, &orow, &oq);
  expect(ofound && first && first->first == orow && first->second == oq, "GpuMatcher::find first == oracle_match_first");
  auto ff = m.find_first(needles, n_needles);
  expect(ff && ff->first == orow && ff->second == oq, "GpuMatcher::find_first == oracle_match_first");
  // a callback that never returns visits every hit
  size_t visited = 0;
  auto none = m.find<int>(needles, n_needles, [&](size_t, size_t) {
    visited++;
    return std::optional<int>();
  });
  expect(!none && visited == n_exp, "GpuMatcher::find visits every hit");
  // no hits: needles far from everything at threshold 0
  std::vector<uint8_t> miss(32 * 4);
  for (auto& b : miss) b = (uint8_t)rng();
  GpuMatcher m0(rows, n_rows, 0);
  expect(!m0.scan(reinterpret_cast<const Row*>(miss.data()), 4), "GpuMatcher::scan negative control");

  GpuBatchMatcher bm(needles, n_needles, 31);
  expect(same(bm.find_all(rows, n_rows)), "GpuBatchMatcher::find_all == oracle");
  expect(bm.scan(rows, n_rows), "GpuBatchMatcher::scan");
}

}  // namespace

int main() {
  if (!yume_pdq::GpuKernel::required_hardware_features_met()) {
    std::printf("FAIL no GPU visible\n");
    return 2;
  }
  try {
    std::vector<float> img(512 * 512);
    for (uint32_t seed : {1u, 2u, 3u}) {
      make_image(seed, img.data());
      std::printf("[host api] image seed %u\n", seed);
      yume_pdq::GpuKernel k;
      test_kernel_methods(k, img.data());
      test_free_functions(k, img.data());
    }
    yume_pdq::GpuKernel k;
    test_luma(k);
    std::printf("[host api] matchers\n");
    test_matchers();
  } catch (const std::exception& e) {
    std::printf("FAIL exception: %s\n", e.what());
    return 1;
  }
  std::printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "PASSED", g_failures);
  return g_failures ? 1 : 0;
}
