// Native driver that launches each measured kernel alone, for profiler
// counter runs (few dispatches, no framework kernels around them) and for
// back-to-back kernel timing.  Cases:
//   copy_db         device-to-device copy of the database (known size: the
//                   scale for the FETCH_SIZE / WRITE_SIZE counters)
//   match_x8        exhaustive match of one image's 8 dihedral hashes
//   match_x1024     exhaustive match of 128 images' 1024 dihedral hashes
//   match_plane_x8, match_plane_x256, match_plane_x1024  the same over the
//                   database's prefix plane (256 queries: the per-wave layout)
//   build_plane     building the prefix plane (not in the default set)
//   hash_u8         fused hash of --batch u8 planes (automatic layout)
//   hash_u8_wave, hash_u8_block4, hash_u8_block8   the same in one layout
//   hash_u8_bf16, hash_u8_bf16x3, hash_u8_f16, hash_u8_f16x3   the same with
//                   the matrix-core DCT (eight waves per image)
// The database holds 50,000 planted near-duplicates (U{0..40} distinct bit
// flips) drawn from a pool of 8192 hashes whose first 1024 are the queries.
// Each case: W untimed launches, then K launches between two HIP events on
// the launch stream; prints one JSON object with the mean ms per launch.
//
// usage: kernel_workloads [--rows N] [--steps K] [--warmup W] [--batch B] [--cases a,b,...]
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "yume_pdq_amd.h"

namespace {

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}
void api_ok(int rc, const char* what) {
  if (rc != YUME_PDQ_AMD_OK) {
    std::fprintf(stderr, "%s: %s\n", what, yume_pdq_amd_strerror(rc));
    std::exit(1);
  }
}
void say(const std::string& s) {
  std::fprintf(stderr, "[kernel_workloads] %s\n", s.c_str());
  std::fflush(stderr);
}

// Smooth gradient + blobs + noise, rounded to u8 (the synthetic image family).
void make_plane(uint32_t k, uint8_t* out) {
  std::mt19937 rng(0x59554D45u + k);
  std::uniform_real_distribution<float> u(0.0f, 1.0f);
  const float c0 = 40.0f + 175.0f * u(rng), cx = -0.2f + 0.4f * u(rng), cy = -0.2f + 0.4f * u(rng);
  const int nb = 4 + (int)(13 * u(rng)) % 13;
  float bx[16], by[16], ba[16], bs[16];
  for (int b = 0; b < nb; b++) {
    bx[b] = 512 * u(rng);
    by[b] = 512 * u(rng);
    ba[b] = (u(rng) < 0.5f ? -1.0f : 1.0f) * (20.0f + 100.0f * u(rng));
    bs[b] = 10.0f + 70.0f * u(rng);
  }
  for (int y = 0; y < 512; y++)
    for (int x = 0; x < 512; x++) {
      float v = c0 + cx * x + cy * y;
      for (int b = 0; b < nb; b++) {
        const float dx = x - bx[b], dy = y - by[b];
        v += ba[b] * std::exp(-(dx * dx + dy * dy) / (2 * bs[b] * bs[b]));
      }
      v += -8.0f + 16.0f * u(rng);
      out[y * 512 + x] = (uint8_t)std::lround(std::fmin(255.0f, std::fmax(0.0f, v)));
    }
}

}  // namespace

int main(int argc, char** argv) {
  uint64_t rows = 10'000'000;
  int steps = 20, warmup = 3;
  uint32_t batch = 128;
  std::string cases = "copy_db,match_x8,match_x1024,match_plane_x8,match_plane_x1024,hash_u8";
  for (int i = 1; i + 1 < argc; i += 2) {
    const std::string a = argv[i];
    if (a == "--rows") rows = std::strtoull(argv[i + 1], nullptr, 10);
    else if (a == "--steps") steps = std::atoi(argv[i + 1]);
    else if (a == "--warmup") warmup = std::atoi(argv[i + 1]);
    else if (a == "--batch") batch = (uint32_t)std::atoi(argv[i + 1]);
    else if (a == "--cases") cases = argv[i + 1];
    else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    }
  }
  if (steps < 1 || warmup < 0 || batch < 1 || batch > 4096 || rows < 1) {
    std::fprintf(stderr, "bad sizes\n");
    return 2;
  }
  const uint32_t n_images = batch < 128 ? 128 : batch;  // 128 images give the 1024 queries

  say("host data: " + std::to_string(rows) + " rows, " + std::to_string(n_images) + " images");
  std::vector<uint8_t> db(rows * 32);
  {
    std::mt19937_64 rng(7);
    uint64_t* w = reinterpret_cast<uint64_t*>(db.data());
    for (uint64_t i = 0; i < rows * 4; i++) w[i] = rng();
  }
  // 256 distinct planes; larger batches repeat them (the kernels' cost does
  // not depend on the pixels)
  std::vector<uint8_t> planes((size_t)n_images * 512 * 512);
  for (uint32_t k = 0; k < n_images; k++) {
    uint8_t* dst = planes.data() + (size_t)k * 512 * 512;
    if (k < 256)
      make_plane(k, dst);
    else
      std::memcpy(dst, planes.data() + (size_t)(k % 256) * 512 * 512, 512 * 512);
  }

  hipStream_t s;
  hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  uint8_t *d_db = nullptr, *d_db2 = nullptr, *d_planes = nullptr, *d_hashes = nullptr;
  float *d_q = nullptr, *d_t = nullptr;
  YumePdqAmdHit* d_hits = nullptr;
  unsigned long long* d_count = nullptr;
  const uint64_t cap = 1ull << 22;
  hip_ok(hipMalloc(&d_db, rows * 32), "hipMalloc db");
  hip_ok(hipMalloc(&d_db2, rows * 32), "hipMalloc db copy");
  hip_ok(hipMalloc(&d_planes, planes.size()), "hipMalloc planes");
  hip_ok(hipMalloc(&d_hashes, (size_t)n_images * 256), "hipMalloc hashes");
  hip_ok(hipMalloc(&d_q, (size_t)n_images * 32), "hipMalloc quality");
  hip_ok(hipMalloc(&d_t, (size_t)n_images * 32), "hipMalloc threshold");
  hip_ok(hipMalloc(&d_hits, cap * sizeof(YumePdqAmdHit)), "hipMalloc hits");
  hip_ok(hipMalloc(&d_count, 8), "hipMalloc count");
  hip_ok(hipMemcpy(d_planes, planes.data(), planes.size(), hipMemcpyHostToDevice), "upload planes");

  // queries = the dihedral hashes of the images.  50k near-duplicates
  // (k ~ U{0..40} distinct flips) are planted from a pool of 8192 hashes (the
  // dihedral hashes of 1024 images; past the 1024 computed here the pool is
  // random), so 1024 queries meet about 6,250 plants.
  api_ok(yume_pdq_amd_hash_batch(d_planes, YUME_PDQ_AMD_INPUT_U8, 512 * 512, n_images, YUME_PDQ_AMD_DIHEDRAL_AVX2,
                                 d_hashes, d_q, d_t, nullptr, s),
         "hash_batch");
  hip_ok(hipStreamSynchronize(s), "sync");
  std::vector<uint8_t> queries((size_t)n_images * 256);
  hip_ok(hipMemcpy(queries.data(), d_hashes, queries.size(), hipMemcpyDeviceToHost), "download hashes");
  {
    std::mt19937_64 rng(11);
    const uint64_t plants = rows / 200 < 50'000 ? rows / 200 : 50'000;
    for (uint64_t p = 0; p < plants; p++) {
      const uint64_t r = rng() % rows;
      const uint64_t q = rng() % 8192;
      uint8_t* row = db.data() + r * 32;
      if (q < 1024) {
        std::memcpy(row, queries.data() + q * 32, 32);
      } else {
        for (int b = 0; b < 32; b++) row[b] = (uint8_t)rng();
      }
      const int k = (int)(rng() % 41);
      int perm[256];
      for (int b = 0; b < 256; b++) perm[b] = b;
      for (int f = 0; f < k; f++) {
        const int j = f + (int)(rng() % (256 - f));
        std::swap(perm[f], perm[j]);
        row[perm[f] >> 3] ^= (uint8_t)(1u << (perm[f] & 7));
      }
    }
  }
  hip_ok(hipMemcpy(d_db, db.data(), db.size(), hipMemcpyHostToDevice), "upload db");

  struct Case {
    std::string name;
    std::function<void()> fn;
  };
  std::vector<Case> all;
  all.push_back({"copy_db", [&] {
                   hip_ok(hipMemcpyAsync(d_db2, d_db, rows * 32, hipMemcpyDeviceToDevice, s), "copy");
                 }});
  all.push_back({"match_x8", [&] {
                   api_ok(yume_pdq_amd_match(d_db, rows, d_hashes, 8, 31, d_hits, cap, d_count, 0, 0, s), "match");
                 }});
  all.push_back({"match_x1024", [&] {
                   api_ok(yume_pdq_amd_match(d_db, rows, d_hashes, 1024, 31, d_hits, cap, d_count, 0, 0, s),
                          "match");
                 }});
  void* d_plane = nullptr;
  hip_ok(hipMalloc(&d_plane, yume_pdq_amd_prefix_plane_bytes(rows)), "hipMalloc plane");
  api_ok(yume_pdq_amd_build_prefix_plane(d_db, rows, d_plane, s), "build_prefix_plane");
  all.push_back({"build_plane", [&] {
                   api_ok(yume_pdq_amd_build_prefix_plane(d_db, rows, d_plane, s), "build_prefix_plane");
                 }});
  all.push_back({"match_plane_x8", [&] {
                   api_ok(yume_pdq_amd_match_with_plane(d_db, d_plane, rows, d_hashes, 8, 31, d_hits, cap, d_count, 0,
                                                        0, s),
                          "match_with_plane");
                 }});
  all.push_back({"match_plane_x256", [&] {
                   api_ok(yume_pdq_amd_match_with_plane(d_db, d_plane, rows, d_hashes, 256, 31, d_hits, cap, d_count,
                                                        0, 0, s),
                          "match_with_plane");
                 }});
  all.push_back({"match_plane_x1024", [&] {
                   api_ok(yume_pdq_amd_match_with_plane(d_db, d_plane, rows, d_hashes, 1024, 31, d_hits, cap, d_count,
                                                        0, 0, s),
                          "match_with_plane");
                 }});
  all.push_back({"hash_u8", [&] {
                   api_ok(yume_pdq_amd_hash_batch(d_planes, YUME_PDQ_AMD_INPUT_U8, 512 * 512, batch,
                                                  YUME_PDQ_AMD_DIHEDRAL_AVX2, d_hashes, d_q, d_t, nullptr, s),
                          "hash_batch");
                 }});
  struct Variant {
    const char* name;
    uint32_t dct, layout;
  };
  const Variant variants[] = {{"hash_u8_wave", YUME_PDQ_AMD_DCT_EXACT, YUME_PDQ_AMD_LAYOUT_WAVE},
                              {"hash_u8_block4", YUME_PDQ_AMD_DCT_EXACT, YUME_PDQ_AMD_LAYOUT_BLOCK4},
                              {"hash_u8_block8", YUME_PDQ_AMD_DCT_EXACT, YUME_PDQ_AMD_LAYOUT_BLOCK8},
                              {"hash_u8_bf16", YUME_PDQ_AMD_DCT_BF16, YUME_PDQ_AMD_LAYOUT_BLOCK8},
                              {"hash_u8_bf16x3", YUME_PDQ_AMD_DCT_BF16X3, YUME_PDQ_AMD_LAYOUT_BLOCK8},
                              {"hash_u8_f16", YUME_PDQ_AMD_DCT_F16, YUME_PDQ_AMD_LAYOUT_BLOCK8},
                              {"hash_u8_f16x3", YUME_PDQ_AMD_DCT_F16X3, YUME_PDQ_AMD_LAYOUT_BLOCK8}};
  for (const Variant& v : variants) {
    all.push_back({v.name, [&, v] {
                     YumePdqAmdHashOptions o = {YUME_PDQ_AMD_DIHEDRAL_AVX2, v.dct, v.layout, nullptr};
                     api_ok(yume_pdq_amd_hash_batch_ex(d_planes, YUME_PDQ_AMD_INPUT_U8, 512 * 512, batch, &o,
                                                       d_hashes, d_q, d_t, nullptr, nullptr, s),
                            "hash_batch_ex");
                   }});
  }

  hipEvent_t e0, e1;
  hip_ok(hipEventCreate(&e0), "event");
  hip_ok(hipEventCreate(&e1), "event");
  std::string json = "{\"rows\": " + std::to_string(rows) + ", \"batch\": " + std::to_string(batch) +
                     ", \"steps\": " + std::to_string(steps) + ", \"mean_ms\": {";
  bool first = true;
  for (const Case& c : all) {
    if (("," + cases + ",").find("," + c.name + ",") == std::string::npos) continue;
    say(c.name + ": " + std::to_string(warmup) + " + " + std::to_string(steps) + " launches");
    for (int i = 0; i < warmup; i++) c.fn();
    hip_ok(hipStreamSynchronize(s), "sync");
    hip_ok(hipEventRecord(e0, s), "record");
    for (int i = 0; i < steps; i++) c.fn();
    hip_ok(hipEventRecord(e1, s), "record");
    hip_ok(hipEventSynchronize(e1), "sync");
    float ms = 0;
    hip_ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    unsigned long long count = 0;
    hip_ok(hipMemcpy(&count, d_count, 8, hipMemcpyDeviceToHost), "count");
    char buf[160];
    std::snprintf(buf, sizeof buf, "%s\"%s\": %.5f", first ? "" : ", ", c.name.c_str(), ms / steps);
    json += buf;
    first = false;
    say(c.name + " done, hits in last launch " + std::to_string(count));
  }
  json += "}}";
  std::printf("%s\n", json.c_str());
  hip_ok(hipStreamSynchronize(s), "sync");
  (void)hipFree(d_db);
  (void)hipFree(d_db2);
  (void)hipFree(d_planes);
  (void)hipFree(d_hashes);
  (void)hipFree(d_q);
  (void)hipFree(d_t);
  (void)hipFree(d_hits);
  (void)hipFree(d_count);
  (void)hipFree(d_plane);
  return 0;
}
