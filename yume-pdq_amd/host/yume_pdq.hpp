// yume_pdq.hpp -- C++17 host API of the MI355X implementation, mirroring the
// reference crate's Rust interface for the hash and match paths:
//
//   trait Kernel           (src/kernel/mod.rs:268-497)  -> yume_pdq::GpuKernel
//   hash, hash_get_threshold, hash_float, visit_dihedrals
//                          (src/lib.rs:322-479)         -> free function templates
//   threshold_2d_f32       (src/kernel/threshold.rs:37-119)
//   struct Dihedrals       (src/lib.rs:85-134)          -> yume_pdq::Dihedrals
//   trait PDQMatcher       (src/matching/mod.rs:50-105) -> GpuMatcher (DB resident,
//                          like VulkanMatcher) and GpuBatchMatcher (needles
//                          resident, haystack streamed, like CpuMatcher)
//
// Each Kernel method is one GPU launch on the object's stream with the
// reference's AVX2 arithmetic; buffers are host arrays as in the trait.  The
// batched device API (yume_pdq_amd.h) is what throughput work should use.
// Header-only; link with -lyume_pdq -lamdhip64.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "yume_pdq.h"
#include "yume_pdq_amd.h"

namespace yume_pdq {

// Array shapes of the reference's associated types for the smart kernel.
using Image = float[512][512];    // InputDimension = U512
using Buffer1 = float[128][128];  // Buffer1WidthX = Buffer1LengthY = U128
using RowTmp = float[128];
using PDQHashF = float[16][16];   // OutputDimension = U16
using PDQHash = uint8_t[16][2];

class Error : public std::runtime_error {
 public:
  explicit Error(const std::string& what) : std::runtime_error(what) {}
};

namespace detail {
inline void check(int rc, const char* what) {
  if (rc < 0) throw Error(std::string(what) + ": " + yume_pdq_amd_strerror(rc));
}
inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}

// Owning device allocation.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) : bytes_(bytes) {
    if (bytes) check_hip(hipMalloc(&p_, bytes), "hipMalloc");
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(bytes_, o.bytes_);
    return *this;
  }
  ~DeviceBuffer() {
    if (p_) (void)hipFree(p_);
  }
  template <typename T = void>
  T* get() const { return static_cast<T*>(p_); }
  size_t size() const { return bytes_; }

 private:
  void* p_ = nullptr;
  size_t bytes_ = 0;
};

class Stream {
 public:
  Stream() { check_hip(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreate"); }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  ~Stream() {
    if (s_) (void)hipStreamDestroy(s_);
  }
  hipStream_t get() const { return s_; }
  void sync() const { check_hip(hipStreamSynchronize(s_), "hipStreamSynchronize"); }

 private:
  hipStream_t s_ = nullptr;
};
}  // namespace detail

// Packed 2x2 dihedral matrix (src/lib.rs:85-134).
struct Dihedrals {
  uint32_t packed;
  static constexpr Dihedrals from_tuples(int8_t dx0, int8_t dx1, int8_t dy0, int8_t dy1) {
    return Dihedrals{(uint32_t)(uint8_t)dx0 << 24 | (uint32_t)(uint8_t)dx1 << 16 | (uint32_t)(uint8_t)dy0 << 8 |
                     (uint32_t)(uint8_t)dy1};
  }
  bool operator==(const Dihedrals& o) const { return packed == o.packed; }
  static const Dihedrals NORMAL, FLIPPED, FLOPPED, ROTATED_180, ROTATED_90, ROTATED_270, ROTATED_90_FLOPPED,
      FLOPPED_ROTATED_270;
};
inline constexpr Dihedrals Dihedrals::NORMAL = Dihedrals::from_tuples(1, 0, 0, 1);
inline constexpr Dihedrals Dihedrals::FLIPPED = Dihedrals::from_tuples(1, 0, 0, -1);
inline constexpr Dihedrals Dihedrals::FLOPPED = Dihedrals::from_tuples(-1, 0, 0, 1);
inline constexpr Dihedrals Dihedrals::ROTATED_180 = Dihedrals::from_tuples(-1, 0, 0, -1);
inline constexpr Dihedrals Dihedrals::ROTATED_90 = Dihedrals::from_tuples(0, 1, -1, 0);
inline constexpr Dihedrals Dihedrals::ROTATED_270 = Dihedrals::from_tuples(0, 1, 1, 0);
inline constexpr Dihedrals Dihedrals::ROTATED_90_FLOPPED = Dihedrals::from_tuples(0, -1, -1, 0);
inline constexpr Dihedrals Dihedrals::FLOPPED_ROTATED_270 = Dihedrals::from_tuples(0, -1, 1, 0);

// BT.601 luma coefficients (src/kernel/constants.rs:23).
inline constexpr float RGB8_TO_LUMA8_TABLE_ITU[3] = {0.299f, 0.587f, 0.114f};

// The Kernel trait on the GPU.  Column negation follows the x86 smart
// kernel (AVX2: columns 0..7, x86.rs:330-376) unless constructed with Exact
// (all 16 columns, as the AVX-512 and scalar kernels).
class GpuKernel {
 public:
  static constexpr int Buffer1WidthX = 128, Buffer1LengthY = 128, InputDimension = 512, OutputDimension = 16;
  using InternalFloat = float;
  enum class ColumnNegation { Avx2, Exact };

  explicit GpuKernel(ColumnNegation neg = ColumnNegation::Avx2)
      : neg_(neg),
        d_img_(512 * 512 * sizeof(float)),
        d_buf1_(128 * 128 * sizeof(float)),
        d_small_(kSmallBytes) {}

  const char* ident() const { return neg_ == ColumnNegation::Avx2 ? "hip_gfx950_f32" : "hip_gfx950_f32_exact"; }

  // RequiredHardwareFeature::met_runtime(): a GPU is visible.
  static bool required_hardware_features_met() {
    int n = 0;
    return hipGetDeviceCount(&n) == hipSuccess && n > 0;
  }

  // Kernel::cvt_rgb8_to_luma8f / cvt_rgba8_to_luma8f on one 512-pixel row.
  void cvt_rgb8_to_luma8f(const uint8_t (&input)[512][3], float (&output)[512],
                          const float (&c)[3] = RGB8_TO_LUMA8_TABLE_ITU) {
    luma(&input[0][0], 3, output, c);
  }
  void cvt_rgba8_to_luma8f(const uint8_t (&input)[512][4], float (&output)[512],
                           const float (&c)[3] = RGB8_TO_LUMA8_TABLE_ITU) {
    luma(&input[0][0], 4, output, c);
  }

  void pdqf_t(PDQHashF& input) { transform(input, YUME_PDQ_AMD_T_TRANSPOSE); }
  template <bool NEGATE>
  void pdqf_negate_alt_cols(PDQHashF& input) {
    const bool avx2 = neg_ == ColumnNegation::Avx2;
    transform(input, NEGATE ? (avx2 ? YUME_PDQ_AMD_T_NEG_COLS_ODD_AVX2 : YUME_PDQ_AMD_T_NEG_COLS_ODD)
                            : (avx2 ? YUME_PDQ_AMD_T_NEG_COLS_EVEN_AVX2 : YUME_PDQ_AMD_T_NEG_COLS_EVEN));
  }
  template <bool NEGATE>
  void pdqf_negate_alt_rows(PDQHashF& input) {
    transform(input, NEGATE ? YUME_PDQ_AMD_T_NEG_ROWS_EVEN : YUME_PDQ_AMD_T_NEG_ROWS_ODD);
  }
  void pdqf_negate_off_diagonals(PDQHashF& input) { transform(input, YUME_PDQ_AMD_T_NEG_OFF_DIAGONALS); }

  void jarosz_compress(const Image& buffer, Buffer1& output) {
    upload(d_img_.get(), buffer, sizeof(Image));
    detail::check(yume_pdq_amd_stage_compress(d_img_.get(), YUME_PDQ_AMD_INPUT_F32, sizeof(Image), 1,
                                              d_buf1_.get<float>(), stream_.get()),
                  "jarosz_compress");
    download(output, d_buf1_.get(), sizeof(Buffer1));
  }

  void dct2d(const Buffer1& buffer, RowTmp& tmp, PDQHashF& output) {
    upload(d_buf1_.get(), buffer, sizeof(Buffer1));
    float* d_tmp = small<float>(kTmpAt);
    float* d_p = small<float>(kPdqfAt);
    detail::check(yume_pdq_amd_stage_dct2d(d_buf1_.get<float>(), 1, d_tmp, d_p, stream_.get()), "dct2d");
    download(tmp, d_tmp, sizeof(RowTmp));
    download(output, d_p, sizeof(PDQHashF));
  }

  float sum_of_gradients(const PDQHashF& input) {
    float* d_p = small<float>(kPdqfAt);
    float* d_g = small<float>(kScalarAt);
    upload(d_p, input, sizeof(PDQHashF));
    detail::check(yume_pdq_amd_stage_sum_of_gradients(d_p, 1, d_g, stream_.get()), "sum_of_gradients");
    float g = 0.0f;
    download(&g, d_g, sizeof g);
    return g;
  }

  // adjust_quality (x86.rs:769-773): min(g / 180, 1), Rust's min ignoring NaN.
  static float adjust_quality(float g) {
    const float s = g / 180.0f;
    return s != s ? 1.0f : (s < 1.0f ? s : 1.0f);
  }

  void quantize(const PDQHashF& input, float& threshold, PDQHash& output) {
    float* d_p = small<float>(kPdqfAt);
    float* d_t = small<float>(kScalarAt);
    uint8_t* d_h = small<uint8_t>(kHashAt);
    upload(d_p, input, sizeof(PDQHashF));
    detail::check(yume_pdq_amd_stage_quantize(d_p, 1, d_t, d_h, stream_.get()), "quantize");
    download(&threshold, d_t, sizeof threshold);
    download(output, d_h, sizeof(PDQHash));
  }

  // threshold_2d_f32 (threshold.rs:37-119).
  void threshold_2d(const PDQHashF& input, PDQHash& output, float threshold) {
    float* d_p = small<float>(kPdqfAt);
    float* d_t = small<float>(kScalarAt);
    uint8_t* d_h = small<uint8_t>(kHashAt);
    upload(d_p, input, sizeof(PDQHashF));
    upload(d_t, &threshold, sizeof threshold);
    detail::check(yume_pdq_amd_stage_threshold(d_p, d_t, 1, d_h, stream_.get()), "threshold_2d");
    download(output, d_h, sizeof(PDQHash));
  }

  hipStream_t stream() const { return stream_.get(); }

 private:
  // Layout of the small device buffer (bytes).
  static constexpr size_t kPdqfAt = 0, kTmpAt = 1024, kScalarAt = 1536, kHashAt = 1600, kSmallBytes = 2048;
  template <typename T>
  T* small(size_t byte_offset) const {
    return reinterpret_cast<T*>(d_small_.get<uint8_t>() + byte_offset);
  }
  void upload(void* d, const void* h, size_t n) {
    detail::check_hip(hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, stream_.get()), "hipMemcpyAsync");
  }
  void download(void* h, const void* d, size_t n) {
    detail::check_hip(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, stream_.get()), "hipMemcpyAsync");
    stream_.sync();
  }
  void transform(PDQHashF& p, int op) {
    float* d_p = small<float>(kPdqfAt);
    upload(d_p, p, sizeof(PDQHashF));
    detail::check(yume_pdq_amd_stage_transform(d_p, 1, op, stream_.get()), "pdqf transform");
    download(p, d_p, sizeof(PDQHashF));
  }
  void luma(const uint8_t* px, int ch, float (&out)[512], const float (&c)[3]) {
    uint8_t* d_px = d_img_.get<uint8_t>();
    float* d_o = d_buf1_.get<float>();
    upload(d_px, px, (size_t)512 * ch);
    detail::check(yume_pdq_amd_stage_luma(d_px, ch, 512, c[0], c[1], c[2], d_o, stream_.get()), "luma");
    download(out, d_o, sizeof out);
  }

  ColumnNegation neg_;
  detail::Stream stream_;
  detail::DeviceBuffer d_img_, d_buf1_, d_small_;
};

template <class K>
void threshold_2d_f32(K& kernel, const PDQHashF& input, PDQHash& output, float threshold) {
  kernel.threshold_2d(input, output, threshold);
}

// hash_get_threshold (src/lib.rs:427-455): the reference's composition of
// kernel methods, in its order.
template <class K>
float hash_get_threshold(K& kernel, const Image& input, float& threshold, PDQHash& output, Buffer1& buf1,
                         RowTmp& tmp, PDQHashF& pdqf) {
  kernel.jarosz_compress(input, buf1);
  kernel.dct2d(buf1, tmp, pdqf);
  const float gradient = kernel.sum_of_gradients(pdqf);
  const float quality = K::adjust_quality(gradient);
  kernel.quantize(pdqf, threshold, output);
  return quality;
}

// hash (src/lib.rs:322-351).
template <class K>
float hash(K& kernel, const Image& input, PDQHash& output, Buffer1& buf1, RowTmp& tmp, PDQHashF& pdqf) {
  float threshold = 0.0f;
  return hash_get_threshold(kernel, input, threshold, output, buf1, tmp, pdqf);
}

// hash_float (src/lib.rs:464-479).
template <class K>
float hash_float(K& kernel, const Image& input, PDQHashF& output, Buffer1& buf1, RowTmp& tmp) {
  kernel.jarosz_compress(input, buf1);
  kernel.dct2d(buf1, tmp, output);
  return K::adjust_quality(kernel.sum_of_gradients(output));
}

// visit_dihedrals (src/lib.rs:358-420): f(dihedral, threshold, quality, pdqf,
// output) is called for the 7 non-identity dihedrals in the reference's
// order, with the buffers holding that dihedral; returning false stops the
// visit (the reference's Err) and visit_dihedrals returns false.
template <class K, class F>
bool visit_dihedrals(K& kernel, PDQHashF& pdqf, PDQHash& output, float threshold, F f) {
  auto call = [&](Dihedrals d, float t) {
    const float quality = K::adjust_quality(kernel.sum_of_gradients(pdqf));
    return f(d, t, quality, pdqf, output);
  };
  float t_cols = threshold, t_rows = threshold, t_diag = threshold;
  kernel.template pdqf_negate_alt_cols<false>(pdqf);
  kernel.quantize(pdqf, t_cols, output);
  if (!call(Dihedrals::FLOPPED, t_cols)) return false;
  kernel.template pdqf_negate_alt_rows<true>(pdqf);
  kernel.quantize(pdqf, t_diag, output);
  if (!call(Dihedrals::ROTATED_180, t_diag)) return false;
  kernel.template pdqf_negate_alt_cols<false>(pdqf);
  kernel.quantize(pdqf, t_rows, output);
  if (!call(Dihedrals::FLIPPED, t_rows)) return false;
  kernel.template pdqf_negate_alt_rows<true>(pdqf);
  kernel.pdqf_t(pdqf);
  threshold_2d_f32(kernel, pdqf, output, threshold);
  if (!call(Dihedrals::ROTATED_90_FLOPPED, threshold)) return false;
  kernel.template pdqf_negate_alt_rows<true>(pdqf);
  threshold_2d_f32(kernel, pdqf, output, t_cols);
  if (!call(Dihedrals::ROTATED_270, t_cols)) return false;
  kernel.template pdqf_negate_alt_cols<false>(pdqf);
  threshold_2d_f32(kernel, pdqf, output, t_diag);
  if (!call(Dihedrals::FLOPPED_ROTATED_270, t_diag)) return false;
  kernel.template pdqf_negate_alt_rows<true>(pdqf);
  threshold_2d_f32(kernel, pdqf, output, t_rows);
  return call(Dihedrals::ROTATED_90, t_rows);
}

// ---------------------------------------------------------------------------
// PDQMatcher
// ---------------------------------------------------------------------------
using Row = uint8_t[32];

struct Hit {
  uint32_t query;
  uint32_t distance;
  uint64_t index;
};

namespace detail {
// Runs the device matcher and returns every hit sorted in the reference's
// visiting order: row ascending, then needle ascending.
inline std::vector<Hit> match_sorted(const void* d_rows, uint64_t n_rows, const void* d_needles, uint32_t n_needles,
                                     uint32_t threshold, DeviceBuffer& d_hits, DeviceBuffer& d_count, hipStream_t s) {
  for (;;) {
    const uint64_t cap = d_hits.size() / sizeof(Hit);
    check(yume_pdq_amd_match(static_cast<const uint8_t*>(d_rows), n_rows, static_cast<const uint8_t*>(d_needles),
                             n_needles, threshold, d_hits.get<YumePdqAmdHit>(), cap,
                             d_count.get<unsigned long long>(), 0, 0, s),
          "match");
    unsigned long long total = 0;
    check_hip(hipMemcpyAsync(&total, d_count.get(), sizeof total, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (total > cap) {  // grow and rerun: the count is exact even when the buffer overflows
      d_hits = DeviceBuffer((size_t)total * sizeof(Hit));
      continue;
    }
    std::vector<Hit> hits(total);
    if (total) {
      check_hip(hipMemcpyAsync(hits.data(), d_hits.get(), total * sizeof(Hit), hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync");
      check_hip(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    std::sort(hits.begin(), hits.end(),
              [](const Hit& a, const Hit& b) { return a.index != b.index ? a.index < b.index : a.query < b.query; });
    return hits;
  }
}
}  // namespace detail

// Database resident on the GPU, queried with needle batches -- the shape of
// VulkanMatcher (src/matching/vulkan.rs:40-473), but exhaustive: find visits
// every hit (FindsAllMatches), not only the first.
class GpuMatcher {
 public:
  static constexpr bool NoFalseNegative = true, NoFalsePositive = true, FindsAllMatches = true;
  static const char* ident() { return "HIP gfx950 matcher (database resident)"; }

  GpuMatcher(const Row* rows, size_t n_rows, uint32_t threshold)
      : n_rows_(n_rows), threshold_(threshold), d_rows_(n_rows * 32), d_hits_(4096 * sizeof(Hit)), d_count_(8) {
    if (n_rows)
      detail::check_hip(hipMemcpyAsync(d_rows_.get(), rows, n_rows * 32, hipMemcpyHostToDevice, stream_.get()),
                        "hipMemcpyAsync");
    stream_.sync();
  }

  // PDQMatcher::scan: is any needle within the threshold of any row?
  bool scan(const Row* needles, size_t n) { return !find_all(needles, n).empty(); }

  // PDQMatcher::find: f(row, needle) -> std::optional<R> over hits in
  // (row, needle) order; the first engaged result is returned.
  template <class R, class F>
  std::optional<R> find(const Row* needles, size_t n, F f) {
    for (const Hit& h : find_all(needles, n)) {
      std::optional<R> r = f((size_t)h.index, (size_t)h.query);
      if (r) return r;
    }
    return std::nullopt;
  }

  std::vector<Hit> find_all(const Row* needles, size_t n) {
    if (n == 0 || n_rows_ == 0) return {};
    upload_needles(needles, n);
    return detail::match_sorted(d_rows_.get(), n_rows_, d_needles_.get(), (uint32_t)n, threshold_, d_hits_, d_count_,
                                stream_.get());
  }

  // VulkanMatcher::find semantics: the lexicographically smallest (row, needle).
  std::optional<std::pair<size_t, size_t>> find_first(const Row* needles, size_t n) {
    if (n == 0 || n_rows_ == 0) return std::nullopt;
    upload_needles(needles, n);
    detail::check(yume_pdq_amd_match_first(d_rows_.get<uint8_t>(), n_rows_, d_needles_.get<uint8_t>(), (uint32_t)n,
                                           threshold_, d_count_.get<unsigned long long>(), stream_.get()),
                  "match_first");
    unsigned long long key = 0;
    detail::check_hip(hipMemcpyAsync(&key, d_count_.get(), sizeof key, hipMemcpyDeviceToHost, stream_.get()),
                      "hipMemcpyAsync");
    stream_.sync();
    if (key == ~0ull) return std::nullopt;
    return std::make_pair((size_t)(key / n), (size_t)(key % n));
  }

  size_t size() const { return n_rows_; }

 private:
  void upload_needles(const Row* needles, size_t n) {
    if (d_needles_.size() < n * 32) d_needles_ = detail::DeviceBuffer(n * 32);
    detail::check_hip(hipMemcpyAsync(d_needles_.get(), needles, n * 32, hipMemcpyHostToDevice, stream_.get()),
                      "hipMemcpyAsync");
  }
  size_t n_rows_;
  uint32_t threshold_;
  detail::Stream stream_;
  detail::DeviceBuffer d_rows_, d_needles_, d_hits_, d_count_;
};

// Needles fixed at construction, haystack batches passed to scan/find -- the
// shape of CpuMatcher (src/matching/mod.rs:117-436): find visits hits in
// (row-in-batch, needle) order.
class GpuBatchMatcher {
 public:
  static constexpr bool NoFalseNegative = true, NoFalsePositive = true, FindsAllMatches = true;
  static const char* ident() { return "HIP gfx950 matcher (needles resident)"; }

  GpuBatchMatcher(const Row* needles, size_t n_needles, uint32_t threshold)
      : n_needles_(n_needles), threshold_(threshold), d_needles_(n_needles * 32), d_hits_(4096 * sizeof(Hit)),
        d_count_(8) {
    if (n_needles)
      detail::check_hip(
          hipMemcpyAsync(d_needles_.get(), needles, n_needles * 32, hipMemcpyHostToDevice, stream_.get()),
          "hipMemcpyAsync");
    stream_.sync();
  }

  bool scan(const Row* batch, size_t n) { return !find_all(batch, n).empty(); }

  template <class R, class F>
  std::optional<R> find(const Row* batch, size_t n, F f) {
    for (const Hit& h : find_all(batch, n)) {
      std::optional<R> r = f((size_t)h.index, (size_t)h.query);
      if (r) return r;
    }
    return std::nullopt;
  }

  std::vector<Hit> find_all(const Row* batch, size_t n) {
    if (n == 0 || n_needles_ == 0) return {};
    if (d_batch_.size() < n * 32) d_batch_ = detail::DeviceBuffer(n * 32);
    detail::check_hip(hipMemcpyAsync(d_batch_.get(), batch, n * 32, hipMemcpyHostToDevice, stream_.get()),
                      "hipMemcpyAsync");
    return detail::match_sorted(d_batch_.get(), n, d_needles_.get(), (uint32_t)n_needles_, threshold_, d_hits_,
                                d_count_, stream_.get());
  }

 private:
  size_t n_needles_;
  uint32_t threshold_;
  detail::Stream stream_;
  detail::DeviceBuffer d_needles_, d_batch_, d_hits_, d_count_;
};

}  // namespace yume_pdq
