// Batching layer: database shards resident in HBM, and batch contexts that
// run "hash a batch of images -> match its 8n dihedral hashes against every
// shard" as one HIP-graph launch.
//
//   ShardSet      any number of shards of any size (rows + prefix plane in
//                 HBM, global indices from each shard's index_base), a raw
//                 N x 32-byte file loader, in-place row replacement.  The
//                 reference keeps its database as an in-memory &[[u8; 32]]
//                 (src/matching/vulkan.rs:40-64) and replaces entries by
//                 substitution (examples/end_to_end.rs:52-53).
//   BatchContext  fixed device buffers for up to max_images images, and one
//                 captured graph per matcher choice: the hash kernel (image
//                 count read from device memory), the hit-counter reset (a
//                 kernel, not a memset), and a persistent matcher walking the
//                 work items of every shard (query count read from device
//                 memory).  A batch of any size up to max_images replays a
//                 graph after one 8-byte upload of its size.
//   Autotuner     per query-count bucket, times the vector-ALU kernel and the
//                 matrix-core kernel at 2/4/8/16 query tiles per wave on the
//                 context's own shards and keeps the fastest in a cache that
//                 every context of the process reads, keyed by GPU,
//                 threshold and shard-set shape; the cache can be saved to
//                 and loaded from a text file.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <fstream>
#include <map>
#include <sstream>
#include <memory>
#include <mutex>
#include <vector>

#include "engine.h"

namespace ypdq {

namespace {
__global__ void fill_u64_kernel(unsigned long long* p, unsigned long long v) { *p = v; }

// In-place replacement: rows[idx[i]] = new_rows[i], and the same row's plane
// entry.  One thread per replaced row.
__global__ void replace_rows_kernel(uint4* __restrict__ rows, uint4* __restrict__ plane, uint64_t n_rows,
                                    const uint64_t* __restrict__ idx, const uint4* __restrict__ new_rows, uint32_t n,
                                    unsigned long long* __restrict__ bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t r = idx[i];
  if (r >= n_rows) {
    atomicAdd(bad, 1ull);
    return;
  }
  const uint4 lo = new_rows[2 * i], hi = new_rows[2 * i + 1];
  rows[2 * r] = lo;
  rows[2 * r + 1] = hi;
  if (plane) plane[r] = lo;
}

}  // namespace

// ---------------------------------------------------------------------------
// Tuning cache
// ---------------------------------------------------------------------------
namespace {
std::mutex g_tune_mu;
std::map<TuneKey, TuneChoices> g_tune;
std::atomic<uint64_t> g_tune_gen{0};

std::string choice_text(const MatcherChoice& c) { return c.mfma ? "mfma:" + std::to_string(c.qt) : "valu"; }

bool parse_choice(const std::string& t, MatcherChoice* c) {
  if (t == "valu") {
    c->mfma = false;
    return true;
  }
  if (t.rfind("mfma:", 0) != 0) return false;
  const int qt = atoi(t.c_str() + 5);
  if (qt != 2 && qt != 4 && qt != 8 && qt != 16) return false;
  c->mfma = true;
  c->qt = qt;
  return true;
}
}  // namespace

bool tune_lookup(const TuneKey& k, TuneChoices* out) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  auto it = g_tune.find(k);
  if (it == g_tune.end()) return false;
  *out = it->second;
  return true;
}

void tune_store(const TuneKey& k, const TuneChoices& c) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  TuneChoices& d = g_tune[k];
  for (const auto& kv : c) d[kv.first] = kv.second;
  g_tune_gen++;
}

uint64_t tune_generation() { return g_tune_gen.load(); }

size_t tune_clear() {
  std::lock_guard<std::mutex> g(g_tune_mu);
  const size_t n = g_tune.size();
  g_tune.clear();
  g_tune_gen++;
  return n;
}

int tune_save(const char* path) {
  if (!path) return -1;
  std::ostringstream o;
  int lines = 0;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    o << "# yume_pdq_amd matcher tuning: arch cus threshold n_shards rows... : bucket=choice ...\n";
    for (const auto& kv : g_tune) {
      const TuneKey& k = kv.first;
      o << k.arch << ' ' << k.cus << ' ' << k.threshold << ' ' << k.shard_rows.size();
      for (uint64_t r : k.shard_rows) o << ' ' << r;
      o << " :";
      for (const auto& b : kv.second) o << ' ' << b.first << '=' << choice_text(b.second);
      o << '\n';
      lines++;
    }
  }
  std::ofstream f(path, std::ios::trunc);
  if (!f) return -1;
  f << o.str();
  f.close();
  return f ? lines : -1;
}

int tune_load(const char* path) {
  if (!path) return -1;
  std::ifstream f(path);
  if (!f) return -1;
  std::vector<std::pair<TuneKey, TuneChoices>> got;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream in(line);
    TuneKey k;
    size_t n = 0;
    if (!(in >> k.arch >> k.cus >> k.threshold >> n) || n > 4096) return -1;
    k.shard_rows.resize(n);
    for (size_t i = 0; i < n; i++)
      if (!(in >> k.shard_rows[i])) return -1;
    std::string colon;
    if (!(in >> colon) || colon != ":") return -1;
    TuneChoices c;
    std::string item;
    while (in >> item) {
      const size_t eq = item.find('=');
      if (eq == std::string::npos || eq == 0) return -1;
      char* end = nullptr;
      const unsigned long nq = strtoul(item.c_str(), &end, 10);
      if (end != item.c_str() + eq || nq == 0 || nq > (1ul << 20)) return -1;
      MatcherChoice mc;
      if (!parse_choice(item.substr(eq + 1), &mc)) return -1;
      c[(uint32_t)nq] = mc;
    }
    if (c.empty()) return -1;
    got.emplace_back(std::move(k), std::move(c));
  }
  for (const auto& kc : got) tune_store(kc.first, kc.second);
  return (int)got.size();
}

// ---------------------------------------------------------------------------
// ShardSet
// ---------------------------------------------------------------------------
ShardSet::~ShardSet() {
  for (auto& s : shards_) {
    if (s.owns_rows && s.rows) (void)hipFree(s.rows);
    if (s.plane) (void)hipFree(s.plane);
  }
  if (d_views_) (void)hipFree(d_views_);
}

hipError_t ShardSet::add(const uint8_t* rows, uint64_t n_rows, uint64_t index_base, bool host, uint32_t* id) {
  std::lock_guard<std::mutex> g(mu_);
  if (n_rows == 0) return hipErrorInvalidValue;
  Shard s;
  s.n_rows = n_rows;
  s.index_base = index_base;
  hipError_t e;
  if (host) {
    if ((e = hipMalloc(&s.rows, n_rows * 32)) != hipSuccess) return e;
    s.owns_rows = true;
    if ((e = hipMemcpy(s.rows, rows, n_rows * 32, hipMemcpyHostToDevice)) != hipSuccess) {
      (void)hipFree(s.rows);
      return e;
    }
  } else {
    if ((uintptr_t)rows % 16) return hipErrorInvalidValue;
    s.rows = const_cast<uint8_t*>(rows);
  }
  if ((e = hipMalloc(&s.plane, prefix_plane_bytes(n_rows))) != hipSuccess ||
      (e = launch_build_prefix_plane(s.rows, n_rows, s.plane, nullptr)) != hipSuccess ||
      (e = hipDeviceSynchronize()) != hipSuccess) {
    if (s.owns_rows) (void)hipFree(s.rows);
    if (s.plane) (void)hipFree(s.plane);
    return e;
  }
  shards_.push_back(s);
  if (id) *id = (uint32_t)(shards_.size() - 1);
  version_++;
  return rebuild_views_locked();
}

hipError_t ShardSet::load_raw_file(const char* path, uint64_t rows_per_shard, uint64_t index_base, uint32_t* n_added) {
  if (!path || rows_per_shard == 0) return hipErrorInvalidValue;
  struct stat st;
  if (stat(path, &st) != 0 || st.st_size % 32) return hipErrorInvalidValue;
  FILE* f = fopen(path, "rb");
  if (!f) return hipErrorInvalidValue;
  const uint64_t total = (uint64_t)st.st_size / 32;
  // staged through a pinned buffer, one shard at a time
  const uint64_t chunk = std::min<uint64_t>(rows_per_shard, total ? total : 1);
  uint8_t* pinned = nullptr;
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&pinned), chunk * 32, hipHostMallocDefault);
  if (e != hipSuccess) {
    fclose(f);
    return e;
  }
  uint32_t added = 0;
  for (uint64_t first = 0; first < total && e == hipSuccess; first += rows_per_shard) {
    const uint64_t n = std::min<uint64_t>(rows_per_shard, total - first);
    if (fread(pinned, 32, n, f) != n) {
      e = hipErrorInvalidValue;
      break;
    }
    e = add(pinned, n, index_base + first, true, nullptr);
    if (e == hipSuccess) added++;
  }
  (void)hipHostFree(pinned);
  fclose(f);
  if (n_added) *n_added = added;
  return e;
}

hipError_t ShardSet::replace(uint32_t shard, const uint64_t* d_local_rows, const uint8_t* d_new_rows, uint32_t n,
                             unsigned long long* d_bad, hipStream_t stream) {
  std::lock_guard<std::mutex> g(mu_);
  if (shard >= shards_.size()) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  if (!d_local_rows || !d_new_rows || !d_bad || (uintptr_t)d_new_rows % 16) return hipErrorInvalidValue;
  const Shard& s = shards_[shard];
  hipLaunchKernelGGL(replace_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, reinterpret_cast<uint4*>(s.rows),
                     reinterpret_cast<uint4*>(s.plane), s.n_rows, d_local_rows,
                     reinterpret_cast<const uint4*>(d_new_rows), n, d_bad);
  return hipGetLastError();
}

hipError_t ShardSet::rebuild_views_locked() {
  std::vector<ShardView> v(shards_.size());
  uint64_t chunk = 0, rblock = 0;
  const MatchTuning vt = default_plane_tuning(0, 0);
  for (size_t i = 0; i < shards_.size(); i++) {
    v[i].rows = reinterpret_cast<const uint4*>(shards_[i].rows);
    v[i].plane = reinterpret_cast<const uint4*>(shards_[i].plane);
    v[i].n_rows = shards_[i].n_rows;
    v[i].index_base = shards_[i].index_base;
    v[i].chunk_begin = chunk;
    v[i].rblock_begin = rblock;
    chunk += mfma_chunks(shards_[i].n_rows);
    rblock += valu_rblocks(vt, shards_[i].n_rows);
  }
  total_chunks_ = chunk;
  total_rblocks_ = rblock;
  if (d_views_) (void)hipFree(d_views_);
  d_views_ = nullptr;
  if (v.empty()) return hipSuccess;
  hipError_t e = hipMalloc(&d_views_, v.size() * sizeof(ShardView));
  if (e != hipSuccess) return e;
  return hipMemcpy(d_views_, v.data(), v.size() * sizeof(ShardView), hipMemcpyHostToDevice);
}

TuneKey ShardSet::tune_key(uint32_t threshold) const {
  static std::mutex mu;
  static std::map<int, std::string> arch;  // device -> architecture name
  TuneKey k;
  int dev = 0;
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = arch.find(dev);
    if (it == arch.end()) {
      hipDeviceProp_t p;
      std::string a = hipGetDeviceProperties(&p, dev) == hipSuccess ? std::string(p.gcnArchName) : "unknown";
      std::replace(a.begin(), a.end(), ' ', '_');
      if (a.empty()) a = "unknown";
      it = arch.emplace(dev, a).first;
    }
    k.arch = it->second;
  }
  k.cus = cus_;
  k.threshold = threshold;
  for (auto& s : shards_) k.shard_rows.push_back(s.n_rows);
  return k;
}

uint64_t ShardSet::total_rows() const {
  uint64_t t = 0;
  for (auto& s : shards_) t += s.n_rows;
  return t;
}

uint64_t ShardSet::device_bytes() const {
  uint64_t t = 0;
  for (auto& s : shards_) t += (s.owns_rows ? s.n_rows * 32 : 0) + prefix_plane_bytes(s.n_rows);
  return t;
}

hipError_t ShardSet::launch_match(const MatcherChoice& c, const void* d_queries, const uint32_t* d_nq, uint32_t nq_cap,
                                  uint32_t threshold, MatchHit* hits, uint64_t capacity, unsigned long long* counter,
                                  hipStream_t stream, void* record_area) const {
  if (shards_.empty()) return hipSuccess;
  const int cus = cus_;
  if (c.mfma && c.shared && record_area) {
    // resident blocks of the streaming kernel as below, and of the recheck
    // kernel
    uint64_t grid = (uint64_t)cus * (c.qt >= 16 ? 2 : 4);
    const uint64_t groups = (nq_cap + c.qt * 16 - 1) / (c.qt * 16);
    const uint64_t items = total_chunks_ * 4 * ((groups + 3) / 4);
    if (grid > items) grid = (items + 7) / 8 * 8;
    grid = std::max<uint64_t>(8, grid / 8 * 8);
    return launch_match_mfma_shards_shared(c.qt, d_views_, (uint32_t)shards_.size(), total_chunks_, true, d_queries,
                                           d_nq, nq_cap, threshold, hits, capacity, counter, record_area,
                                           (uint32_t)grid, match_resolve_blocks(cus), stream);
  }
  if (c.mfma) {
    // resident blocks: 4 waves per SIMD at 2..8 query tiles, 2 at 16
    uint64_t grid = (uint64_t)cus * (c.qt >= 16 ? 2 : 4);
    const uint64_t groups = (nq_cap + c.qt * 16 - 1) / (c.qt * 16);
    const uint64_t items = total_chunks_ * groups;
    if (grid > items) grid = (items + 7) / 8 * 8;
    grid = std::max<uint64_t>(8, grid / 8 * 8);
    return launch_match_mfma_shards(c.qt, d_views_, (uint32_t)shards_.size(), total_chunks_, true, d_queries, d_nq,
                                    nq_cap, 0, threshold, hits, capacity, counter, (uint32_t)grid, stream);
  }
  const MatchTuning vt = default_plane_tuning(0, 0);
  uint64_t grid = (uint64_t)cus * 16;
  if (grid > total_rblocks_) grid = total_rblocks_;
  return launch_match_valu_shards(vt, d_views_, (uint32_t)shards_.size(), total_rblocks_, true, d_queries, d_nq,
                                  nq_cap, 0xffffffffu, threshold, hits, capacity, counter,
                                  (uint32_t)std::max<uint64_t>(1, grid), stream);
}

// ---------------------------------------------------------------------------
// BatchContext
// ---------------------------------------------------------------------------
BatchContext::~BatchContext() {
  for (auto& kv : graphs_) {
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
  }
  void* bufs[] = {d_input_, d_out_, d_sizes_, d_records_};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (h_sizes_) (void)hipHostFree(h_sizes_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

hipError_t BatchContext::init(std::shared_ptr<ShardSet> shards, uint32_t max_images, int input_kind, uint32_t mode,
                              uint32_t threshold, uint64_t hit_capacity) {
  if (max_images == 0 || (input_kind != kInputF32 && input_kind != kInputU8) || mode > 1 || threshold > 256 ||
      hit_capacity == 0)
    return hipErrorInvalidValue;
  shards_ = std::move(shards);
  max_images_ = max_images;
  input_kind_ = input_kind;
  mode_ = mode;
  threshold_ = threshold;
  hit_capacity_ = hit_capacity;
  plane_stride_ = 512ull * 512 * (input_kind == kInputF32 ? 4 : 1);
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking)) != hipSuccess) return e;
  if ((e = hipMalloc(&d_input_, plane_stride_ * max_images)) != hipSuccess) return e;
  if ((e = hipMalloc(&d_out_, out_bytes(hit_capacity))) != hipSuccess) return e;
  d_count_ = reinterpret_cast<unsigned long long*>(d_out_ + out_offset(kOutCount));
  d_hashes_ = d_out_ + out_offset(kOutHashes);
  d_quality_ = reinterpret_cast<float*>(d_out_ + out_offset(kOutQuality));
  d_threshold_ = reinterpret_cast<float*>(d_out_ + out_offset(kOutThreshold));
  d_hits_ = d_out_ + out_offset(kOutHits);
  if ((e = hipMalloc(&d_sizes_, 2 * sizeof(uint32_t))) != hipSuccess) return e;
  if ((e = hipMalloc(&d_records_, match_record_area_bytes())) != hipSuccess) return e;
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&h_sizes_), 2 * sizeof(uint32_t), hipHostMallocDefault)) !=
      hipSuccess)
    return e;
  // choices come from the tuning cache (refresh_choices); without an entry
  // for the shard set's shape, choice_for's defaults apply
  choices_.clear();
  gen_ = ~0ull;
  return hipSuccess;
}

uint64_t BatchContext::out_offset(OutPart p) const {
  const uint64_t m = max_images_;
  switch (p) {
    case kOutCount: return 0;
    case kOutHashes: return 256;
    case kOutQuality: return 256 + m * 256;
    case kOutThreshold: return 256 + m * 256 + m * 32;
    default: return (256 + m * 256 + m * 64 + 255) / 256 * 256;
  }
}

uint64_t BatchContext::device_bytes() const {
  return plane_stride_ * max_images_ + out_bytes(hit_capacity_) + 16 + match_record_area_bytes();
}

// Rows shared by a block's waves pay off from 4 query groups (one per wave).
static bool shared_pays(const MatcherChoice& c, uint32_t nq) {
  return c.mfma && (nq + c.qt * 16 - 1) / (c.qt * 16) >= 4;
}

void BatchContext::refresh_choices() const {
  const uint64_t gen = tune_generation();
  TuneKey k = shards_->tune_key(threshold_);
  if (gen == gen_ && k == key_) return;
  choices_.clear();
  (void)tune_lookup(k, &choices_);
  key_ = std::move(k);
  gen_ = gen;
}

bool BatchContext::tuned() const {
  refresh_choices();
  return !choices_.empty();
}

// Relative cost per query of the matrix-core matcher at qt query tiles per
// wave (a wave's row loads and expansions are shared by its 16 qt queries):
// 10^8 rows at 512 queries take 3.75 / 2.52 / 2.04 / 2.05 ms at 2 / 4 / 8 / 16.
static double mfma_query_cost(int qt) { return qt >= 8 ? 1.0 : qt == 4 ? 1.25 : 1.8; }

// Queries the matrix-core matcher computes for nq: whole groups of 16 qt.
static uint64_t padded_queries(uint32_t nq, int qt) {
  const uint64_t g = (uint64_t)qt * 16;
  return (nq + g - 1) / g * g;
}

MatcherChoice BatchContext::choice_for(uint32_t nq) const {
  refresh_choices();
  // without tuning: vector ALU below 32 queries, eight query tiles per wave
  // on the matrix cores from 32
  auto it = choices_.upper_bound(nq);
  MatcherChoice c;
  if (it != choices_.begin()) {
    --it;
    c = it->second;
  } else {
    c.mfma = threshold_ < 64 && nq >= 32;
    c.qt = 8;
  }
  if (c.mfma) {
    // a bucket is tuned at its first query count, where every candidate's
    // groups are full; above it, the last group of a wide tile can be mostly
    // padding (nq = 320 at 16 tiles computes 512 queries).  Take the narrower
    // tile when its padded work costs less.
    int best = c.qt;
    double best_cost = (double)padded_queries(nq, c.qt) * mfma_query_cost(c.qt);
    for (int qt = c.qt / 2; qt >= 2; qt /= 2) {
      const double cost = (double)padded_queries(nq, qt) * mfma_query_cost(qt);
      if (cost < best_cost) {
        best = qt;
        best_cost = cost;
      }
    }
    c.qt = best;
  }
  c.shared = shared_pays(c, nq);
  return c;
}

static int choice_key(const MatcherChoice& c) { return c.mfma ? c.qt * 2 + (c.shared ? 1 : 0) : 0; }

hipError_t BatchContext::enqueue_body(const MatcherChoice& c, hipStream_t s) {
  HashOutputs o;
  memset(&o, 0, sizeof o);
  o.hashes = d_hashes_;
  o.quality = d_quality_;
  o.threshold = d_threshold_;
  hipError_t e = launch_hash_device_count(d_input_, input_kind_, plane_stride_, max_images_, d_sizes_, mode_, o, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(1), 0, s, d_count_, 0ull);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return shards_->launch_match(c, d_hashes_, d_sizes_ + 1, max_images_ * 8, threshold_,
                               reinterpret_cast<MatchHit*>(d_hits_), hit_capacity_, d_count_, s, d_records_);
}

hipError_t BatchContext::graph_for(const MatcherChoice& c, hipGraphExec_t* exec) {
  const int key = choice_key(c);
  auto it = graphs_.find(key);
  if (it != graphs_.end() && it->second.version == shards_->version()) {
    *exec = it->second.exec;
    return hipSuccess;
  }
  if (it != graphs_.end()) {
    (void)hipGraphExecDestroy(it->second.exec);
    (void)hipGraphDestroy(it->second.graph);
    graphs_.erase(it);
  }
  Graph g;
  hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t body = enqueue_body(c, stream_);
  e = hipStreamEndCapture(stream_, &g.graph);
  if (body != hipSuccess) {
    if (g.graph) (void)hipGraphDestroy(g.graph);
    return body;
  }
  if (e != hipSuccess) return e;
  if ((e = hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0)) != hipSuccess) {
    (void)hipGraphDestroy(g.graph);
    return e;
  }
  g.version = shards_->version();
  graphs_[key] = g;
  *exec = g.exec;
  return hipSuccess;
}

hipError_t BatchContext::prepare_graphs() {
  std::map<int, MatcherChoice> distinct;
  for (uint32_t n = 1; n <= max_images_; n++) {
    const MatcherChoice c = choice_for(n * 8);
    distinct.emplace(choice_key(c), c);
  }
  for (auto& kv : distinct) {
    hipGraphExec_t exec = nullptr;
    const hipError_t e = graph_for(kv.second, &exec);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t BatchContext::run(uint32_t n_images, bool use_graph, bool wait_previous) {
  if (n_images > max_images_) return hipErrorInvalidValue;
  const MatcherChoice c = choice_for(n_images * 8);
  // the batch size goes to device memory through pinned staging, which the
  // previous batch's upload may still be reading: wait for that batch before
  // rewriting it
  hipError_t e = wait_previous ? hipStreamSynchronize(stream_) : hipSuccess;
  if (e != hipSuccess) return e;
  h_sizes_[0] = n_images;
  h_sizes_[1] = n_images * 8;
  if ((e = hipMemcpyAsync(d_sizes_, h_sizes_, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, stream_)) != hipSuccess)
    return e;
  if (!use_graph) return enqueue_body(c, stream_);
  hipGraphExec_t exec = nullptr;
  if ((e = graph_for(c, &exec)) != hipSuccess) return e;
  return hipGraphLaunch(exec, stream_);
}

hipError_t BatchContext::autotune(int iterations, std::vector<TuneRecord>* records) {
  if (iterations < 1) iterations = 1;
  TuneChoices tuned;
  hipEvent_t a, b;
  hipError_t e;
  if ((e = hipEventCreate(&a)) != hipSuccess) return e;
  if ((e = hipEventCreate(&b)) != hipSuccess) return e;
  // queries: whatever the hash buffer holds; timing does not depend on values
  std::vector<uint32_t> buckets;
  for (uint32_t nq = 8; nq <= max_images_ * 8; nq *= 2) buckets.push_back(nq);
  if (buckets.empty() || buckets.back() != max_images_ * 8) buckets.push_back(max_images_ * 8);
  for (uint32_t nq : buckets) {
    std::vector<MatcherChoice> cands;
    MatcherChoice v;
    v.mfma = false;
    cands.push_back(v);
    if (threshold_ < 64)
      for (int qt : {2, 4, 8, 16}) {
        if (qt > 2 && (uint32_t)(qt / 2) * 16 >= nq) break;  // no more than one tile of padding
        MatcherChoice m;
        m.mfma = true;
        m.qt = qt;
        m.shared = shared_pays(m, nq);
        cands.push_back(m);
      }
    float best = 1e30f;
    MatcherChoice best_c = cands[0];
    for (const MatcherChoice& c : cands) {
      h_sizes_[0] = nq / 8;
      h_sizes_[1] = nq;
      if ((e = hipMemcpyAsync(d_sizes_, h_sizes_, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, stream_)) !=
          hipSuccess)
        return e;
      auto launch = [&]() -> hipError_t {
        hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(1), 0, stream_, d_count_, 0ull);
        return shards_->launch_match(c, d_hashes_, d_sizes_ + 1, nq, threshold_,
                                     reinterpret_cast<MatchHit*>(d_hits_), hit_capacity_, d_count_, stream_,
                                     d_records_);
      };
      if ((e = launch()) != hipSuccess) return e;  // warm-up
      if ((e = hipEventRecord(a, stream_)) != hipSuccess) return e;
      for (int i = 0; i < iterations; i++)
        if ((e = launch()) != hipSuccess) return e;
      if ((e = hipEventRecord(b, stream_)) != hipSuccess) return e;
      if ((e = hipEventSynchronize(b)) != hipSuccess) return e;
      float ms = 0;
      if ((e = hipEventElapsedTime(&ms, a, b)) != hipSuccess) return e;
      ms /= iterations;
      if (records) records->push_back(TuneRecord{nq, c, ms});
      if (ms < best) {
        best = ms;
        best_c = c;
      }
    }
    tuned[nq] = best_c;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  tune_store(shards_->tune_key(threshold_), tuned);
  gen_ = ~0ull;
  return hipSuccess;
}

}  // namespace ypdq
