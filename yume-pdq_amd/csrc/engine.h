// Batching layer (host side): database shards in HBM and batch contexts that
// replay one captured hash -> match graph per batch.  See engine.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "pdq_device.h"

namespace ypdq {

// Which matcher a batch uses: the vector-ALU scan, or the matrix-core filter
// with qt 16-query tiles per wave, in the per-wave layout or (shared, with a
// record area) with a block's waves sharing their rows.
struct MatcherChoice {
  bool mfma = false;
  int qt = 8;
  bool shared = false;
};

struct TuneRecord {
  uint32_t n_queries;
  MatcherChoice choice;
  float ms;
};

// Autotuner results shared by every batch context of the process: the
// matcher choice per query-count bucket (first query count of the bucket ->
// choice), keyed by the GPU (architecture name and compute units), the
// threshold and the shard set's shape (its shards' row counts, in order).
struct TuneKey {
  std::string arch;
  int cus = 0;
  uint32_t threshold = 0;
  std::vector<uint64_t> shard_rows;
  bool operator<(const TuneKey& o) const {
    return std::tie(arch, cus, threshold, shard_rows) < std::tie(o.arch, o.cus, o.threshold, o.shard_rows);
  }
  bool operator==(const TuneKey& o) const {
    return arch == o.arch && cus == o.cus && threshold == o.threshold && shard_rows == o.shard_rows;
  }
};
using TuneChoices = std::map<uint32_t, MatcherChoice>;

// The cached choices of `k` into *out (false if none).
bool tune_lookup(const TuneKey& k, TuneChoices* out);
// Adds (replacing bucket by bucket) the choices of `k`.
void tune_store(const TuneKey& k, const TuneChoices& c);
// Bumped by every store, load and clear: contexts re-read the cache when it moves.
uint64_t tune_generation();
size_t tune_clear();
// Text file: one line per key, "arch cus threshold n_shards rows... :
// bucket=valu|mfma:qt ...".  save returns the lines written, load the lines
// read (merged into the cache), -1 on an I/O or format error.
int tune_save(const char* path);
int tune_load(const char* path);

class ShardSet {
 public:
  ShardSet() : cus_(0) {
    int dev = 0;
    hipDeviceProp_t p;
    cus_ = (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) ? p.multiProcessorCount
                                                                                             : 256;
  }
  ~ShardSet();
  ShardSet(const ShardSet&) = delete;
  ShardSet& operator=(const ShardSet&) = delete;

  // rows: host (copied into HBM, owned) or device (referenced, must outlive
  // the set; 16-byte aligned).  The prefix plane is built either way.
  hipError_t add(const uint8_t* rows, uint64_t n_rows, uint64_t index_base, bool host, uint32_t* id);
  // A raw file of N x 32-byte rows, split into shards of rows_per_shard.
  hipError_t load_raw_file(const char* path, uint64_t rows_per_shard, uint64_t index_base, uint32_t* n_added);
  // rows[d_local_rows[i]] = d_new_rows[i] (and the prefix plane), on `stream`;
  // out-of-range indices are counted in *d_bad and skipped.
  hipError_t replace(uint32_t shard, const uint64_t* d_local_rows, const uint8_t* d_new_rows, uint32_t n,
                     unsigned long long* d_bad, hipStream_t stream);

  // One persistent launch over every shard.  d_nq: device query count (or
  // null for nq_cap).  A shared-rows choice needs record_area
  // (match_record_area_bytes(), used by nothing else in flight); without it
  // the per-wave layout runs.
  hipError_t launch_match(const MatcherChoice& c, const void* d_queries, const uint32_t* d_nq, uint32_t nq_cap,
                          uint32_t threshold, MatchHit* hits, uint64_t capacity, unsigned long long* counter,
                          hipStream_t stream, void* record_area = nullptr) const;

  uint32_t count() const { return (uint32_t)shards_.size(); }
  // The set's autotuning key at `threshold` on the current device.
  TuneKey tune_key(uint32_t threshold) const;
  uint64_t total_rows() const;
  uint64_t device_bytes() const;
  uint64_t version() const { return version_; }
  const uint8_t* rows(uint32_t i) const { return i < shards_.size() ? shards_[i].rows : nullptr; }
  uint64_t n_rows(uint32_t i) const { return i < shards_.size() ? shards_[i].n_rows : 0; }
  uint64_t index_base(uint32_t i) const { return i < shards_.size() ? shards_[i].index_base : 0; }
  // the shard's prefix plane: 16 bytes per row, padded to whole 16-row tiles
  const uint8_t* plane(uint32_t i) const { return i < shards_.size() ? shards_[i].plane : nullptr; }

 private:
  struct Shard {
    uint8_t* rows = nullptr;
    uint8_t* plane = nullptr;
    uint64_t n_rows = 0;
    uint64_t index_base = 0;
    bool owns_rows = false;
  };
  hipError_t rebuild_views_locked();

  int cus_;
  std::mutex mu_;
  std::vector<Shard> shards_;
  ShardView* d_views_ = nullptr;
  uint64_t total_chunks_ = 0, total_rblocks_ = 0;
  uint64_t version_ = 0;
};

class BatchContext {
 public:
  BatchContext() = default;
  ~BatchContext();
  BatchContext(const BatchContext&) = delete;
  BatchContext& operator=(const BatchContext&) = delete;

  hipError_t init(std::shared_ptr<ShardSet> shards, uint32_t max_images, int input_kind, uint32_t mode,
                  uint32_t threshold, uint64_t hit_capacity);
  // Hash input[0 .. n_images) and match the 8n hashes against every shard,
  // enqueued on stream() (a graph replay unless use_graph is false).  Waits
  // for this context's previous batch first unless the caller knows it is
  // done (wait_previous = false).
  hipError_t run(uint32_t n_images, bool use_graph, bool wait_previous = true);
  // Times the matchers per query-count bucket and stores the fastest in the
  // process-wide tuning cache, under the shard set's current shape.
  hipError_t autotune(int iterations, std::vector<TuneRecord>* records);
  // The matcher for n_queries: the cached choice for the shard set's current
  // shape if there is one, else a default.
  MatcherChoice choice_for(uint32_t n_queries) const;
  // True if the tuning cache holds choices for the shard set's current shape.
  bool tuned() const;
  // Instantiate now the graph of every matcher choice that some batch size
  // up to max_images() selects, so that no batch pays for it later.
  hipError_t prepare_graphs();

  hipStream_t stream() const { return stream_; }
  void* input() const { return d_input_; }
  uint64_t plane_stride() const { return plane_stride_; }
  uint8_t* hashes() const { return d_hashes_; }
  float* quality() const { return d_quality_; }
  float* threshold_out() const { return d_threshold_; }
  void* hits() const { return d_hits_; }
  unsigned long long* count() const { return d_count_; }
  uint32_t max_images() const { return max_images_; }
  uint64_t hit_capacity() const { return hit_capacity_; }
  uint64_t device_bytes() const;
  // The outputs live in one device block -- hit count, hashes, qualities,
  // thresholds, hits, each at out_offset() -- so that one copy can read a
  // batch back: out_bytes(k) covers everything up to the first k hits.
  const void* out_block() const { return d_out_; }
  enum OutPart { kOutCount = 0, kOutHashes, kOutQuality, kOutThreshold, kOutHits };
  uint64_t out_offset(OutPart p) const;
  uint64_t out_bytes(uint64_t hits_prefix) const { return out_offset(kOutHits) + hits_prefix * sizeof(MatchHit); }

 private:
  struct Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t version = 0;
  };
  hipError_t enqueue_body(const MatcherChoice& c, hipStream_t s);
  hipError_t graph_for(const MatcherChoice& c, hipGraphExec_t* exec);
  void refresh_choices() const;

  std::shared_ptr<ShardSet> shards_;
  uint32_t max_images_ = 0;
  int input_kind_ = kInputU8;
  uint32_t mode_ = kDihedralAvx2;
  uint32_t threshold_ = 31;
  uint64_t hit_capacity_ = 0;
  uint64_t plane_stride_ = 0;
  hipStream_t stream_ = nullptr;
  void* d_input_ = nullptr;
  uint8_t* d_out_ = nullptr;  // the output block (see out_offset)
  uint8_t* d_hashes_ = nullptr;
  float* d_quality_ = nullptr;
  float* d_threshold_ = nullptr;
  void* d_hits_ = nullptr;
  unsigned long long* d_count_ = nullptr;
  uint32_t* d_sizes_ = nullptr;  // device {n_images, n_queries}
  uint32_t* h_sizes_ = nullptr;  // pinned staging of the same
  void* d_records_ = nullptr;    // the shared-rows matcher's candidate records
  // the cache's choices for key_, read at tune generation gen_
  mutable TuneChoices choices_;
  mutable TuneKey key_;
  mutable uint64_t gen_ = ~0ull;
  std::map<int, Graph> graphs_;
};

}  // namespace ypdq
