// Serving loop: continuous batching of hash + match requests over a shard set.
// See server.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "engine.h"

namespace ypdq {

struct ServerConfig {
  uint32_t max_batch_images = 64;
  uint64_t hbm_budget_bytes = 0;  // device memory the batch contexts may hold
  uint32_t max_contexts = 4;      // batches in flight at most
  uint32_t threshold = 31;
  uint32_t dihedral_mode = kDihedralAvx2;
  uint64_t hit_capacity = 1 << 20;  // per batch; a batch with more hits is re-matched into a larger buffer
  bool use_graph = true;
  bool busy_poll = true;  // spin (yielding) while batches are in flight instead of sleeping between polls
  // Run the batches' kernels in launch order (each waits for the previous
  // batch's kernels) rather than side by side, and upload each request as it
  // is admitted: the next batch fills and uploads while the current one
  // computes, and starts when it ends (or when it is full).  Concurrent
  // batches would share the GPU and all finish late.
  bool in_order_compute = true;
  // With no tuned choices cached for the shard set's shape, the first
  // context runs the autotuner with this many timed launches per candidate
  // at start (0: use the cached choices or the defaults).
  uint32_t autotune_iterations = 0;
  // > 0: shared circular scan instead of batches.  The shard set is matched
  // chunk by chunk (rows per chunk, rounded up to 1024), round and round; a
  // request's hashes join the active queries at the next chunk and leave
  // after one full turn, so a request never waits for a batch to finish:
  // its latency is one turn of the scan over the requests active with it.
  // Up to max_batch_images x max_contexts images are active at once.
  uint64_t scan_chunk_rows = 0;
};

struct RequestResult {
  int status = 0;  // 0 ok, negative: error
  uint32_t n_images = 0;
  std::vector<uint8_t> hashes;   // n x 8 x 32
  std::vector<float> quality;    // n x 8
  std::vector<float> threshold;  // n x 8
  std::vector<MatchHit> hits;    // query = 8 * image-in-request + dihedral, index global
  double latency_ms = 0;         // submit -> result ready
  uint64_t batch = 0;            // sequence number of the batch that served it
  uint32_t batch_images = 0;     // images in that batch
};

class Server {
 public:
  Server() = default;
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  hipError_t start(std::shared_ptr<ShardSet> shards, const ServerConfig& cfg);
  void stop();
  // Copies n u8 planes (n x 512 x 512) into pinned staging; blocks while the
  // staging ring is full.  Returns the request id (> 0) or 0 on bad input.
  uint64_t submit(const uint8_t* planes, uint32_t n);
  // Zero-copy form: the planes already sit in page-locked host memory that
  // the caller keeps unchanged until the request's result is ready.
  uint64_t submit_pinned(const uint8_t* planes, uint32_t n);
  // Waits up to timeout_ms (< 0: forever) for a request; true when done.
  bool wait(uint64_t id, int timeout_ms, RequestResult* out);
  uint32_t contexts() const { return (uint32_t)ctx_.size(); }
  uint64_t device_bytes() const;
  uint64_t batches() const { return batches_.load(); }
  uint64_t images_served() const { return images_.load(); }

 private:
  struct Request {
    uint64_t id;
    uint32_t n;
    uint64_t slot;       // first staging slot (staged requests)
    const uint8_t* src;  // where the planes are read from (staging or the caller's pinned memory)
    bool staged;
    std::chrono::steady_clock::time_point t_submit;
  };
  struct Slot {  // one batch context and the requests it is running
    std::unique_ptr<BatchContext> ctx;
    hipEvent_t done = nullptr;
    hipEvent_t computed = nullptr;  // the batch's kernels have finished
    hipEvent_t uploaded = nullptr;  // the upload stream has copied the batch's planes
    bool busy = false;     // launched: kernels and readback queued
    bool open = false;     // admitting requests, their uploads queued
    uint64_t batch_no = 0;
    uint32_t n_images = 0;
    std::vector<Request> reqs;
    std::vector<uint32_t> offsets;  // first image of each request in the batch
    // pinned readback: the context's output block up to the first
    // kHitPrefix hits (written by a kernel at the end of the batch), and the
    // rest of the hits
    uint8_t* h_out = nullptr;
    void* d_h_out = nullptr;  // h_out as the GPU addresses it
    MatchHit* h_hits = nullptr;
  };
  void loop();
  bool launch_batch(Slot& s);
  // in-order mode: admit pending requests into the open slot (uploading
  // each); true when the next pending request would not fit
  bool admit(Slot& s, bool* admitted);
  void launch_open(Slot& s);
  void fail_batch(Slot& s);
  // upload_ -> context stream dependency, then the batch, then the readback
  hipError_t enqueue_batch(Slot& s, bool after_last);
  void harvest(Slot& s);
  void free_slots(uint64_t first, uint32_t n);

  // shared circular scan (scan_chunk_rows > 0)
  struct ScanChunk {
    const uint8_t* rows;
    const uint8_t* plane;
    uint64_t n_rows;
    uint64_t index_base;
  };
  struct ScanReq {
    Request r;
    uint32_t first = 0;          // first image slot of the active area
    uint32_t launched = 0;       // chunk launches it is part of
    uint32_t harvested = 0;      // of those, harvested
    uint64_t first_launch = 0;   // sequence number of its first launch
    uint32_t images_at_join = 0;  // active images in that launch
    bool failed = false;
    std::vector<MatchHit> hits;
  };
  struct ScanLaunch {
    uint32_t chunk = 0;
    uint32_t nq = 0;  // queries matched: every active request's, gathered
    uint64_t seq = 0;
    int buf = 0;      // gathered queries, readback and hit buffers
    std::vector<std::shared_ptr<ScanReq>> reqs;  // the requests active in it
    std::vector<uint32_t> at;                    // first gathered image of each
  };
  hipError_t start_scan();
  void stop_scan();
  void loop_scan();
  bool scan_admit();
  hipError_t scan_launch();
  void scan_harvest(ScanLaunch& l);
  void scan_fail_all();
  std::vector<ScanChunk> chunks_;
  uint32_t scan_cap_ = 0;
  std::vector<uint8_t> scan_used_;  // per image slot of the active area
  std::vector<std::shared_ptr<ScanReq>> active_;
  std::deque<ScanLaunch> inflight_;
  uint32_t next_chunk_ = 0;
  uint64_t launch_seq_ = 0;
  hipStream_t scan_stream_ = nullptr;
  hipEvent_t scan_uploaded_ = nullptr;
  hipEvent_t scan_done_[2] = {nullptr, nullptr};
  uint8_t* d_img_ = nullptr;   // active area: planes
  uint8_t* d_q_ = nullptr;     // their hashes, the queries (8 x 32 bytes per image)
  float* d_qual_ = nullptr;    // 8 per image
  float* d_thr_ = nullptr;     // 8 per image
  uint8_t* h_out_ = nullptr;   // page-locked copy of hashes, qualities, thresholds
  void* dh_out_ = nullptr;
  uint8_t* d_lq_[2] = {nullptr, nullptr};  // a launch's gathered queries
  uint4* h_tab_[2] = {nullptr, nullptr};    // page-locked gather table: {first slot, first image, images}
  void* dh_tab_[2] = {nullptr, nullptr};
  MatchHit* d_shits_[2] = {nullptr, nullptr};
  unsigned long long* d_scount_[2] = {nullptr, nullptr};
  uint8_t* h_rb_[2] = {nullptr, nullptr};  // page-locked: count, then the first hits
  void* dh_rb_[2] = {nullptr, nullptr};
  uint64_t scan_bytes_ = 0;

  ServerConfig cfg_;
  std::shared_ptr<ShardSet> shards_;
  std::vector<std::unique_ptr<Slot>> ctx_;
  uint64_t slot_bytes_ = 512ull * 512;
  uint8_t* staging_ = nullptr;  // pinned ring of ring_slots_ images
  uint64_t ring_slots_ = 0;
  std::vector<uint8_t> slot_used_;
  uint64_t ring_head_ = 0;  // next slot to hand out
  hipStream_t upload_ = nullptr;          // every host -> device copy of request planes
  hipEvent_t last_computed_ = nullptr;  // kernels of the last launched batch (in-order compute)
  int open_ = -1;                       // the slot admitting requests (in-order compute)

  std::mutex mu_;
  std::condition_variable cv_submit_;  // worker: new requests
  std::condition_variable cv_space_;   // submitters: staging freed
  std::condition_variable cv_done_;    // waiters: results
  std::deque<Request> pending_;
  std::map<uint64_t, RequestResult> results_;
  uint64_t next_id_ = 1;
  bool stop_ = false;
  std::thread worker_;
  std::atomic<uint64_t> batches_{0}, images_{0};
};

}  // namespace ypdq
