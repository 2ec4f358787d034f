// Serving loop: continuous batching of hash + match requests.
//
// Requests (1..max_batch_images u8 planes each) are copied by the submitting
// thread into a pinned staging ring and queued.  One worker thread owns every
// GPU call: whenever a batch context is free it admits queued requests in
// arrival order until the batch is full, uploads their planes, replays the
// context's hash -> match graph and records an event; whenever a context's
// event has fired it hands each request its hashes, qualities, thresholds and
// hits.  Contexts run on their own streams, so requests that arrive while
// batches are on the GPU form the next batch and overlap with them.  The
// number of contexts -- batches in flight -- is what the HBM budget allows
// (each holds max_batch_images planes, outputs and a hit buffer), capped at
// max_contexts.
//
// Every upload goes on one stream of its own, and a batch's outputs reach the
// host through a kernel at the end of its graph that stores them into page-
// locked memory: no DMA copy is queued behind a batch's kernels.  Such a
// device-to-host copy on the context's stream, enqueued while batches were in
// flight, blocked the enqueuing thread for 6-10 ms a few times a second, and
// with it every request behind it.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "server.h"

namespace ypdq {

namespace {
// Hits stored to the host with the batch's hashes by the readback kernel,
// which writes only the hits there are (1 MiB of page-locked memory per
// context); a batch with more takes a second, synchronous copy for the rest.
// Near-duplicate traffic reaches tens of hits per image, so a few thousand
// would not cover a full batch.
constexpr uint64_t kHitPrefix = 65536;

// Copies a batch's output block into the slot's page-locked readback buffer:
// the fixed part (count, hashes, qualities, thresholds: fixed16 16-byte
// words) and the hits up to max_hits (a MatchHit is 16 bytes).
__global__ void readback_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t fixed16,
                                const unsigned long long* __restrict__ count, uint64_t max_hits) {
  const unsigned long long c = *count;
  const uint64_t n16 = fixed16 + (c < max_hits ? c : max_hits);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
static_assert(sizeof(MatchHit) == 16, "one hit per 16-byte word");
}  // namespace

Server::~Server() { stop(); }

uint64_t Server::device_bytes() const {
  uint64_t t = scan_bytes_;
  for (auto& s : ctx_) t += s->ctx->device_bytes();
  return t;
}

hipError_t Server::start(std::shared_ptr<ShardSet> shards, const ServerConfig& cfg) {
  if (worker_.joinable()) return hipErrorInvalidValue;
  if (cfg.max_batch_images == 0 || cfg.max_batch_images > 4096 || cfg.threshold > 256 || cfg.dihedral_mode > 1 ||
      cfg.hit_capacity == 0 || cfg.max_contexts == 0)
    return hipErrorInvalidValue;
  cfg_ = cfg;
  shards_ = std::move(shards);
  hipError_t e;
  if (cfg.scan_chunk_rows > 0) {
    if ((e = start_scan()) != hipSuccess) {
      stop_scan();
      return e;
    }
    if ((e = hipStreamCreateWithFlags(&upload_, hipStreamNonBlocking)) != hipSuccess) return e;
    ring_slots_ = 2ull * scan_cap_ + cfg.max_batch_images;
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&staging_), ring_slots_ * slot_bytes_, hipHostMallocDefault)) !=
        hipSuccess)
      return e;
    slot_used_.assign(ring_slots_, 0);
    stop_ = false;
    worker_ = std::thread([this] { loop_scan(); });
    return hipSuccess;
  }
  for (uint32_t k = 0; k < std::min<uint32_t>(cfg.max_contexts, 16); k++) {
    std::unique_ptr<Slot> s(new Slot());
    s->ctx.reset(new BatchContext());
    if ((e = s->ctx->init(shards_, cfg.max_batch_images, kInputU8, cfg.dihedral_mode, cfg.threshold,
                          cfg.hit_capacity)) != hipSuccess)
      return e;
    // admission by HBM budget: a context is kept only if it fits
    if (device_bytes() + s->ctx->device_bytes() > cfg.hbm_budget_bytes) break;
    // tuned choices are shared through the process-wide cache, so the first
    // context tunes for all (before any graph is built on a choice)
    if (k == 0 && cfg.autotune_iterations > 0 && !s->ctx->tuned() &&
        (e = s->ctx->autotune((int)cfg.autotune_iterations, nullptr)) != hipSuccess)
      return e;
    // every graph a batch size can select is built before the first request
    if (cfg.use_graph && (e = s->ctx->prepare_graphs()) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&s->done, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&s->computed, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&s->uploaded, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&s->h_out),
                           s->ctx->out_bytes(std::min<uint64_t>(kHitPrefix, cfg.hit_capacity)),
                           hipHostMallocDefault)) ||
        (e = hipHostGetDevicePointer(&s->d_h_out, s->h_out, 0)) ||
        (e = hipHostMalloc(reinterpret_cast<void**>(&s->h_hits), cfg.hit_capacity * sizeof(MatchHit),
                           hipHostMallocDefault)))
      return e;
    ctx_.push_back(std::move(s));
  }
  if (ctx_.empty()) return hipErrorOutOfMemory;  // the budget does not hold one batch
  if ((e = hipStreamCreateWithFlags(&upload_, hipStreamNonBlocking)) != hipSuccess) return e;
  // room for every context's batch plus as much again queued, so that
  // requests rarely wait for contiguous staging
  ring_slots_ = (uint64_t)(2 * ctx_.size() + 2) * cfg.max_batch_images;
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&staging_), ring_slots_ * slot_bytes_, hipHostMallocDefault)) !=
      hipSuccess)
    return e;
  slot_used_.assign(ring_slots_, 0);
  stop_ = false;
  worker_ = std::thread([this] { loop(); });
  return hipSuccess;
}

void Server::stop() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_ = true;
  }
  cv_submit_.notify_all();
  if (worker_.joinable()) worker_.join();
  stop_scan();
  scan_bytes_ = 0;
  for (auto& s : ctx_) {
    if (s->done) (void)hipEventDestroy(s->done);
    if (s->computed) (void)hipEventDestroy(s->computed);
    if (s->uploaded) (void)hipEventDestroy(s->uploaded);
    void* h[] = {s->h_out, s->h_hits};
    for (void* p : h)
      if (p) (void)hipHostFree(p);
  }
  ctx_.clear();
  if (upload_) (void)hipStreamDestroy(upload_);
  upload_ = nullptr;
  last_computed_ = nullptr;
  open_ = -1;
  if (staging_) (void)hipHostFree(staging_);
  staging_ = nullptr;
}

uint64_t Server::submit(const uint8_t* planes, uint32_t n) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!planes || n == 0 || n > cfg_.max_batch_images || !worker_.joinable()) return 0;
  uint64_t first, id;
  {
    std::unique_lock<std::mutex> g(mu_);
    for (;;) {
      if (stop_) return 0;
      first = ring_head_ + n <= ring_slots_ ? ring_head_ : 0;
      bool free = true;
      for (uint64_t k = first; k < first + n; k++) free = free && !slot_used_[k];
      if (free) break;
      cv_space_.wait(g);
    }
    for (uint64_t k = first; k < first + n; k++) slot_used_[k] = 1;
    ring_head_ = (first + n) % ring_slots_;
    id = next_id_++;
  }
  memcpy(staging_ + first * slot_bytes_, planes, (size_t)n * slot_bytes_);
  {
    std::lock_guard<std::mutex> g(mu_);
    pending_.push_back(Request{id, n, first, staging_ + first * slot_bytes_, true, t0});
  }
  cv_submit_.notify_one();
  return id;
}

uint64_t Server::submit_pinned(const uint8_t* planes, uint32_t n) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!planes || n == 0 || n > cfg_.max_batch_images || !worker_.joinable()) return 0;
  uint64_t id;
  {
    std::lock_guard<std::mutex> g(mu_);
    if (stop_) return 0;
    id = next_id_++;
    pending_.push_back(Request{id, n, 0, planes, false, t0});
  }
  cv_submit_.notify_one();
  return id;
}

bool Server::wait(uint64_t id, int timeout_ms, RequestResult* out) {
  std::unique_lock<std::mutex> g(mu_);
  auto ready = [&] { return results_.count(id) != 0; };
  if (timeout_ms < 0)
    cv_done_.wait(g, ready);
  else if (!cv_done_.wait_for(g, std::chrono::milliseconds(timeout_ms), ready))
    return false;
  auto it = results_.find(id);
  if (out) *out = std::move(it->second);
  results_.erase(it);
  return true;
}

void Server::free_slots(uint64_t first, uint32_t n) {
  for (uint64_t k = first; k < first + n; k++) slot_used_[k] = 0;
}

bool Server::launch_batch(Slot& s) {
  s.reqs.clear();
  s.offsets.clear();
  uint32_t total = 0;
  {
    std::lock_guard<std::mutex> g(mu_);
    while (!pending_.empty() && total + pending_.front().n <= cfg_.max_batch_images) {
      s.offsets.push_back(total);
      total += pending_.front().n;
      s.reqs.push_back(pending_.front());
      pending_.pop_front();
    }
  }
  if (total == 0) return false;
  BatchContext& c = *s.ctx;
  hipError_t e = hipSuccess;
  // one upload per run of requests whose planes are adjacent in host memory
  for (size_t i = 0; i < s.reqs.size() && e == hipSuccess;) {
    size_t j = i + 1;
    uint64_t bytes = (uint64_t)s.reqs[i].n * slot_bytes_;
    while (j < s.reqs.size() && s.reqs[j].src == s.reqs[i].src + bytes) bytes += (uint64_t)s.reqs[j++].n * slot_bytes_;
    e = hipMemcpyAsync(static_cast<uint8_t*>(c.input()) + (uint64_t)s.offsets[i] * slot_bytes_, s.reqs[i].src, bytes,
                       hipMemcpyHostToDevice, upload_);
    i = j;
  }
  s.n_images = total;
  if (e == hipSuccess) e = enqueue_batch(s, /*in_order=*/false);
  s.batch_no = batches_.fetch_add(1) + 1;
  s.busy = true;
  if (e != hipSuccess) fail_batch(s);
  return true;
}

// After the batch's uploads (and, in order, after the last batch's kernels):
// the hash -> match graph, then the readback kernel and the done event.  The
// context's previous batch was harvested, so nothing else is waited for.
hipError_t Server::enqueue_batch(Slot& s, bool in_order) {
  BatchContext& c = *s.ctx;
  hipError_t e = hipEventRecord(s.uploaded, upload_);
  if (e == hipSuccess) e = hipStreamWaitEvent(c.stream(), s.uploaded, 0);
  if (e == hipSuccess && in_order && last_computed_) e = hipStreamWaitEvent(c.stream(), last_computed_, 0);
  if (e == hipSuccess) e = c.run(s.n_images, cfg_.use_graph, /*wait_previous=*/false);
  if (e == hipSuccess) e = hipEventRecord(s.computed, c.stream());
  if (e == hipSuccess && in_order) last_computed_ = s.computed;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(readback_kernel, dim3(64), dim3(256), 0, c.stream(), static_cast<const uint4*>(c.out_block()),
                       static_cast<uint4*>(s.d_h_out), c.out_offset(BatchContext::kOutHits) / 16, c.count(),
                       std::min<uint64_t>(kHitPrefix, c.hit_capacity()));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(s.done, c.stream());
  return e;
}

// Report a failure to every request of the slot's batch and free the slot.
void Server::fail_batch(Slot& s) {
  // no upload may still read the staging ring or the caller's memory
  (void)hipStreamSynchronize(upload_);
  (void)hipStreamSynchronize(s.ctx->stream());
  std::lock_guard<std::mutex> g(mu_);
  for (auto& r : s.reqs) {
    RequestResult rr;
    rr.status = -2;
    rr.n_images = r.n;
    results_[r.id] = std::move(rr);
    if (r.staged) free_slots(r.slot, r.n);
  }
  s.reqs.clear();
  s.offsets.clear();
  s.n_images = 0;
  s.busy = false;
  s.open = false;
  cv_done_.notify_all();
  cv_space_.notify_all();
}

bool Server::admit(Slot& s, bool* admitted) {
  *admitted = false;
  BatchContext& c = *s.ctx;
  for (;;) {
    Request r;
    {
      std::lock_guard<std::mutex> g(mu_);
      if (pending_.empty()) return false;
      if (s.n_images + pending_.front().n > cfg_.max_batch_images) return true;  // full for the next one
      r = pending_.front();
      pending_.pop_front();
    }
    s.offsets.push_back(s.n_images);
    s.reqs.push_back(r);
    const hipError_t e = hipMemcpyAsync(static_cast<uint8_t*>(c.input()) + (uint64_t)s.n_images * slot_bytes_, r.src,
                                        (uint64_t)r.n * slot_bytes_, hipMemcpyHostToDevice, upload_);
    s.n_images += r.n;
    *admitted = true;
    if (e != hipSuccess) {
      fail_batch(s);
      open_ = -1;
      return false;
    }
    if (s.n_images == cfg_.max_batch_images) return true;
  }
}

void Server::launch_open(Slot& s) {
  const hipError_t e = enqueue_batch(s, /*in_order=*/true);
  s.batch_no = batches_.fetch_add(1) + 1;
  s.open = false;
  s.busy = true;
  open_ = -1;
  if (e != hipSuccess) fail_batch(s);
}

void Server::harvest(Slot& s) {
  BatchContext& c = *s.ctx;
  const uint32_t n = s.n_images;
  hipStream_t st = c.stream();
  hipError_t e = hipSuccess;
  const uint8_t* h_hashes = s.h_out + c.out_offset(BatchContext::kOutHashes);
  const float* h_quality = reinterpret_cast<const float*>(s.h_out + c.out_offset(BatchContext::kOutQuality));
  const float* h_threshold = reinterpret_cast<const float*>(s.h_out + c.out_offset(BatchContext::kOutThreshold));
  const uint64_t total = *reinterpret_cast<const unsigned long long*>(s.h_out + c.out_offset(BatchContext::kOutCount));
  const uint64_t prefix = std::min<uint64_t>(kHitPrefix, c.hit_capacity());
  std::vector<MatchHit> overflow;  // hits re-matched when the batch buffer was too small
  const MatchHit* hits = reinterpret_cast<const MatchHit*>(s.h_out + c.out_offset(BatchContext::kOutHits));
  if (total > prefix && total <= c.hit_capacity()) {
    // the hits past the prefix
    memcpy(s.h_hits, hits, prefix * sizeof(MatchHit));
    e = hipMemcpyAsync(s.h_hits + prefix, static_cast<const MatchHit*>(c.hits()) + prefix,
                       (total - prefix) * sizeof(MatchHit), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    hits = s.h_hits;
  } else if (total > c.hit_capacity()) {
    // rerun the match of this batch into a buffer that holds every hit
    MatchHit* d_big = nullptr;
    unsigned long long* d_cnt = nullptr;
    e = hipMalloc(&d_big, total * sizeof(MatchHit));
    if (e == hipSuccess) e = hipMalloc(&d_cnt, 8);
    if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, 8, st);
    if (e == hipSuccess)
      e = shards_->launch_match(c.choice_for(n * 8), c.hashes(), nullptr, n * 8, cfg_.threshold, d_big, total, d_cnt,
                                st);
    if (e == hipSuccess) {
      overflow.resize(total);
      e = hipMemcpyAsync(overflow.data(), d_big, total * sizeof(MatchHit), hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_big) (void)hipFree(d_big);
    if (d_cnt) (void)hipFree(d_cnt);
    hits = overflow.data();
  }
  const auto t_done = std::chrono::steady_clock::now();
  std::vector<RequestResult> rr(s.reqs.size());
  for (size_t i = 0; i < s.reqs.size(); i++) {
    const uint32_t off = s.offsets[i], m = s.reqs[i].n;
    RequestResult& r = rr[i];
    r.status = e == hipSuccess ? 0 : -2;
    r.n_images = m;
    r.hashes.assign(h_hashes + (size_t)off * 256, h_hashes + (size_t)(off + m) * 256);
    r.quality.assign(h_quality + (size_t)off * 8, h_quality + (size_t)(off + m) * 8);
    r.threshold.assign(h_threshold + (size_t)off * 8, h_threshold + (size_t)(off + m) * 8);
    r.latency_ms = std::chrono::duration<double, std::milli>(t_done - s.reqs[i].t_submit).count();
    r.batch = s.batch_no;
    r.batch_images = n;
  }
  if (e == hipSuccess) {
    for (uint64_t h = 0; h < total; h++) {
      const uint32_t img = hits[h].query / 8;
      // the request whose image range holds img
      const size_t i = (size_t)(std::upper_bound(s.offsets.begin(), s.offsets.end(), img) - s.offsets.begin()) - 1;
      MatchHit x = hits[h];
      x.query -= 8 * s.offsets[i];
      rr[i].hits.push_back(x);
    }
    for (auto& r : rr)
      std::sort(r.hits.begin(), r.hits.end(), [](const MatchHit& a, const MatchHit& b) {
        return a.index != b.index ? a.index < b.index : a.query < b.query;
      });
  }
  images_ += n;
  {
    std::lock_guard<std::mutex> g(mu_);
    for (size_t i = 0; i < s.reqs.size(); i++) {
      if (s.reqs[i].staged) free_slots(s.reqs[i].slot, s.reqs[i].n);
      results_[s.reqs[i].id] = std::move(rr[i]);
    }
  }
  s.reqs.clear();
  s.offsets.clear();
  s.n_images = 0;
  s.busy = false;
  cv_done_.notify_all();
  cv_space_.notify_all();
}

void Server::loop() {
  for (;;) {
    bool progress = false, any_busy = false;
    for (auto& sp : ctx_) {
      Slot& s = *sp;
      if (s.busy && hipEventQuery(s.done) == hipSuccess) {
        harvest(s);
        progress = true;
      }
      any_busy = any_busy || s.busy;
    }
    if (cfg_.in_order_compute) {
      // open a free slot for arriving requests
      if (open_ < 0) {
        bool have;
        {
          std::lock_guard<std::mutex> g(mu_);
          have = !pending_.empty();
        }
        for (size_t k = 0; have && k < ctx_.size(); k++)
          if (!ctx_[k]->busy) {
            open_ = (int)k;
            ctx_[k]->open = true;
            break;
          }
      }
      if (open_ >= 0) {
        Slot& s = *ctx_[open_];
        bool admitted = false;
        const bool full = admit(s, &admitted);
        progress = progress || admitted;
        if (open_ >= 0 && s.n_images > 0) {
          // start the batch when the previous batch's kernels are done (the
          // GPU would otherwise idle) or when it cannot take the next request
          const bool idle = !last_computed_ || hipEventQuery(last_computed_) == hipSuccess;
          if (idle || full) {
            launch_open(s);
            progress = true;
            any_busy = true;
          }
        }
      }
    } else {
      for (auto& sp : ctx_) {
        if (sp->busy) continue;
        if (launch_batch(*sp)) {
          progress = true;
          any_busy = true;
        }
      }
    }
    if (progress) continue;
    const bool open_work = open_ >= 0 && ctx_[open_]->n_images > 0;
    std::unique_lock<std::mutex> g(mu_);
    if (stop_ && pending_.empty() && !any_busy && !open_work) break;
    if (!pending_.empty() && !any_busy) continue;  // a context is free: go launch
    if ((any_busy || open_work) && cfg_.busy_poll) {
      g.unlock();
      std::this_thread::yield();  // poll the in-flight batches
    } else if (any_busy || open_work)
      cv_submit_.wait_for(g, std::chrono::microseconds(20));
    else
      cv_submit_.wait_for(g, std::chrono::milliseconds(5), [&] { return stop_ || !pending_.empty(); });
  }
}

}  // namespace ypdq
