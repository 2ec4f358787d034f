// Serving loop, shared circular scan (ServerConfig::scan_chunk_rows > 0).
//
// The batch form makes a request that arrives while a batch computes wait
// for that batch and then for its own: two full passes over the database.
// Here the database is cut into chunks and matched chunk after chunk, round
// and round, against the hashes of every active request at once.  A request
// is hashed into a free slot of the active area as soon as it arrives and
// joins at the next chunk; after one full turn over the chunks it has met
// every row once and leaves.  Its latency is one turn of the scan (over the
// requests active with it) plus its hashing, whenever it arrives.
//
// One worker thread owns every GPU call.  Chunk launches run on one stream,
// two in flight: the next chunk's launch is queued before the previous one
// is harvested.  Each launch gathers the hashes of the requests that have
// not met the chunk yet into one contiguous query array (the active area
// has holes as requests come and go), runs the single-database matcher over
// the chunk (rows, prefix plane, global index base) and stores the hit
// count and the first hits into page-locked memory.  A launch records which
// request each gathered query belongs to.  Slots are freed only when their
// request's last launch has been harvested, after which no queued launch
// reads them.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "../../include/yume_pdq_amd.h"
#include "pdq_device.h"
#include "server.h"

namespace ypdq {

namespace {
// hits stored to the host by the readback kernel after each launch; a launch
// with more takes a synchronous copy for the rest
constexpr uint64_t kScanHitPrefix = 65536;

// dst[0] = the count, dst[1..] = the first min(count, max_hits) hits
__global__ void scan_readback_kernel(const uint4* __restrict__ hits, const unsigned long long* __restrict__ count,
                                     uint4* __restrict__ dst, uint64_t max_hits) {
  const unsigned long long c = *count;
  const uint64_t n = c < max_hits ? c : max_hits;
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 == 0) dst[0] = make_uint4((uint32_t)c, (uint32_t)(c >> 32), 0u, 0u);
  for (uint64_t i = i0; i < n; i += (uint64_t)gridDim.x * blockDim.x) dst[1 + i] = hits[i];
}

// Block b copies the hashes of table entry b (first slot, first gathered
// image, images; 16 16-byte words per image) from the active area to the
// launch's query array.
__global__ void scan_gather_kernel(const uint4* __restrict__ area, uint4* __restrict__ dst,
                                   const uint4* __restrict__ tab) {
  const uint4 t = tab[blockIdx.x];
  const uint32_t n16 = t.z * 16;
  for (uint32_t i = threadIdx.x; i < n16; i += blockDim.x) dst[t.y * 16 + i] = area[t.x * 16 + i];
}

// dst[i] = src[i] for n 16-byte words (device to page-locked host memory)
__global__ void scan_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

constexpr uint64_t kPlane = 512ull * 512;  // bytes of a u8 plane
constexpr uint64_t kHashBytes = 256;       // 8 hashes of 32 bytes per image
constexpr uint64_t kOutPerImage = kHashBytes + 2 * 8 * sizeof(float);
}  // namespace

hipError_t Server::start_scan() {
  hipError_t e;
  const uint64_t chunk_rows = (cfg_.scan_chunk_rows + 1023) / 1024 * 1024;
  for (uint32_t i = 0; i < shards_->count(); i++) {
    const uint64_t n = shards_->n_rows(i);
    for (uint64_t o = 0; o < n; o += chunk_rows)
      chunks_.push_back(ScanChunk{shards_->rows(i) + o * 32, shards_->plane(i) + o * 16,
                                  std::min<uint64_t>(chunk_rows, n - o), shards_->index_base(i) + o});
  }
  if (chunks_.empty()) return hipErrorInvalidValue;
  scan_cap_ = std::min<uint32_t>(cfg_.max_batch_images * std::min<uint32_t>(cfg_.max_contexts, 16), 4096);
  const uint64_t hit_bytes = cfg_.hit_capacity * sizeof(MatchHit);
  scan_bytes_ = scan_cap_ * (kPlane + kOutPerImage + 2 * kHashBytes) + 2 * (hit_bytes + 256);
  if (scan_bytes_ > cfg_.hbm_budget_bytes) return hipErrorOutOfMemory;
  if ((e = hipMalloc(&d_img_, scan_cap_ * kPlane)) != hipSuccess) return e;
  if ((e = hipMalloc(&d_q_, scan_cap_ * kHashBytes)) != hipSuccess) return e;
  if ((e = hipMalloc(&d_qual_, scan_cap_ * 8 * sizeof(float))) != hipSuccess) return e;
  if ((e = hipMalloc(&d_thr_, scan_cap_ * 8 * sizeof(float))) != hipSuccess) return e;
  if ((e = hipHostMalloc(reinterpret_cast<void**>(&h_out_), scan_cap_ * kOutPerImage, hipHostMallocDefault)) !=
          hipSuccess ||
      (e = hipHostGetDevicePointer(&dh_out_, h_out_, 0)) != hipSuccess)
    return e;
  for (int b = 0; b < 2; b++) {
    if ((e = hipMalloc(&d_lq_[b], scan_cap_ * kHashBytes)) != hipSuccess) return e;
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&h_tab_[b]), scan_cap_ * sizeof(uint4), hipHostMallocDefault)) !=
            hipSuccess ||
        (e = hipHostGetDevicePointer(&dh_tab_[b], h_tab_[b], 0)) != hipSuccess)
      return e;
    if ((e = hipMalloc(&d_shits_[b], hit_bytes)) != hipSuccess) return e;
    if ((e = hipMalloc(&d_scount_[b], 256)) != hipSuccess) return e;
    const uint64_t rb = (1 + std::min<uint64_t>(kScanHitPrefix, cfg_.hit_capacity)) * 16;
    if ((e = hipHostMalloc(reinterpret_cast<void**>(&h_rb_[b]), rb, hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostGetDevicePointer(&dh_rb_[b], h_rb_[b], 0)) != hipSuccess)
      return e;
    if ((e = hipEventCreateWithFlags(&scan_done_[b], hipEventDisableTiming)) != hipSuccess) return e;
  }
  if ((e = hipStreamCreateWithFlags(&scan_stream_, hipStreamNonBlocking)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&scan_uploaded_, hipEventDisableTiming)) != hipSuccess) return e;
  scan_used_.assign(scan_cap_, 0);
  return hipSuccess;
}

void Server::stop_scan() {
  if (scan_stream_) (void)hipStreamSynchronize(scan_stream_);
  void* d[] = {d_img_, d_q_, d_qual_, d_thr_, d_lq_[0], d_lq_[1], d_shits_[0], d_shits_[1], d_scount_[0], d_scount_[1]};
  for (void* p : d)
    if (p) (void)hipFree(p);
  void* h[] = {h_out_, h_rb_[0], h_rb_[1], h_tab_[0], h_tab_[1]};
  for (void* p : h)
    if (p) (void)hipHostFree(p);
  for (hipEvent_t ev : {scan_done_[0], scan_done_[1], scan_uploaded_})
    if (ev) (void)hipEventDestroy(ev);
  if (scan_stream_) (void)hipStreamDestroy(scan_stream_);
  d_img_ = d_q_ = nullptr;
  d_qual_ = d_thr_ = nullptr;
  h_out_ = nullptr;
  dh_out_ = nullptr;
  for (int b = 0; b < 2; b++) {
    d_lq_[b] = nullptr;
    h_tab_[b] = nullptr;
    dh_tab_[b] = nullptr;
    d_shits_[b] = nullptr;
    d_scount_[b] = nullptr;
    h_rb_[b] = nullptr;
    dh_rb_[b] = nullptr;
    scan_done_[b] = nullptr;
  }
  scan_uploaded_ = nullptr;
  scan_stream_ = nullptr;
  chunks_.clear();
  active_.clear();
  inflight_.clear();
  scan_used_.clear();
}

// Hash pending requests into free slots of the active area (first fit), in
// arrival order; they join the scan at its next chunk.  True if any joined.
bool Server::scan_admit() {
  bool any = false;
  for (;;) {
    Request r;
    uint32_t first = 0;
    {
      std::lock_guard<std::mutex> g(mu_);
      if (pending_.empty()) break;
      r = pending_.front();
      bool found = false;
      for (uint32_t s = 0; s + r.n <= scan_cap_ && !found;) {
        uint32_t k = 0;
        while (k < r.n && !scan_used_[s + k]) k++;
        if (k == r.n) {
          first = s;
          found = true;
        } else {
          s += k + 1;
        }
      }
      if (!found) break;  // waits for slots: arrival order is kept
      pending_.pop_front();
    }
    for (uint32_t k = 0; k < r.n; k++) scan_used_[first + k] = 1;
    HashOutputs o;
    memset(&o, 0, sizeof o);
    o.hashes = d_q_ + (uint64_t)first * kHashBytes;
    o.quality = d_qual_ + (uint64_t)first * 8;
    o.threshold = d_thr_ + (uint64_t)first * 8;
    hipError_t e = hipMemcpyAsync(d_img_ + (uint64_t)first * kPlane, r.src, (uint64_t)r.n * kPlane,
                                  hipMemcpyHostToDevice, upload_);
    if (e == hipSuccess) e = hipEventRecord(scan_uploaded_, upload_);
    if (e == hipSuccess) e = hipStreamWaitEvent(scan_stream_, scan_uploaded_, 0);
    if (e == hipSuccess)
      e = launch_hash(d_img_ + (uint64_t)first * kPlane, kInputU8, kPlane, r.n, cfg_.dihedral_mode, o, scan_stream_);
    // the hashes, qualities and thresholds into page-locked memory, read
    // when the request leaves (its launches follow on the same stream)
    uint8_t* h = static_cast<uint8_t*>(dh_out_);
    const struct {
      const void* src;
      uint64_t off, bytes;
    } parts[3] = {{o.hashes, 0, kHashBytes}, {o.quality, scan_cap_ * kHashBytes, 32}, {o.threshold, scan_cap_ * (kHashBytes + 32), 32}};
    for (const auto& p : parts) {
      if (e != hipSuccess) break;
      hipLaunchKernelGGL(scan_copy_kernel, dim3(8), dim3(256), 0, scan_stream_, static_cast<const uint4*>(p.src),
                         reinterpret_cast<uint4*>(h + p.off + (uint64_t)first * p.bytes), (uint64_t)r.n * p.bytes / 16);
      e = hipGetLastError();
    }
    auto q = std::make_shared<ScanReq>();
    q->r = r;
    q->first = first;
    active_.push_back(q);
    any = true;
    if (e != hipSuccess) {
      scan_fail_all();
      return true;
    }
  }
  return any;
}

// The next chunk against every request that has not met it yet.
hipError_t Server::scan_launch() {
  ScanLaunch l;
  l.chunk = next_chunk_;
  l.seq = ++launch_seq_;
  l.buf = (int)(l.seq & 1);
  const uint32_t n_chunks = (uint32_t)chunks_.size();
  // the gather table (its buffer's previous launch has been harvested)
  uint32_t images = 0;
  for (auto& q : active_)
    if (q->launched < n_chunks) {
      h_tab_[l.buf][l.reqs.size()] = make_uint4(q->first, images, q->r.n, 0u);
      l.reqs.push_back(q);
      l.at.push_back(images);
      images += q->r.n;
    }
  for (auto& q : l.reqs) {
    if (q->launched++ == 0) {
      q->first_launch = l.seq;
      q->images_at_join = images;
    }
  }
  l.nq = images * 8;
  hipLaunchKernelGGL(scan_gather_kernel, dim3((uint32_t)l.reqs.size()), dim3(256), 0, scan_stream_,
                     reinterpret_cast<const uint4*>(d_q_), reinterpret_cast<uint4*>(d_lq_[l.buf]),
                     static_cast<const uint4*>(dh_tab_[l.buf]));
  if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
  const ScanChunk& c = chunks_[l.chunk];
  const int rc = yume_pdq_amd_match_with_plane(c.rows, c.plane, c.n_rows, d_lq_[l.buf], l.nq, cfg_.threshold,
                                               reinterpret_cast<YumePdqAmdHit*>(d_shits_[l.buf]), cfg_.hit_capacity,
                                               d_scount_[l.buf], c.index_base, 0, scan_stream_);
  if (rc != YUME_PDQ_AMD_OK) return hipErrorLaunchFailure;
  hipLaunchKernelGGL(scan_readback_kernel, dim3(16), dim3(256), 0, scan_stream_,
                     reinterpret_cast<const uint4*>(d_shits_[l.buf]), d_scount_[l.buf], static_cast<uint4*>(dh_rb_[l.buf]),
                     std::min<uint64_t>(kScanHitPrefix, cfg_.hit_capacity));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipEventRecord(scan_done_[l.buf], scan_stream_);
  if (e != hipSuccess) return e;
  next_chunk_ = (next_chunk_ + 1) % n_chunks;
  batches_++;
  inflight_.push_back(std::move(l));
  return hipSuccess;
}

// Hand a finished launch's hits to the requests active in it; a request that
// has now met every chunk leaves with its results.
void Server::scan_harvest(ScanLaunch& l) {
  const uint64_t total = *reinterpret_cast<const unsigned long long*>(h_rb_[l.buf]);
  const uint64_t prefix = std::min<uint64_t>(kScanHitPrefix, cfg_.hit_capacity);
  const MatchHit* hits = reinterpret_cast<const MatchHit*>(h_rb_[l.buf] + 16);
  std::vector<MatchHit> more;
  hipError_t e = hipSuccess;
  if (total > prefix && total <= cfg_.hit_capacity) {
    // the hits past the prefix (this launch's buffer: the launch in flight
    // uses the other one)
    more.resize(total);
    memcpy(more.data(), hits, prefix * sizeof(MatchHit));
    e = hipMemcpyAsync(more.data() + prefix, d_shits_[l.buf] + prefix, (total - prefix) * sizeof(MatchHit),
                       hipMemcpyDeviceToHost, upload_);
    if (e == hipSuccess) e = hipStreamSynchronize(upload_);
    hits = more.data();
  } else if (total > cfg_.hit_capacity) {
    // the chunk matched again into a buffer that holds every hit, from the
    // launch's gathered queries (the next launch gathers into the other
    // buffer)
    const ScanChunk& c = chunks_[l.chunk];
    MatchHit* d_big = nullptr;
    unsigned long long* d_cnt = nullptr;
    e = hipMalloc(&d_big, total * sizeof(MatchHit));
    if (e == hipSuccess) e = hipMalloc(&d_cnt, 8);
    if (e == hipSuccess &&
        yume_pdq_amd_match_with_plane(c.rows, c.plane, c.n_rows, d_lq_[l.buf], l.nq, cfg_.threshold,
                                      reinterpret_cast<YumePdqAmdHit*>(d_big), total, d_cnt, c.index_base, 0,
                                      upload_) != YUME_PDQ_AMD_OK)
      e = hipErrorLaunchFailure;
    if (e == hipSuccess) {
      more.resize(total);
      e = hipMemcpyAsync(more.data(), d_big, total * sizeof(MatchHit), hipMemcpyDeviceToHost, upload_);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(upload_);
    if (d_big) (void)hipFree(d_big);
    if (d_cnt) (void)hipFree(d_cnt);
    hits = more.data();
  }
  if (e == hipSuccess) {
    for (uint64_t h = 0; h < total; h++) {
      // the request whose gathered images hold the query's image
      const uint32_t img = hits[h].query / 8;
      const size_t i = (size_t)(std::upper_bound(l.at.begin(), l.at.end(), img) - l.at.begin()) - 1;
      MatchHit x = hits[h];
      x.query -= 8 * l.at[i];
      l.reqs[i]->hits.push_back(x);
    }
  }
  const uint32_t n_chunks = (uint32_t)chunks_.size();
  const auto t_done = std::chrono::steady_clock::now();
  std::vector<std::pair<uint64_t, RequestResult>> done;
  for (auto& qp : l.reqs) {
    ScanReq& q = *qp;
    q.failed = q.failed || e != hipSuccess;
    if (++q.harvested < n_chunks) continue;
    RequestResult rr;
    rr.status = q.failed ? -2 : 0;
    rr.n_images = q.r.n;
    const uint8_t* hh = h_out_ + (uint64_t)q.first * kHashBytes;
    const float* hq = reinterpret_cast<const float*>(h_out_ + scan_cap_ * kHashBytes) + (uint64_t)q.first * 8;
    const float* ht = reinterpret_cast<const float*>(h_out_ + scan_cap_ * (kHashBytes + 32)) + (uint64_t)q.first * 8;
    rr.hashes.assign(hh, hh + (uint64_t)q.r.n * kHashBytes);
    rr.quality.assign(hq, hq + (uint64_t)q.r.n * 8);
    rr.threshold.assign(ht, ht + (uint64_t)q.r.n * 8);
    rr.hits = std::move(q.hits);
    std::sort(rr.hits.begin(), rr.hits.end(), [](const MatchHit& a, const MatchHit& b) {
      return a.index != b.index ? a.index < b.index : a.query < b.query;
    });
    rr.latency_ms = std::chrono::duration<double, std::milli>(t_done - q.r.t_submit).count();
    rr.batch = q.first_launch;
    rr.batch_images = q.images_at_join;
    for (uint32_t k = 0; k < q.r.n; k++) scan_used_[q.first + k] = 0;
    images_ += q.r.n;
    done.emplace_back(q.r.id, std::move(rr));
    active_.erase(std::find(active_.begin(), active_.end(), qp));
    std::lock_guard<std::mutex> g(mu_);
    if (q.r.staged) free_slots(q.r.slot, q.r.n);
  }
  if (!done.empty()) {
    std::lock_guard<std::mutex> g(mu_);
    for (auto& d : done) results_[d.first] = std::move(d.second);
    cv_done_.notify_all();
    cv_space_.notify_all();
  }
}

// After a failed call: every active request ends with an error.
void Server::scan_fail_all() {
  (void)hipStreamSynchronize(upload_);
  (void)hipStreamSynchronize(scan_stream_);
  inflight_.clear();
  std::lock_guard<std::mutex> g(mu_);
  for (auto& q : active_) {
    RequestResult rr;
    rr.status = -2;
    rr.n_images = q->r.n;
    results_[q->r.id] = std::move(rr);
    if (q->r.staged) free_slots(q->r.slot, q->r.n);
    for (uint32_t k = 0; k < q->r.n; k++) scan_used_[q->first + k] = 0;
  }
  active_.clear();
  cv_done_.notify_all();
  cv_space_.notify_all();
}

void Server::loop_scan() {
  const uint32_t n_chunks = (uint32_t)chunks_.size();
  for (;;) {
    bool progress = false;
    while (!inflight_.empty() && hipEventQuery(scan_done_[inflight_.front().buf]) == hipSuccess) {
      scan_harvest(inflight_.front());
      inflight_.pop_front();
      progress = true;
    }
    if (scan_admit()) progress = true;
    bool need = false;
    for (auto& q : active_) need = need || q->launched < n_chunks;
    if (need && inflight_.size() < 2) {
      if (scan_launch() != hipSuccess)
        scan_fail_all();
      progress = true;
    }
    if (progress) continue;
    std::unique_lock<std::mutex> g(mu_);
    if (stop_ && pending_.empty() && inflight_.empty() && active_.empty()) break;
    if (!inflight_.empty() && cfg_.busy_poll) {
      g.unlock();
      std::this_thread::yield();  // poll the launches in flight
    } else if (!inflight_.empty()) {
      cv_submit_.wait_for(g, std::chrono::microseconds(20));
    } else {
      cv_submit_.wait_for(g, std::chrono::milliseconds(5), [&] { return stop_ || !pending_.empty(); });
    }
  }
}

}  // namespace ypdq
