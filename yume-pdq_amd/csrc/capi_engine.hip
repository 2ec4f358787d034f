// C ABI of the batching layer (shard sets, batch contexts, autotuner); see
// include/yume_pdq_amd.h.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/yume_pdq_amd.h"
#include "engine.h"
#include "server.h"

struct YumePdqAmdShards {
  std::shared_ptr<ypdq::ShardSet> set;
};

struct YumePdqAmdServer {
  ypdq::Server server;
  std::mutex mu;
  std::map<uint64_t, ypdq::RequestResult> held;  // results handed out, until released
};

struct YumePdqAmdBatcher {
  std::unique_ptr<ypdq::BatchContext> ctx;
  std::vector<ypdq::TuneRecord> tune;
};

namespace {
__global__ void reset_count_kernel(unsigned long long* p) { *p = 0ull; }

int status(hipError_t e) {
  switch (e) {
    case hipSuccess: return YUME_PDQ_AMD_OK;
    case hipErrorInvalidValue: return YUME_PDQ_AMD_EINVAL;
    case hipErrorOutOfMemory: return YUME_PDQ_AMD_ENOMEM;
    case hipErrorNoDevice:
    case hipErrorInvalidDevice: return YUME_PDQ_AMD_ENODEV;
    default: return YUME_PDQ_AMD_EHIP;
  }
}
}  // namespace

extern "C" {

int yume_pdq_amd_shards_create(YumePdqAmdShards** out) {
  if (!out) return YUME_PDQ_AMD_EINVAL;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return YUME_PDQ_AMD_ENODEV;
  YumePdqAmdShards* s = new (std::nothrow) YumePdqAmdShards();
  if (!s) return YUME_PDQ_AMD_ENOMEM;
  s->set = std::make_shared<ypdq::ShardSet>();
  *out = s;
  return YUME_PDQ_AMD_OK;
}

void yume_pdq_amd_shards_destroy(YumePdqAmdShards* s) { delete s; }

int yume_pdq_amd_shards_add(YumePdqAmdShards* s, const uint8_t* rows, uint64_t n_rows, uint64_t index_base,
                            int rows_on_device, uint32_t* shard_id) {
  if (!s || !rows || n_rows == 0) return YUME_PDQ_AMD_EINVAL;
  return status(s->set->add(rows, n_rows, index_base, !rows_on_device, shard_id));
}

int yume_pdq_amd_shards_load_raw(YumePdqAmdShards* s, const char* path, uint64_t rows_per_shard,
                                 uint64_t index_base, uint32_t* n_added) {
  if (!s) return YUME_PDQ_AMD_EINVAL;
  return status(s->set->load_raw_file(path, rows_per_shard, index_base, n_added));
}

int yume_pdq_amd_shards_replace(YumePdqAmdShards* s, uint32_t shard, const uint64_t* d_local_rows,
                                const uint8_t* d_new_rows, uint32_t n, unsigned long long* d_bad, void* stream) {
  if (!s) return YUME_PDQ_AMD_EINVAL;
  return status(s->set->replace(shard, d_local_rows, d_new_rows, n, d_bad, static_cast<hipStream_t>(stream)));
}

int yume_pdq_amd_shards_info(const YumePdqAmdShards* s, uint32_t* n_shards, uint64_t* total_rows,
                             uint64_t* device_bytes) {
  if (!s) return YUME_PDQ_AMD_EINVAL;
  if (n_shards) *n_shards = s->set->count();
  if (total_rows) *total_rows = s->set->total_rows();
  if (device_bytes) *device_bytes = s->set->device_bytes();
  return YUME_PDQ_AMD_OK;
}

int yume_pdq_amd_shards_match(const YumePdqAmdShards* s, const uint8_t* d_queries, uint32_t n_queries,
                              uint32_t threshold, YumePdqAmdHit* d_hits, uint64_t capacity,
                              unsigned long long* d_count, uint32_t flags, void* stream) {
  if (!s || !d_count || (capacity && !d_hits) || (n_queries && !d_queries)) return YUME_PDQ_AMD_EINVAL;
  if ((flags & YUME_PDQ_AMD_MATCH_VALU) && (flags & YUME_PDQ_AMD_MATCH_MFMA)) return YUME_PDQ_AMD_EINVAL;
  if ((flags & YUME_PDQ_AMD_MATCH_MFMA) && threshold >= 64) return YUME_PDQ_AMD_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!(flags & YUME_PDQ_AMD_MATCH_APPEND)) {
    hipLaunchKernelGGL(reset_count_kernel, dim3(1), dim3(1), 0, st, d_count);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return status(e);
  }
  if (n_queries == 0) return YUME_PDQ_AMD_OK;
  ypdq::MatcherChoice c;
  c.mfma = (flags & YUME_PDQ_AMD_MATCH_MFMA) ||
           (!(flags & YUME_PDQ_AMD_MATCH_VALU) && ypdq::match_mfma_applicable(n_queries, threshold));
  c.qt = 8;
  return status(s->set->launch_match(c, d_queries, nullptr, n_queries, threshold,
                                     reinterpret_cast<ypdq::MatchHit*>(d_hits), capacity, d_count, st));
}

int yume_pdq_amd_batcher_create(YumePdqAmdShards* s, uint32_t max_images, int input_kind, uint32_t dihedral_mode,
                                uint32_t threshold, uint64_t hit_capacity, YumePdqAmdBatcher** out) {
  if (!s || !out) return YUME_PDQ_AMD_EINVAL;
  YumePdqAmdBatcher* b = new (std::nothrow) YumePdqAmdBatcher();
  if (!b) return YUME_PDQ_AMD_ENOMEM;
  b->ctx.reset(new (std::nothrow) ypdq::BatchContext());
  if (!b->ctx) {
    delete b;
    return YUME_PDQ_AMD_ENOMEM;
  }
  const hipError_t e = b->ctx->init(s->set, max_images, input_kind, dihedral_mode, threshold, hit_capacity);
  if (e != hipSuccess) {
    delete b;
    return status(e);
  }
  *out = b;
  return YUME_PDQ_AMD_OK;
}

void yume_pdq_amd_batcher_destroy(YumePdqAmdBatcher* b) { delete b; }

int yume_pdq_amd_batcher_buffers(const YumePdqAmdBatcher* b, YumePdqAmdBatchBuffers* out) {
  if (!b || !out) return YUME_PDQ_AMD_EINVAL;
  const ypdq::BatchContext& c = *b->ctx;
  out->d_input = c.input();
  out->input_stride_bytes = c.plane_stride();
  out->d_hashes = c.hashes();
  out->d_quality = c.quality();
  out->d_threshold = c.threshold_out();
  out->d_hits = static_cast<YumePdqAmdHit*>(c.hits());
  out->hit_capacity = c.hit_capacity();
  out->d_count = c.count();
  out->stream = c.stream();
  out->max_images = c.max_images();
  out->device_bytes = c.device_bytes();
  return YUME_PDQ_AMD_OK;
}

int yume_pdq_amd_batcher_run(YumePdqAmdBatcher* b, uint32_t n_images, int use_graph) {
  if (!b) return YUME_PDQ_AMD_EINVAL;
  return status(b->ctx->run(n_images, use_graph != 0));
}

int yume_pdq_amd_batcher_autotune(YumePdqAmdBatcher* b, int iterations) {
  if (!b) return YUME_PDQ_AMD_EINVAL;
  b->tune.clear();
  const hipError_t e = b->ctx->autotune(iterations, &b->tune);
  if (e == hipSuccess) return status(hipStreamSynchronize(b->ctx->stream()));
  return status(e);
}

int yume_pdq_amd_batcher_tune_records(const YumePdqAmdBatcher* b, YumePdqAmdTuneRecord* out, uint32_t cap,
                                      uint32_t* n) {
  if (!b || !n) return YUME_PDQ_AMD_EINVAL;
  *n = (uint32_t)b->tune.size();
  for (uint32_t i = 0; i < cap && i < b->tune.size(); i++) {
    out[i].n_queries = b->tune[i].n_queries;
    out[i].path = b->tune[i].choice.mfma ? YUME_PDQ_AMD_MATCH_MFMA : YUME_PDQ_AMD_MATCH_VALU;
    out[i].query_tiles = b->tune[i].choice.mfma ? (uint32_t)b->tune[i].choice.qt : 0;
    out[i].ms = b->tune[i].ms;
  }
  return YUME_PDQ_AMD_OK;
}

int yume_pdq_amd_batcher_choice(const YumePdqAmdBatcher* b, uint32_t n_queries, uint32_t* path,
                                uint32_t* query_tiles) {
  if (!b || !path || !query_tiles) return YUME_PDQ_AMD_EINVAL;
  const ypdq::MatcherChoice c = b->ctx->choice_for(n_queries);
  *path = c.mfma ? YUME_PDQ_AMD_MATCH_MFMA : YUME_PDQ_AMD_MATCH_VALU;
  *query_tiles = c.mfma ? (uint32_t)c.qt : 0;
  return YUME_PDQ_AMD_OK;
}

int yume_pdq_amd_tune_save(const char* path) {
  if (!path) return YUME_PDQ_AMD_EINVAL;
  const int n = ypdq::tune_save(path);
  return n < 0 ? YUME_PDQ_AMD_EIO : n;
}

int yume_pdq_amd_tune_load(const char* path) {
  if (!path) return YUME_PDQ_AMD_EINVAL;
  const int n = ypdq::tune_load(path);
  return n < 0 ? YUME_PDQ_AMD_EINVAL : n;
}

int yume_pdq_amd_tune_clear(void) { return (int)ypdq::tune_clear(); }

int yume_pdq_amd_server_create(YumePdqAmdShards* s, const YumePdqAmdServerConfig* cfg, YumePdqAmdServer** out) {
  if (!s || !cfg || !out) return YUME_PDQ_AMD_EINVAL;
  YumePdqAmdServer* v = new (std::nothrow) YumePdqAmdServer();
  if (!v) return YUME_PDQ_AMD_ENOMEM;
  ypdq::ServerConfig c;
  c.max_batch_images = cfg->max_batch_images;
  c.hbm_budget_bytes = cfg->hbm_budget_bytes;
  c.max_contexts = cfg->max_contexts;
  c.threshold = cfg->threshold;
  c.dihedral_mode = cfg->dihedral_mode;
  c.hit_capacity = cfg->hit_capacity;
  c.use_graph = cfg->use_graph != 0;
  c.busy_poll = cfg->busy_poll != 0;
  c.in_order_compute = cfg->in_order_compute != 0;
  c.autotune_iterations = cfg->autotune_iterations;
  c.scan_chunk_rows = cfg->scan_chunk_rows;
  const hipError_t e = v->server.start(s->set, c);
  if (e != hipSuccess) {
    delete v;
    return status(e);
  }
  *out = v;
  return YUME_PDQ_AMD_OK;
}

void yume_pdq_amd_server_destroy(YumePdqAmdServer* v) { delete v; }

int yume_pdq_amd_server_submit(YumePdqAmdServer* v, const uint8_t* planes, uint32_t n_images, uint64_t* id) {
  if (!v || !planes || !id) return YUME_PDQ_AMD_EINVAL;
  *id = v->server.submit(planes, n_images);
  return *id ? YUME_PDQ_AMD_OK : YUME_PDQ_AMD_EINVAL;
}

int yume_pdq_amd_server_submit_pinned(YumePdqAmdServer* v, const uint8_t* planes, uint32_t n_images, uint64_t* id) {
  if (!v || !planes || !id) return YUME_PDQ_AMD_EINVAL;
  *id = v->server.submit_pinned(planes, n_images);
  return *id ? YUME_PDQ_AMD_OK : YUME_PDQ_AMD_EINVAL;
}

int yume_pdq_amd_server_wait(YumePdqAmdServer* v, uint64_t id, int timeout_ms, YumePdqAmdResult* out) {
  if (!v || !out) return YUME_PDQ_AMD_EINVAL;
  ypdq::RequestResult r;
  if (!v->server.wait(id, timeout_ms, &r)) return YUME_PDQ_AMD_ETIMEOUT;
  std::lock_guard<std::mutex> g(v->mu);
  ypdq::RequestResult& h = v->held[id];
  h = std::move(r);
  out->status = h.status;
  out->n_images = h.n_images;
  out->hashes = h.hashes.data();
  out->quality = h.quality.data();
  out->threshold = h.threshold.data();
  out->hits = reinterpret_cast<const YumePdqAmdHit*>(h.hits.data());
  out->n_hits = h.hits.size();
  out->latency_ms = h.latency_ms;
  out->batch = h.batch;
  out->batch_images = h.batch_images;
  return YUME_PDQ_AMD_OK;
}

int yume_pdq_amd_server_release(YumePdqAmdServer* v, uint64_t id) {
  if (!v) return YUME_PDQ_AMD_EINVAL;
  std::lock_guard<std::mutex> g(v->mu);
  return v->held.erase(id) ? YUME_PDQ_AMD_OK : YUME_PDQ_AMD_EINVAL;
}

int yume_pdq_amd_server_stats(const YumePdqAmdServer* v, YumePdqAmdServerStats* out) {
  if (!v || !out) return YUME_PDQ_AMD_EINVAL;
  out->contexts = v->server.contexts();
  out->device_bytes = v->server.device_bytes();
  out->batches = v->server.batches();
  out->images = v->server.images_served();
  return YUME_PDQ_AMD_OK;
}

}  // extern "C"
