// Stream ingestion: raw 512x512 frames (luma8, rgb8 or rgba8) from a file
// descriptor are hashed on the GPU in batches and written to another file
// descriptor in the output formats of the reference's `pipe` subcommand
// (src/main.rs:1091-1304, record layout src/main.rs:662-773).
//
// Pipeline: a reader thread fills a ring of page-locked batch buffers with
// whole frames; the processing thread uploads a batch (in chunks of 64
// frames, each as soon as the reader has read it), runs the fused hash
// kernel (input conversion fused into its row loader) and reads the hashes,
// qualities and thresholds back, and while the GPU works on batch i + 1 it
// formats and writes batch i.
//
// The page-locked ring, the device buffers, the streams and the events of a
// run are kept for the next run of the same shape (batch size, frame size,
// diagnostic output): pinning a ring of batches of 256 frames costs tens of
// milliseconds, more than hashing a few thousand frames.  At most two idle
// sets stay cached; trim_pipe_cache frees them.
#include <errno.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "pdq_device.h"
#include "pipe.h"

namespace ypdq {

// ---------------------------------------------------------------------------
// Output format
// ---------------------------------------------------------------------------
bool parse_output_format(const char* spec, OutputFormat* f) {
  if (!spec || !f) return false;
  std::string s(spec);
  // The last '+'-separated component may name a record separator
  // (src/main.rs:1106-1133); the format is what precedes it.
  const size_t plus = s.rfind('+');
  const std::string last = plus == std::string::npos ? s : s.substr(plus + 1);
  std::string sep;
  bool has_sep = true;
  if (last == "nl" || last == "NL" || last == "lf" || last == "LF") sep = "\n";
  else if (last == "crlf" || last == "CRLF") sep = "\r\n";
  else if (last == "cr" || last == "CR") sep = "\r";
  else if (last == "nul" || last == "NUL" || last == "null" || last == "NULL") sep = std::string(1, '\0');
  else has_sep = false;
  if (has_sep) s = plus == std::string::npos ? std::string() : s.substr(0, plus);
  OutputFormat o;
  o.separator = sep;
  if (s == "raw" || s == "Raw" || s == "RAW") o.type = OutputFormat::kRaw;
  else if (s == "q+raw" || s == "q+Raw" || s == "q+RAW") o.type = OutputFormat::kRawQuality;
  else if (s == "hex") o.type = OutputFormat::kHex;
  else if (s == "HEX") o.type = OutputFormat::kHex, o.upper = true;
  else if (s == "bin" || s == "BIN") o.type = OutputFormat::kBin;
  else if (s == "q+hex") o.type = OutputFormat::kHexQuality;
  else if (s == "q+HEX") o.type = OutputFormat::kHexQuality, o.upper = true;
  else if (s == "q+bin" || s == "q+BIN") o.type = OutputFormat::kBinQuality;
  else if (s == "diagnostic") o.type = OutputFormat::kDiagnostic;
  else return false;
  *f = o;
  return true;
}

namespace {

// Packed Dihedrals codes of orientations 1..7 in the order visit_dihedrals
// reports them (src/lib.rs:119-133, 391-418).
constexpr uint32_t kOrientationCode[8] = {0x01000001u, 0xFF000001u, 0xFF0000FFu, 0x010000FFu,
                                          0x00FFFF00u, 0x00010100u, 0x00FF0100u, 0x0001FF00u};

// A float with `prec` decimals as Rust's `{:.prec}` prints it: the exact
// value correctly rounded (printf does the same for finite values), "NaN"
// for any NaN, "inf" / "-inf".
void put_fixed(std::string& out, float v, int prec) {
  if (v != v) {
    out += "NaN";
    return;
  }
  char buf[80];
  if (v == __builtin_inff()) {
    out += "inf";
    return;
  }
  if (v == -__builtin_inff()) {
    out += "-inf";
    return;
  }
  const int n = snprintf(buf, sizeof buf, "%.*f", prec, (double)v);
  out.append(buf, buf + (n > 0 ? (n < (int)sizeof buf ? n : (int)sizeof buf - 1) : 0));
}

// `{0:02.3}` of quality * 100 (an f32 product): at least 5 characters, so the
// width never pads.
void put_quality(std::string& out, float quality) { put_fixed(out, quality * 100.0f, 3); }

void put_bits(std::string& out, const uint8_t* hash) {
  // every byte of the 32 in order, most significant bit first ({:08b})
  for (int b = 0; b < 32; b++)
    for (int k = 7; k >= 0; k--) out += (hash[b] >> k) & 1 ? '1' : '0';
}

void put_record(std::string& out, const OutputFormat& f, const uint8_t* hash, float quality, float threshold,
                const float* pdqf, float orig_threshold, int orientation, int index) {
  static const char kLower[] = "0123456789abcdef";
  static const char kUpper[] = "0123456789ABCDEF";
  switch (f.type) {
    case OutputFormat::kRaw:
      out.append(reinterpret_cast<const char*>(hash), 32);
      break;
    case OutputFormat::kRawQuality: {
      uint32_t bits;
      memcpy(&bits, &quality, 4);
      const char le[4] = {(char)(bits & 0xff), (char)((bits >> 8) & 0xff), (char)((bits >> 16) & 0xff),
                          (char)(bits >> 24)};
      out.append(le, 4);
      out.append(reinterpret_cast<const char*>(hash), 32);
      break;
    }
    case OutputFormat::kHex:
    case OutputFormat::kHexQuality: {
      if (f.type == OutputFormat::kHexQuality) {
        put_quality(out, quality);
        out += ':';
      }
      const char* digits = f.upper ? kUpper : kLower;
      for (int b = 0; b < 32; b++) {
        out += digits[hash[b] >> 4];
        out += digits[hash[b] & 15];
      }
      break;
    }
    case OutputFormat::kBin:
    case OutputFormat::kBinQuality:
      if (f.type == OutputFormat::kBinQuality) {
        put_quality(out, quality);
        out += ':';
      }
      put_bits(out, hash);
      break;
    case OutputFormat::kDiagnostic: {
      // src/main.rs:738-770
      if (orientation == 0) {
        out += "[\"original\"]\n";
      } else {
        out += "[Dihedrals { packed: " + std::to_string(kOrientationCode[orientation]) + " }]\n";
      }
      out += "Index: " + std::to_string(index) + "\n";
      out += "Quality: ";
      put_quality(out, quality);
      out += "\nThreshold: ";
      put_fixed(out, threshold, 7);
      out += "\nHash:\n";
      put_bits(out, hash);
      out += "\nPDQF:\n";
      for (int i = 0; i < 16; i++) {
        for (int j = 0; j < 16; j++) {
          if (j > 0) out += '\t';
          // marked against the threshold of the original orientation
          if (pdqf[i * 16 + j] > orig_threshold) out += '*';
          put_fixed(out, pdqf[i * 16 + j], 5);
        }
        out += '\n';
      }
      break;
    }
  }
  out += f.separator;
}

}  // namespace

void format_frames(const OutputFormat& f, uint32_t n_frames, bool dihedrals, const uint8_t* hashes,
                   const float* quality, const float* threshold, const float* dihedral_pdqf, std::string& out) {
  const int per = dihedrals ? 8 : 1;
  for (uint32_t i = 0; i < n_frames; i++) {
    for (int d = 0; d < per; d++) {
      const size_t k = (size_t)i * 8 + d;
      put_record(out, f, hashes + k * 32, quality[k], threshold[k],
                 dihedral_pdqf ? dihedral_pdqf + k * 256 : nullptr, threshold[(size_t)i * 8], d, d);
    }
  }
}

uint64_t format_bound(const OutputFormat& f, uint32_t n_frames, bool dihedrals) {
  uint64_t rec;
  switch (f.type) {
    case OutputFormat::kRaw: rec = 32; break;
    case OutputFormat::kRawQuality: rec = 36; break;
    case OutputFormat::kHex: rec = 64; break;
    case OutputFormat::kBin: rec = 256; break;
    case OutputFormat::kHexQuality: rec = 64 + 64; break;
    case OutputFormat::kBinQuality: rec = 256 + 64; break;
    default: rec = 256 * 64 + 1024; break;  // diagnostic: at most 64 characters per PDQF value
  }
  return (rec + f.separator.size()) * (uint64_t)n_frames * (dihedrals ? 8 : 1);
}

// ---------------------------------------------------------------------------
// The stream
// ---------------------------------------------------------------------------
namespace {

// Reads exactly `len` bytes unless the stream ends first; returns the bytes
// read, or -1 (errno set) on a read error.
int64_t read_full(int fd, uint8_t* dst, uint64_t len) {
  uint64_t got = 0;
  while (got < len) {
    const ssize_t r = read(fd, dst + got, (size_t)std::min<uint64_t>(len - got, 1u << 30));
    if (r == 0) break;
    if (r < 0) {
      if (errno == EINTR) continue;
      return -1;
    }
    got += (uint64_t)r;
  }
  return (int64_t)got;
}

bool write_full(int fd, const char* src, size_t len) {
  size_t put = 0;
  while (put < len) {
    const ssize_t w = write(fd, src + put, len - put);
    if (w < 0) {
      if (errno == EINTR) continue;
      return false;
    }
    put += (size_t)w;
  }
  return true;
}

struct Slot {
  uint8_t* frames = nullptr;  // page-locked, batch x frame bytes
  uint32_t n = 0;             // whole frames read
  uint32_t ready = 0;         // frames of the current fill read so far (0 once released)
  int status = 0;             // 0, or the error that ended the stream after these frames
  bool last = false;          // the stream ends after this slot
  bool full = false;          // filled by the reader, not yet released
};

// The reader publishes a slot's progress every kUploadChunk frames and the
// processing thread uploads each chunk as it is published, so a batch of
// 256 frames (64 MiB) leaves for the GPU in 16 MiB pieces right behind the
// reader instead of after it: batches of 256 then run as fast as batches of
// 64 (whose upload is one piece, as before).
constexpr uint32_t kUploadChunk = 64;

struct Outputs {
  uint8_t* d_hashes = nullptr;
  float* d_quality = nullptr;
  float* d_threshold = nullptr;
  float* d_pdqf = nullptr;
  uint8_t* h_hashes = nullptr;
  float* h_quality = nullptr;
  float* h_threshold = nullptr;
  float* h_pdqf = nullptr;
  hipEvent_t uploaded = nullptr;  // the batch's frames are in device memory
  hipEvent_t done = nullptr;
};

// Up to four device -> page-locked host regions, each a whole number of
// 16-byte words, stored by one kernel: the batch's outputs reach the host
// without a DMA copy queued behind the hash kernel (such copies blocked the
// enqueuing thread for milliseconds at a time while work was in flight).
struct HostStores {
  const uint4* src[4];
  uint4* dst[4];
  uint64_t n16[4];
};

__global__ void store_to_host_kernel(HostStores r) {
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
#pragma unroll
  for (int k = 0; k < 4; k++)
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < r.n16[k]; i += step) r.dst[k][i] = r.src[k][i];
}

template <typename T>
hipError_t device_view(T* host, T** dev) {
  void* p = nullptr;
  const hipError_t e = hipHostGetDevicePointer(&p, host, 0);
  *dev = static_cast<T*>(p);
  return e;
}

// Everything a run needs besides its threads: the ring of page-locked
// batches, two sets of device inputs and outputs, the streams and events.
constexpr int kSlots = 3;

struct PipeRes {
  int device = 0;
  uint32_t batch = 0;
  uint64_t frame_bytes = 0;
  bool diag = false;
  Slot slots[kSlots];
  Outputs outs[2];
  void* d_in[2] = {nullptr, nullptr};
  HostStores stores[2];
  hipStream_t stream = nullptr;  // hashing and readback
  hipStream_t upload = nullptr;  // frame uploads
};

void res_destroy(PipeRes* r) {
  if (!r) return;
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  if (r->upload) (void)hipStreamSynchronize(r->upload);
  for (int s = 0; s < kSlots; s++)
    if (r->slots[s].frames) (void)hipHostFree(r->slots[s].frames);
  for (int k = 0; k < 2; k++) {
    Outputs& x = r->outs[k];
    if (r->d_in[k]) (void)hipFree(r->d_in[k]);
    if (x.d_hashes) (void)hipFree(x.d_hashes);
    if (x.d_quality) (void)hipFree(x.d_quality);
    if (x.d_threshold) (void)hipFree(x.d_threshold);
    if (x.d_pdqf) (void)hipFree(x.d_pdqf);
    if (x.h_hashes) (void)hipHostFree(x.h_hashes);
    if (x.h_quality) (void)hipHostFree(x.h_quality);
    if (x.h_threshold) (void)hipHostFree(x.h_threshold);
    if (x.h_pdqf) (void)hipHostFree(x.h_pdqf);
    if (x.done) (void)hipEventDestroy(x.done);
    if (x.uploaded) (void)hipEventDestroy(x.uploaded);
  }
  if (r->stream) (void)hipStreamDestroy(r->stream);
  if (r->upload) (void)hipStreamDestroy(r->upload);
  delete r;
}

hipError_t res_create(int device, uint32_t B, uint64_t frame_bytes, bool diag, PipeRes** out) {
  PipeRes* r = new PipeRes();
  r->device = device;
  r->batch = B;
  r->frame_bytes = frame_bytes;
  r->diag = diag;
  memset(r->stores, 0, sizeof r->stores);
  hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->upload, hipStreamNonBlocking);
  for (int s = 0; s < kSlots && e == hipSuccess; s++) e = hipHostMalloc((void**)&r->slots[s].frames, B * frame_bytes, 0);
  for (int k = 0; k < 2 && e == hipSuccess; k++) {
    Outputs& x = r->outs[k];
    if (e == hipSuccess) e = hipMalloc(&r->d_in[k], B * frame_bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&x.d_hashes, (size_t)B * 256);
    if (e == hipSuccess) e = hipMalloc((void**)&x.d_quality, (size_t)B * 32);
    if (e == hipSuccess) e = hipMalloc((void**)&x.d_threshold, (size_t)B * 32);
    if (e == hipSuccess && diag) e = hipMalloc((void**)&x.d_pdqf, (size_t)B * 8 * 256 * 4);
    if (e == hipSuccess) e = hipHostMalloc((void**)&x.h_hashes, (size_t)B * 256, 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&x.h_quality, (size_t)B * 32, 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&x.h_threshold, (size_t)B * 32, 0);
    if (e == hipSuccess && diag) e = hipHostMalloc((void**)&x.h_pdqf, (size_t)B * 8 * 256 * 4, 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&x.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&x.uploaded, hipEventDisableTiming);
    HostStores& st = r->stores[k];
    if (e == hipSuccess) e = device_view(reinterpret_cast<uint4*>(x.h_hashes), &st.dst[0]);
    if (e == hipSuccess) e = device_view(reinterpret_cast<uint4*>(x.h_quality), &st.dst[1]);
    if (e == hipSuccess) e = device_view(reinterpret_cast<uint4*>(x.h_threshold), &st.dst[2]);
    if (e == hipSuccess && diag) e = device_view(reinterpret_cast<uint4*>(x.h_pdqf), &st.dst[3]);
    st.src[0] = reinterpret_cast<const uint4*>(x.d_hashes);
    st.src[1] = reinterpret_cast<const uint4*>(x.d_quality);
    st.src[2] = reinterpret_cast<const uint4*>(x.d_threshold);
    st.src[3] = reinterpret_cast<const uint4*>(x.d_pdqf);
  }
  if (e != hipSuccess) {
    res_destroy(r);
    r = nullptr;
  }
  *out = r;
  return e;
}

// Idle resource sets of finished runs, the most recently used last.
std::mutex g_res_mu;
std::vector<PipeRes*> g_res_idle;
constexpr size_t kIdleMax = 2;

hipError_t res_acquire(uint32_t B, uint64_t frame_bytes, bool diag, PipeRes** out) {
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return e;
  {
    std::lock_guard<std::mutex> g(g_res_mu);
    for (size_t i = g_res_idle.size(); i-- > 0;) {
      PipeRes* r = g_res_idle[i];
      if (r->device == device && r->batch == B && r->frame_bytes == frame_bytes && r->diag == diag) {
        g_res_idle.erase(g_res_idle.begin() + i);
        *out = r;
        return hipSuccess;
      }
    }
  }
  return res_create(device, B, frame_bytes, diag, out);
}

// Back to the cache after a run whose GPU work has completed.
void res_release(PipeRes* r) {
  std::vector<PipeRes*> evict;
  {
    std::lock_guard<std::mutex> g(g_res_mu);
    g_res_idle.push_back(r);
    while (g_res_idle.size() > kIdleMax) {
      evict.push_back(g_res_idle.front());
      g_res_idle.erase(g_res_idle.begin());
    }
  }
  for (PipeRes* x : evict) res_destroy(x);
}

}  // namespace

int run_pipe(const PipeOptions& o, uint64_t* frames_out, std::string* error) {
  if (frames_out) *frames_out = 0;
  const uint64_t frame_bytes = input_plane_bytes(o.input_kind);
  if (frame_bytes == 0 || o.input_kind == kInputF32 || o.batch == 0) return -1;
  const uint32_t B = o.batch;
  const bool diag = o.format.type == OutputFormat::kDiagnostic;
  auto fail_hip = [&](hipError_t err) {
    if (error) *error = std::string("HIP: ") + hipGetErrorString(err);
    return -2;
  };
  PipeRes* res = nullptr;
  hipError_t e = res_acquire(B, frame_bytes, diag, &res);
  if (e != hipSuccess) return fail_hip(e);
  Slot* slots = res->slots;
  for (int s = 0; s < kSlots; s++) {
    slots[s].n = 0;
    slots[s].ready = 0;
    slots[s].status = 0;
    slots[s].last = false;
    slots[s].full = false;
  }
  Outputs* outs = res->outs;
  void** d_in = res->d_in;
  const HostStores* stores = res->stores;
  const hipStream_t stream = res->stream, upload = res->upload;
  bool hip_failed = false;
  int rc = 0;
  std::mutex mu;
  std::condition_variable cv;
  bool stop_reader = false;
  std::thread reader;
  {
    reader = std::thread([&] {
      for (uint64_t k = 0;; k++) {
        Slot& s = slots[k % kSlots];
        {
          std::unique_lock<std::mutex> lk(mu);
          cv.wait(lk, [&] { return !s.full || stop_reader; });
          if (stop_reader) return;
        }
        uint32_t n = 0;
        int status = 0;
        bool last = false;
        while (n < B) {
          const int64_t got = read_full(o.input_fd, s.frames + (uint64_t)n * frame_bytes, frame_bytes);
          if (got < 0) {
            status = kPipeEio;
            last = true;
            break;
          }
          if ((uint64_t)got < frame_bytes) {
            // the end of the stream: clean at a frame boundary, an error
            // inside a frame (src/main.rs:482-489)
            if (got > 0) status = kPipeEof;
            last = true;
            break;
          }
          n++;
          if (n % kUploadChunk == 0 && n < B) {
            {
              std::lock_guard<std::mutex> lk(mu);
              s.ready = n;
            }
            cv.notify_all();
          }
        }
        {
          std::lock_guard<std::mutex> lk(mu);
          s.n = n;
          s.ready = n;
          s.status = status;
          s.last = last;
          s.full = true;
        }
        cv.notify_all();
        if (last) return;
      }
    });
  }

  uint64_t frames = 0;
  auto t_start = std::chrono::steady_clock::now();
  auto t_stat = t_start;
  uint64_t frames_at_stat = 0;
  std::string text;
  int pending = -1;  // slot whose results are in flight
  int pending_out = 0;
  uint32_t pending_n = 0;

  // Format and write the batch whose GPU work was queued last.
  auto drain = [&]() -> int {
    if (pending < 0) return 0;
    Outputs& x = outs[pending_out];
    hipError_t err = hipEventSynchronize(x.done);
    {
      std::lock_guard<std::mutex> lk(mu);
      slots[pending].full = false;
      slots[pending].ready = 0;
    }
    cv.notify_all();
    pending = -1;
    if (err != hipSuccess) {
      hip_failed = true;
      return fail_hip(err);
    }
    text.clear();
    format_frames(o.format, pending_n, o.dihedrals, x.h_hashes, x.h_quality, x.h_threshold, x.h_pdqf, text);
    if (!write_full(o.output_fd, text.data(), text.size())) {
      if (error) *error = std::string("write: ") + strerror(errno);
      return kPipeEio;
    }
    frames += pending_n;
    if (o.stats) {
      const auto now = std::chrono::steady_clock::now();
      const double dt = std::chrono::duration<double>(now - t_stat).count();
      if (dt >= 1.0) {
        const double all = std::chrono::duration<double>(now - t_start).count();
        fprintf(stderr, "%llu new frames processed (%.0f fps), %llu total frames processed (%.0f fps overall)\n",
                (unsigned long long)(frames - frames_at_stat), (frames - frames_at_stat) / dt,
                (unsigned long long)frames, frames / all);
        t_stat = now;
        frames_at_stat = frames;
      }
    }
    return 0;
  };

  for (uint64_t k = 0; rc == 0 && reader.joinable(); k++) {
    const int si = (int)(k % kSlots);
    Slot& s = slots[si];
    const int oi = (int)(k & 1);
    // the slot's frames go up in chunks as the reader publishes them, until
    // the slot is full.  d_in[oi] and outs[oi] were last used two batches
    // ago, whose done event the host has already waited for
    hipError_t up_err = hipSuccess;
    for (uint32_t up = 0;;) {
      uint32_t avail = 0;
      bool full = false;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return s.full || s.ready >= up + kUploadChunk; });
        full = s.full;
        avail = full ? s.n : s.ready;
      }
      if (avail > up) {
        up_err = hipMemcpyAsync(static_cast<uint8_t*>(d_in[oi]) + (uint64_t)up * frame_bytes,
                                s.frames + (uint64_t)up * frame_bytes, (uint64_t)(avail - up) * frame_bytes,
                                hipMemcpyHostToDevice, upload);
        up = avail;
      }
      if (full || up_err != hipSuccess) break;
    }
    if (up_err != hipSuccess) {
      hip_failed = true;
      rc = fail_hip(up_err);
      break;
    }
    if (s.n > 0) {
      Outputs& x = outs[oi];
      HashOutputs ho;
      ho.hashes = x.d_hashes;
      ho.quality = x.d_quality;
      ho.threshold = x.d_threshold;
      ho.pdqf = nullptr;
      ho.buf1 = nullptr;
      ho.tmp = nullptr;
      ho.dihedral_pdqf = diag ? x.d_pdqf : nullptr;
      hipError_t err = hipEventRecord(x.uploaded, upload);
      if (err == hipSuccess) err = hipStreamWaitEvent(stream, x.uploaded, 0);
      if (err == hipSuccess) err = launch_hash(d_in[oi], o.input_kind, frame_bytes, s.n, o.dihedral_mode, ho, stream);
      if (err == hipSuccess) {
        HostStores r = stores[oi];
        r.n16[0] = (uint64_t)s.n * 256 / 16;
        r.n16[1] = (uint64_t)s.n * 32 / 16;
        r.n16[2] = (uint64_t)s.n * 32 / 16;
        r.n16[3] = diag ? (uint64_t)s.n * 8 * 256 * 4 / 16 : 0;
        hipLaunchKernelGGL(store_to_host_kernel, dim3(diag ? 256 : 32), dim3(256), 0, stream, r);
        err = hipGetLastError();
      }
      if (err == hipSuccess) err = hipEventRecord(x.done, stream);
      if (err != hipSuccess) {
        hip_failed = true;
        rc = fail_hip(err);
        break;
      }
    }
    // while the GPU works on this batch, write the previous one
    if ((rc = drain()) != 0) break;
    if (s.n > 0) {
      pending = si;
      pending_out = oi;
      pending_n = s.n;
    } else {
      std::lock_guard<std::mutex> lk(mu);
      s.full = false;
      s.ready = 0;
    }
    if (s.last) {
      if ((rc = drain()) != 0) break;
      if (s.status != 0) {
        rc = s.status;
        if (error) *error = s.status == kPipeEof ? "unexpected end of input in the middle of a frame" : "read error";
      }
      break;
    }
  }
  if (rc != 0 && pending >= 0) (void)hipEventSynchronize(outs[pending_out].done);
  {
    std::lock_guard<std::mutex> lk(mu);
    stop_reader = true;
  }
  cv.notify_all();
  if (reader.joinable()) reader.join();
  hipError_t se = hipStreamSynchronize(stream);
  if (se == hipSuccess) se = hipStreamSynchronize(upload);
  // a set whose streams saw an error is not reused
  if (hip_failed || se != hipSuccess) res_destroy(res);
  else res_release(res);
  if (frames_out) *frames_out = frames;
  return rc;
}

size_t trim_pipe_cache() {
  std::vector<PipeRes*> idle;
  {
    std::lock_guard<std::mutex> g(g_res_mu);
    idle.swap(g_res_idle);
  }
  for (PipeRes* r : idle) res_destroy(r);
  return idle.size();
}

}  // namespace ypdq
