// The reference's precise kernel, ReferenceKernel<f32> (src/kernel/mod.rs:
// 691-821 with the box filter of src/kernel/ref.rs), on the GPU: the
// "forensic" mode whose hashes follow the officially endorsed PDQ resampler
// (two passes of a 3-wide box filter along rows and columns, then nearest
// sampling at floor((2i + 1) 512 / 254)) instead of the optimised tent.
//
// One block of 512 threads per image:
//   1. four box passes (rows, columns, rows, columns) over the 512 x 512 plane,
//      thread t walking row t or column t with the reference's running sum
//      (sum += in; sum -= out; out = sum / window), through two scratch planes
//      in HBM (the plane does not fit LDS);
//   2. the 127 x 127 samples into LDS;
//   3. the reference's sequential DCT (mul then add, k ascending, one
//      accumulator per output), 2032 + 256 independent sums spread over the
//      threads;
//   4. the gradient sum, quality, Torben's median (one wave, exact
//      tie-breaking) and the bits.
// Every rounding happens where the reference's scalar code rounds, so the
// results are its bits.  Only the NORMAL orientation is computed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pdq_hash_common.h"

namespace ypdq {

namespace {

constexpr int kRefThreads = 512;

// Box filter of one line (src/kernel/ref.rs:49-124 with window 3: half
// window 2, one leading accumulation, two short-window writes, full-window
// writes, one trailing short-window write).  The line is read and written in
// chunks of 16 (load(k, buf): inputs 16k .. 16k+15; store(m, buf): outputs
// 16m .. 16m+15), the next chunk loading while this one is summed.
template <typename Load, typename Store>
__device__ __forceinline__ void box3_line(Load load, Store store) {
  float cur[16], nxt[16], o[16];
  load(0, cur);
  load(1, nxt);
  float sum = 0.0f, cws = 0.0f;
  sum += cur[0];
  cws += 1.0f;
  sum += cur[1];
  cws += 1.0f;
  o[0] = sum / cws;
  sum += cur[2];
  cws += 1.0f;
  o[1] = sum / cws;
  float w0 = cur[0], w1 = cur[1], w2 = cur[2];  // inputs ri - 3, ri - 2, ri - 1
  // output j = 16m + e comes from input ri = j + 1: sum += in[ri],
  // sum -= in[ri - 3], out[j] = sum / 3
  auto step = [&](float x, float& out) {
    sum += x;
    sum -= w0;
    out = sum / cws;
    w0 = w1;
    w1 = w2;
    w2 = x;
  };
#pragma unroll
  for (int e = 2; e < 15; e++) step(cur[e + 1], o[e]);
  step(nxt[0], o[15]);
  store(0, o);
#pragma unroll 1
  for (int m = 1; m < 31; m++) {
#pragma unroll
    for (int e = 0; e < 16; e++) cur[e] = nxt[e];
    load(m + 1, nxt);
#pragma unroll
    for (int e = 0; e < 15; e++) step(cur[e + 1], o[e]);
    step(nxt[0], o[15]);
    store(m, o);
  }
#pragma unroll
  for (int e = 0; e < 16; e++) cur[e] = nxt[e];
#pragma unroll
  for (int e = 0; e < 15; e++) step(cur[e + 1], o[e]);
  // the last output: one input leaves the window (w0 = input 509)
  sum -= w0;
  cws += -1.0f;
  o[15] = sum / cws;
  store(31, o);
}

// Chunk loaders and storers of the four passes.
template <typename T>
__device__ __forceinline__ void load_row_chunk(const T* __restrict__ plane, int r, int k, float (&b)[16]);
template <>
__device__ __forceinline__ void load_row_chunk<float>(const float* __restrict__ plane, int r, int k, float (&b)[16]) {
  const f32x4* p = reinterpret_cast<const f32x4*>(plane + r * 512 + 16 * k);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const f32x4 v = p[q];
    b[4 * q] = v.x;
    b[4 * q + 1] = v.y;
    b[4 * q + 2] = v.z;
    b[4 * q + 3] = v.w;
  }
}
template <>
__device__ __forceinline__ void load_row_chunk<uint8_t>(const uint8_t* __restrict__ plane, int r, int k,
                                                        float (&b)[16]) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(plane + r * 512 + 16 * k);
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int q = 0; q < 16; q++) b[q] = (float)((w[q >> 2] >> (8 * (q & 3))) & 0xffu);
}

__device__ __forceinline__ void store_row_chunk(float* __restrict__ dst, int r, int m, const float (&o)[16]) {
  f32x4* p = reinterpret_cast<f32x4*>(dst + r * 512 + 16 * m);
#pragma unroll
  for (int q = 0; q < 4; q++) p[q] = f32x4{o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]};
}

__device__ __forceinline__ void load_col_chunk(const float* __restrict__ src, int c, int k, float (&b)[16]) {
#pragma unroll
  for (int e = 0; e < 16; e++) b[e] = src[(16 * k + e) * 512 + c];
}

__device__ __forceinline__ void store_col_chunk(float* __restrict__ dst, int c, int m, const float (&o)[16]) {
#pragma unroll
  for (int e = 0; e < 16; e++) dst[(16 * m + e) * 512 + c] = o[e];
}

struct RefLds {
  float c[127 * 128];  // samples, row stride 128
  float t[16 * 128];   // DCT stage-1 rows
  float p[256];        // PDQF
};

// Torben's median (src/kernel/mod.rs:212-253) over P by one wave: lane l
// holds elements 4l .. 4l+3, so lane order is the reference's flat order.
// Counts are order-free; maxltguess / mingtguess keep the reference's
// strict-comparison updates: the first element (in flat order) of the
// extreme value among those strictly beyond the initial bound, else the
// bound itself.
__device__ __forceinline__ float torben_median_wave(const float* __restrict__ P, int lane) {
  float v[4];
#pragma unroll
  for (int e = 0; e < 4; e++) v[e] = P[4 * lane + e];
  float mn = v[0], mx = v[0];
#pragma unroll
  for (int e = 1; e < 4; e++) {
    mn = fminf(mn, v[e]);
    mx = fmaxf(mx, v[e]);
  }
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    mn = fminf(mn, __shfl_xor(mn, m, 64));
    mx = fmaxf(mx, __shfl_xor(mx, m, 64));
  }
  const int half = (16 * 16 + 1) / 2;
  // Every iteration either returns or moves min or max to an element
  // strictly inside [min, max], so 257 iterations always suffice.
  float result = 0.0f;
#pragma unroll 1
  for (int iter = 0; iter < 300; iter++) {
    const float guess = (mn + mx) / 2.0f;
    uint32_t less = 0, greater = 0, equal = 0;
    // best element strictly above mn among v < guess: (found, value, index)
    bool lt_found = false, gt_found = false;
    float lt_v = 0.0f, gt_v = 0.0f;
    int lt_i = 0, gt_i = 0;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const float x = v[e];
      if (x < guess) {
        less++;
        if (x > (lt_found ? lt_v : mn)) {
          lt_found = true;
          lt_v = x;
          lt_i = 4 * lane + e;
        }
      } else if (x > guess) {
        greater++;
        if (x < (gt_found ? gt_v : mx)) {
          gt_found = true;
          gt_v = x;
          gt_i = 4 * lane + e;
        }
      } else {
        equal++;
      }
    }
    uint32_t packed = less | (greater << 10) | (equal << 20);
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
      packed += (uint32_t)__shfl_xor((int)packed, m, 64);
      const bool of = __shfl_xor((int)lt_found, m, 64) != 0;
      const float ov = __shfl_xor(lt_v, m, 64);
      const int oi = __shfl_xor(lt_i, m, 64);
      if (of && (!lt_found || ov > lt_v || (ov == lt_v && oi < lt_i))) {
        lt_found = true;
        lt_v = ov;
        lt_i = oi;
      }
      const bool gf = __shfl_xor((int)gt_found, m, 64) != 0;
      const float gv = __shfl_xor(gt_v, m, 64);
      const int gi = __shfl_xor(gt_i, m, 64);
      if (gf && (!gt_found || gv < gt_v || (gv == gt_v && gi < gt_i))) {
        gt_found = true;
        gt_v = gv;
        gt_i = gi;
      }
    }
    less = packed & 1023u;
    greater = (packed >> 10) & 1023u;
    equal = packed >> 20;
    const float maxltguess = lt_found ? lt_v : mn;
    const float mingtguess = gt_found ? gt_v : mx;
    if (less <= (uint32_t)half && greater <= (uint32_t)half) {
      result = less >= (uint32_t)half ? maxltguess : (less + equal >= (uint32_t)half ? guess : mingtguess);
      break;
    } else if (less > greater) {
      mx = maxltguess;
    } else {
      mn = mingtguess;
    }
  }
  return result;
}

template <typename T>
__global__ void __launch_bounds__(kRefThreads) pdq_hash_ref_kernel(const T* __restrict__ planes, uint64_t plane_stride,
                                                                   uint32_t n_images, float* __restrict__ work,
                                                                   uint8_t* __restrict__ hashes,
                                                                   float* __restrict__ quality,
                                                                   float* __restrict__ threshold,
                                                                   float* __restrict__ pdqf) {
  const uint32_t img = blockIdx.x;
  if (img >= n_images) return;
  const int t = threadIdx.x;
  const T* plane = reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(planes) + (size_t)img * plane_stride);
  float* A = work + (size_t)img * 2 * 512 * 512;
  float* B = A + 512 * 512;

  // 1. jarosz_filter_float(.., 512, 512, 3, 3, 2): rows then columns, twice
  //    (ref.rs:7-47); the input plane is not modified.
  box3_line([&](int k, float(&b)[16]) { load_row_chunk<T>(plane, t, k, b); },
            [&](int m, const float(&o)[16]) { store_row_chunk(A, t, m, o); });
  __syncthreads();
  box3_line([&](int k, float(&b)[16]) { load_col_chunk(A, t, k, b); },
            [&](int m, const float(&o)[16]) { store_col_chunk(B, t, m, o); });
  __syncthreads();
  box3_line([&](int k, float(&b)[16]) { load_row_chunk<float>(B, t, k, b); },
            [&](int m, const float(&o)[16]) { store_row_chunk(A, t, m, o); });
  __syncthreads();
  box3_line([&](int k, float(&b)[16]) { load_col_chunk(A, t, k, b); },
            [&](int m, const float(&o)[16]) { store_col_chunk(B, t, m, o); });
  __syncthreads();

  // 2. samples at ((2o + 1) 512) / 254 (mod.rs:808-818)
  __shared__ RefLds L;
  for (int e = t; e < 127 * 127; e += kRefThreads) {
    const int oi = e / 127, oj = e - oi * 127;
    const int ii = ((oi * 2 + 1) * 512) / 254, jj = ((oj * 2 + 1) * 512) / 254;
    L.c[oi * 128 + oj] = B[ii * 512 + jj];
  }
  __syncthreads();

  // 3. dct2d (mod.rs:711-735): tmp[i][j] = sum_k D[i][k] * c[k][j], then
  //    out[i][j] = sum_k tmp[i][k] * D[j][k]; each sum from 0.0, k ascending,
  //    a product then an addition per term.
  {
    const int j = t & 127;
    const int i0 = 4 * __builtin_amdgcn_readfirstlane(t >> 7);
    if (j < 127) {
      float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
      for (int k = 0; k < 127; k++) {
        const float cv = L.c[k * 128 + j];
#pragma unroll
        for (int q = 0; q < 4; q++) s[q] += kDct[(i0 + q) * 127 + k] * cv;
      }
#pragma unroll
      for (int q = 0; q < 4; q++) L.t[(i0 + q) * 128 + j] = s[q];
    }
  }
  __syncthreads();
  if (t < 256) {
    const int i = t >> 4, j = t & 15;
    float s = 0.0f;
#pragma unroll 4
    for (int k = 0; k < 127; k++) s += L.t[i * 128 + k] * kDct[j * 127 + k];
    L.p[t] = s;
    if (pdqf) pdqf[(size_t)img * 256 + t] = s;
  }
  __syncthreads();
  if (t >= 64) return;

  // 4. sum_of_gradients (mod.rs:737-761, the order of sum_of_gradients16),
  //    adjust_quality, quantize (mod.rs:764-788)
  const float med = torben_median_wave(L.p, t);
  if (t == 0) {
    quality[img] = adjust_quality(sum_of_gradients16(L.p));
    threshold[img] = med;
  }
  if (t < 32) {
    // byte 31 - t holds elements 8t .. 8t+7, bit j = element 8t + j > median
    uint32_t byte = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) byte |= (L.p[8 * t + j] > med ? 1u : 0u) << j;
    hashes[(size_t)img * 32 + (31 - t)] = (uint8_t)byte;
  }
}

}  // namespace

uint64_t hash_ref_workspace_bytes(uint32_t n_images) { return (uint64_t)n_images * 2 * 512 * 512 * sizeof(float); }

hipError_t launch_hash_ref(const void* planes, int input_kind, uint64_t plane_stride, uint32_t n_images, void* work,
                           uint8_t* hashes, float* quality, float* threshold, float* pdqf, hipStream_t stream) {
  if (n_images == 0) return hipSuccess;
  if (input_kind == kInputF32)
    hipLaunchKernelGGL(pdq_hash_ref_kernel<float>, dim3(n_images), dim3(kRefThreads), 0, stream,
                       static_cast<const float*>(planes), plane_stride, n_images, static_cast<float*>(work), hashes,
                       quality, threshold, pdqf);
  else if (input_kind == kInputU8)
    hipLaunchKernelGGL(pdq_hash_ref_kernel<uint8_t>, dim3(n_images), dim3(kRefThreads), 0, stream,
                       static_cast<const uint8_t*>(planes), plane_stride, n_images, static_cast<float*>(work), hashes,
                       quality, threshold, pdqf);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace ypdq
