// Staged hash kernels: one launch per method of the reference's Kernel trait
// (src/kernel/mod.rs:268-497), each with the arithmetic of the reference's
// AVX2 smart kernel.  They back the per-call trait mirror of the C++ host API
// and serve as the plain, unfused GPU path the fused kernel is measured
// against.  Every buffer is device memory, one entry per image.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pdq_hash_common.h"

namespace ypdq {

// jarosz_compress (x86.rs:626-682): one thread per element of the 128x128
// buffer; row and column 127 are written as zero padding.
template <typename T>
__global__ void __launch_bounds__(256) stage_compress_kernel(const T* __restrict__ planes, uint64_t stride,
                                                             uint32_t n, float* __restrict__ buf1) {
  const uint32_t img = blockIdx.y;
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (img >= n || e >= 128 * 128) return;
  const T* plane = reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(planes) + (size_t)img * stride);
  const int i = e >> 7, j = e & 127;
  buf1[(size_t)img * 128 * 128 + e] = (i == 127 || j == 127) ? 0.0f : compress_one_literal(plane, i, j);
}

// dct2d (x86.rs:100-281): one 256-thread block per image.  Stage 1: thread
// (k, column block) accumulates 8 columns over k2 = 0..127 with the (1,0,0,1)
// accumulator alternation and NaN-zeroing; stage 2: one output per thread.
__global__ void __launch_bounds__(256) stage_dct2d_kernel(const float* __restrict__ buf1, uint32_t n,
                                                          float* __restrict__ tmp_out, float* __restrict__ pdqf) {
  __shared__ float t[16 * 128];
  const uint32_t img = blockIdx.x;
  if (img >= n) return;
  const float* B = buf1 + (size_t)img * 128 * 128;
  const int tid = threadIdx.x;
  {
    const int k = tid >> 4, jb = (tid & 15) * 8;
    float s[2][8];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int l = 0; l < 8; l++) s[p][l] = 0.0f;
#pragma unroll 4
    for (int k2 = 0; k2 < 128; k2++) {
      const float d = kDct[k * 127 + k2];  // k2 = 127 reads the next row's first entry (or the pad)
      const int p = alt(k2);
#pragma unroll
      for (int l = 0; l < 8; l++) {
        const float v = B[k2 * 128 + jb + l];
        if (p) s[1][l] = __builtin_fmaf(v, d, s[1][l]);
        else s[0][l] = __builtin_fmaf(v, d, s[0][l]);
      }
    }
#pragma unroll
    for (int l = 0; l < 8; l++) {
      const float a0 = __builtin_isnan(s[0][l]) ? 0.0f : s[0][l];
      const float a1 = __builtin_isnan(s[1][l]) ? 0.0f : s[1][l];
      t[k * 128 + jb + l] = a0 + a1;
    }
  }
  __syncthreads();
  if (tmp_out && tid < 128) tmp_out[(size_t)img * 128 + tid] = t[15 * 128 + tid];
  {
    const int k = tid >> 4, j = tid & 15;
    float h[2][8];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int l = 0; l < 8; l++) h[p][l] = 0.0f;
#pragma unroll
    for (int c = 0; c < 16; c++) {
      const int p = alt(c);
#pragma unroll
      for (int l = 0; l < 8; l++) h[p][l] = __builtin_fmaf(t[k * 128 + 8 * c + l], kDct[j * 127 + 8 * c + l], h[p][l]);
    }
    float v[8];
#pragma unroll
    for (int l = 0; l < 8; l++) v[l] = h[0][l] + h[1][l];
    pdqf[(size_t)img * 256 + k * 16 + j] = ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
  }
}

// sum_of_gradients (x86.rs:840-865), one thread per image.
__global__ void stage_gradient_kernel(const float* __restrict__ pdqf, uint32_t n, float* __restrict__ out) {
  const uint32_t img = blockIdx.x * blockDim.x + threadIdx.x;
  if (img >= n) return;
  out[img] = sum_of_gradients16(pdqf + (size_t)img * 256);
}

// quantize (x86.rs:440-621): one wave per image, the search on lanes 0..7
// (the other groups repeat it), bits written by lanes 0..31.
__global__ void __launch_bounds__(64) stage_quantize_kernel(const float* __restrict__ pdqf, uint32_t n,
                                                            float* __restrict__ threshold, uint8_t* __restrict__ hash) {
  const uint32_t img = blockIdx.x;
  if (img >= n) return;
  const int lane = threadIdx.x;
  const float* P = pdqf + (size_t)img * 256;
  float thr, bthr;
  median_search8(P, lane, thr, bthr);
  if (lane == 0) threshold[img] = thr;
  if (lane < 32) hash[(size_t)img * 32 + lane] = (uint8_t)hash_byte(P, lane, bthr);
}

// threshold_2d_f32 (threshold.rs:37-119) with a given threshold per image.
__global__ void __launch_bounds__(64) stage_threshold_kernel(const float* __restrict__ pdqf,
                                                             const float* __restrict__ thr, uint32_t n,
                                                             uint8_t* __restrict__ hash) {
  const uint32_t img = blockIdx.x;
  const int lane = threadIdx.x;
  if (img >= n || lane >= 32) return;
  hash[(size_t)img * 32 + lane] = (uint8_t)hash_byte(pdqf + (size_t)img * 256, lane, thr[img]);
}

// In-place PDQF transforms, one thread per element.  op:
//   0 pdqf_t (mod.rs:345-361)
//   1 / 2 pdqf_negate_alt_cols::<false / true>, AVX2 extent (x86.rs:330-376: columns 0..7)
//   3 / 4 pdqf_negate_alt_cols::<false / true>, all 16 columns (mod.rs:368-382)
//   5 / 6 pdqf_negate_alt_rows::<false / true> (x86.rs:285-326: odd / even rows)
//   7 pdqf_negate_off_diagonals (x86.rs:380-436)
__global__ void __launch_bounds__(256) stage_transform_kernel(float* __restrict__ pdqf, uint32_t n, int op) {
  const uint32_t img = blockIdx.x;
  if (img >= n) return;
  float* P = pdqf + (size_t)img * 256;
  const int e = threadIdx.x, i = e >> 4, j = e & 15;
  const float v = P[e];
  float r = v;
  switch (op) {
    case 0: r = P[j * 16 + i]; break;
    case 1: r = (j < 8 && (j & 1) == 0) ? v * -1.0f : v; break;
    case 2: r = (j < 8 && (j & 1) == 1) ? v * -1.0f : v; break;
    case 3: r = ((j & 1) == 0) ? v * -1.0f : v; break;
    case 4: r = ((j & 1) == 1) ? v * -1.0f : v; break;
    case 5: r = ((i & 1) == 1) ? v * -1.0f : v; break;
    case 6: r = ((i & 1) == 0) ? v * -1.0f : v; break;
    case 7: r = (((i + j) & 1) == 1) ? v * -1.0f : v; break;
    default: break;
  }
  __syncthreads();
  P[e] = r;
}

// RGB8 / RGBA8 to luma f32 with coefficients (R, G, B).  RGB follows the
// trait's default method (mod.rs:303-322: R*r + G*g + B*b, no fusion); RGBA
// follows the AVX2 override (x86.rs:686-753: r*R, then fma(g, G, .), then
// fma(b, B, .)).
__global__ void stage_luma_kernel(const uint8_t* __restrict__ px, int channels, uint64_t n, float cr, float cg,
                                  float cb, float* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float r = (float)px[i * channels], g = (float)px[i * channels + 1], b = (float)px[i * channels + 2];
  float v;
  if (channels == 3) {
    const float a = cr * r;
    const float c = cg * g;
    const float d = cb * b;
    v = (a + c) + d;
  } else {
    v = r * cr;
    v = __builtin_fmaf(g, cg, v);
    v = __builtin_fmaf(b, cb, v);
  }
  out[i] = v;
}

hipError_t launch_stage_compress(const void* planes, int input_kind, uint64_t stride, uint32_t n, float* buf1,
                                 hipStream_t s) {
  if (!n) return hipSuccess;
  const dim3 grid(128 * 128 / 256, n);
  if (input_kind == kInputF32)
    hipLaunchKernelGGL(stage_compress_kernel<float>, grid, dim3(256), 0, s, static_cast<const float*>(planes), stride,
                       n, buf1);
  else
    hipLaunchKernelGGL(stage_compress_kernel<uint8_t>, grid, dim3(256), 0, s, static_cast<const uint8_t*>(planes),
                       stride, n, buf1);
  return hipGetLastError();
}

hipError_t launch_stage_dct2d(const float* buf1, uint32_t n, float* tmp, float* pdqf, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(stage_dct2d_kernel, dim3(n), dim3(256), 0, s, buf1, n, tmp, pdqf);
  return hipGetLastError();
}

hipError_t launch_stage_gradient(const float* pdqf, uint32_t n, float* out, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(stage_gradient_kernel, dim3((n + 63) / 64), dim3(64), 0, s, pdqf, n, out);
  return hipGetLastError();
}

hipError_t launch_stage_quantize(const float* pdqf, uint32_t n, float* thr, uint8_t* hash, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(stage_quantize_kernel, dim3(n), dim3(64), 0, s, pdqf, n, thr, hash);
  return hipGetLastError();
}

hipError_t launch_stage_threshold(const float* pdqf, const float* thr, uint32_t n, uint8_t* hash, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(stage_threshold_kernel, dim3(n), dim3(64), 0, s, pdqf, thr, n, hash);
  return hipGetLastError();
}

hipError_t launch_stage_transform(float* pdqf, uint32_t n, int op, hipStream_t s) {
  if (!n) return hipSuccess;
  if (op < 0 || op > 7) return hipErrorInvalidValue;
  hipLaunchKernelGGL(stage_transform_kernel, dim3(n), dim3(256), 0, s, pdqf, n, op);
  return hipGetLastError();
}

hipError_t launch_stage_luma(const uint8_t* px, int channels, uint64_t n, const float coeff[3], float* out,
                             hipStream_t s) {
  if (!n) return hipSuccess;
  if (channels != 3 && channels != 4) return hipErrorInvalidValue;
  const uint64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(stage_luma_kernel, dim3((unsigned)blocks), dim3(256), 0, s, px, channels, n, coeff[0], coeff[1],
                     coeff[2], out);
  return hipGetLastError();
}

}  // namespace ypdq
