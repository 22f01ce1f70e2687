// Ingest kernels: 8-bit holograms to Float32 on the device (load_gray2float, utils.jl:23-25,
// with the tutorial's background removal, preprocessings.md:21-23) and the per-pixel
// background of an image stack (cu_make_background_mode, utils.jl:243-317; make_background
// :mean, utils.jl:218-224).
//
// Images are uint8 planes of npx = nx * ny pixels, image m of a stack at m * npx.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace holo {

// v / 255 rounded to Float32 (the N0f8 -> Float32 conversion). v / 255 is the byte v repeated
// in base 256, so for 0 < v < 255 its bits past the 24th are never within 2^-32 (relative) of
// a Float32 rounding midpoint; the FP64 product v * fl(1/255), within 2^-52 of v / 255, rounds
// to the same Float32 as the exact quotient (checked for all 256 bytes).
__device__ __forceinline__ float n0f8_to_f32(unsigned int v) {
  return static_cast<float>(static_cast<double>(v) * (1.0 / 255.0));
}

// out = Float32((Float64(v / 255) - bkg) + offset), v / 255 rounded to Float32 first;
// bkg NULL means 0, so with offset 0 this is exactly v / 255. Four pixels per thread, loaded
// as one uchar4 and stored as one float4 when VEC (npx % 4 == 0 and the host found src
// 4-byte and out 16-byte aligned). All loads read valid (clamped) addresses without a branch
// each, so the image bytes and the background values are in flight together.
template <bool VEC, bool BKG>
__global__ void k_ingest_u8(size_t npx, int count, const uint8_t* __restrict__ src,
                            const double* __restrict__ bkg, double offset, float* __restrict__ out) {
#pragma clang fp contract(off)
  const size_t q = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t m = blockIdx.y;
  if (static_cast<int>(m) >= count) return;
  const uint8_t* s = src + m * npx;
  float* o = out + m * npx;
  const size_t i0 = q * 4;
  if (i0 >= npx) return;
  const int n = npx - i0 >= 4 ? 4 : static_cast<int>(npx - i0);
  uint8_t v[4];
  double b[4] = {0.0, 0.0, 0.0, 0.0};
  if constexpr (VEC) {
    const uchar4 w = *reinterpret_cast<const uchar4*>(s + i0);
    v[0] = w.x; v[1] = w.y; v[2] = w.z; v[3] = w.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = s[min(i0 + k, npx - 1)];
  }
  if constexpr (BKG) {
#pragma unroll
    for (int k = 0; k < 4; ++k) b[k] = bkg[min(i0 + k, npx - 1)];
  }
  float r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float f = n0f8_to_f32(v[k]);
    r[k] = static_cast<float>((static_cast<double>(f) - b[k]) + offset);
  }
  if constexpr (VEC) {
    *reinterpret_cast<float4*>(o + i0) = make_float4(r[0], r[1], r[2], r[3]);
  } else {
    for (int k = 0; k < n; ++k) o[i0 + k] = r[k];
  }
}

// Per-pixel mode of a stack of `count` images (count < 65536). A workgroup of 256 threads
// owns 128 consecutive pixels: LDS holds 256 bins x 64 words, each word the two 16-bit
// counters of a pixel pair (even pixel in the low half). Thread (pair p = t % 64, phase
// q = t / 64) reads the pair's two bytes of images q, q+4, ... and adds 1 (even pixel) or
// 1 << 16 (odd pixel) to the bin words, so each byte of the stack is read once. The mode is
// the smallest value among the most frequent (the reference's first strict maximum over
// b = 1..256, utils.jl:259-267); out = mode / 255.0 in FP64 (utils.jl:315).
// Every load reads a valid address (pixel and image indices clamped; the values they bring
// past the stack's end are dropped at the counter update), so the loads carry no branch and
// stay in flight together: loads under per-load branches were each followed by a full wait
// for memory, and the kernel took 0.47 ms for 256 x 1024^2 whatever the unroll depth.
// VEC: npx even and the stack 2-byte aligned, one uchar2 load per pair and image.
constexpr int kModePairs = 64;
constexpr int kModeThreads = 256;
constexpr int kModeUnroll = 32;
template <bool VEC>
__global__ __launch_bounds__(kModeThreads) void k_background_mode(size_t npx, int count,
                                                                   const uint8_t* __restrict__ stack,
                                                                   double* __restrict__ out) {
  __shared__ unsigned int hist[256 * kModePairs];
  __shared__ unsigned int best[4][2 * kModePairs];
  const int t = threadIdx.x;
  for (int k = t; k < 256 * kModePairs; k += kModeThreads) hist[k] = 0u;
  __syncthreads();
  const int p = t % kModePairs;
  const int q = t / kModePairs;
  const size_t i0 = static_cast<size_t>(blockIdx.x) * (2 * kModePairs) + 2 * p;
  const bool has0 = i0 < npx, has1 = i0 + 1 < npx;
  // VEC: npx is even and i0 even, so has0 implies has1 and ix is 2-byte aligned
  const size_t ix = has0 ? i0 : 0, iy = has1 ? i0 + 1 : 0;
  // kModeUnroll loads in flight per thread before their counter updates
  for (int m0 = q; m0 < count; m0 += 4 * kModeUnroll) {
    uchar2 w[kModeUnroll];
#pragma unroll
    for (int u = 0; u < kModeUnroll; ++u) {
      const uint8_t* s = stack + static_cast<size_t>(min(m0 + 4 * u, count - 1)) * npx;
      if constexpr (VEC) {
        w[u] = *reinterpret_cast<const uchar2*>(s + ix);
      } else {
        w[u] = make_uchar2(s[ix], s[iy]);
      }
    }
#pragma unroll
    for (int u = 0; u < kModeUnroll; ++u) {
      if (m0 + 4 * u < count) {
        if (has0) atomicAdd(&hist[w[u].x * kModePairs + p], 1u);
        if (has1) atomicAdd(&hist[w[u].y * kModePairs + p], 1u << 16);
      }
    }
  }
  __syncthreads();
  // phase q scans bins [64 q, 64 q + 64) for both pixels of its pair
  unsigned int c0 = 0, c1 = 0, b0 = 0, b1 = 0;
  for (int b = 64 * q; b < 64 * q + 64; ++b) {
    const unsigned int w = hist[b * kModePairs + p];
    const unsigned int lo = w & 0xffffu, hi = w >> 16;
    if (lo > c0) { c0 = lo; b0 = b; }
    if (hi > c1) { c1 = hi; b1 = b; }
  }
  best[q][2 * p] = (c0 << 8) | b0;
  best[q][2 * p + 1] = (c1 << 8) | b1;
  __syncthreads();
  if (t < 2 * kModePairs) {
    const size_t i = static_cast<size_t>(blockIdx.x) * (2 * kModePairs) + t;
    if (i < npx) {
      unsigned int cb = 0, bb = 0;
      for (int r = 0; r < 4; ++r) {  // earlier phases hold smaller bins: strict > keeps them
        const unsigned int e = best[r][t];
        if ((e >> 8) > cb) { cb = e >> 8; bb = e & 0xffu; }
      }
      out[i] = static_cast<double>(bb) / 255.0;
    }
  }
}

// The same mode for stacks of at most 256 images with 8-bit counters: a word holds the counters
// of four consecutive pixels, so a 128-pixel tile needs 32 KB of LDS and four workgroups fit a
// CU (twice the loads in flight of the 16-bit kernel). Thread (quad p = t % 32, phase
// q = t / 32) counts images q, q+8, ... below 255, so no counter passes 255; the 256th image,
// if any, is added in registers while phase q scans bins [32 q, 32 q + 32). VEC4: npx % 4 == 0
// and the stack 4-byte aligned, one uchar4 load per quad and image.
constexpr int kMode8Quads = 32;
constexpr int kMode8Phases = kModeThreads / kMode8Quads;  // 8
constexpr int kMode8Unroll = 16;
template <bool VEC4>
__global__ __launch_bounds__(kModeThreads) void k_background_mode8(size_t npx, int count,
                                                                    const uint8_t* __restrict__ stack,
                                                                    double* __restrict__ out) {
  __shared__ unsigned int hist[256 * kMode8Quads];
  __shared__ unsigned int best[kMode8Phases][4 * kMode8Quads];
  const int t = threadIdx.x;
  for (int k = t; k < 256 * kMode8Quads; k += kModeThreads) hist[k] = 0u;
  __syncthreads();
  const int p = t % kMode8Quads;
  const int q = t / kMode8Quads;
  const size_t i0 = static_cast<size_t>(blockIdx.x) * (4 * kMode8Quads) + 4 * p;
  bool has[4];
  size_t ix[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    has[c] = i0 + c < npx;
    ix[c] = has[c] ? i0 + c : 0;  // loads read valid addresses without a branch each
  }
  const int counted = min(count, 255);
  for (int m0 = q; m0 < counted; m0 += kMode8Phases * kMode8Unroll) {
    uchar4 w[kMode8Unroll];
#pragma unroll
    for (int u = 0; u < kMode8Unroll; ++u) {
      const uint8_t* s = stack + static_cast<size_t>(min(m0 + kMode8Phases * u, counted - 1)) * npx;
      if constexpr (VEC4) {
        w[u] = *reinterpret_cast<const uchar4*>(s + ix[0]);  // i0 % 4 == 0: has[0] implies has[3]
      } else {
        w[u] = make_uchar4(s[ix[0]], s[ix[1]], s[ix[2]], s[ix[3]]);
      }
    }
#pragma unroll
    for (int u = 0; u < kMode8Unroll; ++u) {
      if (m0 + kMode8Phases * u < counted) {
        if (has[0]) atomicAdd(&hist[w[u].x * kMode8Quads + p], 1u);
        if (has[1]) atomicAdd(&hist[w[u].y * kMode8Quads + p], 1u << 8);
        if (has[2]) atomicAdd(&hist[w[u].z * kMode8Quads + p], 1u << 16);
        if (has[3]) atomicAdd(&hist[w[u].w * kMode8Quads + p], 1u << 24);
      }
    }
  }
  // the 256th image (count == 256), added during the scan
  unsigned last[4] = {256u, 256u, 256u, 256u};  // 256: no bin
  if (count > 255) {
    const uint8_t* s = stack + static_cast<size_t>(255) * npx;
#pragma unroll
    for (int c = 0; c < 4; ++c) last[c] = s[ix[c]];
  }
  __syncthreads();
  unsigned cb[4] = {0u, 0u, 0u, 0u}, bb[4] = {0u, 0u, 0u, 0u};
  for (int b = 32 * q; b < 32 * q + 32; ++b) {
    const unsigned int w = hist[b * kMode8Quads + p];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const unsigned n = ((w >> (8 * c)) & 0xffu) + (last[c] == static_cast<unsigned>(b) ? 1u : 0u);
      if (n > cb[c]) { cb[c] = n; bb[c] = b; }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) best[q][4 * p + c] = (cb[c] << 8) | bb[c];
  __syncthreads();
  if (t < 4 * kMode8Quads) {
    const size_t i = static_cast<size_t>(blockIdx.x) * (4 * kMode8Quads) + t;
    if (i < npx) {
      unsigned int c = 0, bv = 0;
      for (int r = 0; r < kMode8Phases; ++r) {  // earlier phases hold smaller bins: strict > keeps them
        const unsigned int e = best[r][t];
        if ((e >> 8) > c) { c = e >> 8; bv = e & 0xffu; }
      }
      out[i] = static_cast<double>(bv) / 255.0;
    }
  }
}

// Per-pixel mean of a stack: sum of Float64(Float32(v / 255)) in image order from 0.0, then
// divided by count (make_background :mean, utils.jl:219-224). With `vec` (npx % 4 == 0 and
// the stack 4-byte aligned) a thread owns 4 consecutive pixels and reads one uchar4 per
// image, else one pixel; kMeanUnroll images are loaded before they are added in order.
constexpr int kMeanUnroll = 16;
__global__ void k_background_mean(size_t npx, int count, const uint8_t* __restrict__ stack,
                                  double* __restrict__ out, int vec) {
#pragma clang fp contract(off)
  const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int P = vec ? 4 : 1;
  const size_t i0 = t * P;
  if (i0 >= npx) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int m0 = 0; m0 < count; m0 += kMeanUnroll) {
    uchar4 w[kMeanUnroll];
#pragma unroll
    for (int u = 0; u < kMeanUnroll; ++u) {
      const uint8_t* s = stack + static_cast<size_t>(m0 + u) * npx + i0;
      w[u] = make_uchar4(0, 0, 0, 0);
      if (m0 + u < count) {
        if (vec) w[u] = *reinterpret_cast<const uchar4*>(s);
        else w[u].x = s[0];
      }
    }
#pragma unroll
    for (int u = 0; u < kMeanUnroll; ++u) {
      if (m0 + u < count) {
        acc[0] = acc[0] + static_cast<double>(n0f8_to_f32(w[u].x));
        acc[1] = acc[1] + static_cast<double>(n0f8_to_f32(w[u].y));
        acc[2] = acc[2] + static_cast<double>(n0f8_to_f32(w[u].z));
        acc[3] = acc[3] + static_cast<double>(n0f8_to_f32(w[u].w));
      }
    }
  }
  for (int k = 0; k < P; ++k) out[i0 + k] = acc[k] / static_cast<double>(count);
}

}  // namespace holo
