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
// q = t / 32, PH phases) counts images q, q+PH, ... below 255, so no counter passes 255; the
// 256th image, if any, is added in registers while phase q scans bins [256/PH q, 256/PH (q+1)).
// LOAD 0: byte loads; 1 (npx % 4 == 0, stack 4-byte aligned): one uchar4 load per quad and
// image; 2 (as 1, and the stack within 2 GiB): the same words through a buffer resource, the
// image stride in the scalar offset and one 32-bit offset register per thread, all of a
// thread's images in flight at once (PH phases x 256/PH = the whole 256-image stack); images
// past the stack read past the buffer's range and return 0 (their counts are skipped).
// The LDS histogram bounds the pixels a CU can hold, so the images in flight per pixel set the
// bytes in flight: 16 per thread gave 64 KB per CU (0.11 ms for 256 x 1024^2).
// PH phases share a tile (32 * PH threads per workgroup): 8, or 16 for twice the waves per CU
// at almost the same LDS (each thread then counts and scans half as much).
constexpr int kMode8Quads = 32;
template <int LOAD, int PH>
constexpr int kMode8Unroll = LOAD == 2 ? 256 / PH : 16;
template <int LOAD, int PH = 8>
__global__ __launch_bounds__(32 * PH) void k_background_mode8(size_t npx, int count,
                                                              const uint8_t* __restrict__ stack,
                                                              double* __restrict__ out) {
  constexpr bool VEC4 = LOAD >= 1;
  constexpr int kUnroll = kMode8Unroll<LOAD, PH>;
  constexpr int kMode8Phases = PH;
  constexpr int kThreads = 32 * PH;
  constexpr int kBins = 256 / PH;  // bins scanned per phase
  __shared__ __align__(16) unsigned int hist[256 * kMode8Quads];
  __shared__ unsigned int best[kMode8Phases][4 * kMode8Quads];
  const int t = threadIdx.x;
  for (int k = t; k < 256 * kMode8Quads / 4; k += kThreads) reinterpret_cast<uint4*>(hist)[k] = make_uint4(0u, 0u, 0u, 0u);
  __syncthreads();
  const int p = t % kMode8Quads;
  const int q = t / kMode8Quads;
  const size_t i0 = static_cast<size_t>(blockIdx.x) * (4 * kMode8Quads) + 4 * p;
  size_t ix[4];
  unsigned int inc[4];  // the counter increment of each pixel of the quad, 0 past the image
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const bool has = i0 + c < npx;
    ix[c] = has ? i0 + c : 0;  // loads read valid addresses without a branch each
    inc[c] = has ? 1u << (8 * c) : 0u;
  }
  const int counted = min(count, 255);
  for (int m0 = q; m0 < counted; m0 += kMode8Phases * kUnroll) {
    uchar4 w[kUnroll];
    if constexpr (LOAD == 2) {
      const __amdgpu_buffer_rsrc_t r =
          __builtin_amdgcn_make_buffer_rsrc(const_cast<uint8_t*>(stack), 0, static_cast<int>(count * npx), 0x00020000);
      const unsigned voff = static_cast<unsigned>(static_cast<size_t>(m0) * npx + ix[0]);
      const unsigned stride = static_cast<unsigned>(kMode8Phases * npx);
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
        w[u] = __builtin_bit_cast(uchar4, __builtin_amdgcn_raw_buffer_load_b32(r, voff, u * stride, 0));
    } else {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const uint8_t* s = stack + static_cast<size_t>(min(m0 + kMode8Phases * u, counted - 1)) * npx;
        if constexpr (VEC4) {
          w[u] = *reinterpret_cast<const uchar4*>(s + ix[0]);  // i0 % 4 == 0: has[0] implies has[3]
        } else {
          w[u] = make_uchar4(s[ix[0]], s[ix[1]], s[ix[2]], s[ix[3]]);
        }
      }
    }
    // pixels past the image add zero (no branch per update); images past the stack are
    // skipped by a branch that is uniform over each half-wave (one phase)
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (m0 + kMode8Phases * u < counted) {
        atomicAdd(&hist[w[u].x * kMode8Quads + p], inc[0]);
        atomicAdd(&hist[w[u].y * kMode8Quads + p], inc[1]);
        atomicAdd(&hist[w[u].z * kMode8Quads + p], inc[2]);
        atomicAdd(&hist[w[u].w * kMode8Quads + p], inc[3]);
      }
    }
  }
  __syncthreads();
  // Phase q scans bins [kBins q, kBins (q + 1)) of its quad. Each count and bin form the key
  // count << 8 | (255 - bin): the largest key is the most frequent value, the smallest bin on
  // ties (the reference's first strict maximum). One byte permute builds the 16-bit keys of two
  // pixels from a counter word and the replicated 255 - bin, one packed 16-bit max keeps them.
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
  u16x2 a01 = {0, 0}, a23 = {0, 0};
  for (int b = kBins * q; b < kBins * (q + 1); ++b) {
    const unsigned int w = hist[b * kMode8Quads + p];
    const unsigned int kk = (255u - static_cast<unsigned>(b)) * 0x01010101u;
    // bytes of {kk, w}: 0-3 the counters of pixels 0-3, 4-7 copies of 255 - bin
    const unsigned int k01 = __builtin_amdgcn_perm(kk, w, 0x01040004u);  // (c1:k, c0:k)
    const unsigned int k23 = __builtin_amdgcn_perm(kk, w, 0x03040204u);  // (c3:k, c2:k)
    a01 = __builtin_elementwise_max(a01, __builtin_bit_cast(u16x2, k01));
    a23 = __builtin_elementwise_max(a23, __builtin_bit_cast(u16x2, k23));
  }
  best[q][4 * p + 0] = a01.x;
  best[q][4 * p + 1] = a01.y;
  best[q][4 * p + 2] = a23.x;
  best[q][4 * p + 3] = a23.y;
  __syncthreads();
  if (t < 4 * kMode8Quads) {
    const size_t i = static_cast<size_t>(blockIdx.x) * (4 * kMode8Quads) + t;
    unsigned int key = 0;
    for (int r = 0; r < kMode8Phases; ++r) key = max(key, best[r][t]);
    if (count > 255) {
      // the 256th image adds one to its own bin L only: the maximum over the updated counts is
      // the larger of the maximum so far and L's updated key (a count of 256 needs the 9th bit)
      const unsigned L = stack[static_cast<size_t>(255) * npx + (i < npx ? i : npx - 1)];
      const unsigned cL = (hist[L * kMode8Quads + t / 4] >> (8 * (t % 4))) & 0xffu;
      key = max(key, ((cL + 1u) << 8) | (255u - L));
    }
    if (i < npx) out[i] = static_cast<double>(255u - (key & 0xffu)) / 255.0;
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
