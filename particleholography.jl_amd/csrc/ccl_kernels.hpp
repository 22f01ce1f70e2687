// Connected-component labelling, bounding boxes and dilation of slices (ccl.jl:186-264,
// particle_detection.jl:12-96).
//
// A slice is an nx x ny plane (x fastest); `count` slices lie back to back. A pixel's class
// is 0 (background) or a nonzero key; 8-neighbours with equal nonzero keys are connected.
// The label of a component is 1 + the smallest memory offset among its pixels, background
// pixels are 0: the labels the reference's union-find produces (every union links the
// larger root under the smaller, so the surviving root is the first pixel in memory order).
//
// The passes over a slice, on a grid of 64 x 16 tiles:
//   0. k_ccl_scan: one wave per tile reads it with 16-byte loads and votes; occupied tiles are
//      flagged and listed, so a sparse volume costs one streaming read of the image;
//   1. k_ccl_local: a workgroup per listed tile unions its pixels in LDS and writes for every
//      foreground pixel the global offset of its tile-local root (P), background 0xffffffff;
//   2. k_ccl_border: unions across the edges of listed tiles only (W and NW of the left
//      column, NW/N/NE of the top row, NE of the right column), lock-free on P with atomicMin;
//   3. k_ccl_labels (labels = root + 1), or k_ccl_roots + k_ccl_boxes (one box slot per root,
//      then min / max of the coordinates of every run of equal roots along a row).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace holo {

constexpr int kTileX = 64;
constexpr int kTileY = 16;
constexpr int kTilePx = kTileX * kTileY;
constexpr unsigned kNoLabel = 0xffffffffu;

enum CclKind : int {
  CCL_F32 = 0,     // Float32 image: positive values are foreground, equal values connect
  CCL_BINARY = 1,  // uint8 / Bool image: nonzero is foreground
  CCL_N0F8_LE = 2  // uint8 (N0f8) volume: v < lim is foreground (d_vol .<= threshold: the
                   // bytes whose v/255 <= threshold are 0 .. lim-1)
};

struct CclSrc {
  const void* img;   // slice z at img + z * npx elements
  int kind;
  unsigned lim;      // CCL_N0F8_LE: number of foreground byte values
  int vec;           // nx % 4 == 0 and the image aligned for 4-pixel loads
};

__device__ __forceinline__ unsigned ccl_key(const CclSrc& s, size_t idx) {
  if (s.kind == CCL_F32) {
    const float v = static_cast<const float*>(s.img)[idx];
    return v > 0.0f ? __float_as_uint(v) : 0u;  // positive floats: equal values <=> equal bits
  }
  const unsigned v = static_cast<const uint8_t*>(s.img)[idx];
  if (s.kind == CCL_BINARY) return v != 0u ? 1u : 0u;
  return v < s.lim ? 1u : 0u;
}

__device__ __forceinline__ unsigned lds_find(volatile unsigned* par, unsigned x) {
  unsigned p = par[x];
  while (p != x) {
    x = p;
    p = par[x];
  }
  return x;
}

// Links the roots of a and b, the larger under the smaller (lock-free: a failed atomicMin
// means a's root changed under us, and the loop continues from its new parent).
__device__ __forceinline__ void lds_union(unsigned* par, unsigned a, unsigned b) {
  volatile unsigned* vp = par;
  while (true) {
    a = lds_find(vp, a);
    b = lds_find(vp, b);
    if (a == b) return;
    if (a < b) {
      const unsigned t = a;
      a = b;
      b = t;
    }
    const unsigned old = atomicMin(&par[a], b);
    if (old == a) return;
    a = old;
  }
}

__device__ __forceinline__ unsigned g_load(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned g_find(unsigned* P, unsigned x) {
  unsigned p = g_load(P + x);
  while (p != x) {
    x = p;
    p = g_load(P + x);
  }
  return x;
}

__device__ __forceinline__ void g_union(unsigned* P, unsigned a, unsigned b) {
  while (true) {
    a = g_find(P, a);
    b = g_find(P, b);
    if (a == b) return;
    if (a < b) {
      const unsigned t = a;
      a = b;
      b = t;
    }
    const unsigned old = atomicMin(P + a, b);
    if (old == a) return;
    a = old;
  }
}

// Thread t of a 256-thread tile workgroup owns the 4 adjacent pixels lx = 4*(t % 16) + c,
// c = 0..3, of row ly = t / 16. `vec` (host-checked: nx % 4 == 0 and 4-byte / 16-byte aligned
// image) lets the 4 keys load as one uchar4 / float4 and P store as one uint4.
__device__ __forceinline__ void tile_keys(const CclSrc& s, size_t base, int nx, int ny, int x0, int y,
                                          int vec, unsigned k4[4]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) k4[c] = 0u;
  if (y >= ny) return;
  const size_t o = base + static_cast<size_t>(y) * nx + x0;
  if (vec && x0 + 3 < nx) {
    if (s.kind == CCL_F32) {
      const float4 w = *reinterpret_cast<const float4*>(static_cast<const float*>(s.img) + o);
      const float f[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int c = 0; c < 4; ++c) k4[c] = f[c] > 0.0f ? __float_as_uint(f[c]) : 0u;
    } else {
      const uchar4 w = *reinterpret_cast<const uchar4*>(static_cast<const uint8_t*>(s.img) + o);
      const unsigned b[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int c = 0; c < 4; ++c)
        k4[c] = s.kind == CCL_BINARY ? (b[c] != 0u ? 1u : 0u) : (b[c] < s.lim ? 1u : 0u);
    }
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (x0 + c < nx) k4[c] = ccl_key(s, o + c);
  }
}

// Occupancy scan: one wave per tile, the whole tile read with 16-byte loads (u8: 64 lanes x
// 16 B = the tile's 16 rows x 64 pixels; F32: 4 such sweeps), a wave vote, then lane 0 sets the
// tile's flag and appends an occupied tile to `list`. No LDS and no barrier, so the pass
// streams the volume. `vec16`: nx % 16 == 0 and the image 16-byte aligned (else byte/float
// loads). grid-stride over all tiles of the round.
__device__ __forceinline__ bool key_any16(const CclSrc& s, const uint4 w) {
  if (s.kind == CCL_F32) {
    return __uint_as_float(w.x) > 0.0f || __uint_as_float(w.y) > 0.0f || __uint_as_float(w.z) > 0.0f ||
           __uint_as_float(w.w) > 0.0f;
  }
  const unsigned words[4] = {w.x, w.y, w.z, w.w};
  bool a = false;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const unsigned v = (words[q] >> (8 * b)) & 0xffu;
      a |= s.kind == CCL_BINARY ? v != 0u : v < s.lim;
    }
  return a;
}

__global__ __launch_bounds__(256) void k_ccl_scan(int nx, int ny, int count, CclSrc src, int vec16,
                                                  unsigned* __restrict__ flags, unsigned* __restrict__ list,
                                                  unsigned* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const unsigned nwaves = (gridDim.x * blockDim.x) >> 6;
  const int tiles_x = (nx + kTileX - 1) / kTileX, tiles_y = (ny + kTileY - 1) / kTileY;
  const unsigned ntiles = static_cast<unsigned>(tiles_x) * tiles_y * count;
  const size_t npx = static_cast<size_t>(nx) * ny;
  const int esz = src.kind == CCL_F32 ? 4 : 1;
  const int sweeps = esz;          // 16-byte lanes: 16 pixels (u8) or 4 pixels (f32) each
  const int lanes_per_row = 4 * esz;  // lanes covering the 64 pixels of a row
  const int rows_per_sweep = 64 / lanes_per_row;
  for (unsigned tid = wave; tid < ntiles; tid += nwaves) {
    const int tx = static_cast<int>(tid % tiles_x);
    const int ty = static_cast<int>((tid / tiles_x) % tiles_y);
    const size_t z = tid / (static_cast<unsigned>(tiles_x) * tiles_y);
    const size_t base = z * npx;
    bool any = false;
    for (int sw = 0; sw < sweeps; ++sw) {
      const int row = sw * rows_per_sweep + lane / lanes_per_row;
      const int y = ty * kTileY + row;
      const int px0 = (lane % lanes_per_row) * (16 / esz);
      const int x = tx * kTileX + px0;
      if (y >= ny || x >= nx) continue;
      const size_t o = base + static_cast<size_t>(y) * nx + x;
      if (vec16 && x + 16 / esz <= nx) {
        const uint4 w = *reinterpret_cast<const uint4*>(static_cast<const uint8_t*>(src.img) + o * esz);
        any |= key_any16(src, w);
      } else {
        for (int c = 0; c < 16 / esz && x + c < nx; ++c) any |= ccl_key(src, o + c) != 0u;
      }
    }
    const bool tile_any = __any(any);
    if (lane == 0) {
      flags[tid] = tile_any ? 1u : 0u;
      if (tile_any) list[atomicAdd(cnt, 1u)] = tid;
    }
  }
}

// A tile of the list: slice z, tile column tx, tile row ty.
struct TileId {
  int z, tx, ty;
};
__device__ __forceinline__ TileId tile_of(unsigned id, int nx, int ny) {
  const int tiles_x = (nx + kTileX - 1) / kTileX, tiles_y = (ny + kTileY - 1) / kTileY;
  TileId r;
  r.tx = static_cast<int>(id % tiles_x);
  id /= tiles_x;
  r.ty = static_cast<int>(id % tiles_y);
  r.z = static_cast<int>(id / tiles_y);
  return r;
}

// Pass 1 over the occupied tiles of the list (from k_ccl_scan): the tile's pixels are unioned
// in LDS and P receives each foreground pixel's tile-local root (background kNoLabel).
// grid-stride over the list.
__global__ __launch_bounds__(256) void k_ccl_local(int nx, int ny, CclSrc src, unsigned* __restrict__ P,
                                                   const unsigned* __restrict__ list,
                                                   const unsigned* __restrict__ cnt) {
  __shared__ unsigned key[kTilePx];
  __shared__ unsigned par[kTilePx];
  const unsigned n = *cnt;
  const int t = threadIdx.x;
  const int ly = t >> 4, lx0 = 4 * (t & 15);
  const size_t npx = static_cast<size_t>(nx) * ny;
  const int i0 = ly * kTileX + lx0;
  for (unsigned k = blockIdx.x; k < n; k += gridDim.x) {
    const TileId tl = tile_of(list[k], nx, ny);
    const int x0 = tl.tx * kTileX + lx0;
    const int y0 = tl.ty * kTileY;
    const int y = y0 + ly;
    const size_t base = static_cast<size_t>(tl.z) * npx;
    unsigned k4[4];
    tile_keys(src, base, nx, ny, x0, y, src.vec, k4);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      key[i0 + c] = k4[c];
      par[i0 + c] = i0 + c;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int lx = lx0 + c, i = i0 + c;
      const unsigned kk = k4[c];
      if (kk == 0u) continue;
      if (lx > 0 && key[i - 1] == kk) lds_union(par, i, i - 1);
      if (ly > 0) {
        if (lx > 0 && key[i - kTileX - 1] == kk) lds_union(par, i, i - kTileX - 1);
        if (key[i - kTileX] == kk) lds_union(par, i, i - kTileX);
        if (lx < kTileX - 1 && key[i - kTileX + 1] == kk) lds_union(par, i, i - kTileX + 1);
      }
    }
    __syncthreads();
    if (y < ny) {
      unsigned out[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        out[c] = kNoLabel;
        if (k4[c] != 0u) {
          const unsigned rt = lds_find(par, i0 + c);
          out[c] = static_cast<unsigned>((y0 + static_cast<int>(rt) / kTileX) * nx + tl.tx * kTileX + rt % kTileX);
        }
      }
      unsigned* dst = P + base + static_cast<size_t>(y) * nx + x0;
      if (src.vec && x0 + 3 < nx) {
        *reinterpret_cast<uint4*>(dst) = make_uint4(out[0], out[1], out[2], out[3]);
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (x0 + c < nx) dst[c] = out[c];
      }
    }
    __syncthreads();  // the LDS of this tile is read before the next tile overwrites it
  }
}

// Pass 2 on one tile, 128 threads: t < 64 the top-row pixel lx = t (NW, N, NE across the top
// edge), 64 <= t < 80 the left-column pixel ly = t - 64 (W, NW), 80 <= t < 96 the
// right-column pixel ly = t - 80 (NE).
__device__ __forceinline__ void border_tile(int nx, int ny, const CclSrc& src, unsigned* __restrict__ P,
                                            const TileId& tl) {
  const int t = threadIdx.x;
  const int x0 = tl.tx * kTileX, y0 = tl.ty * kTileY;
  int x, y;
  if (t < 64) {
    x = x0 + t;
    y = y0;
  } else if (t < 80) {
    x = x0;
    y = y0 + (t - 64);
  } else if (t < 96) {
    x = x0 + kTileX - 1;
    y = y0 + (t - 80);
  } else {
    return;
  }
  if (x >= nx || y >= ny) return;
  const size_t npx = static_cast<size_t>(nx) * ny;
  const size_t base = static_cast<size_t>(tl.z) * npx;
  unsigned* Ps = P + base;
  const unsigned me = static_cast<unsigned>(y) * nx + x;
  const unsigned c = ccl_key(src, base + me);
  if (c == 0u) return;
  auto link = [&](int qx, int qy) {
    if (qx < 0 || qx >= nx || qy < 0) return;
    const unsigned q = static_cast<unsigned>(qy) * nx + qx;
    if (ccl_key(src, base + q) == c) g_union(Ps, me, q);
  };
  if (t < 64) {
    if (y > 0) {
      link(x - 1, y - 1);
      link(x, y - 1);
      link(x + 1, y - 1);
    }
  } else if (t < 80) {
    link(x - 1, y);
    link(x - 1, y - 1);
  } else {
    link(x + 1, y - 1);
  }
}

// Pass 2 over the tiles of the list: a small grid walks it (tiles without foreground have no
// edges to join and are not in it).
__global__ __launch_bounds__(128) void k_ccl_border(int nx, int ny, CclSrc src, unsigned* __restrict__ P,
                                                    const unsigned* __restrict__ list,
                                                    const unsigned* __restrict__ count) {
  const unsigned n = *count;
  for (unsigned k = blockIdx.x; k < n; k += gridDim.x) border_tile(nx, ny, src, P, tile_of(list[k], nx, ny));
}

// Pass 3 (labels), tile grid: labels = root + 1 for foreground pixels, 0 for background (a
// whole empty tile is written with zeros without reading P). P is compressed on the way.
__global__ __launch_bounds__(256) void k_ccl_labels(int nx, int ny, unsigned* __restrict__ P,
                                                    const unsigned* __restrict__ flags,
                                                    uint32_t* __restrict__ labels, int vec) {
  const size_t tile = (static_cast<size_t>(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  const int t = threadIdx.x;
  const int x0 = blockIdx.x * kTileX + 4 * (t & 15);
  const int y = blockIdx.y * kTileY + (t >> 4);
  if (y >= ny) return;
  const size_t npx = static_cast<size_t>(nx) * ny;
  const size_t row = static_cast<size_t>(blockIdx.z) * npx + static_cast<size_t>(y) * nx;
  unsigned out[4] = {0u, 0u, 0u, 0u};
  if (flags[tile] != 0u) {
    unsigned* Ps = P + static_cast<size_t>(blockIdx.z) * npx;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (x0 + c >= nx) continue;
      const size_t idx = static_cast<size_t>(y) * nx + x0 + c;
      const unsigned p = Ps[idx];
      if (p != kNoLabel) {
        const unsigned r = g_find(Ps, p);
        Ps[idx] = r;
        out[c] = r + 1u;
      }
    }
  }
  uint32_t* dst = labels + row + x0;
  if (vec && x0 + 3 < nx) {
    *reinterpret_cast<uint4*>(dst) = make_uint4(out[0], out[1], out[2], out[3]);
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (x0 + c < nx) dst[c] = out[c];
  }
}

// P from a label image. A pixel joins label l only when l is a valid label -- the pixel at
// offset l - 1 carries l itself (cu_find_valid_labels' test, ccl.jl:240-248) -- so every entry
// of P is a self-pointing root or kNoLabel, whatever the image holds, and no find can cycle.
// Every tile is flagged (the label image is not scanned for empty tiles).
__global__ void k_ccl_from_labels(size_t npx, int count, const uint32_t* __restrict__ labels,
                                  unsigned* __restrict__ P) {
  const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t z = blockIdx.y;
  if (idx >= npx || static_cast<int>(z) >= count) return;
  const uint32_t* lz = labels + z * npx;
  const uint32_t l = lz[idx];
  const bool valid = l != 0u && l <= npx && lz[l - 1u] == l;
  P[z * npx + idx] = valid ? l - 1u : kNoLabel;
}

// list[k] = k, *count = n: every tile listed (boxes of a given label image).
__global__ void k_ccl_list_all(unsigned n, unsigned* __restrict__ list, unsigned* __restrict__ count) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) list[k] = k;
  if (k == 0) *count = n;
}

// A box of one component of one slice: label = root offset + 1, coordinates 0-based
// (i fast, j slow), inclusive.
struct CclBox {
  unsigned label;
  int i0, j0, i1, j1;
};

// Pass 3a (boxes) over the list: every foreground pixel of a listed tile points P at its
// root; a root claims a box slot (slot order is arbitrary) and initialises it; M[root] = slot.
__global__ __launch_bounds__(256) void k_ccl_roots(int nx, int ny, unsigned* __restrict__ P,
                                                   const unsigned* __restrict__ list,
                                                   const unsigned* __restrict__ count, unsigned* __restrict__ M,
                                                   CclBox* __restrict__ boxes, int max_boxes, int* __restrict__ nbox) {
  const unsigned n = *count;
  const int t = threadIdx.x;
  const size_t npx = static_cast<size_t>(nx) * ny;
  for (unsigned k = blockIdx.x; k < n; k += gridDim.x) {
    const TileId tl = tile_of(list[k], nx, ny);
    const int x0 = tl.tx * kTileX + 4 * (t & 15);
    const int y = tl.ty * kTileY + (t >> 4);
    if (y >= ny) continue;
    const size_t z = tl.z;
    unsigned* Ps = P + z * npx;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (x0 + c >= nx) continue;
      const unsigned idx = static_cast<unsigned>(y) * nx + x0 + c;
      const unsigned p = Ps[idx];
      if (p == kNoLabel) continue;
      const unsigned r = g_find(Ps, p);
      if (r == idx) {
        const int slot = atomicAdd(nbox + z, 1);
        M[z * npx + idx] = static_cast<unsigned>(slot);
        if (slot < max_boxes) {
          CclBox b;
          b.label = r + 1u;
          b.i0 = 0x7fffffff;
          b.j0 = 0x7fffffff;
          b.i1 = -1;
          b.j1 = -1;
          boxes[z * max_boxes + slot] = b;
        }
      }
      Ps[idx] = r;  // every foreground entry now points at its root
    }
  }
}

// Pass 3b (boxes) over the list: a thread's 4 pixels issue one set of min / max atomics per
// run of equal roots.
__global__ __launch_bounds__(256) void k_ccl_boxes(int nx, int ny, const unsigned* __restrict__ P,
                                                   const unsigned* __restrict__ list,
                                                   const unsigned* __restrict__ count,
                                                   const unsigned* __restrict__ M, CclBox* __restrict__ boxes,
                                                   int max_boxes) {
  const unsigned n = *count;
  const int t = threadIdx.x;
  const size_t npx = static_cast<size_t>(nx) * ny;
  for (unsigned k = blockIdx.x; k < n; k += gridDim.x) {
    const TileId tl = tile_of(list[k], nx, ny);
    const int x0 = tl.tx * kTileX + 4 * (t & 15);
    const int j = tl.ty * kTileY + (t >> 4);
    if (j >= ny) continue;
    const size_t z = tl.z;
    const unsigned* Ps = P + z * npx + static_cast<size_t>(j) * nx;
    const unsigned* Ms = M + z * npx;
    CclBox* bz = boxes + z * max_boxes;
    unsigned cur = kNoLabel;
    int first = 0, last = 0;
    auto flush = [&]() {
      if (cur == kNoLabel) return;
      const unsigned s = Ms[cur];
      if (s >= static_cast<unsigned>(max_boxes)) return;
      atomicMin(&bz[s].i0, first);
      atomicMax(&bz[s].i1, last);
      atomicMin(&bz[s].j0, j);
      atomicMax(&bz[s].j1, j);
    };
    const int i_end = min(x0 + 4, nx);
    for (int i = x0; i < i_end; ++i) {
      const unsigned r = Ps[i];
      if (r != cur) {
        flush();
        cur = r;
        first = i;
      }
      last = i;
    }
    flush();
  }
}

// cu_dilate (particle_detection.jl:12-42): out = OR of the 3 x 3 in-plane neighbourhood for
// interior pixels (1 <= i <= nx-2, 1 <= j <= ny-2), false on the border.
__global__ void k_dilate3x3(int nx, int ny, int count, const uint8_t* __restrict__ in, uint8_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  const size_t z = blockIdx.z;
  if (i >= nx || j >= ny || static_cast<int>(z) >= count) return;
  const size_t npx = static_cast<size_t>(nx) * ny;
  const uint8_t* s = in + z * npx;
  uint8_t v = 0;
  if (i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2) {
    for (int dj = -1; dj <= 1; ++dj) {
      const uint8_t* r = s + static_cast<size_t>(j + dj) * nx + i;
      v |= static_cast<uint8_t>((r[-1] != 0) | (r[0] != 0) | (r[1] != 0));
    }
  }
  out[z * npx + static_cast<size_t>(j) * nx + i] = v;
}

}  // namespace holo
