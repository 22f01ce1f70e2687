// HIP kernels of the angular-spectrum / Gerchberg-Saxton path (gfx950).
//
// Memory layout: an nx x ny plane is stored with the x axis contiguous (the reference's Julia
// column-major CuArray layout, dim 1 = x): element (i, j) at j*nx + i. "Row" passes run the
// FFT along x (contiguous lines), "column" passes along y (stride nx; a workgroup owns W
// adjacent columns so that every global access is a W*8-byte contiguous segment).
//
// Spectra are kept in natural FFT order. The reference multiplies fftshift(fft(psi)) by a
// centred transfer array H and applies ifftshift before the inverse FFT; since
// ifftshift(fftshift(X) .* H) == X .* ifftshift(H) for every size, the kernels either
// generate ifftshift(H) directly or read the centred array at the shifted index.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "fft_core.hpp"

namespace holo {

struct Geom {
  int nx, ny;
  LinePlan px;  // plan along x (row passes)
  LinePlan py;  // plan along y (column passes)
  const float2* twx;
  const float2* twy;
  int W;        // columns per workgroup in column passes (power of two)
  int wshift;   // log2(W)
  int rows;     // rows per workgroup in row passes
  int rtiles;   // row-pass workgroups along x: ceil(ny / rows) (the grid's x extent)
  int gblk;     // the reconstruction's column-transformed slices G are stored in 4 x 4 blocks
};

// Element (row j, column i) of a reconstruction slice G: natural order j * nx + i, or, with
// gblk, in blocks of 4 rows x 4 columns (16 consecutive elements, one 128-byte line), so that a
// column pass over 4-column tiles writes whole lines instead of 32-byte pieces of four
// (2048^2 pass A: 2.1 -> 1.5 ms per 192 slices). Needs nx and ny multiples of 4, and threads
// per line (both passes) multiples of 4: element t + T*e of a line then sits at the thread's
// base plus T*nx*e (columns) or 4*T*e (rows).
__device__ __forceinline__ unsigned g_index(const Geom& g, unsigned j, unsigned i) {
  return g.gblk ? ((j >> 2) * (static_cast<unsigned>(g.nx) >> 2) + (i >> 2)) * 16u + (j & 3u) * 4u + (i & 3u)
                : j * static_cast<unsigned>(g.nx) + i;
}

// Parameters of the FP64 transfer-function generator (natural FFT order).
//   a(n, N, c) = (((ic - N/2) * wavlen) / N / dx)^2 with ic = (n + N/2) mod N  (holofunc.jl:12)
//   d = (1 - a_x) - a_y;  H = exp(2 pi i frac(z/wavlen * sqrt(d)))
struct GenParams {
  double wavlen, dx;
  double zl0;   // z / wavlen of a single-plane propagation (cycles per unit sqrt(d))
};

__device__ __forceinline__ double axis_a(int n, int N, double wavlen, double dx) {
#pragma clang fp contract(off)
  const int ic = (n + (N >> 1)) % N;  // natural index -> centred index
  const double t = (static_cast<double>(ic) - 0.5 * N) * wavlen / N / dx;
  return t * t;
}

// The pass kernels evaluate the same axis term with one product by s = wavlen / N / dx (computed
// once per thread): t = (ic - N/2) * s differs from the reference's three-rounding form by at
// most two FP64 ulps, a phase change below 1e-10 cycles at |z| = 2e5 um. The two FP64
// divisions per element of the reference form were a fifth of pass A's instructions.
__device__ __forceinline__ double axis_scale(int N, double wavlen, double dx) { return wavlen / N / dx; }
// n in [0, N): the centring shift is one conditional subtraction (no integer division).
__device__ __forceinline__ double axis_a_s(int n, int N, double s) {
  int ic = n + (N >> 1);
  if (ic >= N) ic -= N;
  const double t = (static_cast<double>(ic) - 0.5 * N) * s;
  return t * t;
}

// exp(2 pi i u) for u in cycles, with u wrapped to f = u - rint(u) in [-1/2, 1/2] in FP64, then
// the hardware sine and cosine (v_sin_f32 / v_cos_f32 take their argument in revolutions; over
// [-1/2, 1/2] measured on MI355X against FP64: max abs error 1.25e-7, RMS 3.5e-8; sincospif
// gave 5.1e-8 at about 35 instructions per pair instead of 2).
__device__ __forceinline__ float2 cis_cycles(double u) {
  const float f = static_cast<float>(u - rint(u));
  return make_float2(__builtin_amdgcn_cosf(f), __builtin_amdgcn_sinf(f));
}

// ------------------------------------------------------------------------------------------
// A2 / A3 / A4: elementwise constructors (centred layout, the reference's arrays)
// ------------------------------------------------------------------------------------------
// d[x,y] = 1 - ax - ay in FP64 with the reference's op order, stored as F32 or F64
// (holofunc.jl:8-15).
template <typename T>
__global__ void k_transfer_sqrt(int nx, int ny, double wavlen, double dx, T* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const double tx = (static_cast<double>(i) - 0.5 * nx) * wavlen / nx / dx;
  const double ty = (static_cast<double>(j) - 0.5 * ny) * wavlen / ny / dx;
  const double d = (1.0 - tx * tx) - ty * ty;
  out[static_cast<size_t>(j) * nx + i] = static_cast<T>(d);
}

// H = exp(i * ((2 pi z)/wavlen) * s) rounded to ComplexF32 (holofunc.jl:38-45).
//   mode 0 (reference): s = Float32 sqrt of the Float32 d array, FP64 sincos of the full
//                       angle (the reference's arithmetic, including its FP32 sqrt).
//   mode 1: s = FP64 sqrt of the given d values (F32 or F64), phase wrapped in FP64.
//   mode 2: s = FP64 sqrt of d recomputed in FP64 from (wavlen, dx), phase wrapped in FP64.
template <typename DT>
__global__ void k_transfer(int nx, int ny, int mode, double two_pi_z_over_l, double z_over_l,
                           double wavlen, double dx, const DT* __restrict__ dsq,
                           float2* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  const size_t idx = static_cast<size_t>(j) * nx + i;
  float2 h;
  if (mode == 0) {
    const float d32 = static_cast<float>(dsq[idx]);
    const float s32 = sqrtf(d32);  // NaN for d < 0, as in the reference
    const double th = two_pi_z_over_l * static_cast<double>(s32);
    double sn, cs;
    sincos(th, &sn, &cs);
    h = make_float2(static_cast<float>(cs), static_cast<float>(sn));
  } else {
    double d;
    if (mode == 1) {
      d = static_cast<double>(dsq[idx]);
    } else {
      const double tx = (static_cast<double>(i) - 0.5 * nx) * wavlen / nx / dx;
      const double ty = (static_cast<double>(j) - 0.5 * ny) * wavlen / ny / dx;
      d = (1.0 - tx * tx) - ty * ty;
    }
    h = d >= 0.0 ? cis_cycles(z_over_l * sqrt(d)) : make_float2(0.f, 0.f);
  }
  out[idx] = h;
}

// Natural-order (ifftshifted) H(z) from FP64-wrapped phases; evanescent components 0. Writes
// columns i < w of rows j < gridDim.y with row stride w: w = nx over ny rows for the whole
// array, or the quadrant i <= nx/2, j <= ny/2 of an even-sided one. H depends on each axis
// through the square of (ic - N/2), and for even N the natural indices n and N - n carry
// opposite values of it (exactly, in FP64), so H(i, j) = H(min(i, nx - i), min(j, ny - j)).
__global__ void k_transfer_natural(int nx, int ny, int w, double zl, double wavlen, double dx,
                                   float2* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= w) return;
  const double d = (1.0 - axis_a(i, nx, wavlen, dx)) - axis_a(j, ny, wavlen, dx);
  out[static_cast<size_t>(j) * w + i] = d >= 0.0 ? cis_cycles(zl * sqrt(d)) : make_float2(0.f, 0.f);
}

// sqrt.(holo): a GS plane's amplitude, taken once per call (holofunc.jl:107-108)
__global__ void k_sqrt_plane(size_t n, const float* __restrict__ in, float* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sqrtf(in[i]);
}

// psi = ComplexF32(sqrt(holo) + 0im)  (holofunc.jl:80-82)
__global__ void k_gabor(size_t n, const float* __restrict__ holo, float2* __restrict__ out) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_float2(sqrtf(holo[i]), 0.f);
}

// ------------------------------------------------------------------------------------------
// Low-pass filters (frequency_filters.jl), square n x n, centred layout. Element (i, j) of
// the row-major array (i fast) is the reference's arr[y, x] with y = i + 1, x = j + 1. All
// arithmetic is FP64 in the reference's order, without contraction; the result is rounded
// to Float32 on store.
// ------------------------------------------------------------------------------------------
// Rectangle window (frequency_filters.jl:5-15, 31-38): 0 where
// (x - n/2 + 1)^2 + (y - n/2)^2 >= maxi^2, else 1. The two axes are centred one sample
// apart, as in the reference.
__global__ void k_lpf_rect(int n, double maxi, float* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n || j >= n) return;
  const double hx = static_cast<double>(j + 1) - 0.5 * n + 1.0;
  const double hy = static_cast<double>(i + 1) - 0.5 * n;
  out[static_cast<size_t>(j) * n + i] = (hx * hx + hy * hy >= maxi * maxi) ? 0.f : 1.f;
}

// Super-Gaussian window of order 3 (frequency_filters.jl:40-50, 66-74):
// exp(-1/2 * (ax^2 + ay^2)^3) with a = (index - n/2 + 1) / n / dx / sigma on both axes.
__global__ void k_lpf_super_gaussian(int n, double sigma, double dx, float* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n || j >= n) return;
  const double nd = static_cast<double>(n);
  const double ax = (static_cast<double>(j + 1) - 0.5 * n + 1.0) / nd / dx / sigma;
  const double ay = (static_cast<double>(i + 1) - 0.5 * n + 1.0) / nd / dx / sigma;
  const double s = ax * ax + ay * ay;
  out[static_cast<size_t>(j) * n + i] = static_cast<float>(exp(-0.5 * (s * s * s)));
}

// ------------------------------------------------------------------------------------------
// helpers shared by the pass kernels
// ------------------------------------------------------------------------------------------
// Loads of out-of-range lines (the last row or column tile of a plane) read a clamped, valid
// line instead of being skipped: a load under a per-thread branch is followed by a full wait
// for memory, which serialises a thread's loads. Those lines' results are never stored.
// Single-pass kernels issue their data loads before copying the twiddle table into LDS, so that
// one memory latency covers both; the slice-looping row passes (projection, store) stage the
// table first (staging inside their slice loop cost the projection pass 35%).
// LDS of a row pass: `rows` padded lines, then the x plan's twiddle table.
// STAGE false leaves the twiddle copy to the kernel, which then issues its global loads
// first: the copy's LDS writes wait for the table, and a single-plane launch (two waves per
// SIMD) cannot hide that latency behind other waves.
// The twiddle table is staged by every thread of the workgroup. A row FFT whose lines lie in one
// wavefront orders its LDS exchanges at wavefront scope only, so it needs one workgroup barrier
// after the staging before a stage reads another wave's part of the table; other plans get that
// barrier from their first exchange.
template <int E, int L>
__device__ __forceinline__ void twiddles_visible() {
  if constexpr (kWaveRow<E, L>) __syncthreads();
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own L2) by linear
// workgroup id. Column tiles are renumbered so that each XCD walks a contiguous range of
// tiles: neighbouring tiles, which share cache lines when W*8 bytes is below a line, meet in
// the same L2. Row tiles of a blocked G (which share its 4-row blocks when a tile holds
// fewer than 4 rows) are renumbered the same way.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int q = n / kXcds;
  return b < q * kXcds ? (b % kXcds) * q + b / kXcds : b;
}

struct RowCtx {
  int t, w, j;
  int jl;  // j clamped into the plane: loads read row jl unconditionally (see below)
  bool ok;
  float2* line;
  float2* tw;
};
template <int E, int L, bool STAGE = true>
__device__ __forceinline__ RowCtx row_ctx(const Geom& g, float2* lds) {
  RowCtx c;
  const int T = line_threads<E, L>(g.px);
  c.t = threadIdx.x % T;
  c.w = threadIdx.x / T;
  // the tile count comes from the host: reading gridDim.x here makes the compiler fetch the
  // workgroup size for a partial last workgroup with a vector load it must wait for
  c.j = (g.gblk ? xcd_tile(blockIdx.x, g.rtiles) : static_cast<int>(blockIdx.x)) * g.rows + c.w;
  c.ok = c.j < g.ny;
  c.jl = c.ok ? c.j : g.ny - 1;
  c.line = lds + c.w * lds_padded_len(g.nx, kRowPad);
  float2* tw = lds + g.rows * lds_padded_len(g.nx, kRowPad);
  if (STAGE) {
    stage_twiddles(tw, g.twx, g.px.ntw);
    twiddles_visible<E, L>();
  }
  c.tw = tw;
  return c;
}

// LDS of a column pass: W interleaved padded columns, then the y plan's twiddle table.
struct ColCtx {
  int t, w, i;
  float2* col;
  float2* tw;
};

// Column-tile accesses: element (row t + T*e, column ci) of a plane at a uniform base. The row
// term of each element folds into a scalar base pointer and the thread keeps one 32-bit offset
// t*nx + ci, so the loads and stores use scalar-base addressing instead of one 64-bit address
// register pair per element (16 pairs per thread at E = 16 held 60-120 extra VGPRs).
template <int E>
__device__ __forceinline__ void col_load(float2 (&v)[E], const float2* __restrict__ base, int nx, int T,
                                         unsigned off) {
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = (base + static_cast<size_t>(e) * T * nx)[off];
}
template <int E>
__device__ __forceinline__ void col_store(const float2 (&v)[E], float2* __restrict__ base, int nx, int T,
                                          unsigned off) {
#pragma unroll
  for (int e = 0; e < E; ++e) (base + static_cast<size_t>(e) * T * nx)[off] = v[e];
}

// The same accesses through a buffer resource (gfx9 raw buffer: base, byte range, no stride):
// one 32-bit offset register per thread, the element's row term in the scalar offset, and
// accesses past `bytes` dropped by the hardware (stores) or returning 0 (loads).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
template <int E>
__device__ __forceinline__ void col_load_buf(float2 (&v)[E], __amdgpu_buffer_rsrc_t r, unsigned voff,
                                             unsigned row_bytes) {
#pragma unroll
  for (int e = 0; e < E; ++e)
    v[e] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, e * row_bytes, 0));
}
template <int E>
__device__ __forceinline__ void col_store_buf(const float2 (&v)[E], __amdgpu_buffer_rsrc_t r, unsigned voff,
                                              unsigned row_bytes) {
#pragma unroll
  for (int e = 0; e < E; ++e)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v[e]), r, voff, e * row_bytes, 0);
}

// NBUF exchange buffers of the tile's W columns, then the twiddle table.
// Thread layout: column w = thread & (W - 1) (W adjacent columns per 64-bit access segment). By
// default the line's thread index is t = thread >> log2(W). SWZ (64-thread lines, W <= 16): the
// lane bits above log2(W) carry the same bits of t and the wave index the low log2(W) bits, so
// t's bits 4 and 5 sit in lane bits 4 and 5 and an exchange of the lane-swap form stays in
// registers (fft_col_xs). A wave then covers 64 / W rows t0, t0 + W, ... of W columns each, as
// many W*8-byte segments per access as the default layout's 64 / W consecutive rows.
// WC: the tile width as a compile-time constant (0: g.W at run time).
template <int E, int L, bool STAGE = true, int NBUF = 1, bool SWZ = false, int WC = 0>
__device__ __forceinline__ ColCtx col_ctx(const Geom& g, float2* lds) {
  ColCtx c;
  const int W = WC > 0 ? WC : g.W;
  c.w = threadIdx.x & (W - 1);
  if constexpr (SWZ) {
    static_assert(L != 0 && L / E == 64, "the swizzled layout is for 64-thread lines");
    c.t = (threadIdx.x & 63 & ~(W - 1)) | (threadIdx.x >> 6);
  } else {
    c.t = WC > 0 ? static_cast<int>(threadIdx.x) / WC : static_cast<int>(threadIdx.x >> g.wshift);
  }
  c.i = xcd_tile(blockIdx.x, gridDim.x) * W + c.w;
  c.col = lds + c.w;
  float2* tw = lds + NBUF * W * lds_padded_len(L != 0 ? L : g.ny, kColPad);
  if (STAGE) stage_twiddles(tw, g.twy, g.py.ntw);
  c.tw = tw;
  return c;
}

// ------------------------------------------------------------------------------------------
// generic 2-D FFT passes (A12; forward spectrum of A5)
// ------------------------------------------------------------------------------------------
// Row pass with an optional scale on the output. SRC 0: complex input; SRC 1: real
// intensity I, input sqrt(I) (the Gabor wavefront, holofunc.jl:81, fused into the first FFT
// pass).
template <int E, int L, int DIR, int SRC>
__global__ __launch_bounds__(512) void k_rows_fft(const void* __restrict__ in, size_t in_stride,
                                                  float2* __restrict__ out, size_t out_stride,
                                                  float scale, Geom g) {
  extern __shared__ float2 lds[];
  const RowCtx c = row_ctx<E, L, false>(g, lds);
  const int T = line_threads<E, L>(g.px);
  const size_t rowoff = static_cast<size_t>(c.jl) * g.nx;
  const size_t zin = blockIdx.y * in_stride, zout = blockIdx.y * out_stride;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = c.t + T * e;
    if (SRC == 0) {
      v[e] = static_cast<const float2*>(in)[zin + rowoff + i];
    } else {
      v[e] = make_float2(sqrtf(static_cast<const float*>(in)[zin + rowoff + i]), 0.f);
    }
  }
  stage_twiddles(c.tw, g.twx, g.px.ntw);
  twiddles_visible<E, L>();
  fft_row<E, DIR, L>(v, c.line, c.t, g.px, c.tw);
  if (c.ok) {
#pragma unroll
    for (int e = 0; e < E; ++e) out[zout + rowoff + c.t + T * e] = cscale(v[e], scale);
  }
}

// Column pass with an optional scale on the output.
template <int E, int L, int DIR>
__global__ __launch_bounds__(512) void k_cols_fft(const float2* __restrict__ in, size_t in_stride,
                                                  float2* __restrict__ out, size_t out_stride,
                                                  float scale, Geom g) {
  extern __shared__ float2 lds[];
  const ColCtx c = col_ctx<E, L, false>(g, lds);
  const int T = line_threads<E, L>(g.py);
  const bool ok = c.i < g.nx;
  const int ci = ok ? c.i : 0;
  const size_t zin = blockIdx.y * in_stride, zout = blockIdx.y * out_stride;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = in[zin + static_cast<size_t>(c.t + T * e) * g.nx + ci];
  stage_twiddles(c.tw, g.twy, g.py.ntw);
  fft_col<E, DIR, L>(v, c.col, g.W, c.t, g.py, c.tw);
  if (ok) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      out[zout + static_cast<size_t>(c.t + T * e) * g.nx + c.i] = cscale(v[e], scale);
  }
}

// ------------------------------------------------------------------------------------------
// Reconstruction pass A (column inverse FFT of F .* H_k), A6-A9
// ------------------------------------------------------------------------------------------
// HMODE 0 (generated): H_k = exp(2 pi i frac((zl0 + k*dzl) * sqrt(d))) for the hologram's slice k.
//   A workgroup owns
//   `spg` consecutive slices of a W x ny tile: it evaluates H at its first slice and the
//   per-slice rotation exp(2 pi i frac(dzl * sqrt(d))) from FP64-wrapped phases, then advances
//   F .* H_k by one ComplexF32 multiply per slice (relative drift <= spg * 1.2e-7).
// HMODE 1 (arrays, the reference's running product, holofunc.jl:254-260): the ComplexF32
//   running spectrum R_k = R_{k-1} .* Hdz (R_0 = F .* H0) is advanced in registers; the WG
//   covers every slice of the batch in order and stores R back to `run` at the end.
// G receives slice k of the batch at G + k*gstride.
struct RecArgs {
  const float2* F;       // natural-order spectrum of the wavefront
  float2* G;             // batch of column-transformed slices
  size_t gstride;        // elements between slices of G
  int nbatch;            // slices in this batch
  int spg;               // slices per workgroup (HMODE 0)
  GenParams gp;          // HMODE 0
  const float2* H0;      // HMODE 1: centred front transfer
  const float2* Hdz;     // HMODE 1: centred slice-step transfer
  float2* run;           // HMODE 1: running spectrum (natural order)
  int first;             // HMODE 1: this batch starts at slice 0 (R_0 = F .* H0)
};

// Per-hologram parameters of one round of the generated-H reconstruction. A launch covers
// up to kMaxJobs holograms of one size (blockIdx.z = entry); the table travels in the kernel
// arguments. Entries with nbatch = 0 have no slices left in this round.
constexpr int kMaxJobs = 32;
struct JobDesc {
  const float2* F;  // natural-order spectrum
  float2* G;        // this hologram's batch of column-transformed slices
  float* proj;      // its projection
  double zl0;       // z / wavlen of the hologram's slice 0
  double dzl;       // dz / wavlen
  int nbatch;       // its slices in this round
  int kbase;        // index of the round's first slice in the hologram's stack
};
struct JobTable {
  int n;
  JobDesc d[kMaxJobs];
};

__device__ __forceinline__ size_t centred_index(int i, int j, int nx, int ny) {
  const int ic = (i + (nx >> 1)) % nx;
  const int jc = (j + (ny >> 1)) % ny;
  return static_cast<size_t>(jc) * nx + ic;
}

// XB, the slice loop's column exchanges (HMODE 0 only for XB > 0):
//   0: one LDS buffer, two workgroup barriers per exchange (four per slice at 1024 = 16*16*4);
//   1: two buffers alternating by exchange, one barrier per exchange, twiddles in registers;
//   2: the last exchange by lane swaps (swizzled thread layout, fft_col_xs) and the other in a
//      buffer alternating by slice: one barrier per slice; twiddles from the LDS table;
//   3: as 2 with the twiddles in registers.
// WC: the tile width as a compile-time constant (0: g.W), which folds the LDS addresses of the
// exchange into immediate offsets.
template <int E, int L, int HMODE, int XB = 0, int WC = 0>
__global__ __launch_bounds__(512) void k_recon_cols(RecArgs a, JobTable jt, Geom g) {
  static_assert(XB == 0 || HMODE == 0, "the exchange variants are for the generated-H slice loop");
  static_assert(XB < 2 || kColXS<L, E>, "lane-swap columns need a plan of that form");
  constexpr bool DB = XB == 1;
  constexpr bool XS = XB >= 2;
  constexpr bool TWR = XB == 1 || XB == 3;
  extern __shared__ float2 lds[];
  // HMODE 0: hologram blockIdx.z of the table; workgroups past its slice count leave at once
  const JobDesc& jd = jt.d[HMODE == 0 ? blockIdx.z : 0];
  if (HMODE == 0 && static_cast<int>(blockIdx.y) * a.spg >= jd.nbatch) return;
  const float2* __restrict__ Fsrc = HMODE == 0 ? jd.F : a.F;
  static_assert(WC == 0 || XS, "compile-time tile widths are for the lane-swap variants");
  const ColCtx c = col_ctx<E, L, false, XB >= 1 ? 2 : 1, XS, WC>(g, lds);
  const int W = WC > 0 ? WC : g.W;
  const int T = line_threads<E, L>(g.py);
  const bool ok = c.i < g.nx;
  const int ci = ok ? c.i : 0;
  const unsigned off = static_cast<unsigned>(c.t * g.nx + ci);
  const unsigned offg = g_index(g, static_cast<unsigned>(c.t), static_cast<unsigned>(ci));  // G's element t, ci
  const unsigned pbytes = static_cast<unsigned>(static_cast<size_t>(g.nx) * g.ny * sizeof(float2));
  const unsigned rowb = static_cast<unsigned>(T * g.nx * sizeof(float2));
  float2 f[E];
  // TWR: the twiddles live in registers for the whole slice loop (the column exchanges are the
  // kernel's LDS traffic)
  float2 twr[TWR ? kTwRegs<L, E> : 1];
  if constexpr (HMODE == 0) {
    col_load_buf(f, plane_rsrc(Fsrc, pbytes), off * 8u, rowb);
    if constexpr (TWR) {
      load_tw_regs<E, L>(twr, c.t, g.twy);
    } else {
      stage_twiddles(c.tw, g.twy, g.py.ntw);
    }
  } else {
    col_load(f, Fsrc, g.nx, T, off);
  }

  if constexpr (HMODE == 0) {
    // f <- F .* H_{k0} and r = exp(2 pi i dzl sqrt(d)), both from FP64-wrapped phases; then
    // every further slice of the group advances f <- f .* r (one complex multiply).
    const int k0 = blockIdx.y * a.spg;
    const int k1 = min(k0 + a.spg, jd.nbatch);
    // the group's phase origin depends only on its absolute slice index, so round sizes that
    // are multiples of `spg` give bit-identical slices
    const double zl = jd.zl0 + static_cast<double>(jd.kbase + k0) * jd.dzl;
    const double dzl = jd.dzl;
    float2* __restrict__ G = jd.G;
    float2 r[E];
    {
      const double ax = axis_a_s(ci, g.nx, axis_scale(g.nx, a.gp.wavlen, a.gp.dx));
      const double sy = axis_scale(g.ny, a.gp.wavlen, a.gp.dx);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const double d = (1.0 - ax) - axis_a_s(c.t + T * e, g.ny, sy);
        if (d >= 0.0) {
          const double sq = sqrt(d);
          f[e] = cmul(f[e], cis_cycles(zl * sq));
          r[e] = cis_cycles(dzl * sq);
        } else {  // evanescent component: H = 0
          f[e] = make_float2(0.f, 0.f);
          r[e] = make_float2(0.f, 0.f);
        }
      }
    }
    if constexpr (XS) {
      // consecutive slices alternate between the two exchange buffers
      const int bo = W * lds_padded_len(L, kColPad);
      for (int k = k0; k < k1; ++k) {
        float2 v[E];
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = f[e];
        // unconditional (the group's last rotation is unused): a branch here made the compiler
        // merge two versions of f with a register copy per element and slice
#pragma unroll
        for (int e = 0; e < E; ++e) f[e] = cmul(f[e], r[e]);
        fft_col_xs<E, +1, L, TWR>(v, c.col + ((k & 1) ? bo : 0), W, c.t, c.tw, twr);
        col_store_buf(v, plane_rsrc(G + static_cast<size_t>(k) * a.gstride, pbytes), ok ? offg * 8u : pbytes, rowb);
      }
    } else {
    for (int k = k0; k < k1; ++k) {
      float2 v[E];
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = f[e];
      if (k + 1 < k1) {
#pragma unroll
        for (int e = 0; e < E; ++e) f[e] = cmul(f[e], r[e]);
      }
      if constexpr (DB) {
        fft_col_db<E, +1, L>(v, c.col, g.W, c.t, twr, g.W * lds_padded_len(g.ny, kColPad));
      } else {
        fft_col<E, +1, L>(v, c.col, g.W, c.t, g.py, c.tw);
      }
      // columns past the plane (the last tile) get an offset past the buffer: dropped stores
      col_store_buf(v, plane_rsrc(G + static_cast<size_t>(k) * a.gstride, pbytes), ok ? offg * 8u : pbytes, rowb);
    }
    }
  } else {
    float2 hdz[E];
#pragma unroll
    for (int e = 0; e < E; ++e) hdz[e] = a.Hdz[centred_index(ci, c.t + T * e, g.nx, g.ny)];
    // one branch around all loads of each kind (a branch per load costs a wait per load)
    if (a.first) {
      float2 h0[E];
#pragma unroll
      for (int e = 0; e < E; ++e) h0[e] = a.H0[centred_index(ci, c.t + T * e, g.nx, g.ny)];
      stage_twiddles(c.tw, g.twy, g.py.ntw);
#pragma unroll
      for (int e = 0; e < E; ++e) f[e] = cmul(f[e], h0[e]);
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) f[e] = a.run[static_cast<size_t>(c.t + T * e) * g.nx + ci];
      stage_twiddles(c.tw, g.twy, g.py.ntw);
    }
    for (int k = 0; k < a.nbatch; ++k) {
      if (k > 0 || !a.first) {
#pragma unroll
        for (int e = 0; e < E; ++e) f[e] = cmul(f[e], hdz[e]);
      }
      float2 v[E];
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = f[e];
      fft_col<E, +1, L>(v, c.col, g.W, c.t, g.py, c.tw);
      if (ok) {
        float2* __restrict__ dst = a.G + static_cast<size_t>(k) * a.gstride;
#pragma unroll
        for (int e = 0; e < E; ++e) dst[offg + static_cast<size_t>(T) * g.nx * e] = v[e];
      }
    }
    if (ok) {
#pragma unroll
      for (int e = 0; e < E; ++e) a.run[static_cast<size_t>(c.t + T * e) * g.nx + c.i] = f[e];
    }
  }
}

// ------------------------------------------------------------------------------------------
// Reconstruction pass B (row inverse FFT + epilogue)
// ------------------------------------------------------------------------------------------
// Pass B workgroups split the slices of a batch among themselves (blockIdx.y = slice group of
// `spg` slices) and merge their minima with an atomic min on the bit pattern. Intensities are
// >= +0, for which the IEEE bit pattern orders like an unsigned integer; the projection is
// initialised to 0x7fffffff (a NaN that compares above every other non-negative pattern), so
// the merged result equals fminf over all slices: a NaN is ignored unless every slice is NaN.
__device__ __forceinline__ void atomic_min_nonneg(float* p, float v) {
  atomicMin(reinterpret_cast<unsigned int*>(p), __float_as_uint(v));
}

__global__ void k_fill_u32(unsigned int* __restrict__ p, size_t n, unsigned int v) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Projection epilogue (A6): proj = min over slices of |psi_k|^2, psi_k = IFFT/(nx*ny). The
// minimum follows CUDA.jl's device min for Float32 (fminf: a NaN operand is ignored). The
// workgroup keeps the minimum of the unscaled |IFFT|^2 as an unsigned bit pattern (the order of
// non-negative floats, NaN above them, as in the merge below) and scales it once by
// sc2 = 1/(nx*ny)^2 before the merge; for power-of-two sizes the scaling is exact and the result
// equals the per-slice scaled form bit for bit.
// PF: the next slice's row is loaded while the current one is transformed (two row buffers
// used alternately, two slices per loop iteration), so a wave waits for memory once per
// workgroup instead of once per slice. TWR: the twiddles in registers; XS false (with TWR):
// every exchange through LDS instead of the last one by lane swaps.
template <int E, int L, bool PF = false, bool TWR = false, bool XS = true>
__global__ __launch_bounds__(512) void k_recon_rows_proj(JobTable jt, size_t gstride, int spg,
                                                         float sc2, Geom g) {
  static_assert(!TWR || L != 0, "register twiddles need a compile-time plan");
  extern __shared__ float2 lds[];
  const JobDesc& jd = jt.d[blockIdx.z];
  const int nbatch = jd.nbatch;
  if (static_cast<int>(blockIdx.y) * spg >= nbatch) return;
  const float2* __restrict__ G = jd.G;
  float* __restrict__ proj = jd.proj;
  const RowCtx c = row_ctx<E, L, !TWR>(g, lds);
  float2 twr[TWR ? kTwRegs<L, E> : 1];
  if constexpr (TWR) load_tw_regs<E, L>(twr, c.t, g.twx);
  const int T = line_threads<E, L>(g.px);
  const size_t rowoff = static_cast<size_t>(c.jl) * g.nx;
  // this thread's elements t + T*e of row jl in G: base + estride*e
  const unsigned gbase = g_index(g, static_cast<unsigned>(c.jl), static_cast<unsigned>(c.t));
  const unsigned estride = g.gblk ? 4u * T : static_cast<unsigned>(T);
  const int k0 = blockIdx.y * spg;
  const int k1 = min(k0 + spg, nbatch);
  unsigned mn[E];
#pragma unroll
  for (int e = 0; e < E; ++e) mn[e] = 0x7fffffffu;
  auto fold = [&](float2 (&v)[E]) {
    if constexpr (TWR) {
      fft_row_twr<E, +1, L, XS>(v, c.line, c.t, twr);
    } else {
      fft_row<E, +1, L>(v, c.line, c.t, g.px, c.tw);
    }
#pragma unroll
    for (int e = 0; e < E; ++e) mn[e] = min(mn[e], __float_as_uint(v[e].x * v[e].x + v[e].y * v[e].y));
  };
  if constexpr (PF) {
    // rows through a buffer resource: the slice's base in SGPRs, one 32-bit offset per thread
    const unsigned pbytes = static_cast<unsigned>(gstride * sizeof(float2));
    const unsigned voff = gbase * static_cast<unsigned>(sizeof(float2));
    auto load = [&](float2 (&v)[E], int k) {
      const __amdgpu_buffer_rsrc_t r = plane_rsrc(G + static_cast<size_t>(k) * gstride, pbytes);
#pragma unroll
      for (int e = 0; e < E; ++e)
        v[e] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, e * estride * 8, 0));
    };
    float2 a[E], b[E];
    load(a, k0);
    int k = k0;
    for (; k + 1 < k1; k += 2) {
      load(b, k + 1);
      fold(a);
      if (k + 2 < k1) load(a, k + 2);
      fold(b);
    }
    if (k < k1) fold(a);
  } else {
    for (int k = k0; k < k1; ++k) {
      float2 v[E];
      const float2* __restrict__ src = G + static_cast<size_t>(k) * gstride + gbase;
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = src[estride * e];
      fold(v);
    }
  }
  if (c.ok && k1 > k0) {
#pragma unroll
    for (int e = 0; e < E; ++e) atomic_min_nonneg(proj + rowoff + c.t + T * e, __uint_as_float(mn[e]) * sc2);
  }
}

// Store epilogue (A7 volume, A8 complex volume, A9 projection of the stored values, and
// the padded variant's centre window). OUT: 0 N0f8, 1 Float16, 2 Float32, 3 Float64,
// 4 ComplexF32. Intensities are clamp(|psi|^2, 0, 1) (holofunc.jl:158-166); N0f8 is
// round(UInt8, c*255) for c in [0, 1] with ties to even, as the reference computes it in FP64
// (holofunc.jl:152-154): the FP64 product of a float and 255 is exact. Here in FP32: p = c*255
// rounded; a half-integer is a float, so an exact product never rounds across one, and only a
// product that rounds onto k + 1/2 can round differently from the exact one; the exact FMA
// residual then tells on which side the exact product lies. Equal to the FP64 rule for every
// float in [0, 1] (checked exhaustively on the host).
__device__ __forceinline__ unsigned n0f8_round(float c) {
#pragma clang fp contract(off)
  const float p = c * 255.0f;
  float r = rintf(p);
  const float fl = floorf(p);
  if (p - fl == 0.5f) {
    const float e = fmaf(c, 255.0f, -p);
    if (e > 0.0f) r = fl + 1.0f;
    else if (e < 0.0f) r = fl;
  }
  return static_cast<unsigned>(r);
}

// A NaN intensity is stored as 0 in N0f8 (the reference raises InexactError there).
struct StoreArgs {
  const float2* G;
  size_t gstride;
  int nbatch;
  int spg;             // slices per workgroup
  void* vol;           // output volume, slice k of this batch at vol + (slice0 + k) * vstride
  size_t vstride;      // elements per output slice
  int slice0;
  float* proj;         // optional: min over slices of the stored value (A9)
  float sc;            // 1 / (nx * ny)
  int ox, oy, onx, ony;  // output window (crop) in the transform's pixel grid
};

template <int OUT>
__device__ __forceinline__ float store_value(void* vol, size_t idx, float2 psi) {
  if constexpr (OUT == 4) {
    static_cast<float2*>(vol)[idx] = psi;
    return 0.f;
  } else {
    const float in = psi.x * psi.x + psi.y * psi.y;
    const float cl = in != in ? in : fminf(fmaxf(in, 0.f), 1.f);
    if constexpr (OUT == 0) {
      // returns the stored byte: k/255 is monotonic in k, so the projection takes the min of
      // the bytes and divides once (k_recon_rows_store)
      const unsigned u = cl != cl ? 0u : n0f8_round(cl);
      static_cast<unsigned char*>(vol)[idx] = static_cast<unsigned char>(u);
      return static_cast<float>(u);
    } else if constexpr (OUT == 1) {
      const __half h = __float2half_rn(cl);
      static_cast<__half*>(vol)[idx] = h;
      return __half2float(h);
    } else if constexpr (OUT == 2) {
      static_cast<float*>(vol)[idx] = cl;
      return cl;
    } else {
      static_cast<double*>(vol)[idx] = static_cast<double>(cl);
      return cl;
    }
  }
}

// CROP: the output is a window of the transform (the padded reconstruction); without it the
// window is the whole plane and every element is stored at a constant offset from its row,
// which keeps the kernel at 4 waves per SIMD (per-element window offsets and checks cost it one
// workgroup per CU: 153 VGPRs).
template <int E, int L, int OUT, bool CROP>
__global__ __launch_bounds__(512) void k_recon_rows_store(StoreArgs a, Geom g) {
  extern __shared__ float2 lds[];
  const RowCtx c = row_ctx<E, L>(g, lds);
  const int T = line_threads<E, L>(g.px);
  const int oj = CROP ? c.j - a.oy : c.j;
  const bool rowin = CROP ? (c.ok && oj >= 0 && oj < a.ony) : c.ok;
  const int k0 = blockIdx.y * a.spg;
  const int k1 = min(k0 + a.spg, a.nbatch);
  float mn[E];
#pragma unroll
  for (int e = 0; e < E; ++e) mn[e] = __uint_as_float(0x7fffffffu);
  // this thread's elements t + T*e of row jl in G: base + estride*e
  const unsigned gbase = g_index(g, static_cast<unsigned>(c.jl), static_cast<unsigned>(c.t));
  const unsigned estride = g.gblk ? 4u * T : static_cast<unsigned>(T);
  for (int k = k0; k < k1; ++k) {
    float2 v[E];
    const float2* __restrict__ src = a.G + static_cast<size_t>(k) * a.gstride + gbase;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = c.ok ? src[estride * e] : make_float2(0.f, 0.f);
    fft_row<E, +1, L>(v, c.line, c.t, g.px, c.tw);
    if (rowin) {
      const size_t sbase = static_cast<size_t>(a.slice0 + k) * a.vstride + static_cast<size_t>(oj) * a.onx;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if constexpr (CROP) {
          const int oi = c.t + T * e - a.ox;
          if (oi >= 0 && oi < a.onx) mn[e] = fminf(mn[e], store_value<OUT>(a.vol, sbase + oi, cscale(v[e], a.sc)));
        } else {
          mn[e] = fminf(mn[e], store_value<OUT>(a.vol, sbase + c.t + T * e, cscale(v[e], a.sc)));
        }
      }
    }
  }
  if (OUT != 4 && a.proj != nullptr && rowin && k1 > k0) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int oi = CROP ? c.t + T * e - a.ox : c.t + T * e;
      // N0f8: mn holds the smallest stored byte; its value is byte / 255
      const float pv = OUT == 0 ? mn[e] / 255.0f : mn[e];
      if (!CROP || (oi >= 0 && oi < a.onx)) atomic_min_nonneg(a.proj + static_cast<size_t>(oj) * a.onx + oi, pv);
    }
  }
}

// ------------------------------------------------------------------------------------------
// Propagation column pass (GS, cu_asm_prop!): forward column FFT, .* H, inverse column FFT.
// HMODE 0: H generated (FP64-wrapped, natural order); HMODE 1: centred ComplexF32 array.
// ------------------------------------------------------------------------------------------
// Per-pair parameters of a Gerchberg-Saxton pass; a launch covers up to kMaxJobs pairs of one
// size (blockIdx.z = entry), the table travelling in the kernel arguments.
struct GsDesc {
  float2* B;          // the pair's field (transformed in place)
  const float2* H;    // natural-order transfer array of this propagation
  const float* A;     // amplitude sqrt(I) imposed by the row pass (holofunc.jl:107-108)
  float2* out;        // the retrieved wavefront (row pass with final = 1)
  int conj;           // multiply by conj(H)
  int final;          // row pass of the last iteration: store psi instead of its row FFT
  int hq;             // H holds only the quadrant i <= nx/2, j <= ny/2 (row stride nx/2 + 1) of
                      // an even-sided generated transfer (see k_transfer_natural)
};
struct GsTable {
  int n;
  GsDesc d[kMaxJobs];
};

// HMODE 2: natural-order ComplexF32 array from the GS table (precomputed once per GS call),
// multiplied as is or conjugated (H(-z) = conj(H(z)) for propagating components).
// HMODE 3: centred real Float32 filter (low-pass filter, frequency_filters.jl:89), each
// component of the spectrum scaled by it.
struct PropArgs {
  const float2* in;
  float2* out;
  GenParams gp;        // HMODE 0: zl0 = z / wavlen
  const float2* H;     // HMODE 1: centred array; HMODE 2: natural-order array
  int conj;            // HMODE 2: multiply by conj(H)
  const float* lpf;    // HMODE 3: centred real filter
};

// XB: 0 every exchange in LDS; 1 and 2 (columns of the lane-swap form, swizzled thread layout)
// each transform's last exchange by lane swaps, its other one in LDS: one buffer shared by the
// two transforms (three barriers) or one buffer each (two barriers).
template <int E, int L, int HMODE, int XB = 0>
__global__ __launch_bounds__(512) void k_prop_cols(PropArgs a, GsTable gt, Geom g) {
  static_assert(XB == 0 || kColXS<L, E>, "lane-swap columns need a plan of that form");
  extern __shared__ float2 lds[];
  const GsDesc& pd = gt.d[HMODE == 2 ? blockIdx.z : 0];
  const float2* __restrict__ src = HMODE == 2 ? pd.B : a.in;
  float2* __restrict__ dst = HMODE == 2 ? pd.B : a.out;
  const ColCtx c = col_ctx<E, L, false, XB == 2 ? 2 : 1, (XB > 0)>(g, lds);
  const int T = line_threads<E, L>(g.py);
  const bool ok = c.i < g.nx;
  const int ci = ok ? c.i : 0;
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = src[static_cast<size_t>(c.t + T * e) * g.nx + ci];
  // the transfer (or filter) values are loaded with the column, in the same memory latency
  constexpr bool kHArr = HMODE == 1 || HMODE == 2;
  float2 hv[kHArr ? E : 1];
  float lv[HMODE == 3 ? E : 1];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if constexpr (HMODE == 1) hv[e] = a.H[centred_index(ci, c.t + T * e, g.nx, g.ny)];
    if constexpr (HMODE == 2) {
      const int r = c.t + T * e;
      hv[e] = pd.hq ? pd.H[min(r, g.ny - r) * ((g.nx >> 1) + 1) + min(ci, g.nx - ci)]
                    : pd.H[static_cast<size_t>(r) * g.nx + ci];
    }
    if constexpr (HMODE == 3) lv[e] = a.lpf[centred_index(ci, c.t + T * e, g.nx, g.ny)];
  }
  stage_twiddles(c.tw, g.twy, g.py.ntw);
  if constexpr (XB > 0) {
    fft_col_xs<E, -1, L, false>(v, c.col, g.W, c.t, c.tw, nullptr);
  } else {
    fft_col<E, -1, L>(v, c.col, g.W, c.t, g.py, c.tw);
  }
  if constexpr (HMODE == 0) {
    const double ax = axis_a_s(ci, g.nx, axis_scale(g.nx, a.gp.wavlen, a.gp.dx));
    const double sy = axis_scale(g.ny, a.gp.wavlen, a.gp.dx);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const double d = (1.0 - ax) - axis_a_s(c.t + T * e, g.ny, sy);
      const float2 h = d >= 0.0 ? cis_cycles(a.gp.zl0 * sqrt(d)) : make_float2(0.f, 0.f);
      v[e] = cmul(v[e], h);
    }
  } else if constexpr (HMODE == 1) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = cmul(v[e], hv[e]);
  } else if constexpr (HMODE == 2) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = pd.conj ? cmulconj(v[e], hv[e]) : cmul(v[e], hv[e]);
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = make_float2(lv[e] * v[e].x, lv[e] * v[e].y);
  }
  if constexpr (XB == 1) {
    __syncthreads();  // every thread has read the first transform's exchange
    fft_col_xs<E, +1, L, false>(v, c.col, g.W, c.t, c.tw, nullptr);
  } else if constexpr (XB == 2) {
    fft_col_xs<E, +1, L, false>(v, c.col + g.W * lds_padded_len(g.ny, kColPad), g.W, c.t, c.tw, nullptr);
  } else {
    fft_col<E, +1, L>(v, c.col, g.W, c.t, g.py, c.tw);
  }
  if (ok) {
#pragma unroll
    for (int e = 0; e < E; ++e) dst[static_cast<size_t>(c.t + T * e) * g.nx + c.i] = v[e];
  }
}

// GS row pass (holofunc.jl:115-125) for pair blockIdx.z of the table: inverse row FFT (scaled
// 1/(nx*ny)), amplitude replacement psi <- sqrt(I) * psi/|psi| (|psi| = 0 -> sqrt(I), i.e.
// angle(0) = 0), then either the forward row FFT of the result back into B (first half of
// the next propagation) or, in the last iteration, the result itself into `out`.
template <int E, int L>
__global__ __launch_bounds__(512) void k_gs_rows(GsTable gt, float sc, Geom g) {
  extern __shared__ float2 lds[];
  const GsDesc& rd = gt.d[blockIdx.z];
  const float2* __restrict__ U = rd.B;
  const float* __restrict__ Am = rd.A;
  const bool final = rd.final != 0;
  float2* __restrict__ out = final ? rd.out : rd.B;
  const RowCtx c = row_ctx<E, L, false>(g, lds);
  const int T = line_threads<E, L>(g.px);
  const size_t rowoff = static_cast<size_t>(c.jl) * g.nx;
  float2 v[E];
  float av[E];  // the amplitude is loaded with the row, in the same memory latency
#pragma unroll
  for (int e = 0; e < E; ++e) {
    v[e] = U[rowoff + c.t + T * e];
    av[e] = Am[rowoff + c.t + T * e];
  }
  stage_twiddles(c.tw, g.twx, g.px.ntw);
  twiddles_visible<E, L>();
  fft_row<E, +1, L>(v, c.line, c.t, g.px, c.tw);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float A = av[e];
    const float a = v[e].x * sc, b = v[e].y * sc;
    const float r2 = a * a + b * b;
    if (r2 >= 1e-30f) {  // v_rsq_f32 (1 ulp) away from the denormal range
      const float f = A * __builtin_amdgcn_rsqf(r2);
      v[e] = make_float2(a * f, b * f);
    } else if (r2 > 0.f) {
      const float f = A / sqrtf(r2);
      v[e] = make_float2(a * f, b * f);
    } else {
      v[e] = make_float2(A, 0.f);
    }
  }
  if (!final) fft_row<E, -1, L>(v, c.line, c.t, g.px, c.tw);
  if (c.ok) {
#pragma unroll
    for (int e = 0; e < E; ++e) out[rowoff + c.t + T * e] = v[e];
  }
}

// Centred (fftshifted) array -> natural FFT order (ifftshift), for array-backed transfers.
__global__ void k_to_natural(int nx, int ny, const float2* __restrict__ src, float2* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= nx) return;
  dst[static_cast<size_t>(j) * nx + i] = src[centred_index(i, j, nx, ny)];
}

// ------------------------------------------------------------------------------------------
// f2: cu_2d_pad (holofunc.jl:139-149): mean-filled 2N x 2N with the input at offset N/2.
// ------------------------------------------------------------------------------------------
__global__ void k_partial_sum(const float2* __restrict__ in, size_t n, double2* __restrict__ part) {
  __shared__ double sx[256], sy[256];
  double ax = 0.0, ay = 0.0;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    ax += in[i].x;
    ay += in[i].y;
  }
  sx[threadIdx.x] = ax;
  sy[threadIdx.x] = ay;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sx[threadIdx.x] += sx[threadIdx.x + s];
      sy[threadIdx.x] += sy[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = make_double2(sx[0], sy[0]);
}

__global__ void k_pad_fill(const float2* __restrict__ in, int nx, int ny,
                           const double2* __restrict__ part, int nparts, float2* __restrict__ out) {
  double mx = 0.0, my = 0.0;
  for (int p = 0; p < nparts; ++p) {
    mx += part[p].x;
    my += part[p].y;
  }
  const double inv = 1.0 / (static_cast<double>(nx) * ny);
  const float2 mean = make_float2(static_cast<float>(mx * inv), static_cast<float>(my * inv));
  const int onx = 2 * nx, ony = 2 * ny;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= onx) return;
  const int ox = onx / 4, oy = ony / 4;
  const int si = i - ox, sj = j - oy;
  const bool inside = si >= 0 && si < nx && sj >= 0 && sj < ny;
  out[static_cast<size_t>(j) * onx + i] = inside ? in[static_cast<size_t>(sj) * nx + si] : mean;
}

}  // namespace holo
