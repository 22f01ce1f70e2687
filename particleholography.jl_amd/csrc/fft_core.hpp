// Register/LDS Stockham FFT engine for CDNA4 (gfx950), complex FP32.
//
// A line of length L is owned by T = L/E threads of one workgroup; thread t holds the
// elements n = t + T*e (e = 0..E-1) in registers. Every radix-R stage performs E/R
// butterflies per thread in registers. Between stages the values go through LDS once
// (write at the Stockham autosort position, barrier, read back in the t + T*e pattern), so a
// plan with S stages costs S-1 LDS exchanges. The first stage reads its inputs in the
// t + T*e pattern and the last stage leaves its outputs in the same pattern, which lets a
// kernel load from global memory straight into the first stage, apply an elementwise
// operation to the last stage's outputs in registers, and even chain a second FFT without an
// extra LDS round trip.
//
// Stockham autosort recursion (Van Loan, "Computational Frameworks for the FFT", 1992,
// section 1.7): stage with radix R and span p (product of earlier radices) takes butterfly b
// (0 <= b < L/R) with inputs x[b + m*L/R], k = b mod p, multiplies input m by w^(m*k),
// w = exp(-+2 pi i/(p*R)), does an R-point DFT and writes output m to
// y[(b - k)*R + k + m*p]. After the last stage the output is in natural order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace holo {

constexpr int kMaxStages = 12;

struct StagePlan {
  int R;            // radix
  int p;            // span: product of the radices of earlier stages
  uint32_t pmagic;  // ceil(2^32 / p) for p >= 2: b / p == umulhi(b, pmagic) for b < 2^16
  int twoff;        // offset (complex elements) of this stage's twiddles in the table
};

struct LinePlan {
  int L;    // line length
  int T;    // threads per line (L / E)
  int ns;   // number of stages
  int ntw;  // twiddle-table entries: sum over stages s >= 1 of (R_s - 1) * p_s = L - R_0
  StagePlan st[kMaxStages];
};

// Copies a plan's twiddle table into LDS (every thread of the workgroup participates). Kernels
// call it after issuing their first data loads; up to 8 loads per thread are in flight
// together. The first FFT stage needs no twiddles and ends with a barrier before its LDS
// writes, so the copy is complete and visible before any stage reads it.
__device__ __forceinline__ void stage_twiddles(float2* __restrict__ dst, const float2* __restrict__ src, int n) {
  const int nt = blockDim.x;
  float2 tmp[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = threadIdx.x + k * nt;
    if (i < n) tmp[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = threadIdx.x + k * nt;
    if (i < n) dst[i] = tmp[k];
  }
  for (int i = threadIdx.x + 8 * nt; i < n; i += nt) dst[i] = src[i];
}

// ------------------------------------------------------------------------------------------
// complex helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulconj(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by exp(DIR * i pi / 2): DIR = -1 -> -i, DIR = +1 -> +i
template <int DIR>
__device__ __forceinline__ float2 mul_w4(float2 a) {
  return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// ------------------------------------------------------------------------------------------
// compile-time twiddles exp(DIR * 2 pi i m / R), exact at multiples of pi/2
// ------------------------------------------------------------------------------------------
namespace cx {
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double sin_taylor(double x) {
  double t = x, s = x;
  for (int k = 1; k < 16; ++k) { t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0)); s += t; }
  return s;
}
constexpr double cos_taylor(double x) {
  double t = 1.0, s = 1.0;
  for (int k = 1; k < 16; ++k) { t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k)); s += t; }
  return s;
}
}  // namespace cx

template <int R, int DIR>
struct TwTab {
  float c[R];
  float s[R];
  constexpr TwTab() : c(), s() {
    for (int m = 0; m < R; ++m) {
      // angle = 2 pi m / R = (pi/2) * q + 2 pi f with q = round(4m/R), |f| <= 1/8
      const int q = (8 * m + R) / (2 * R);  // round(4m/R)
      const double f = (4.0 * m - q * R) / (4.0 * R);
      const double x = 2.0 * cx::kPi * f;
      double cc = cx::cos_taylor(x), ss = cx::sin_taylor(x);
      if (4 * m - q * R == 0) { cc = 1.0; ss = 0.0; }
      double cr = cc, sr = ss;  // rotate by q quarter turns
      switch (q & 3) {
        case 0: cr = cc; sr = ss; break;
        case 1: cr = -ss; sr = cc; break;
        case 2: cr = -cc; sr = -ss; break;
        default: cr = ss; sr = -cc; break;
      }
      c[m] = static_cast<float>(cr);
      s[m] = static_cast<float>(DIR < 0 ? -sr : sr);
    }
  }
};
template <int R, int DIR>
struct Tw {
  static constexpr TwTab<R, DIR> t{};
};

// multiply by exp(DIR * 2 pi i e / R); e is a constant after unrolling
template <int R, int DIR>
__device__ __forceinline__ float2 twmul(float2 v, int e) {
  e %= R;
  if (e == 0) return v;
  if (4 * e == R) return mul_w4<DIR>(v);
  if (2 * e == R) return make_float2(-v.x, -v.y);
  if (4 * e == 3 * R) return mul_w4<-DIR>(v);
  return cmul(v, make_float2(Tw<R, DIR>::t.c[e], Tw<R, DIR>::t.s[e]));
}

// ------------------------------------------------------------------------------------------
// small DFTs in registers (in place, natural-order output)
// ------------------------------------------------------------------------------------------
template <int R, int DIR>
struct DFT;

template <int DIR>
struct DFT<1, DIR> {
  __device__ __forceinline__ static void run(float2*) {}
};

template <int DIR>
struct DFT<2, DIR> {
  __device__ __forceinline__ static void run(float2* x) {
    const float2 a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  }
};

template <int DIR>
struct DFT<3, DIR> {
  __device__ __forceinline__ static void run(float2* x) {
    constexpr float kC = -0.5f;
    constexpr float kS = 0.866025403784438646763723f;  // sin(2 pi / 3)
    const float2 s = cadd(x[1], x[2]);
    const float2 d = csub(x[1], x[2]);
    const float2 m = make_float2(x[0].x + kC * s.x, x[0].y + kC * s.y);
    // DIR * i * kS * d
    const float2 r = make_float2(-DIR * kS * d.y, DIR * kS * d.x);
    x[0] = cadd(x[0], s);
    x[1] = cadd(m, r);
    x[2] = csub(m, r);
  }
};

template <int DIR>
struct DFT<4, DIR> {
  __device__ __forceinline__ static void run(float2* x) {
    const float2 t0 = cadd(x[0], x[2]);
    const float2 t1 = csub(x[0], x[2]);
    const float2 t2 = cadd(x[1], x[3]);
    const float2 t3 = mul_w4<DIR>(csub(x[1], x[3]));
    x[0] = cadd(t0, t2);
    x[2] = csub(t0, t2);
    x[1] = cadd(t1, t3);
    x[3] = csub(t1, t3);
  }
};

template <int DIR>
struct DFT<5, DIR> {
  __device__ __forceinline__ static void run(float2* x) {
    constexpr float c1 = 0.309016994374947424102293f;   // cos(2 pi / 5)
    constexpr float c2 = -0.809016994374947424102293f;  // cos(4 pi / 5)
    constexpr float s1 = 0.951056516295153572116439f;   // sin(2 pi / 5)
    constexpr float s2 = 0.587785252292473129168706f;   // sin(4 pi / 5)
    const float2 a1 = cadd(x[1], x[4]), b1 = csub(x[1], x[4]);
    const float2 a2 = cadd(x[2], x[3]), b2 = csub(x[2], x[3]);
    const float2 x0 = x[0];
    const float2 m1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    const float2 m2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    const float2 n1 = make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    const float2 n2 = make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    // DIR * i * n
    const float2 r1 = make_float2(-DIR * n1.y, DIR * n1.x);
    const float2 r2 = make_float2(-DIR * n2.y, DIR * n2.x);
    x[0] = make_float2(x0.x + a1.x + a2.x, x0.y + a1.y + a2.y);
    x[1] = cadd(m1, r1);
    x[4] = csub(m1, r1);
    x[2] = cadd(m2, r2);
    x[3] = csub(m2, r2);
  }
};

// R = A * B Cooley-Tukey in registers: n = A*n2 + n1, k = B*k1 + k2.
template <int A, int B, int DIR>
__device__ __forceinline__ void dft_composite(float2* x) {
  constexpr int R = A * B;
  float2 y[R];
#pragma unroll
  for (int n1 = 0; n1 < A; ++n1) {
    float2 t[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) t[n2] = x[A * n2 + n1];
    DFT<B, DIR>::run(t);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) y[n1 * B + k2] = twmul<R, DIR>(t[k2], n1 * k2);
  }
#pragma unroll
  for (int k2 = 0; k2 < B; ++k2) {
    float2 t[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) t[n1] = y[n1 * B + k2];
    DFT<A, DIR>::run(t);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) x[B * k1 + k2] = t[k1];
  }
}

template <int DIR>
struct DFT<8, DIR> {
  __device__ __forceinline__ static void run(float2* x) { dft_composite<2, 4, DIR>(x); }
};
template <int DIR>
struct DFT<16, DIR> {
  __device__ __forceinline__ static void run(float2* x) { dft_composite<4, 4, DIR>(x); }
};
template <int DIR>
struct DFT<6, DIR> {
  __device__ __forceinline__ static void run(float2* x) { dft_composite<2, 3, DIR>(x); }
};
template <int DIR>
struct DFT<10, DIR> {
  __device__ __forceinline__ static void run(float2* x) { dft_composite<2, 5, DIR>(x); }
};
template <int DIR>
struct DFT<12, DIR> {
  __device__ __forceinline__ static void run(float2* x) { dft_composite<4, 3, DIR>(x); }
};
template <int DIR>
struct DFT<20, DIR> {
  __device__ __forceinline__ static void run(float2* x) { dft_composite<4, 5, DIR>(x); }
};

// ------------------------------------------------------------------------------------------
// LDS addressing: one padding slot every 2^PS elements breaks the strides of the autosort
// writes across banks. Rows (one line per thread group, stride 1) use one slot in 32: their
// radix-4 stages of the 768 and 1280 plans write with element stride 4, which one slot in 16
// maps onto repeating banks (measured on the 1280 row pass: 11.1M bank-conflict cycles per
// launch with 1/16, 6.0M with 1/32, 12% shorter). Sparser slots did not help it (1/64 and
// 1/128: the same 1280 row pass) and cost the 1024 one 4% and 30%. Columns (W interleaved lines)
// keep one in 16, under which the 1024 column pass has no conflicts at all.
// ------------------------------------------------------------------------------------------
constexpr int kRowPad = 5;
constexpr int kColPad = 4;
template <int PS>
__device__ __forceinline__ int lds_pad(int n) { return n + (n >> PS); }
__host__ __device__ constexpr int lds_padded_len(int L, int PS) { return L + (L >> PS) + 1; }

// Exchange ordering between the LDS writes and reads of one stage. A line owned by the lanes of
// a single wavefront (64 % T == 0 with T threads per line, lines packed in thread order) needs
// no workgroup barrier: LDS instructions of a wave execute in issue order, so a compiler fence
// at wavefront scope keeps the reads after the writes. Lines that span waves use the workgroup
// barrier.
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham stage: E/R butterflies per thread, outputs to LDS unless this is the last.
template <int R, int E, int DIR, int PS>
__device__ __forceinline__ void fft_stage(float2 (&v)[E], const int t, const int T,
                                          const StagePlan sp, const float2* __restrict__ tw,
                                          float2* __restrict__ lds, const int es, const bool last) {
  static_assert(E % R == 0, "radix must divide the per-thread element count");
  constexpr int Q = E / R;
  int base[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    float2 x[R];
#pragma unroll
    for (int m = 0; m < R; ++m) x[m] = v[q + Q * m];
    const int b = t + T * q;
    if (sp.p > 1) {
      const int j = static_cast<int>(__umulhi(static_cast<uint32_t>(b), sp.pmagic));
      const int k = b - j * sp.p;
      const float2* __restrict__ w = tw + sp.twoff + k;
#pragma unroll
      for (int m = 1; m < R; ++m) {
        const float2 wm = w[(m - 1) * sp.p];
        x[m] = DIR < 0 ? cmul(x[m], wm) : cmulconj(x[m], wm);
      }
      base[q] = j * sp.p * R + k;
    } else {
      base[q] = b * R;
    }
    DFT<R, DIR>::run(x);
#pragma unroll
    for (int m = 0; m < R; ++m) v[q + Q * m] = x[m];
  }
  if (!last) {
    __syncthreads();  // every thread has read the previous stage's LDS contents
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int m = 0; m < R; ++m) lds[lds_pad<PS>(base[q] + m * sp.p) * es] = v[q + Q * m];
    __syncthreads();
  }
}

template <int R, int E, int DIR, int PS>
__device__ __forceinline__ void fft_stage_if(float2 (&v)[E], int t, int T, const StagePlan& sp,
                                             const float2* __restrict__ tw, float2* __restrict__ lds,
                                             int es, bool last) {
  if constexpr (E % R == 0) fft_stage<R, E, DIR, PS>(v, t, T, sp, tw, lds, es, last);
}

// Full line FFT (unnormalized). DIR = -1 forward, +1 inverse. Must be called by every thread
// of the workgroup (it contains barriers). On entry and exit v[e] holds element t + T*e.
// lds: this line's LDS base; element n lives at lds[lds_pad<PS>(n) * es].
template <int E, int DIR, int PS>
__device__ __forceinline__ void fft_line(float2 (&v)[E], float2* __restrict__ lds, const int es,
                                         const int t, const LinePlan& pl,
                                         const float2* __restrict__ tw) {
  const int T = pl.T;
  for (int s = 0; s < pl.ns; ++s) {
    const StagePlan sp = pl.st[s];
    if (s > 0) {
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = lds[lds_pad<PS>(t + T * e) * es];
    }
    const bool last = (s + 1 == pl.ns);
    switch (sp.R) {
      case 2: fft_stage_if<2, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      case 3: fft_stage_if<3, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      case 4: fft_stage_if<4, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      case 5: fft_stage_if<5, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      case 6: fft_stage_if<6, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      case 8: fft_stage_if<8, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      case 10: fft_stage_if<10, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      case 12: fft_stage_if<12, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      case 16: fft_stage_if<16, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      case 20: fft_stage_if<20, E, DIR, PS>(v, t, T, sp, tw, lds, es, last); break;
      default: break;
    }
  }
}

// ------------------------------------------------------------------------------------------
// Radix plans: at each stage the largest radix dividing both E and the remaining length. The
// host planner takes its schedule from the same function, so compile-time and runtime plans
// agree on the stages and the twiddle-table layout. (Putting the radices with an odd factor
// last was measured too: it helps the 1280 and 768 row passes with 1-in-16 padding but slows
// the interleaved 768 column pass by 20%; with the row padding below it gains nothing.)
// ------------------------------------------------------------------------------------------
constexpr int kNumRadices = 10;
constexpr int kRadixOrder[kNumRadices] = {20, 16, 12, 10, 8, 6, 5, 4, 3, 2};

struct CPlanData {
  int ns;
  int R[kMaxStages];
  int P[kMaxStages];
  int off[kMaxStages];
  int ntw;
  bool complete;  // the radices multiply to L (false: no plan within kMaxStages stages)
};

constexpr CPlanData make_cplan(int L, int E) {
  CPlanData d{};
  int radix[kMaxStages] = {};
  int n = 0, rem = L;
  while (rem > 1 && n < kMaxStages) {
    int R = 0;
    for (int i = 0; i < kNumRadices && R == 0; ++i)
      if (E % kRadixOrder[i] == 0 && rem % kRadixOrder[i] == 0) R = kRadixOrder[i];
    if (R == 0) break;
    radix[n++] = R;
    rem /= R;
  }
  d.complete = rem == 1;
  for (int i = 0; i < n; ++i) d.R[d.ns++] = radix[i];
  int span = 1, off = 0;
  for (int s = 0; s < d.ns; ++s) {
    d.P[s] = span;
    d.off[s] = off;
    if (span > 1) off += (d.R[s] - 1) * span;
    span *= d.R[s];
  }
  d.ntw = off;
  return d;
}

template <int L, int E>
struct CPlan {
  static constexpr CPlanData d = make_cplan(L, E);
  static constexpr int T = L / E;
};

// ------------------------------------------------------------------------------------------
// Exchanges by cross-lane swaps (gfx950 v_permlane32_swap_b32 / v_permlane16_swap_b32).
// v_permlane32_swap(A, B) swaps lanes 32-63 of A with lanes 0-31 of B; v_permlane16_swap(A, B)
// swaps the odd 16-lane rows of A with the even rows of B. For registers A = v[r] and
// B = v[r | 1 << a] that is a transpose of lane bit 5 (resp. 4) with register-index bit a: after
// it, register r' at lane l holds what register r' with bit a := l5 held at lane l with
// l5 := bit a of r'. In a line owned by one wavefront (T = 64: lane = t; T = 32: lane bit 5 is
// the line), an exchange whose sources differ from their destinations only in lane bits 4 and 5
// is such a transpose followed by a renaming of registers: no LDS traffic and no ordering. For
// 1024 = 16*16*4 and 512 = 16*16*2 at 16 elements per thread, the exchange in front of the last
// stage is of this kind (the first one moves lane bits 0-3 and stays in LDS).
// ------------------------------------------------------------------------------------------
constexpr int kMaxLaneSwapE = 32;
struct LaneSwap {
  bool ok;
  int rb5, rb4;                // register-index bit transposed with lane bit 5 / 4 (-1: none)
  int sigma[kMaxLaneSwapE];    // after the swaps, v[e] takes register sigma[e]
};

// Which (thread, register) holds element n after a stage of radix R and span P with Q = E/R
// butterflies per thread: the stage writes output m of butterfly b = t + T*q (k = b mod P,
// j = b / P) to element j*P*R + k + m*P from register q + Q*m.
constexpr void stage_source(int n, int R, int P, int T, int Q, int& tp, int& reg) {
  const int k = n % P, m = (n / P) % R, j = n / (P * R);
  const int b = j * P + k;
  tp = b % T;
  reg = b / T + Q * m;
}

// The swap form of the exchange into stage S (S >= 1) of the compile-time plan for (L, E), if
// there is one; found and verified for every (thread, element) at compile time.
constexpr LaneSwap make_lane_swap(int L, int E, int S) {
  LaneSwap best{};
  best.ok = false;
  const CPlanData d = make_cplan(L, E);
  const int T = L / E;
  if (!d.complete || S < 1 || S >= d.ns || E > kMaxLaneSwapE || (E & (E - 1)) != 0) return best;
  if (T != 64 && T != 32) return best;
  const int R = d.R[S - 1], P = d.P[S - 1], Q = E / R;
  int ebits = 0;
  while ((1 << ebits) < E) ++ebits;
  for (int a = -1; a < ebits; ++a) {
    for (int c = -1; c < ebits; ++c) {
      if ((a < 0 && c < 0) || (a >= 0 && a == c) || (T == 32 && a >= 0)) continue;
      const int moved = (a >= 0 ? 0x20 : 0) | (c >= 0 ? 0x10 : 0);
      const int rmask = (a >= 0 ? 1 << a : 0) | (c >= 0 ? 1 << c : 0);
      LaneSwap s{};
      s.ok = true;
      s.rb5 = a;
      s.rb4 = c;
      unsigned long long used = 0;  // registers already taken as sources (E <= 32)
      for (int e = 0; e < E && s.ok; ++e) {
        int t0 = 0, r0 = 0;
        stage_source(T * e, R, P, T, Q, t0, r0);
        if ((t0 & ~moved) != 0 || (r0 & rmask) != 0) { s.ok = false; break; }
        int sg = r0;
        if (a >= 0) sg |= ((t0 >> 5) & 1) << a;
        if (c >= 0) sg |= ((t0 >> 4) & 1) << c;
        if ((used >> sg) & 1ull) { s.ok = false; break; }
        used |= 1ull << sg;
        s.sigma[e] = sg;
        for (int t = 0; t < T && s.ok; ++t) {
          int tp = 0, r = 0;
          stage_source(t + T * e, R, P, T, Q, tp, r);
          int want_t = t & ~moved, want_r = sg & ~rmask;
          if (a >= 0) { want_t |= ((sg >> a) & 1) << 5; want_r |= ((t >> 5) & 1) << a; }
          if (c >= 0) { want_t |= ((sg >> c) & 1) << 4; want_r |= ((t >> 4) & 1) << c; }
          if (tp != want_t || r != want_r) s.ok = false;
        }
      }
      if (s.ok) return s;
    }
  }
  return best;
}

template <int L, int E, int S>
struct LaneSwapPlan {
  static constexpr LaneSwap s = make_lane_swap(L, E, S);
};

__device__ __forceinline__ void swap_lanes32(float& a, float& b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap_lanes16(float& a, float& b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  a = __uint_as_float(r[0]);
  b = __uint_as_float(r[1]);
}

// The exchange into stage S by lane swaps: on exit v[e] holds element t + T*e, as after the LDS
// read. Every lane of the wavefront must be active.
template <int L, int E, int S>
__device__ __forceinline__ void lane_swap_exchange(float2 (&v)[E]) {
  constexpr LaneSwap sw = LaneSwapPlan<L, E, S>::s;
  static_assert(sw.ok, "exchange has no lane-swap form");
  if constexpr (sw.rb5 >= 0) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      if ((r >> sw.rb5) & 1) continue;
      swap_lanes32(v[r].x, v[r | (1 << sw.rb5)].x);
      swap_lanes32(v[r].y, v[r | (1 << sw.rb5)].y);
    }
  }
  if constexpr (sw.rb4 >= 0) {
#pragma unroll
    for (int r = 0; r < E; ++r) {
      if ((r >> sw.rb4) & 1) continue;
      swap_lanes16(v[r].x, v[r | (1 << sw.rb4)].x);
      swap_lanes16(v[r].y, v[r | (1 << sw.rb4)].y);
    }
  }
  float2 u[E];
#pragma unroll
  for (int e = 0; e < E; ++e) u[e] = v[sw.sigma[e]];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = u[e];
}

// The exchange into stage S goes by lane swaps when the line's layout allows it (XS: a line owned
// by one wavefront, or a column line whose thread bits 4 and 5 sit in lane bits 4 and 5) and
// the exchange has that form.
template <int L, int E, int S, bool XS>
constexpr bool kSwapExchange = XS && L != 0 && LaneSwapPlan<L, E, S>::s.ok;

// TWR: the stage's twiddles come from registers (`twr`, R - 1 factors per butterfly q) instead of
// the LDS table.
template <int R, int E, int DIR, int T, int P, int OFF, bool LAST, int PS, bool WAVE, bool DB, bool TWR>
__device__ __forceinline__ void fft_stage_ct(float2 (&v)[E], const int t, const float2* __restrict__ tw,
                                             float2* __restrict__ lds, const int es,
                                             const float2* __restrict__ twr) {
  static_assert(E % R == 0, "radix must divide the per-thread element count");
  constexpr int Q = E / R;
  int base[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    float2 x[R];
#pragma unroll
    for (int m = 0; m < R; ++m) x[m] = v[q + Q * m];
    const int b = t + T * q;
    if constexpr (P > 1) {
      const int j = b / P;
      const int k = b - j * P;
      const float2* __restrict__ w = tw + OFF + k;
#pragma unroll
      for (int m = 1; m < R; ++m) {
        const float2 wm = TWR ? twr[q * (R - 1) + (m - 1)] : w[(m - 1) * P];
        x[m] = DIR < 0 ? cmul(x[m], wm) : cmulconj(x[m], wm);
      }
      base[q] = j * P * R + k;
    } else {
      base[q] = b * R;
    }
    DFT<R, DIR>::run(x);
#pragma unroll
    for (int m = 0; m < R; ++m) v[q + Q * m] = x[m];
  }
  if constexpr (!LAST) {
    // every thread of the line has read the buffer's previous contents (with two buffers the
    // barrier after the previous exchange's writes already ordered those reads)
    if constexpr (!DB) lds_sync<WAVE>();
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int m = 0; m < R; ++m) lds[lds_pad<PS>(base[q] + m * P) * es] = v[q + Q * m];
    lds_sync<WAVE>();
  }
}

// DB: exchanges alternate between `lds` and `lds + bo` (exchange s uses buffer s mod 2), so a
// buffer is rewritten only after a barrier that follows every read of it. With an even number
// of exchanges per line, consecutive lines of one loop keep that order too.
// Register twiddles of a compile-time plan: stage s >= 1 holds (E / R_s) * (R_s - 1) factors per
// thread, stages in order. tw_regs_before(S) is the offset of stage S's block.
template <int L, int E>
constexpr int tw_regs_before(int S) {
  constexpr CPlanData d = CPlan<L, E>::d;
  int o = 0;
  for (int s = 1; s < S && s < d.ns; ++s) o += (E / d.R[s]) * (d.R[s] - 1);
  return o;
}
template <int L, int E>
constexpr int kTwRegs = tw_regs_before<L, E>(kMaxStages) > 0 ? tw_regs_before<L, E>(kMaxStages) : 1;

// Loads thread t's register twiddles from a plan table `tw` (global or LDS).
template <int E, int L, int N>
__device__ __forceinline__ void load_tw_regs(float2 (&twr)[N], const int t, const float2* __restrict__ tw) {
  static_assert(N == kTwRegs<L, E>, "register twiddle array of the plan's size");
  constexpr CPlanData d = CPlan<L, E>::d;
  constexpr int T = CPlan<L, E>::T;
  int o = 0;
#pragma unroll
  for (int s = 1; s < d.ns; ++s) {
#pragma unroll
    for (int q = 0; q < E / d.R[s]; ++q) {
      const int k = (t + T * q) % d.P[s];
#pragma unroll
      for (int m = 1; m < d.R[s]; ++m) twr[o++] = tw[d.off[s] + k + (m - 1) * d.P[s]];
    }
  }
}

// XS: exchanges of the swap form go by lane swaps (default: lines owned by one wavefront).
template <int E, int DIR, int L, int PS, bool WAVE = false, int S = 0, bool DB = false, bool TWR = false,
          bool XS = WAVE>
__device__ __forceinline__ void fft_line_ct(float2 (&v)[E], float2* __restrict__ lds, const int es,
                                            const int t, const float2* __restrict__ tw, const int bo = 0,
                                            const float2* __restrict__ twr = nullptr) {
  constexpr CPlanData d = CPlan<L, E>::d;
  if constexpr (S < d.ns) {
    constexpr int T = CPlan<L, E>::T;
    if constexpr (S > 0 && kSwapExchange<L, E, S, XS>) {
      lane_swap_exchange<L, E, S>(v);
    } else if constexpr (S > 0) {
      const float2* __restrict__ rd = DB && ((S - 1) & 1) ? lds + bo : lds;
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = rd[lds_pad<PS>(t + T * e) * es];
    }
    float2* __restrict__ wr = DB && (S & 1) ? lds + bo : lds;
    // the stage skips its LDS writes when it is the last or the next exchange goes by lane swaps
    constexpr bool kNoWrite = S + 1 == d.ns || kSwapExchange<L, E, S + 1, XS>;
    fft_stage_ct<d.R[S], E, DIR, T, d.P[S], d.off[S], kNoWrite, PS, WAVE, DB, TWR>(
        v, t, tw, wr, es, TWR ? twr + tw_regs_before<L, E>(S) : nullptr);
    fft_line_ct<E, DIR, L, PS, WAVE, S + 1, DB, TWR, XS>(v, lds, es, t, tw, bo, twr);
  }
}

// Column plans that can run double-buffered in a slice loop: compile-time, complete, and an
// even number of exchanges (an odd stage count).
template <int L, int E>
constexpr bool kColDB = L != 0 && CPlan<L, E>::d.complete && (CPlan<L, E>::d.ns % 2) == 1;

// Line FFT with a compile-time plan when L != 0, else the runtime plan `pl`. WAVE: the line
// lies within one wavefront (wave-scope ordering instead of workgroup barriers).
template <int E, int DIR, int L, int PS, bool WAVE = false>
__device__ __forceinline__ void fft_run(float2 (&v)[E], float2* __restrict__ lds, const int es,
                                        const int t, const LinePlan& pl, const float2* __restrict__ tw) {
  if constexpr (L != 0) {
    fft_line_ct<E, DIR, L, PS, WAVE>(v, lds, es, t, tw);
  } else {
    fft_line<E, DIR, PS>(v, lds, es, t, pl, tw);
  }
}
// Rows with a compile-time plan whose T = L/E threads divide a wavefront: every line lies in
// one wave (row workgroups hold whole lines in thread order).
template <int E, int L>
constexpr bool kWaveRow = L != 0 && (64 % (L / E)) == 0;

// A row line (stride 1 in LDS) and a column line of a tile of W interleaved columns.
template <int E, int DIR, int L>
__device__ __forceinline__ void fft_row(float2 (&v)[E], float2* __restrict__ line, const int t,
                                        const LinePlan& pl, const float2* __restrict__ tw) {
  fft_run<E, DIR, L, kRowPad, kWaveRow<E, L>>(v, line, 1, t, pl, tw);
}
// A row line with the twiddles in registers (`twr`, from load_tw_regs; compile-time plans only).
// XS false keeps every exchange in LDS.
template <int E, int DIR, int L, bool XS = true>
__device__ __forceinline__ void fft_row_twr(float2 (&v)[E], float2* __restrict__ line, const int t,
                                            const float2* __restrict__ twr) {
  static_assert(L != 0, "register twiddles need a compile-time plan");
  fft_line_ct<E, DIR, L, kRowPad, kWaveRow<E, L>, 0, false, true, XS && kWaveRow<E, L>>(v, line, 1, t, nullptr, 0,
                                                                                        twr);
}
template <int E, int DIR, int L>
__device__ __forceinline__ void fft_col(float2 (&v)[E], float2* __restrict__ col, const int W, const int t,
                                        const LinePlan& pl, const float2* __restrict__ tw) {
  fft_run<E, DIR, L, kColPad>(v, col, W, t, pl, tw);
}
// A column line with two exchange buffers, the second `bo` elements after the first, and the
// twiddles in registers (`twr`, from load_tw_regs).
template <int E, int DIR, int L>
__device__ __forceinline__ void fft_col_db(float2 (&v)[E], float2* __restrict__ col, const int W, const int t,
                                           const float2* __restrict__ twr, const int bo) {
  static_assert(kColDB<L, E>, "double-buffered columns need a compile-time plan with an even number of exchanges");
  fft_line_ct<E, DIR, L, kColPad, false, 0, true, true, false>(v, col, W, t, nullptr, bo, twr);
}

// Column lines whose last exchange goes by lane swaps and whose single LDS exchange uses the
// buffer `col` chosen by the caller: compile-time plans with one line per 64 threads, one LDS
// exchange and the swap form for the rest (1024 = 16*16*4 at 16 elements per thread). The tile's
// threads must be laid out by col_ctx's swizzled mapping (thread bits 4 and 5 in lane bits 4 and
// 5). The exchange is not preceded by a barrier: a caller that alternates two buffers between
// consecutive transforms has every read of a buffer ordered before its next write by the
// barrier of the transform in between.
template <int L, int E>
constexpr bool colxs_ok() {
  if constexpr (L == 0) {
    return false;
  } else {
    constexpr CPlanData d = CPlan<L, E>::d;
    if (!d.complete || L / E != 64 || d.ns != 3) return false;
    return !LaneSwapPlan<L, E, 1>::s.ok && LaneSwapPlan<L, E, 2>::s.ok;
  }
}
template <int L, int E>
constexpr bool kColXS = colxs_ok<L, E>();
template <int E, int DIR, int L, bool TWR>
__device__ __forceinline__ void fft_col_xs(float2 (&v)[E], float2* __restrict__ col, const int W, const int t,
                                           const float2* __restrict__ tw, const float2* __restrict__ twr) {
  static_assert(kColXS<L, E>, "no lane-swap column form for this plan");
  fft_line_ct<E, DIR, L, kColPad, false, 0, true, TWR, true>(v, col, W, t, tw, 0, twr);
}

// Threads per line: a constant when the plan is compile-time.
template <int E, int L>
__device__ __forceinline__ int line_threads(const LinePlan& pl) {
  if constexpr (L != 0) return L / E;
  else return pl.T;
}

}  // namespace holo
