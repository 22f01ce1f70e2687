// libholo: host runtime and C ABI of the angular-spectrum / Gerchberg-Saxton path.
//
// The handle owns the FFT plans (radix schedule + FP64-rounded twiddle tables per line
// length), a grow-only device workspace, a capture stream and a cache of instantiated
// hipGraphs for the Gerchberg-Saxton iteration. Every entry point validates its arguments
// on the host (the reference's @assert / ArgumentError cases become error codes), then
// enqueues kernels on the caller's stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <limits>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "holo.h"
#include "holo_kernels.hpp"
#include "ingest_kernels.hpp"
#include "ccl_kernels.hpp"
#include "link_host.hpp"

using namespace holo;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HOLO_HIP(x)                                                                        \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail(HOLO_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define HOLO_TRY(x)              \
  do {                           \
    int r_ = (x);                \
    if (r_ != HOLO_OK) return r_; \
  } while (0)

struct Plan1D {
  LinePlan lp;
  int E = 0;
  float2* dtw = nullptr;
};

struct Tuning {
  int batch = 0, spg_gen = 0, spg_out = 0, cols = 0, rows = 0;
  int erow = 0, ecol = 0;  // elements per thread of row / column passes (0: the length's default)
  int col_db = 0;          // pass A exchanges: 1 two buffers, -1 one, 0 two where occupancy allows; 2 / 3 lane swaps + buffers by slice (twiddles in LDS / registers)
  int row_pf = 0;          // pass B rows loaded one slice ahead: 1 yes, 2 yes with register twiddles, 3 as 2 with every exchange in LDS, -1 / 0 no
};

// E = elements per thread; the set below is what the kernels are instantiated for.
int choose_E(int L) {
  if (L < 1 || L > 4096) return 0;
  int o = L;
  while ((o & 1) == 0) o >>= 1;
  const int cap = o == 1 ? 16 : o == 3 ? 12 : o == 5 ? 20 : 0;
  if (cap == 0) return 0;
  return L <= cap ? L : cap;
}

template <class F>
bool dispatch_E(int E, F&& f) {
  switch (E) {
#define HOLO_E(e) \
  case e:         \
    f(std::integral_constant<int, e>{}); \
    return true;
    HOLO_E(1) HOLO_E(2) HOLO_E(4) HOLO_E(8) HOLO_E(16)
    HOLO_E(3) HOLO_E(6) HOLO_E(12)
    HOLO_E(5) HOLO_E(10) HOLO_E(20)
#undef HOLO_E
    default: return false;
  }
}

// Lengths with a compile-time plan (the configured hologram sides at their default E, and the
// Gerchberg-Saxton line shapes); others use the runtime plan (512^2 GS: 32.5k iterations/s on
// runtime plans, 42.0k on compile-time ones). f receives (integral_constant<E>,
// integral_constant<L or 0>).
template <class F>
bool dispatch_EL(int E, int L, F&& f) {
#define HOLO_L(l, e)                                                          \
  if (L == l && E == e) {                                                     \
    f(std::integral_constant<int, e>{}, std::integral_constant<int, l>{});    \
    return true;                                                              \
  }
  HOLO_L(256, 16) HOLO_L(512, 16) HOLO_L(1024, 16) HOLO_L(2048, 16) HOLO_L(768, 12) HOLO_L(1280, 20)
  HOLO_L(1024, 8) HOLO_L(1024, 4) HOLO_L(512, 8) HOLO_L(512, 4) HOLO_L(2048, 8)
#undef HOLO_L
  return dispatch_E(E, [&](auto ec) { f(ec, std::integral_constant<int, 0>{}); });
}

size_t env_size(const char* name, size_t dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return static_cast<size_t>(std::strtoull(v, nullptr, 10));
}

}  // namespace

struct holo_context {
  int device = 0;
  int max_lds = 65536;
  std::map<int, Plan1D> plans;  // key (L << 8) | E
  int col_e = 0;                 // elements per thread of column passes (0: the length's default)
  int col_db = 1;                // pass A exchange buffers: HOLO_COL_DB 0 one, 1 per size, 2 two
  int row_pf = 1;                // pass B rows one slice ahead: HOLO_ROW_PF 0 never, 1 per size, 2 always
  int gs_xs = 1;                 // batched GS column passes over 1024-tall columns: HOLO_GS_XS 0 the
                                 // plane tiles' 8 elements per thread, 1 / 2 16 elements with lane-swap
                                 // exchanges and one / two LDS buffers
  int gblk_w = 4;                // widest pass-A tile whose slices are stored in 4 x 4 blocks
                                 // (HOLO_GBLK_W; 0 keeps every slice row-major)
  std::map<std::tuple<const void*, int, size_t>, int> occ;  // resident workgroups per CU, by kernel
  void* ws = nullptr;
  size_t ws_bytes = 0;
  Tuning tune;                                    // global override (tests, sweeps)
  std::map<std::pair<int, int>, Tuning> size_tune;  // per-size choices (autotuner)
  size_t batch_ws_limit = 0;                      // batched calls' workspace cap (0 = default)
  hipStream_t cap = nullptr;
  double2* partials = nullptr;  // 256 partial sums for cu_2d_pad's mean
  // The allocation-free entry point: this device's address of the static scratch array, and
  // the host halves of the plans it built there (kept so their twiddle tables outlive the
  // asynchronous copies into the scratch).
  unsigned char* scratch = nullptr;
  std::map<int, std::pair<LinePlan, std::vector<float2>>> host_plans;
  // A GS graph reads and writes only the workspace (field, transfer tables, copies of the two
  // intensities), so its key holds no caller pointer and one graph serves every call of a size.
  struct GsKey {
    int nx, ny;
    const void *buf, *hb;
    int conj, W, rows, ex, ey;
    int n;       // iterations captured
    int npairs;  // batched calls: pairs of the chunk (0 for the single-pair call)
    unsigned long long conjmask;  // batched calls: pairs whose backward table is the conjugate
    unsigned long long hqmask;    // bits 2q, 2q + 1: pair q's forward / backward table is a quadrant
    bool operator<(const GsKey& o) const {
      return std::tie(nx, ny, buf, hb, conj, W, rows, ex, ey, n, npairs, conjmask, hqmask) <
             std::tie(o.nx, o.ny, o.buf, o.hb, o.conj, o.W, o.rows, o.ex, o.ey, o.n, o.npairs, o.conjmask,
                      o.hqmask);
    }
  };
  std::map<GsKey, hipGraphExec_t> gs_graphs;
  bool gs_batch_graphs = true;  // batched GS runs its uniform iterations as cached hipGraphs
  // per-pass HIP-event timing (holo_set_profiling)
  struct ProfRec {
    int pass;
    hipEvent_t a, b;
    long long units;
  };
  bool prof = false;
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> ev_pool;
  double prof_ms[8] = {};
  long long prof_launch[8] = {}, prof_units[8] = {};
};

namespace {

hipEvent_t ev_get(holo_context* h) {
  if (!h->ev_pool.empty()) {
    hipEvent_t e = h->ev_pool.back();
    h->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Opens a timed region on `st` when profiling is on (never inside a graph capture).
hipEvent_t prof_begin(holo_context* h, hipStream_t st) {
  if (!h->prof || st == h->cap) return nullptr;
  hipEvent_t a = ev_get(h);
  if (a && hipEventRecord(a, st) != hipSuccess) return nullptr;
  return a;
}

void prof_end(holo_context* h, hipStream_t st, hipEvent_t a, int pass, long long units) {
  if (!a) return;
  hipEvent_t b = ev_get(h);
  if (!b || hipEventRecord(b, st) != hipSuccess) return;
  h->prof_pending.push_back({pass, a, b, units});
}

void prof_collect(holo_context* h) {
  for (auto& r : h->prof_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      h->prof_ms[r.pass] += ms;
      h->prof_launch[r.pass] += 1;
      h->prof_units[r.pass] += r.units;
    }
    h->ev_pool.push_back(r.a);
    h->ev_pool.push_back(r.b);
  }
  h->prof_pending.clear();
}

// Host half of a plan: the radix schedule of the device's compile-time planner and the
// FP64-computed twiddle table, rounded to float.
int build_plan(int L, int E, LinePlan* lp, std::vector<float2>* tw) {
  lp->L = L;
  lp->T = L / E;
  lp->ns = 0;
  // the schedule of the device's compile-time plans (fft_core.hpp: make_cplan)
  const CPlanData c = make_cplan(L, E);
  if (!c.complete) return fail(HOLO_ERR_SIZE, "no radix plan for length " + std::to_string(L));
  tw->clear();
  for (int s = 0; s < c.ns; ++s) {
    StagePlan& sp = lp->st[lp->ns++];
    const int R = c.R[s], span = c.P[s];
    sp.R = R;
    sp.p = span;
    sp.pmagic = span >= 2 ? static_cast<uint32_t>((0x100000000ull + span - 1) / span) : 0u;
    sp.twoff = c.off[s];
    if (span > 1) {
      const long long PR = static_cast<long long>(span) * R;
      for (int m = 1; m < R; ++m)
        for (int k = 0; k < span; ++k) {
          const long long mk = (static_cast<long long>(m) * k) % PR;
          const double a = -2.0 * M_PI * static_cast<double>(mk) / static_cast<double>(PR);
          tw->push_back(make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a))));
        }
    }
  }
  lp->ntw = c.ntw;
  if (static_cast<int>(tw->size()) != c.ntw) return fail(HOLO_ERR_SIZE, "internal: twiddle table size");
  if (tw->empty()) tw->push_back(make_float2(1.f, 0.f));
  return HOLO_OK;
}

int plan_E(int L, int E_req) {
  const int E = E_req > 0 ? E_req : choose_E(L);
  if (E == 0)
    fail(HOLO_ERR_SIZE, "FFT length " + std::to_string(L) + " not supported (need 2^a, 3*2^a or 5*2^a, <= 4096)");
  return E;
}

// Plan of a length with E elements per thread (E_req 0: the default E of the length).
int get_plan(holo_context* h, int L, const Plan1D** out, int E_req = 0) {
  const int E = plan_E(L, E_req);
  if (E == 0) return HOLO_ERR_SIZE;
  const int key = (L << 8) | E;
  auto it = h->plans.find(key);
  if (it != h->plans.end()) {
    *out = &it->second;
    return HOLO_OK;
  }
  Plan1D p;
  p.E = E;
  std::vector<float2> tw;
  HOLO_TRY(build_plan(L, E, &p.lp, &tw));
  HOLO_HIP(hipMalloc(&p.dtw, tw.size() * sizeof(float2)));
  HOLO_HIP(hipMemcpy(p.dtw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
  auto ins = h->plans.emplace(key, p);
  *out = &ins.first->second;
  return HOLO_OK;
}

// A plan without device allocation: the handle's cached plan when it has one, otherwise the
// host plan (built once, kept on the handle) with its twiddle table copied into `slot`, a
// region of the static scratch array, on `st`.
int get_plan_noalloc(holo_context* h, int L, int E_req, float2* slot, hipStream_t st, Plan1D* out) {
  const int E = plan_E(L, E_req);
  if (E == 0) return HOLO_ERR_SIZE;
  const int key = (L << 8) | E;
  auto it = h->plans.find(key);
  if (it != h->plans.end()) {
    *out = it->second;
    return HOLO_OK;
  }
  auto hp = h->host_plans.find(key);
  if (hp == h->host_plans.end()) {
    std::pair<LinePlan, std::vector<float2>> e;
    HOLO_TRY(build_plan(L, E, &e.first, &e.second));
    hp = h->host_plans.emplace(key, std::move(e)).first;
  }
  HOLO_HIP(hipMemcpyAsync(slot, hp->second.second.data(), hp->second.second.size() * sizeof(float2),
                          hipMemcpyHostToDevice, st));
  out->lp = hp->second.first;
  out->E = E;
  out->dtw = slot;
  return HOLO_OK;
}

int floor_pow2(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}
int ilog2(int v) {
  int s = 0;
  while ((1 << s) < v) ++s;
  return s;
}

// dynamic LDS: the lines of the tile, then the plan's twiddle table
size_t row_lds(const Geom& g) {
  return (static_cast<size_t>(g.rows) * lds_padded_len(g.nx, kRowPad) + g.px.ntw) * sizeof(float2);
}
size_t col_lds(const Geom& g, int nbuf = 1) {
  return (static_cast<size_t>(nbuf) * g.W * lds_padded_len(g.ny, kColPad) + g.py.ntw) * sizeof(float2);
}

// Effective tuning for a size: the per-size entry, with every nonzero field of the global
// override taking precedence.
Tuning eff_tuning(const holo_context* h, int nx, int ny) {
  Tuning t;
  auto it = h->size_tune.find({nx, ny});
  if (it != h->size_tune.end()) t = it->second;
  if (h->tune.batch > 0) t.batch = h->tune.batch;
  if (h->tune.spg_gen > 0) t.spg_gen = h->tune.spg_gen;
  if (h->tune.spg_out > 0) t.spg_out = h->tune.spg_out;
  if (h->tune.cols > 0) t.cols = h->tune.cols;
  if (h->tune.rows > 0) t.rows = h->tune.rows;
  if (h->tune.col_db != 0) t.col_db = h->tune.col_db;
  if (h->tune.row_pf != 0) t.row_pf = h->tune.row_pf;
  return t;
}

// Tiles of passes over a single plane (Gerchberg-Saxton, cu_asm_prop!, the low-pass filter
// apply, the 2-D FFT). One plane per launch: at the reconstruction's 16 elements per thread a
// 1024^2 row pass is one wave per SIMD, so these passes run their lines with fewer elements
// per thread (more waves, one more LDS exchange) and narrower tiles. Measured on MI355X (E rows / columns, iterations/s): 1024^2 E 4/8 on 4-column and
// 1-row tiles 20.9k (16.2k at the reconstruction's E 16/16, 20.3k at E 8/8); 512^2 E 4/8, 4 x 1
// 42.0k; 2048^2 E 8/16, 4 x 1 8.7k. Global tuning (holo_set_tuning) still sets the tile sizes.
// In batched GS, 1024-tall columns take 16 elements per thread with the lane-swap exchanges
// (HOLO_GS_XS 1, the default): each transform's last exchange in registers, three barriers per
// column pass instead of twelve at 8 elements; 1024^2 x 16 pairs 53.1k -> 56.6k
// pair-iterations/s. A single plane keeps 8 (a single GS pair 21.4k against 20.75k iterations/s
// with the lane-swap form; cu_asm_prop! the same either way).
Tuning plane_tiles(const holo_context* h, int nx, int ny, bool batched_gs = false) {
  Tuning t;
  if (nx == 1024 && ny == 1024) {
    t.erow = 4;
    t.ecol = batched_gs && h->gs_xs > 0 ? 0 : 8;
    t.cols = 4;
    t.rows = 1;
  } else if (nx == 512 && ny == 512) {
    t.erow = 4;
    t.ecol = 8;
    t.cols = 4;
    t.rows = 1;
  } else if (nx == 2048 && ny == 2048) {
    t.erow = 8;
    t.cols = 4;
    t.rows = 1;
  }
  if (h->tune.cols > 0) t.cols = h->tune.cols;
  if (h->tune.rows > 0) t.rows = h->tune.rows;
  return t;
}

// row_threads: target threads per row-pass workgroup when the row count is not tuned (512 for
// the reconstruction's pass B, which loops over slices; 256 elsewhere).
// Where make_geom takes its plans from when it must not allocate: the twiddle tables of plans
// the handle has not cached go to two fixed slots (x, y) of the static scratch array.
struct NoAllocPlans {
  float2* slot[2];
  hipStream_t st;
  Plan1D px, py;
};

int make_geom(holo_context* h, int nx, int ny, Geom* g, int* Ex, int* Ey, int row_threads = 256,
              const Tuning* tiles = nullptr, NoAllocPlans* na = nullptr) {
  if (nx < 1 || ny < 1) return fail(HOLO_ERR_ARG, "sizes must be positive");
  const Plan1D *px, *py;
  // passes may run with fewer elements per thread (more threads per line, fewer registers)
  // when the length allows it: E must divide L and a line must fit a workgroup
  auto fewer = [](int want, int L) {
    return (want > 0 && L % want == 0 && L / want <= 512 && want < choose_E(L)) ? want : 0;
  };
  const Tuning tn = tiles ? *tiles : eff_tuning(h, nx, ny);
  if (na) {
    HOLO_TRY(get_plan_noalloc(h, nx, fewer(tn.erow, nx), na->slot[0], na->st, &na->px));
    HOLO_TRY(get_plan_noalloc(h, ny, fewer(tn.ecol ? tn.ecol : h->col_e, ny), na->slot[1], na->st, &na->py));
    px = &na->px;
    py = &na->py;
  } else {
    HOLO_TRY(get_plan(h, nx, &px, fewer(tn.erow, nx)));
    HOLO_TRY(get_plan(h, ny, &py, fewer(tn.ecol ? tn.ecol : h->col_e, ny)));
  }
  std::memset(g, 0, sizeof(Geom));
  g->nx = nx;
  g->ny = ny;
  g->px = px->lp;
  g->py = py->lp;
  g->twx = px->dtw;
  g->twy = py->dtw;
  *Ex = px->E;
  *Ey = py->E;
  const size_t colbytes = static_cast<size_t>(lds_padded_len(ny, kColPad)) * sizeof(float2);
  const size_t coltw = static_cast<size_t>(py->lp.ntw) * sizeof(float2);
  const size_t maxl = static_cast<size_t>(h->max_lds);
  int W = tn.cols;
  if (W <= 0) {
    W = 16;
    while (W > 1 && (W * py->lp.T > 512 || W * colbytes + coltw > maxl)) W /= 2;
    W = std::min(W, floor_pow2(nx));
  }
  W = floor_pow2(std::max(W, 1));
  if (W * py->lp.T > 512 || W * colbytes + coltw > maxl)
    return fail(HOLO_ERR_ARG, "column tile too large for a workgroup");
  g->W = W;
  g->wshift = ilog2(W);
  const size_t rowbytes = static_cast<size_t>(lds_padded_len(nx, kRowPad)) * sizeof(float2);
  const size_t rowtw = static_cast<size_t>(px->lp.ntw) * sizeof(float2);
  int rows = tn.rows;
  if (rows <= 0) rows = std::max(1, row_threads / px->lp.T);
  rows = std::min(rows, ny);
  while (rows > 1 && (rows * px->lp.T > 512 || rows * rowbytes + rowtw > maxl)) --rows;
  if (rows * px->lp.T > 512 || rows * rowbytes + rowtw > maxl)
    return fail(HOLO_ERR_ARG, "row tile too large for a workgroup");
  g->rows = rows;
  g->rtiles = (ny + rows - 1) / rows;
  // the reconstruction's slices in 4 x 4 blocks where pass A's tiles are at most 4 columns wide
  // (their row segments would be 32-byte pieces of 128-byte lines)
  g->gblk = (W <= h->gblk_w && nx % 4 == 0 && ny % 4 == 0 && px->lp.T % 4 == 0 && py->lp.T % 4 == 0) ? 1 : 0;
  return HOLO_OK;
}

dim3 row_grid(const Geom& g, int y = 1) { return dim3(g.rtiles, y); }
dim3 row_block(const Geom& g) { return dim3(g.rows * g.px.T); }
dim3 col_grid(const Geom& g, int y = 1) { return dim3((g.nx + g.W - 1) / g.W, y); }
dim3 col_block(const Geom& g) { return dim3(g.W * g.py.T); }

// Grow-only workspace. Growing keeps the first `keep` bytes (a caller's data at the start of
// the workspace, e.g. the padded wavefront, across a call that needs more room). The old buffer
// may still be read by work enqueued on any stream the handle served, so growth waits for the
// device before freeing it; holo_reserve_workspace lets a caller grow it before such work.
int ensure_ws(holo_context* h, size_t bytes, size_t keep = 0) {
  if (bytes <= h->ws_bytes) return HOLO_OK;
  void* nw = nullptr;
  HOLO_HIP(hipDeviceSynchronize());
  if (hipMalloc(&nw, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(HOLO_ERR_NOMEM, "workspace allocation of " + std::to_string(bytes) + " bytes failed");
  }
  if (h->ws) {
    if (keep > 0) HOLO_HIP(hipMemcpy(nw, h->ws, std::min(keep, h->ws_bytes), hipMemcpyDeviceToDevice));
    HOLO_HIP(hipFree(h->ws));
  }
  h->ws = nw;
  h->ws_bytes = bytes;
  return HOLO_OK;
}

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(HOLO_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return HOLO_OK;
}

template <typename K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
}

// Workgroups of `kernel` resident per CU at `block` threads and `lds` bytes (cached per handle).
template <class K>
int resident(holo_context* h, K kernel, int block, size_t lds) {
  const auto key = std::make_tuple(reinterpret_cast<const void*>(kernel), block, lds);
  auto it = h->occ.find(key);
  if (it != h->occ.end()) return it->second;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(kernel), block, lds) !=
      hipSuccess)
    n = 0;
  h->occ.emplace(key, n);
  return n;
}

// ---- launch helpers ---------------------------------------------------------------------
// rows: DIR -1 forward / +1 inverse; SRC 0 complex, 1 sqrt(real)
int launch_rows_fft(const Geom& g, int Ex, int dir, int src, const void* in, size_t in_stride,
                    float2* out, size_t out_stride, float scale, int nbatch, hipStream_t st) {
  const size_t lds = row_lds(g);
  bool ok = dispatch_EL(Ex, g.nx, [&](auto ec, auto lc) {
    constexpr int E = decltype(ec)::value;
    constexpr int L = decltype(lc)::value;
    auto launch = [&](auto kern) {
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, row_grid(g, nbatch), row_block(g), lds, st, in, in_stride, out, out_stride, scale, g);
    };
    if (dir < 0 && src == 0) launch(k_rows_fft<E, L, -1, 0>);
    else if (dir < 0) launch(k_rows_fft<E, L, -1, 1>);
    else launch(k_rows_fft<E, L, +1, 0>);
  });
  if (!ok) return fail(HOLO_ERR_SIZE, "no kernel for this row length");
  return check_launch();
}

int launch_cols_fft(const Geom& g, int Ey, int dir, const float2* in, size_t in_stride, float2* out,
                    size_t out_stride, float scale, int nbatch, hipStream_t st) {
  const size_t lds = col_lds(g);
  bool ok = dispatch_EL(Ey, g.ny, [&](auto ec, auto lc) {
    constexpr int E = decltype(ec)::value;
    constexpr int L = decltype(lc)::value;
    auto launch = [&](auto kern) {
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, col_grid(g, nbatch), col_block(g), lds, st, in, in_stride, out, out_stride, scale, g);
    };
    if (dir < 0) launch(k_cols_fft<E, L, -1>);
    else launch(k_cols_fft<E, L, +1>);
  });
  if (!ok) return fail(HOLO_ERR_SIZE, "no kernel for this column length");
  return check_launch();
}

int launch_fill(float* p, size_t n, unsigned int v, hipStream_t st) {
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, st, reinterpret_cast<unsigned int*>(p), n, v);
  return check_launch();
}

int validate_transfer(const holo_tf* t, const char* name) {
  if (!t) return fail(HOLO_ERR_ARG, std::string(name) + " transfer is NULL");
  if (t->generated) {
    if (!(t->wavlen > 0) || !(t->dx > 0) || !std::isfinite(t->z))
      return fail(HOLO_ERR_ARG, std::string(name) + ": generated transfer needs finite z and positive wavlen, dx");
  } else if (!t->data) {
    return fail(HOLO_ERR_ARG, std::string(name) + ": transfer array is NULL");
  }
  return HOLO_OK;
}

// Forward 2-D spectrum of the wavefront (complex) or of sqrt(holo) (real) into F.
int forward_spectrum(const Geom& g, int Ex, int Ey, const void* src, int src_real, float2* F,
                     hipStream_t st) {
  HOLO_TRY(launch_rows_fft(g, Ex, -1, src_real, src, 0, F, 0, 1.0f, 1, st));
  return launch_cols_fft(g, Ey, -1, F, 0, F, 0, 1.0f, 1, st);
}

// 1/(nx*ny)^2: the projection pass takes the minimum of unscaled |IFFT|^2 and scales it once.
float proj_scale(int nx, int ny) {
  const double n = static_cast<double>(nx) * ny;
  return static_cast<float>(1.0 / (n * n));
}

struct ReconPlan {
  int m;        // slices per batch
  int spg_gen;  // slices per workgroup of the pass that generates H_k
  int spg_out;  // slices per workgroup of the pass that forms the output
};

ReconPlan recon_plan(holo_context* h, const Geom& g, int slices) {
  ReconPlan r;
  const size_t plane = static_cast<size_t>(g.nx) * g.ny * sizeof(float2);
  const size_t budget = env_size("HOLO_BATCH_BYTES", 256ull << 20);
  const Tuning tn = eff_tuning(h, g.nx, g.ny);
  int m = tn.batch > 0 ? tn.batch : static_cast<int>(std::max<size_t>(1, budget / plane));
  r.m = std::max(1, std::min(m, slices));
  // Generating workgroups advance up to 16 slices by the ComplexF32 rotation (drift <= 2e-6);
  // output workgroups merge 16 slices per atomic min (1024^2 x 1000: 3.52 ms against 3.58 ms
  // with 8).
  r.spg_gen = tn.spg_gen > 0 ? tn.spg_gen : 16;
  r.spg_out = tn.spg_out > 0 ? tn.spg_out : 16;
  return r;
}

// Core of A6-A9 and the padded variant. kind: 0 projection (fminf, A6); 1 store (A7-A9).
// The workspace holds [reserve bytes | F | G (batch of slices) | running spectrum].
int launch_round(holo_context* h, const Geom& g, int Ex, int Ey, const ReconPlan& rp, double wavlen,
                 double dx, const JobTable& jt, int maxnb, hipStream_t st, bool do_a = true,
                 bool do_b = true, bool timed = true);

// Pass B of a projection round: k_recon_rows_proj over (row tiles, `groups` slice groups,
// holograms of `jt`); rows prefetched one slice ahead where the size's tuning asks for it and the
// plane fits a 32-bit buffer range.
int launch_rows_proj(holo_context* h, const Geom& g, int Ex, const JobTable& jt, int spg, int groups,
                     hipStream_t st) {
  const size_t plane = static_cast<size_t>(g.nx) * g.ny;
  const int tpf = eff_tuning(h, g.nx, g.ny).row_pf;
  const bool pf = (h->row_pf == 2 || (h->row_pf == 1 && tpf > 0)) && plane * sizeof(float2) <= 0xffffffffull;
  const bool twr = pf && h->row_pf == 1 && tpf >= 2;
  const bool xs = tpf != 3;  // 3: register twiddles with every exchange through LDS
  const size_t rlds = row_lds(g);
  const float sc2 = proj_scale(g.nx, g.ny);
  dim3 grid = row_grid(g, groups);
  grid.z = jt.n;
  bool ok = dispatch_EL(Ex, g.nx, [&](auto ec, auto lc) {
    constexpr int E = decltype(ec)::value;
    constexpr int L = decltype(lc)::value;
    if constexpr (L != 0) {
      if (twr && !xs) {
        allow_lds(k_recon_rows_proj<E, L, true, true, false>, rlds);
        hipLaunchKernelGGL((k_recon_rows_proj<E, L, true, true, false>), grid, row_block(g), rlds, st, jt, plane,
                           spg, sc2, g);
        return;
      }
      if (twr) {
        allow_lds(k_recon_rows_proj<E, L, true, true>, rlds);
        hipLaunchKernelGGL((k_recon_rows_proj<E, L, true, true>), grid, row_block(g), rlds, st, jt, plane, spg, sc2, g);
        return;
      }
      if (pf) {
        allow_lds(k_recon_rows_proj<E, L, true>, rlds);
        hipLaunchKernelGGL((k_recon_rows_proj<E, L, true>), grid, row_block(g), rlds, st, jt, plane, spg, sc2, g);
        return;
      }
    }
    allow_lds(k_recon_rows_proj<E, L>, rlds);
    hipLaunchKernelGGL((k_recon_rows_proj<E, L>), grid, row_block(g), rlds, st, jt, plane, spg, sc2, g);
  });
  if (!ok) return fail(HOLO_ERR_SIZE, "no row kernel for this size");
  return check_launch();
}
int launch_gen_cols(holo_context* h, const Geom& g, int Ey, const ReconPlan& rp, double wavlen, double dx,
                    const JobTable& jt, int maxnb, hipStream_t st);

// Fixed workspace of the allocation-free entry point: `buf` (cap bytes) replaces the handle's
// workspace, rounds shrink to fit it, plans come from `plans`, and no pass is timed.
struct FixedWs {
  void* buf;
  size_t cap;
  NoAllocPlans* plans;
};

int reconstruct(holo_context* h, hipStream_t st, int nx, int ny, const void* src, int src_real,
                const holo_tf* front, const holo_tf* dz, int slices, int kind,
                int out_dtype, void* vol, float* proj, int ox, int oy, int onx, int ony,
                size_t reserve = 0, const FixedWs* fixed = nullptr) {
  HOLO_TRY(validate_transfer(front, "front"));
  HOLO_TRY(validate_transfer(dz, "dz"));
  if (front->generated != dz->generated)
    return fail(HOLO_ERR_ARG, "front and dz transfers must both be generated or both be arrays");
  if (front->generated && (front->wavlen != dz->wavlen || front->dx != dz->dx))
    return fail(HOLO_ERR_ARG, "generated transfers must share wavlen and dx");
  if (slices < 1) return fail(HOLO_ERR_ARG, "slices must be >= 1");
  if (!src) return fail(HOLO_ERR_ARG, "input is NULL");
  if ((ox != 0 || oy != 0 || onx != nx || ony != ny) && (kind != 1 || out_dtype != HOLO_N0F8))
    return fail(HOLO_ERR_ARG, "a cropped output window is an N0f8 volume");
  Geom g;
  int Ex, Ey;
  HOLO_TRY(make_geom(h, nx, ny, &g, &Ex, &Ey, 512, nullptr, fixed ? fixed->plans : nullptr));
  ReconPlan rp = recon_plan(h, g, slices);
  const size_t plane = static_cast<size_t>(nx) * ny;
  const bool arrays = !front->generated;
  void* ws = nullptr;
  if (fixed) {
    // the spectrum (and the running spectrum of array transfers) plus at least one slice
    const size_t planes = fixed->cap / (plane * sizeof(float2));
    const size_t other = 1 + (arrays ? 1 : 0);
    if (planes < other + 1)
      return fail(HOLO_ERR_BUDGET, std::to_string(nx) + " x " + std::to_string(ny) + " needs " +
                                       std::to_string((other + 1) * plane * sizeof(float2)) +
                                       " bytes of scratch; the allocation-free entry point holds " +
                                       std::to_string(fixed->cap));
    rp.m = static_cast<int>(std::min<size_t>(rp.m, planes - other));
    ws = fixed->buf;
  } else {
    const size_t need = reserve + plane * sizeof(float2) * (1 + static_cast<size_t>(rp.m) + (arrays ? 1 : 0));
    HOLO_TRY(ensure_ws(h, need, reserve));
    ws = h->ws;
  }
  float2* F = reinterpret_cast<float2*>(static_cast<char*>(ws) + reserve);
  float2* G = F + plane;
  float2* run = G + plane * rp.m;
  const size_t oplane = static_cast<size_t>(onx) * ony;
  if (proj) HOLO_TRY(launch_fill(proj, oplane, 0x7fffffffu, st));
  {
    hipEvent_t pe = fixed ? nullptr : prof_begin(h, st);
    HOLO_TRY(forward_spectrum(g, Ex, Ey, src, src_real, F, st));
    prof_end(h, st, pe, HOLO_PASS_FWD_FFT, 1);
  }
  if (kind == 0 && !arrays) {
    // generated transfers: the projection pair of passes, one round at a time
    for (int b0 = 0; b0 < slices; b0 += rp.m) {
      JobTable jt{};
      jt.n = 1;
      jt.d[0] = JobDesc{F, G, proj, front->z / front->wavlen, dz->z / front->wavlen, std::min(rp.m, slices - b0), b0};
      HOLO_TRY(launch_round(h, g, Ex, Ey, rp, front->wavlen, front->dx, jt, jt.d[0].nbatch, st, true, true,
                            fixed == nullptr));
    }
    return HOLO_OK;
  }
  const float sc = static_cast<float>(1.0 / (static_cast<double>(nx) * ny));
  const size_t clds = col_lds(g), rlds = row_lds(g);
  for (int b0 = 0; b0 < slices; b0 += rp.m) {
    const int nb = std::min(rp.m, slices - b0);
    RecArgs ra{};
    ra.F = F;
    ra.G = G;
    ra.gstride = plane;
    ra.nbatch = nb;
    ra.spg = rp.spg_gen;
    if (arrays) {
      ra.H0 = static_cast<const float2*>(front->data);
      ra.Hdz = static_cast<const float2*>(dz->data);
      ra.run = run;
      ra.first = b0 == 0;
    }
    hipEvent_t pa = fixed ? nullptr : prof_begin(h, st);
    JobTable jt{};
    jt.n = 1;
    jt.d[0].F = F;
    jt.d[0].G = G;
    jt.d[0].proj = proj;
    jt.d[0].zl0 = arrays ? 0.0 : front->z / front->wavlen;
    jt.d[0].dzl = arrays ? 0.0 : dz->z / front->wavlen;
    jt.d[0].nbatch = nb;
    jt.d[0].kbase = b0;
    if (!arrays) {
      HOLO_TRY(launch_gen_cols(h, g, Ey, rp, front->wavlen, front->dx, jt, nb, st));
    } else {
      bool ok = dispatch_EL(Ey, g.ny, [&](auto ec, auto lc) {
        constexpr int E = decltype(ec)::value;
        constexpr int L = decltype(lc)::value;
        allow_lds(k_recon_cols<E, L, 1>, clds);
        hipLaunchKernelGGL((k_recon_cols<E, L, 1>), col_grid(g, 1), col_block(g), clds, st, ra, jt, g);
      });
      if (!ok) return fail(HOLO_ERR_SIZE, "no column kernel for this size");
      HOLO_TRY(check_launch());
    }
    prof_end(h, st, pa, HOLO_PASS_RECON_COLS, nb);
    const int ry = (nb + rp.spg_out - 1) / rp.spg_out;
    hipEvent_t pb = fixed ? nullptr : prof_begin(h, st);
    if (kind == 0) {
      HOLO_TRY(launch_rows_proj(h, g, Ex, jt, rp.spg_out, ry, st));
      prof_end(h, st, pb, HOLO_PASS_RECON_ROWS, nb);
      continue;
    }
    bool ok = dispatch_EL(Ex, g.nx, [&](auto ec, auto lc) {
      constexpr int E = decltype(ec)::value;
      constexpr int L = decltype(lc)::value;
      {
        StoreArgs sa{};
        sa.G = G;
        sa.gstride = plane;
        sa.nbatch = nb;
        sa.spg = rp.spg_out;
        sa.vol = vol;
        sa.vstride = oplane;
        sa.slice0 = b0;
        sa.proj = proj;
        sa.sc = sc;
        sa.ox = ox;
        sa.oy = oy;
        sa.onx = onx;
        sa.ony = ony;
        auto launch = [&](auto kern) {
          allow_lds(kern, rlds);
          hipLaunchKernelGGL(kern, row_grid(g, ry), row_block(g), rlds, st, sa, g);
        };
        // a cropped window only comes from the padded reconstruction, whose volume is N0f8
        const bool crop = ox != 0 || oy != 0 || onx != nx || ony != ny;
        if (crop) {
          launch(k_recon_rows_store<E, L, 0, true>);
        } else {
          switch (out_dtype) {
            case HOLO_N0F8: launch(k_recon_rows_store<E, L, 0, false>); break;
            case HOLO_F16: launch(k_recon_rows_store<E, L, 1, false>); break;
            case HOLO_F32: launch(k_recon_rows_store<E, L, 2, false>); break;
            case HOLO_F64: launch(k_recon_rows_store<E, L, 3, false>); break;
            default: launch(k_recon_rows_store<E, L, 4, false>); break;
          }
        }
      }
    });
    if (!ok) return fail(HOLO_ERR_SIZE, "no row kernel for this size");
    HOLO_TRY(check_launch());
    prof_end(h, st, pb, HOLO_PASS_RECON_ROWS, nb);
  }
  return HOLO_OK;
}

// Whether the column pass over lines of length ny (at the length's default elements per thread)
// has the lane-swap variants (col_db 2 and 3).
bool col_xs_form(int ny) {
  bool xs = false;
  dispatch_EL(choose_E(ny), ny, [&](auto ec, auto lc) {
    xs = kColXS<decltype(lc)::value, decltype(ec)::value>;
  });
  return xs;
}

// Pass A of a generated-H round: the column pass that generates H_k, over (column tiles, slice
// groups, holograms).
int launch_gen_cols(holo_context* h, const Geom& g, int Ey, const ReconPlan& rp, double wavlen, double dx,
                    const JobTable& jt, int maxnb, hipStream_t st) {
  const size_t plane = static_cast<size_t>(g.nx) * g.ny;
  RecArgs ra{};
  ra.gstride = plane;
  ra.spg = rp.spg_gen;
  ra.gp.wavlen = wavlen;
  ra.gp.dx = dx;
  const size_t lds = col_lds(g), lds2 = col_lds(g, 2);
  const Tuning tn = eff_tuning(h, g.nx, g.ny);
  bool ok = dispatch_EL(Ey, g.ny, [&](auto ec, auto lc) {
    constexpr int E = decltype(ec)::value;
    constexpr int L = decltype(lc)::value;
    dim3 grid = col_grid(g, (maxnb + rp.spg_gen - 1) / rp.spg_gen);
    grid.z = jt.n;
    // the last exchange by lane swaps and the other in a buffer alternating by slice, where the
    // size's tuning asks for it (2: twiddles from LDS, 3: in registers) and the plan has that form
    if constexpr (kColXS<L, E>) {
      if (h->col_db == 1 && (tn.col_db == 2 || tn.col_db == 3) && lds2 <= static_cast<size_t>(h->max_lds)) {
        // 8-column tiles (the shipped width) with the width compiled in
        auto launch = [&](auto kern, size_t bytes) {
          allow_lds(kern, bytes);
          hipLaunchKernelGGL(kern, grid, col_block(g), bytes, st, ra, jt, g);
        };
        const size_t lx = lds2 - g.py.ntw * sizeof(float2);  // no table: twiddles in registers
        if (tn.col_db == 2) {
          if (g.W == 8) launch(k_recon_cols<E, L, 0, 2, 8>, lds2);
          else launch(k_recon_cols<E, L, 0, 2>, lds2);
        } else {
          if (g.W == 8) launch(k_recon_cols<E, L, 0, 3, 8>, lx);
          else launch(k_recon_cols<E, L, 0, 3>, lx);
        }
        return;
      }
    }
    // two exchange buffers when the plan allows it, the tile's LDS fits twice and the size's
    // tuning asks for them (by default: when they leave as many workgroups resident per CU)
    if constexpr (kColDB<L, E>) {
      bool db = (h->col_db == 2 || (h->col_db == 1 && tn.col_db >= 0)) && lds2 <= static_cast<size_t>(h->max_lds);
      if (db && h->col_db == 1 && tn.col_db == 0) {
        allow_lds(k_recon_cols<E, L, 0, 1>, lds2);
        allow_lds(k_recon_cols<E, L, 0>, lds);
        const int bs = static_cast<int>(col_block(g).x);
        db = resident(h, k_recon_cols<E, L, 0, 1>, bs, lds2) >= resident(h, k_recon_cols<E, L, 0>, bs, lds);
      }
      if (db) {
        allow_lds(k_recon_cols<E, L, 0, 1>, lds2);
        hipLaunchKernelGGL((k_recon_cols<E, L, 0, 1>), grid, col_block(g), lds2, st, ra, jt, g);
        return;
      }
    }
    allow_lds(k_recon_cols<E, L, 0>, lds);
    hipLaunchKernelGGL((k_recon_cols<E, L, 0>), grid, col_block(g), lds, st, ra, jt, g);
  });
  if (!ok) return fail(HOLO_ERR_SIZE, "no column kernel for this size");
  return check_launch();
}

// One round of the generated-H projection reconstruction for every hologram in `jt` (one size):
// pass A (the H-generating column pass) over (column tiles, slice groups, holograms), then pass B
// (k_recon_rows_proj) over (row tiles, slice groups, holograms). `maxnb` is the largest
// per-hologram slice count of the round.
int launch_round(holo_context* h, const Geom& g, int Ex, int Ey, const ReconPlan& rp, double wavlen,
                 double dx, const JobTable& jt, int maxnb, hipStream_t st, bool do_a, bool do_b, bool timed) {
  long long slices = 0;
  for (int i = 0; i < jt.n; ++i) slices += jt.d[i].nbatch;
  if (do_a) {
    hipEvent_t pa = timed ? prof_begin(h, st) : nullptr;
    HOLO_TRY(launch_gen_cols(h, g, Ey, rp, wavlen, dx, jt, maxnb, st));
    prof_end(h, st, pa, HOLO_PASS_RECON_COLS, slices);
  }
  if (do_b) {
    hipEvent_t pb = timed ? prof_begin(h, st) : nullptr;
    HOLO_TRY(launch_rows_proj(h, g, Ex, jt, rp.spg_out, (maxnb + rp.spg_out - 1) / rp.spg_out, st));
    prof_end(h, st, pb, HOLO_PASS_RECON_ROWS, slices);
  }
  return HOLO_OK;
}

// One propagation column pass over a single plane: in -> FFT_col -> .* H(t) -> IFFT_col -> out
// (H generated per element, or read from the centred array; cu_asm_prop!).
// Column pass of a propagation (transfer `t`) or, when `lpf` is given, of a low-pass filter.
// xb: the column exchanges of k_prop_cols where the plan has the lane-swap form (0 all in LDS,
// 1 lane swaps with one LDS buffer, 2 with two buffers where they fit `max_lds`).
int launch_prop_cols(const Geom& g, int Ey, const holo_tf* t, const float2* in, float2* out,
                     hipStream_t st, const float* lpf = nullptr, int xb = 0, int max_lds = 65536) {
  PropArgs pa{};
  pa.in = in;
  pa.out = out;
  const int mode = lpf ? 3 : t->generated ? 0 : 1;
  if (mode == 0) {
    pa.gp.wavlen = t->wavlen;
    pa.gp.dx = t->dx;
    pa.gp.zl0 = t->z / t->wavlen;
  } else if (mode == 1) {
    pa.H = static_cast<const float2*>(t->data);
  } else {
    pa.lpf = lpf;
  }
  GsTable gt{};
  bool ok = dispatch_EL(Ey, g.ny, [&](auto ec, auto lc) {
    constexpr int E = decltype(ec)::value;
    constexpr int L = decltype(lc)::value;
    auto launch = [&](auto kern, int nbuf) {
      const size_t lds = col_lds(g, nbuf);
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, col_grid(g), col_block(g), lds, st, pa, gt, g);
    };
    if constexpr (kColXS<L, E>) {
      if (xb == 1 || (xb == 2 && col_lds(g, 2) > static_cast<size_t>(max_lds))) {
        if (mode == 0) launch(k_prop_cols<E, L, 0, 1>, 1);
        else if (mode == 1) launch(k_prop_cols<E, L, 1, 1>, 1);
        else launch(k_prop_cols<E, L, 3, 1>, 1);
        return;
      }
      if (xb == 2) {
        if (mode == 0) launch(k_prop_cols<E, L, 0, 2>, 2);
        else if (mode == 1) launch(k_prop_cols<E, L, 1, 2>, 2);
        else launch(k_prop_cols<E, L, 3, 2>, 2);
        return;
      }
    }
    if (mode == 0) launch(k_prop_cols<E, L, 0>, 1);
    else if (mode == 1) launch(k_prop_cols<E, L, 1>, 1);
    else launch(k_prop_cols<E, L, 3>, 1);
  });
  if (!ok) return fail(HOLO_ERR_SIZE, "no column kernel for this size");
  return check_launch();
}

// GS column pass for every pair of the table (natural-order transfer arrays).
int launch_prop_cols_table(const Geom& g, int Ey, const GsTable& gt, hipStream_t st, int xb = 0,
                           int max_lds = 65536) {
  PropArgs pa{};
  bool ok = dispatch_EL(Ey, g.ny, [&](auto ec, auto lc) {
    constexpr int E = decltype(ec)::value;
    constexpr int L = decltype(lc)::value;
    dim3 grid = col_grid(g);
    grid.z = gt.n;
    auto launch = [&](auto kern, int nbuf) {
      const size_t lds = col_lds(g, nbuf);
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, grid, col_block(g), lds, st, pa, gt, g);
    };
    if constexpr (kColXS<L, E>) {
      if (xb == 1 || (xb == 2 && col_lds(g, 2) > static_cast<size_t>(max_lds))) {
        launch(k_prop_cols<E, L, 2, 1>, 1);
        return;
      }
      if (xb == 2) {
        launch(k_prop_cols<E, L, 2, 2>, 2);
        return;
      }
    }
    launch(k_prop_cols<E, L, 2>, 1);
  });
  if (!ok) return fail(HOLO_ERR_SIZE, "no column kernel for this size");
  return check_launch();
}

// GS row pass for every pair of the table.
int launch_gs_rows_table(const Geom& g, int Ex, const GsTable& gt, float sc, hipStream_t st) {
  const size_t lds = row_lds(g);
  bool ok = dispatch_EL(Ex, g.nx, [&](auto ec, auto lc) {
    constexpr int E = decltype(ec)::value;
    constexpr int L = decltype(lc)::value;
    allow_lds(k_gs_rows<E, L>, lds);
    dim3 grid = row_grid(g);
    grid.z = gt.n;
    hipLaunchKernelGGL((k_gs_rows<E, L>), grid, row_block(g), lds, st, gt, sc, g);
  });
  if (!ok) return fail(HOLO_ERR_SIZE, "no row kernel for this size");
  return check_launch();
}

// State of one pair during GS: its field B, natural-order transfers and iteration count.
struct GsPair {
  float2* B;
  const float2* hf;
  const float2* hb;
  int conj_b;
  int hq_f, hq_b;  // hf / hb hold the quadrant of an even-sided generated transfer
  const float* A1;  // amplitudes sqrt(I1), sqrt(I2) in the workspace
  const float* A2;
  float2* out;
  int iters;
};

// Iteration `it` of every pair in [p0, p1) that still has iterations left: column pass
// (+z), row pass with I2, column pass (-z), row pass with I1 (final for a pair's last one).
int gs_round(holo_context* h, const Geom& g, int Ex, int Ey, const GsPair* pairs, int np, int it,
             float sc, hipStream_t st) {
  GsTable f{}, r2{}, b{}, r1{};
  for (int q = 0; q < np; ++q) {
    const GsPair& p = pairs[q];
    if (p.iters <= it) continue;
    const int k = f.n;
    f.d[k] = GsDesc{p.B, p.hf, nullptr, nullptr, 0, 0, p.hq_f};
    r2.d[k] = GsDesc{p.B, nullptr, p.A2, nullptr, 0, 0, 0};
    b.d[k] = GsDesc{p.B, p.hb, nullptr, nullptr, p.conj_b, 0, p.hq_b};
    const bool last = it + 1 == p.iters;
    r1.d[k] = GsDesc{p.B, nullptr, p.A1, last ? p.out : nullptr, 0, last ? 1 : 0, 0};
    f.n = r2.n = b.n = r1.n = k + 1;
  }
  if (f.n == 0) return HOLO_OK;
  const GsTable* cols[2] = {&f, &b};
  const GsTable* rows[2] = {&r2, &r1};
  for (int half = 0; half < 2; ++half) {
    hipEvent_t pc = prof_begin(h, st);
    HOLO_TRY(launch_prop_cols_table(g, Ey, *cols[half], st, h->gs_xs, h->max_lds));
    prof_end(h, st, pc, HOLO_PASS_PROP_COLS, cols[half]->n);
    hipEvent_t pr = prof_begin(h, st);
    HOLO_TRY(launch_gs_rows_table(g, Ex, *rows[half], sc, st));
    prof_end(h, st, pr, HOLO_PASS_GS_ROWS, rows[half]->n);
  }
  return HOLO_OK;
}

// The amplitudes sqrt(I) a GS call imposes, into its workspace: the graphs then reference
// workspace addresses only, whatever the caller passes.
int launch_sqrt_plane(const void* in, float* out, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_sqrt_plane, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, n,
                     static_cast<const float*>(in), out);
  return check_launch();
}

// Natural-order transfer arrays of a pair into hf (and hbuf unless mirrored): generated
// transfers are evaluated once (H(-z) = conj H(z) for a mirrored pair), arrays are ifftshifted.
// A generated transfer of even sides is stored as its quadrant (nx/2 + 1) x (ny/2 + 1), which
// determines the rest (k_transfer_natural): the column pass reads 2 MB instead of 8 at 1024^2,
// and a batch's tables stay in the caches.
int gs_transfers(const Geom& g, const holo_tf* fwd, const holo_tf* bwd, GsPair* p, float2* hf, float2* hbuf,
                 hipStream_t st) {
  const dim3 block(256);
  const bool mirrored = fwd->generated && bwd->generated && bwd->z == -fwd->z &&
                        bwd->wavlen == fwd->wavlen && bwd->dx == fwd->dx;
  const bool even = g.nx % 2 == 0 && g.ny % 2 == 0;
  auto make = [&](const holo_tf* t, float2* dst) {
    if (t->generated && even) {
      const int w = g.nx / 2 + 1;
      hipLaunchKernelGGL(k_transfer_natural, dim3((w + 255) / 256, g.ny / 2 + 1), block, 0, st, g.nx, g.ny, w,
                         t->z / t->wavlen, t->wavlen, t->dx, dst);
      return 1;
    }
    const dim3 grid((g.nx + 255) / 256, g.ny);
    if (t->generated)
      hipLaunchKernelGGL(k_transfer_natural, grid, block, 0, st, g.nx, g.ny, g.nx, t->z / t->wavlen, t->wavlen,
                         t->dx, dst);
    else
      hipLaunchKernelGGL(k_to_natural, grid, block, 0, st, g.nx, g.ny, static_cast<const float2*>(t->data), dst);
    return 0;
  };
  p->hf = hf;
  p->hq_f = make(fwd, hf);
  HOLO_TRY(check_launch());
  if (mirrored) {
    p->hb = p->hf;
    p->hq_b = p->hq_f;
    p->conj_b = 1;
  } else {
    p->hq_b = make(bwd, hbuf);
    HOLO_TRY(check_launch());
    p->hb = hbuf;
    p->conj_b = 0;
  }
  return HOLO_OK;
}

}  // namespace

// Allocation-free reconstruction: a static device array serves as its workspace. Each device
// holds its own copy (kScratchBytes, from the library's load on that device until the process
// ends); the host-side symbol is not a device address, so the address comes from
// hipGetSymbolAddress, once per device with that device current (holo_create), and only that
// address is ever passed to kernels or copies. Layout: [x twiddles | y twiddles | workspace].
constexpr size_t kScratchBytes = 128ull << 20;
constexpr size_t kScratchTwBytes = 4096 * sizeof(float2);  // one axis's table: < 4096 entries
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

int scratch_address(int device, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return fail(HOLO_ERR_ARG, "device index beyond the scratch table");
  if (!g_scratch[device]) {
    void* p = nullptr;
    HOLO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[device];
  return HOLO_OK;
}
}  // namespace

// =========================================================================================
// C ABI
// =========================================================================================
extern "C" {

int holo_version(void) { return 1; }

const char* holo_last_error(void) { return g_err.c_str(); }

int holo_supported_length(int n) { return choose_E(n) != 0 ? 1 : 0; }

int holo_create(holo_handle* out, int device) {
  if (!out) return fail(HOLO_ERR_ARG, "out is NULL");
  *out = nullptr;
  int count = 0;
  HOLO_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(HOLO_ERR_ARG, "device index out of range");
  HOLO_HIP(hipSetDevice(device));
  holo_context* h = new holo_context();
  h->device = device;
  int lds = 65536;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    h->max_lds = lds;
  const size_t cap = env_size("HOLO_MAX_LDS", 0);
  if (cap > 0) h->max_lds = static_cast<int>(std::min<size_t>(cap, h->max_lds));
  h->col_e = static_cast<int>(env_size("HOLO_COL_E", 0));
  h->col_db = static_cast<int>(env_size("HOLO_COL_DB", 1));
  h->row_pf = static_cast<int>(env_size("HOLO_ROW_PF", 1));
  h->gs_xs = static_cast<int>(env_size("HOLO_GS_XS", 1));
  h->gblk_w = static_cast<int>(env_size("HOLO_GBLK_W", 4));
  // Per-size tiles chosen by holo_autotune on an MI355X and confirmed on single-hologram
  // reconstructions (768^2 and 1280x1024 +37%, 512^2 +5% over the generic rule). 2048^2 comes
  // from a direct sweep: no useful round fits the 256 MB Infinity Cache at that size, so long
  // rounds (64 slices, 2 GB) amortise each launch's tail (53.8k slices/s against 48.9k with 16),
  // and its pass A keeps one exchange buffer (two: 3.88 ms against 3.76 per 200 slices). Pass B
  // loads rows one slice ahead at 1024^2 (3.34 ms against 3.46 per 1000 slices; with the
  // twiddles in registers 3.30) and 1280x1024 (2.86 against 3.17 per 500; its register twiddles
  // would spill); at 512^2, 768^2 and 2048^2 it was 0.2-1.4% slower. 1280x1024 re-tuned with
  // both variants available: 8-column tiles, 16-slice groups, two exchange buffers (2.72 ms
  // against 2.92 per 500 slices). Sides whose columns are 1024 tall run pass A with its last
  // exchange by lane swaps and one barrier per slice (col_db 3): 1024^2 3.23 -> 3.11-3.14 ms per
  // 1000 slices (pass A 56.7 -> 52.7 us per 32-slice launch), 1280x1024 2.52 -> 2.39 ms per 500;
  // with the twiddles from LDS instead (col_db 2) 3.19 and 2.45. Pass B at 1024^2 keeps both row
  // exchanges in LDS (row_pf 3: 46.2 against 46.9 us per launch with the last one by lane swaps).
  // Pass A over 1024-tall columns holds one 512-thread workgroup per CU, so its workgroup count
  // should fill whole waves of 256: at 1280 x 1024 the 160 eight-column tiles in 2 groups of 16
  // slices made 320 (1.25 waves); rounds of 24 slices in groups of 8 make 480, and the
  // projection's groups of 12 split the round evenly: 2.51-2.61 -> 2.29-2.35 ms per 500 slices.
  // 2048^2 pass A generates H for its whole 64-slice round in one group (512 workgroups): 3.10-3.15
  // -> 2.92 ms per 192 slices (groups of 32: 2.96-2.99). Shorter rounds that would stay in the
  // Infinity Cache lose more to launch tails (rounds of 8 or 16: 3.66-3.83 ms). 768^2 reads its
  // blocked slices in 4-row tiles (whole blocks: pass B 0.53 -> 0.49-0.50 ms per 500 slices) and
  // runs rounds of 48 in four full groups of 12: 1.09-1.20 -> 1.04-1.05 ms per 500 slices.
  // holo_set_tuning_for_size / holo_autotune replace them.
  {
    struct Shipped {
      int nx, ny, batch, spg_gen, spg_out, cols, rows, erow = 0, ecol = 0, col_db = 0, row_pf = 0;
    };
    static const Shipped kShipped[] = {{512, 512, 0, 0, 0, 8, 0},
                                       {1024, 1024, 0, 0, 0, 0, 0, 0, 0, 3, 3},
                                       {768, 768, 48, 12, 4, 4, 4},
                                       {1280, 1024, 24, 8, 12, 8, 4, 0, 0, 3, 1},
                                       {2048, 2048, 64, 64, 16, 4, 2, 0, 0, -1}};
    for (const Shipped& e : kShipped) {
      Tuning t;
      t.batch = e.batch;
      t.spg_gen = e.spg_gen;
      t.spg_out = e.spg_out;
      t.cols = e.cols;
      t.rows = e.rows;
      t.erow = e.erow;
      t.ecol = e.ecol;
      t.col_db = e.col_db;
      t.row_pf = e.row_pf;
      h->size_tune[{e.nx, e.ny}] = t;
    }
  }
  if (hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(HOLO_ERR_HIP, "capture stream creation failed");
  }
  if (hipMalloc(&h->partials, 256 * sizeof(double2)) != hipSuccess) {
    (void)hipStreamDestroy(h->cap);
    delete h;
    return fail(HOLO_ERR_NOMEM, "partial-sum buffer allocation failed");
  }
  if (const int r = scratch_address(device, &h->scratch); r != HOLO_OK) {
    (void)hipFree(h->partials);
    (void)hipStreamDestroy(h->cap);
    delete h;
    return r;
  }
  *out = h;
  return HOLO_OK;
}

int holo_destroy(holo_handle h) {
  if (!h) return HOLO_OK;
  (void)hipSetDevice(h->device);
  (void)hipDeviceSynchronize();
  for (auto& kv : h->gs_graphs) (void)hipGraphExecDestroy(kv.second);
  prof_collect(h);
  for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
  for (auto& kv : h->plans) (void)hipFree(kv.second.dtw);
  if (h->ws) (void)hipFree(h->ws);
  if (h->cap) (void)hipStreamDestroy(h->cap);
  if (h->partials) (void)hipFree(h->partials);
  delete h;
  return HOLO_OK;
}

int holo_set_tuning(holo_handle h, int batch, int spg_gen, int spg_out, int cols_per_wg,
                    int rows_per_wg) {
  if (!h) return fail(HOLO_ERR_ARG, "handle is NULL");
  if (cols_per_wg > 0 && (cols_per_wg & (cols_per_wg - 1)) != 0)
    return fail(HOLO_ERR_ARG, "cols_per_wg must be a power of two");
  h->tune.batch = batch;
  h->tune.spg_gen = spg_gen;
  h->tune.spg_out = spg_out;
  h->tune.cols = cols_per_wg;
  h->tune.rows = rows_per_wg;  // cached GS graphs are keyed by their tiles and stay valid
  return HOLO_OK;
}

int holo_transfer_sqrt(holo_handle h, void* stream, int nx, int ny, double wavlen, double dx,
                       int dtype, void* out) {
  if (!h || !out) return fail(HOLO_ERR_ARG, "NULL handle or output");
  if (nx < 1 || ny < 1) return fail(HOLO_ERR_ARG, "sizes must be positive");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid((nx + 255) / 256, ny), block(256);
  if (dtype == HOLO_F32)
    hipLaunchKernelGGL(k_transfer_sqrt<float>, grid, block, 0, st, nx, ny, wavlen, dx, static_cast<float*>(out));
  else if (dtype == HOLO_F64)
    hipLaunchKernelGGL(k_transfer_sqrt<double>, grid, block, 0, st, nx, ny, wavlen, dx, static_cast<double*>(out));
  else
    return fail(HOLO_ERR_TYPE, "transfer_sqrt dtype must be HOLO_F32 or HOLO_F64");
  return check_launch();
}

int holo_transfer(holo_handle h, void* stream, int nx, int ny, double z, double wavlen,
                  const void* d_sqr, int d_dtype, int precision, double dx, void* out) {
  if (!h || !out) return fail(HOLO_ERR_ARG, "NULL handle or output");
  if (nx < 1 || ny < 1) return fail(HOLO_ERR_ARG, "sizes must be positive");
  if (precision != HOLO_TF_FP64 && !d_sqr) return fail(HOLO_ERR_ARG, "d_sqr is NULL");
  if (precision == HOLO_TF_FP64 && !(dx > 0)) return fail(HOLO_ERR_ARG, "dx must be positive");
  if (precision < 0 || precision > 2) return fail(HOLO_ERR_ARG, "bad precision");
  if (d_dtype != HOLO_F32 && d_dtype != HOLO_F64) return fail(HOLO_ERR_TYPE, "d_sqr dtype must be HOLO_F32 or HOLO_F64");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid((nx + 255) / 256, ny), block(256);
  const double two_pi_z_over_l = (2.0 * M_PI * z) / wavlen;
  const double z_over_l = z / wavlen;
  if (d_dtype == HOLO_F32)
    hipLaunchKernelGGL(k_transfer<float>, grid, block, 0, st, nx, ny, precision, two_pi_z_over_l, z_over_l,
                       wavlen, dx, static_cast<const float*>(d_sqr), static_cast<float2*>(out));
  else
    hipLaunchKernelGGL(k_transfer<double>, grid, block, 0, st, nx, ny, precision, two_pi_z_over_l, z_over_l,
                       wavlen, dx, static_cast<const double*>(d_sqr), static_cast<float2*>(out));
  return check_launch();
}

int holo_gabor_wavefront(holo_handle h, void* stream, int nx, int ny, const float* holo, void* out) {
  if (!h || !holo || !out) return fail(HOLO_ERR_ARG, "NULL argument");
  if (nx < 1 || ny < 1) return fail(HOLO_ERR_ARG, "sizes must be positive");
  const size_t n = static_cast<size_t>(nx) * ny;
  hipLaunchKernelGGL(k_gabor, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, holo, static_cast<float2*>(out));
  return check_launch();
}

int holo_fft2(holo_handle h, void* stream, int nx, int ny, const void* in, void* out, int inverse) {
  if (!h || !in || !out) return fail(HOLO_ERR_ARG, "NULL argument");
  Geom g;
  int Ex, Ey;
  const Tuning pt = plane_tiles(h, nx, ny);
  HOLO_TRY(make_geom(h, nx, ny, &g, &Ex, &Ey, 256, &pt));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int dir = inverse ? +1 : -1;
  const float scale = inverse ? static_cast<float>(1.0 / (static_cast<double>(nx) * ny)) : 1.0f;
  float2* o = static_cast<float2*>(out);
  HOLO_TRY(launch_rows_fft(g, Ex, dir, 0, in, 0, o, 0, 1.0f, 1, st));
  return launch_cols_fft(g, Ey, dir, o, 0, o, 0, scale, 1, st);
}

int holo_asm_prop(holo_handle h, void* stream, int nx, int ny, void* out, const void* in,
                  const holo_tf* tf) {
  if (!h || !in || !out) return fail(HOLO_ERR_ARG, "NULL argument");
  HOLO_TRY(validate_transfer(tf, "tf"));
  Geom g;
  int Ex, Ey;
  const Tuning pt = plane_tiles(h, nx, ny);
  HOLO_TRY(make_geom(h, nx, ny, &g, &Ex, &Ey, 256, &pt));
  const size_t plane = static_cast<size_t>(nx) * ny;
  HOLO_TRY(ensure_ws(h, plane * sizeof(float2)));
  float2* B = static_cast<float2*>(h->ws);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float sc = static_cast<float>(1.0 / (static_cast<double>(nx) * ny));
  HOLO_TRY(launch_rows_fft(g, Ex, -1, 0, in, 0, B, 0, 1.0f, 1, st));
  HOLO_TRY(launch_prop_cols(g, Ey, tf, B, B, st, nullptr, h->gs_xs, h->max_lds));
  return launch_rows_fft(g, Ex, +1, 0, B, 0, static_cast<float2*>(out), 0, sc, 1, st);
}

int holo_lowpass_filter(holo_handle h, void* stream, int kind, double prop_dist, double wavlen,
                        int n, double dx, float* out) {
#pragma clang fp contract(off)
  if (!h || !out) return fail(HOLO_ERR_ARG, "NULL handle or output");
  if (n < 1) return fail(HOLO_ERR_ARG, "imglen must be positive");
  if (kind != HOLO_LPF_RECTANGLE && kind != HOLO_LPF_SUPER_GAUSSIAN)
    return fail(HOLO_ERR_ARG, "kind must be HOLO_LPF_RECTANGLE or HOLO_LPF_SUPER_GAUSSIAN");
  // cutoff radius in frequency samples (frequency_filters.jl:33, 68), FP64 in that order
  const double n2 = static_cast<double>(static_cast<long long>(n) * n);
  const double p2 = dx * dx;
  const double maxi = 1.0 / wavlen * n2 * p2 / sqrt(4.0 * (prop_dist * prop_dist) + n2 * p2);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 grid((n + 255) / 256, n), block(256);
  if (kind == HOLO_LPF_RECTANGLE) {
    hipLaunchKernelGGL(k_lpf_rect, grid, block, 0, st, n, maxi, out);
  } else {
    // frequency_filters.jl:69
    const double sigma = maxi / (static_cast<double>(n) * dx) / pow(2.0 * log(2.0), 1.0 / 6.0);
    hipLaunchKernelGGL(k_lpf_super_gaussian, grid, block, 0, st, n, sigma, dx, out);
  }
  return check_launch();
}

int holo_apply_low_pass_filter(holo_handle h, void* stream, int nx, int ny, void* out,
                               const void* in, const float* lpf) {
  if (!h || !in || !out || !lpf) return fail(HOLO_ERR_ARG, "NULL argument");
  Geom g;
  int Ex, Ey;
  const Tuning pt = plane_tiles(h, nx, ny);
  HOLO_TRY(make_geom(h, nx, ny, &g, &Ex, &Ey, 256, &pt));
  const size_t plane = static_cast<size_t>(nx) * ny;
  HOLO_TRY(ensure_ws(h, plane * sizeof(float2)));
  float2* B = static_cast<float2*>(h->ws);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const float sc = static_cast<float>(1.0 / (static_cast<double>(nx) * ny));
  HOLO_TRY(launch_rows_fft(g, Ex, -1, 0, in, 0, B, 0, 1.0f, 1, st));
  HOLO_TRY(launch_prop_cols(g, Ey, nullptr, B, B, st, lpf, h->gs_xs, h->max_lds));
  return launch_rows_fft(g, Ex, +1, 0, B, 0, static_cast<float2*>(out), 0, sc, 1, st);
}

int holo_ingest_u8(holo_handle h, void* stream, int nx, int ny, int count, const uint8_t* src,
                   const double* bkg, double offset, float* out) {
  if (!h || !src || !out) return fail(HOLO_ERR_ARG, "NULL argument");
  if (nx < 1 || ny < 1 || count < 0) return fail(HOLO_ERR_ARG, "sizes must be positive");
  const size_t npx = static_cast<size_t>(nx) * ny;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned gx = static_cast<unsigned>((npx + 4 * 256 - 1) / (4 * 256));
  for (int m0 = 0; m0 < count; m0 += 65535) {
    const int c = std::min(65535, count - m0);
    const uint8_t* s = src + static_cast<size_t>(m0) * npx;
    float* o = out + static_cast<size_t>(m0) * npx;
    const int vec = (npx % 4 == 0) && (reinterpret_cast<uintptr_t>(s) % 4 == 0) &&
                    (reinterpret_cast<uintptr_t>(o) % 16 == 0);
    const dim3 grid(gx, c), block(256);
    if (vec && bkg) hipLaunchKernelGGL((k_ingest_u8<true, true>), grid, block, 0, st, npx, c, s, bkg, offset, o);
    else if (vec) hipLaunchKernelGGL((k_ingest_u8<true, false>), grid, block, 0, st, npx, c, s, bkg, offset, o);
    else if (bkg) hipLaunchKernelGGL((k_ingest_u8<false, true>), grid, block, 0, st, npx, c, s, bkg, offset, o);
    else hipLaunchKernelGGL((k_ingest_u8<false, false>), grid, block, 0, st, npx, c, s, bkg, offset, o);
    HOLO_TRY(check_launch());
  }
  return HOLO_OK;
}

int holo_make_background(holo_handle h, void* stream, int nx, int ny, int count,
                         const uint8_t* stack, int kind, double* out) {
  if (!h || !stack || !out) return fail(HOLO_ERR_ARG, "NULL argument");
  if (nx < 1 || ny < 1) return fail(HOLO_ERR_ARG, "sizes must be positive");
  if (count < 1) return fail(HOLO_ERR_ARG, "the image list is empty");
  const size_t npx = static_cast<size_t>(nx) * ny;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (kind == HOLO_BACKGROUND_MODE) {
    if (count > 65535) return fail(HOLO_ERR_SIZE, "at most 65535 images per background-mode call");
    const unsigned gx = static_cast<unsigned>((npx + 2 * kModePairs - 1) / (2 * kModePairs));
    static_assert(2 * kModePairs == 4 * kMode8Quads, "both mode kernels take 128-pixel tiles");
    if (count <= 256) {  // 8-bit counters
      const bool vec4 = (npx % 4 == 0) && (reinterpret_cast<uintptr_t>(stack) % 4 == 0);
      const bool buf = vec4 && npx * static_cast<size_t>(count) <= 0x7fffffffull;
      // buffer loads: 16 phases of 512 threads share a tile (256 x 1024^2: 0.067 ms against
      // 0.071 with 8 phases, twice the waves per CU at almost the same LDS)
      if (buf)
        hipLaunchKernelGGL((k_background_mode8<2, 16>), dim3(gx), dim3(512), 0, st, npx, count, stack, out);
      else if (vec4)
        hipLaunchKernelGGL(k_background_mode8<1>, dim3(gx), dim3(kModeThreads), 0, st, npx, count, stack, out);
      else
        hipLaunchKernelGGL(k_background_mode8<0>, dim3(gx), dim3(kModeThreads), 0, st, npx, count, stack, out);
      return check_launch();
    }
    const int vec = (npx % 2 == 0) && (reinterpret_cast<uintptr_t>(stack) % 2 == 0);
    if (vec)
      hipLaunchKernelGGL(k_background_mode<true>, dim3(gx), dim3(kModeThreads), 0, st, npx, count, stack, out);
    else
      hipLaunchKernelGGL(k_background_mode<false>, dim3(gx), dim3(kModeThreads), 0, st, npx, count, stack, out);
  } else if (kind == HOLO_BACKGROUND_MEAN) {
    const int vec = (npx % 4 == 0) && (reinterpret_cast<uintptr_t>(stack) % 4 == 0);
    const size_t threads = vec ? npx / 4 : npx;
    hipLaunchKernelGGL(k_background_mean, dim3(static_cast<unsigned>((threads + 255) / 256)), dim3(256), 0, st,
                       npx, count, stack, out, vec);
  } else {
    return fail(HOLO_ERR_ARG, "kind must be HOLO_BACKGROUND_MODE or HOLO_BACKGROUND_MEAN");
  }
  return check_launch();
}

// ---- connected components, boxes, dilation (ccl.jl:186-264, particle_detection.jl:12-96) ----
static_assert(sizeof(holo_box) == sizeof(CclBox), "holo_box mirrors CclBox");

namespace {

int ccl_source(int kind, const void* img, double threshold, int nx, CclSrc* src) {
  src->img = img;
  src->lim = 0;
  size_t align = 4;
  if (kind == HOLO_CCL_F32) {
    src->kind = CCL_F32;
    align = 16;
  } else if (kind == HOLO_CCL_BINARY) {
    src->kind = CCL_BINARY;
  } else if (kind == HOLO_CCL_N0F8_LE) {
    // d_vol .<= threshold compares Float64(v/255) (a correctly rounded quotient) with the
    // threshold; the foreground bytes are a prefix 0 .. lim-1
    src->kind = CCL_N0F8_LE;
    unsigned lim = 0;
    while (lim < 256 && static_cast<double>(lim) / 255.0 <= threshold) ++lim;
    src->lim = lim;
  } else {
    return fail(HOLO_ERR_TYPE, "kind must be HOLO_CCL_F32, HOLO_CCL_BINARY or HOLO_CCL_N0F8_LE");
  }
  src->vec = (nx % 4 == 0) && (reinterpret_cast<uintptr_t>(img) % align == 0);
  return HOLO_OK;
}

}  // extern "C"

struct CclGrid {
  dim3 tiles;       // (ceil(nx/64), ceil(ny/16), slices of the round)
  size_t per_slice; // tiles per slice
};

static CclGrid ccl_grid(int nx, int ny, int c) {
  CclGrid g;
  g.tiles = dim3((nx + kTileX - 1) / kTileX, (ny + kTileY - 1) / kTileY, c);
  g.per_slice = static_cast<size_t>(g.tiles.x) * g.tiles.y;
  return g;
}

extern "C" {

// Slices per round so that the union-find state (P and the root slots M, 8 B/px, plus the
// tile flags and list) stays within 1 GiB; at least one. P and M are written only in tiles
// with foreground, so a larger round costs address space, not traffic, and saves launches.
int ccl_round(size_t npx, int count) {
  const size_t per = 8 * npx + npx / 64;
  size_t r = per ? env_size("HOLO_CCL_ROUND_BYTES", size_t(1) << 30) / per : 1;
  if (r < 1) r = 1;
  if (r > 65535) r = 65535;
  return static_cast<int>(std::min<size_t>(r, static_cast<size_t>(count)));
}

struct CclWs {
  unsigned* P;
  unsigned* M;
  unsigned* flags;
  unsigned* list;
  unsigned* count;
};

int ccl_workspace(holo_context* h, int nx, int ny, int round, bool with_m, CclWs* ws) {
  const size_t npx = static_cast<size_t>(nx) * ny;
  const size_t tiles = ccl_grid(nx, ny, 1).per_slice * round;
  const size_t pm = static_cast<size_t>(round) * npx * (with_m ? 2 : 1);
  const size_t words = pm + 2 * tiles + 1;
  HOLO_TRY(ensure_ws(h, words * 4));
  ws->P = static_cast<unsigned*>(h->ws);
  ws->M = with_m ? ws->P + static_cast<size_t>(round) * npx : nullptr;
  ws->flags = ws->P + pm;
  ws->list = ws->flags + tiles;
  ws->count = ws->list + tiles;
  return HOLO_OK;
}

// Workgroups of the passes that walk the tile list (tiles with foreground only).
constexpr unsigned kListGrid = 1024;

// Scan (flags + list of occupied tiles), tile-local unions, cross-tile unions for one round
// of c slices. nbox (boxes path) is zeroed here.
int ccl_passes_1_2(int nx, int ny, int c, const CclSrc& src, const CclWs& ws, int* nbox, hipStream_t st) {
  const CclGrid g = ccl_grid(nx, ny, c);
  HOLO_TRY(launch_fill(reinterpret_cast<float*>(ws.count), 1, 0u, st));
  if (nbox) HOLO_TRY(launch_fill(reinterpret_cast<float*>(nbox), c, 0u, st));
  const int vec16 = (nx % 16 == 0) && (reinterpret_cast<uintptr_t>(src.img) % 16 == 0);
  const size_t ntiles = g.per_slice * c;
  const unsigned sgrid = static_cast<unsigned>(std::min<size_t>(4096, (ntiles + 3) / 4));
  hipLaunchKernelGGL(k_ccl_scan, dim3(sgrid), dim3(256), 0, st, nx, ny, c, src, vec16, ws.flags, ws.list, ws.count);
  HOLO_TRY(check_launch());
  hipLaunchKernelGGL(k_ccl_local, dim3(kListGrid), dim3(256), 0, st, nx, ny, src, ws.P, ws.list, ws.count);
  HOLO_TRY(check_launch());
  hipLaunchKernelGGL(k_ccl_border, dim3(kListGrid), dim3(128), 0, st, nx, ny, src, ws.P, ws.list, ws.count);
  return check_launch();
}

int ccl_boxes_round(int nx, int ny, const CclWs& ws, holo_box* boxes, int max_boxes, int* nbox, hipStream_t st) {
  hipLaunchKernelGGL(k_ccl_roots, dim3(kListGrid), dim3(256), 0, st, nx, ny, ws.P, ws.list, ws.count, ws.M,
                     reinterpret_cast<CclBox*>(boxes), max_boxes, nbox);
  HOLO_TRY(check_launch());
  hipLaunchKernelGGL(k_ccl_boxes, dim3(kListGrid), dim3(256), 0, st, nx, ny, ws.P, ws.list, ws.count, ws.M,
                     reinterpret_cast<CclBox*>(boxes), max_boxes);
  return check_launch();
}

int ccl_check_dims(int nx, int ny, int count) {
  if (nx < 1 || ny < 1 || count < 0) return fail(HOLO_ERR_ARG, "sizes must be positive");
  if (static_cast<unsigned long long>(nx) * ny > 0xfffffffeull)
    return fail(HOLO_ERR_ARG, "Image is too large. Maximum length is 2^32-1.");
  if ((ny + kTileY - 1) / kTileY > 65535) return fail(HOLO_ERR_ARG, "image too tall");
  return HOLO_OK;
}

}  // namespace

int holo_ccl_label(holo_handle h, void* stream, int nx, int ny, int count, const void* img, int kind,
                   double threshold, uint32_t* labels) {
  if (!h || !img || !labels) return fail(HOLO_ERR_ARG, "NULL argument");
  HOLO_TRY(ccl_check_dims(nx, ny, count));
  CclSrc src;
  HOLO_TRY(ccl_source(kind, img, threshold, nx, &src));
  const size_t npx = static_cast<size_t>(nx) * ny;
  const size_t esz = kind == HOLO_CCL_F32 ? 4 : 1;
  const int round = ccl_round(npx, std::max(count, 1));
  CclWs ws;
  HOLO_TRY(ccl_workspace(h, nx, ny, round, false, &ws));
  const int lvec = (nx % 4 == 0) && (reinterpret_cast<uintptr_t>(labels) % 16 == 0);
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (int z0 = 0; z0 < count; z0 += round) {
    const int c = std::min(round, count - z0);
    CclSrc s = src;
    s.img = static_cast<const uint8_t*>(img) + static_cast<size_t>(z0) * npx * esz;
    HOLO_TRY(ccl_passes_1_2(nx, ny, c, s, ws, nullptr, st));
    hipLaunchKernelGGL(k_ccl_labels, ccl_grid(nx, ny, c).tiles, dim3(256), 0, st, nx, ny, ws.P, ws.flags,
                       labels + static_cast<size_t>(z0) * npx, lvec);
    HOLO_TRY(check_launch());
  }
  return HOLO_OK;
}

int holo_ccl_boxes(holo_handle h, void* stream, int nx, int ny, int count, const void* img, int kind,
                   double threshold, int max_boxes, holo_box* boxes, int* nboxes) {
  if (!h || !img || !boxes || !nboxes) return fail(HOLO_ERR_ARG, "NULL argument");
  HOLO_TRY(ccl_check_dims(nx, ny, count));
  if (max_boxes < 1) return fail(HOLO_ERR_ARG, "max_boxes must be positive");
  CclSrc src;
  HOLO_TRY(ccl_source(kind, img, threshold, nx, &src));
  const size_t npx = static_cast<size_t>(nx) * ny;
  const size_t esz = kind == HOLO_CCL_F32 ? 4 : 1;
  const int round = ccl_round(npx, std::max(count, 1));
  CclWs ws;
  HOLO_TRY(ccl_workspace(h, nx, ny, round, true, &ws));
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (int z0 = 0; z0 < count; z0 += round) {
    const int c = std::min(round, count - z0);
    CclSrc s = src;
    s.img = static_cast<const uint8_t*>(img) + static_cast<size_t>(z0) * npx * esz;
    HOLO_TRY(ccl_passes_1_2(nx, ny, c, s, ws, nboxes + z0, st));
    HOLO_TRY(ccl_boxes_round(nx, ny, ws, boxes + static_cast<size_t>(z0) * max_boxes, max_boxes, nboxes + z0, st));
  }
  return HOLO_OK;
}

int holo_label_boxes(holo_handle h, void* stream, int nx, int ny, int count, const uint32_t* labels,
                     int max_boxes, holo_box* boxes, int* nboxes) {
  if (!h || !labels || !boxes || !nboxes) return fail(HOLO_ERR_ARG, "NULL argument");
  HOLO_TRY(ccl_check_dims(nx, ny, count));
  if (max_boxes < 1) return fail(HOLO_ERR_ARG, "max_boxes must be positive");
  const size_t npx = static_cast<size_t>(nx) * ny;
  const int round = ccl_round(npx, std::max(count, 1));
  CclWs ws;
  HOLO_TRY(ccl_workspace(h, nx, ny, round, true, &ws));
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (int z0 = 0; z0 < count; z0 += round) {
    const int c = std::min(round, count - z0);
    hipLaunchKernelGGL(k_ccl_from_labels, dim3(static_cast<unsigned>((npx + 255) / 256), c), dim3(256), 0, st,
                       npx, c, labels + static_cast<size_t>(z0) * npx, ws.P);
    HOLO_TRY(check_launch());
    const unsigned ntiles = static_cast<unsigned>(ccl_grid(nx, ny, c).per_slice * c);
    hipLaunchKernelGGL(k_ccl_list_all, dim3((ntiles + 255) / 256), dim3(256), 0, st, ntiles, ws.list, ws.count);
    HOLO_TRY(check_launch());
    HOLO_TRY(launch_fill(reinterpret_cast<float*>(nboxes + z0), c, 0u, st));
    HOLO_TRY(ccl_boxes_round(nx, ny, ws, boxes + static_cast<size_t>(z0) * max_boxes, max_boxes, nboxes + z0, st));
  }
  return HOLO_OK;
}

int holo_dilate(holo_handle h, void* stream, int nx, int ny, int count, const uint8_t* in, uint8_t* out) {
  if (!h || !in || !out) return fail(HOLO_ERR_ARG, "NULL argument");
  if (nx < 1 || ny < 1 || count < 0) return fail(HOLO_ERR_ARG, "sizes must be positive");
  if (in == out) return fail(HOLO_ERR_ARG, "dilation cannot run in place");
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (int z0 = 0; z0 < count; z0 += 65535) {
    const int c = std::min(65535, count - z0);
    const size_t off = static_cast<size_t>(z0) * nx * ny;
    hipLaunchKernelGGL(k_dilate3x3, dim3((nx + 255) / 256, ny, c), dim3(256), 0, st, nx, ny, c, in + off, out + off);
    HOLO_TRY(check_launch());
  }
  return HOLO_OK;
}

static int link_rectangles(const int32_t* rects, const int64_t* offsets, int nslices, int three_d,
                           int64_t* rows, int32_t* born, uint8_t* alive, int64_t* nrows) {
  if (!nrows || (nslices > 0 && !offsets)) return fail(HOLO_ERR_ARG, "NULL argument");
  if (nslices < 0) return fail(HOLO_ERR_ARG, "nslices must be >= 0");
  *nrows = 0;
  if (nslices == 0) return HOLO_OK;
  if (offsets[0] != 0) return fail(HOLO_ERR_ARG, "offsets[0] must be 0");
  for (int s = 0; s < nslices; ++s)
    if (offsets[s + 1] < offsets[s]) return fail(HOLO_ERR_ARG, "offsets must not decrease");
  const int64_t total = offsets[nslices];
  if (total > 0 && (!rects || !rows || !born || !alive)) return fail(HOLO_ERR_ARG, "NULL argument");
  // rows are indexed with 32-bit integers
  if (total > std::numeric_limits<int32_t>::max())
    return fail(HOLO_ERR_ARG, "more than 2^31 - 1 rectangles");
  LinkTable t;
  t.reserve(static_cast<size_t>(total));
  LinkGrid grid_store;
  const bool use_grid = !three_d && grid_store.init(total, [&](int64_t q, int32_t* r) {
    for (int c = 0; c < 4; ++c) r[c] = rects[4 * q + c];
  });
  LinkGrid* grid = use_grid ? &grid_store : nullptr;
  for (int64_t q = 0; q < offsets[1]; ++q) link_start(t, grid, rects + 4 * q, 1);
  t.rebuild_live();
  std::vector<uint32_t> sub, touched;
  std::vector<int32_t> seen;
  for (uint32_t i = 0; i < t.size(); ++i) sub.push_back(i);  // slice 1's rows end on slice 1
  for (int s = 2; s <= nslices; ++s)
    link_update(t, grid, rects + 4 * offsets[s - 1], offsets[s] - offsets[s - 1], s, three_d != 0, sub, touched, seen);
  link_finalize(t);
  for (size_t i = 0; i < t.size(); ++i) {
    int64_t* o = rows + 6 * i;
    o[0] = t.x0(i);
    o[1] = t.y0(i);
    o[2] = t.z0[i];
    o[3] = t.x1(i);
    o[4] = t.y1(i);
    o[5] = t.z1[i];
    born[i] = t.born[i];
    alive[i] = t.dead[i] ? 0 : 1;
  }
  *nrows = static_cast<int64_t>(t.size());
  return HOLO_OK;
}

int holo_link_rectangles(const int32_t* rects, const int64_t* offsets, int nslices, int three_d,
                         int64_t* rows, int32_t* born, uint8_t* alive, int64_t* nrows) {
  try {
    return link_rectangles(rects, offsets, nslices, three_d, rows, born, alive, nrows);
  } catch (const std::bad_alloc&) {
    return fail(HOLO_ERR_NOMEM, "host allocation failed while linking rectangles");
  }
}

int holo_reconst_xyprojection(holo_handle h, void* stream, int nx, int ny, const void* wavefront,
                              const holo_tf* front, const holo_tf* dz, int slices,
                              float* proj) {
  if (!h || !proj) return fail(HOLO_ERR_ARG, "NULL handle or output");
  return reconstruct(h, static_cast<hipStream_t>(stream), nx, ny, wavefront, 0, front, dz, slices, 0,
                     HOLO_F32, nullptr, proj, 0, 0, nx, ny);
}

int holo_reconst_xyprojection_holo(holo_handle h, void* stream, int nx, int ny, const float* holo,
                                   const holo_tf* front, const holo_tf* dz, int slices,
                                   float* proj) {
  if (!h || !proj) return fail(HOLO_ERR_ARG, "NULL handle or output");
  return reconstruct(h, static_cast<hipStream_t>(stream), nx, ny, holo, 1, front, dz, slices, 0,
                     HOLO_F32, nullptr, proj, 0, 0, nx, ny);
}

int holo_reconst_xyprojection_noalloc(holo_handle h, void* stream, int nx, int ny, const void* wavefront,
                                      const holo_tf* front, const holo_tf* dz, int slices, float* proj) {
  if (!h || !proj) return fail(HOLO_ERR_ARG, "NULL handle or output");
  if (!h->scratch) return fail(HOLO_ERR_ARG, "handle has no scratch address");
  hipStream_t st = static_cast<hipStream_t>(stream);
  NoAllocPlans na{};
  na.slot[0] = reinterpret_cast<float2*>(h->scratch);
  na.slot[1] = reinterpret_cast<float2*>(h->scratch + kScratchTwBytes);
  na.st = st;
  FixedWs fw{h->scratch + 2 * kScratchTwBytes, kScratchBytes - 2 * kScratchTwBytes, &na};
  return reconstruct(h, st, nx, ny, wavefront, 0, front, dz, slices, 0, HOLO_F32, nullptr, proj, 0, 0, nx,
                     ny, 0, &fw);
}

int holo_reconst_vol(holo_handle h, void* stream, int nx, int ny, const void* wavefront,
                     const holo_tf* front, const holo_tf* dz, int slices, int out_dtype,
                     void* vol, float* proj) {
  if (!h || !vol) return fail(HOLO_ERR_ARG, "NULL handle or volume");
  if (out_dtype < HOLO_N0F8 || out_dtype > HOLO_C64)
    return fail(HOLO_ERR_TYPE, "return_type must be N0f8, Float16, Float32, Float64 or ComplexF32");
  if (out_dtype == HOLO_C64 && proj) return fail(HOLO_ERR_ARG, "complex volume has no projection");
  return reconstruct(h, static_cast<hipStream_t>(stream), nx, ny, wavefront, 0, front, dz, slices, 1,
                     out_dtype, vol, proj, 0, 0, nx, ny);
}

int holo_2d_pad(holo_handle h, void* stream, int nx, int ny, const void* in, void* out) {
  if (!h || !in || !out) return fail(HOLO_ERR_ARG, "NULL argument");
  if (nx < 1 || ny < 1) return fail(HOLO_ERR_ARG, "sizes must be positive");
  constexpr int kParts = 256;
  double2* part = h->partials;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t n = static_cast<size_t>(nx) * ny;
  hipLaunchKernelGGL(k_partial_sum, dim3(kParts), dim3(256), 0, st, static_cast<const float2*>(in), n, part);
  HOLO_TRY(check_launch());
  hipLaunchKernelGGL(k_pad_fill, dim3((2 * nx + 255) / 256, 2 * ny), dim3(256), 0, st,
                     static_cast<const float2*>(in), nx, ny, part, kParts, static_cast<float2*>(out));
  return check_launch();
}

int holo_reconst_vol_and_xyprojection_padded(holo_handle h, void* stream, int nx, int ny,
                                             const void* wavefront, const holo_tf* front,
                                             const holo_tf* dz, int slices, uint8_t* vol,
                                             float* proj) {
  if (!h || !wavefront || !vol || !proj) return fail(HOLO_ERR_ARG, "NULL argument");
  if (nx < 2 || ny < 2 || (nx & 1) || (ny & 1))
    return fail(HOLO_ERR_ARG, "padded reconstruction needs even sizes");
  // workspace: [padded 2nx x 2ny wavefront | reconstruction buffers]
  const size_t pbytes = static_cast<size_t>(4) * nx * ny * sizeof(float2);
  HOLO_TRY(ensure_ws(h, pbytes));
  void* padded = h->ws;
  HOLO_TRY(holo_2d_pad(h, stream, nx, ny, wavefront, padded));
  // reconstruct() grows the workspace behind the padded wavefront (growth preserves it)
  const int r = reconstruct(h, static_cast<hipStream_t>(stream), 2 * nx, 2 * ny, padded, 0, front, dz,
                            slices, 1, HOLO_N0F8, vol, proj, nx / 2, ny / 2, nx, ny, pbytes);
  return r;
}

// The cached graph of `n` identical GS iterations for `key` (captured on first use). Keys
// change only with the size, the tiles and the workspace address; past kGsGraphCache graphs
// (workspace reallocated many times) the cache is emptied once the device is idle.
constexpr int kGsGraphIters = 16;
constexpr size_t kGsGraphCache = 64;
int gs_graph(holo_context* h, holo_context::GsKey key, int n, const Geom& g, int Ex, int Ey, GsPair* p,
             float sc, hipGraphExec_t* out, int np = 1) {
  key.n = n;
  auto it = h->gs_graphs.find(key);
  if (it != h->gs_graphs.end()) {
    *out = it->second;
    return HOLO_OK;
  }
  if (h->gs_graphs.size() >= kGsGraphCache) {
    HOLO_HIP(hipDeviceSynchronize());
    for (auto& kv : h->gs_graphs) (void)hipGraphExecDestroy(kv.second);
    h->gs_graphs.clear();
  }
  hipGraph_t graph = nullptr;
  HOLO_HIP(hipStreamBeginCapture(h->cap, hipStreamCaptureModeRelaxed));
  int r = HOLO_OK;
  for (int i = 0; i < n && r == HOLO_OK; ++i) r = gs_round(h, g, Ex, Ey, p, np, 0, sc, h->cap);
  const hipError_t e = hipStreamEndCapture(h->cap, &graph);
  if (r != HOLO_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return r;
  }
  HOLO_HIP(e);
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HOLO_HIP(ei);
  h->gs_graphs.emplace(key, exec);
  *out = exec;
  return HOLO_OK;
}

int holo_phase_retrieval(holo_handle h, void* stream, int nx, int ny, const float* holo1,
                         const float* holo2, const holo_tf* fwd, const holo_tf* bwd,
                         int iters, void* out, int use_graph) {
  if (!h || !holo1 || !holo2 || !out) return fail(HOLO_ERR_ARG, "NULL argument");
  HOLO_TRY(validate_transfer(fwd, "transfer"));
  HOLO_TRY(validate_transfer(bwd, "invtransfer"));
  if (iters < 0) return fail(HOLO_ERR_ARG, "iterations must be >= 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (iters == 0) return holo_gabor_wavefront(h, stream, nx, ny, holo1, out);
  Geom g;
  int Ex, Ey;
  const Tuning gt = plane_tiles(h, nx, ny);
  HOLO_TRY(make_geom(h, nx, ny, &g, &Ex, &Ey, 256, &gt));
  const size_t plane = static_cast<size_t>(nx) * ny;
  // workspace: [field | H fwd | H bwd | I1 | I2]
  HOLO_TRY(ensure_ws(h, plane * (3 * sizeof(float2) + 2 * sizeof(float))));
  GsPair p{};
  p.B = static_cast<float2*>(h->ws);
  float2* hf = p.B + plane;
  float* a1 = reinterpret_cast<float*>(hf + 2 * plane);
  float* a2 = a1 + plane;
  HOLO_TRY(launch_sqrt_plane(holo1, a1, plane, st));
  HOLO_TRY(launch_sqrt_plane(holo2, a2, plane, st));
  p.A1 = a1;
  p.A2 = a2;
  p.out = static_cast<float2*>(out);
  p.iters = iters;
  HOLO_TRY(gs_transfers(g, fwd, bwd, &p, hf, hf + plane, st));
  const float sc = static_cast<float>(1.0 / (static_cast<double>(nx) * ny));
  HOLO_TRY(launch_rows_fft(g, Ex, -1, 1, holo1, 0, p.B, 0, 1.0f, 1, st));  // psi1 = sqrt(I1), row FFT
  const int body = iters - 1;  // iterations 0 .. iters-2 are identical; the last one stores psi
  if (use_graph && body > 0) {
    // A graph holds up to kGsGraphIters iterations (4 launches each): graphs of one iteration
    // leave the device idle between graph launches.
    holo_context::GsKey key{nx, ny, p.B, p.hb, p.conj_b, g.W, g.rows, Ex, Ey, 0, 0, 0,
                            static_cast<unsigned long long>(p.hq_f | p.hq_b << 1)};
    const int chunk = std::min(body, kGsGraphIters);
    hipGraphExec_t full = nullptr, rest = nullptr;
    HOLO_TRY(gs_graph(h, key, chunk, g, Ex, Ey, &p, sc, &full));
    if (body % chunk) HOLO_TRY(gs_graph(h, key, body % chunk, g, Ex, Ey, &p, sc, &rest));
    for (int c = 0; c < body / chunk; ++c) HOLO_HIP(hipGraphLaunch(full, st));
    if (rest) HOLO_HIP(hipGraphLaunch(rest, st));
  } else {
    for (int it2 = 0; it2 < body; ++it2) HOLO_TRY(gs_round(h, g, Ex, Ey, &p, 1, it2, sc, st));
  }
  return gs_round(h, g, Ex, Ey, &p, 1, iters - 1, sc, st);
}

int holo_phase_retrieval_batch(holo_handle h, void* stream, const holo_gs_job* jobs, int njobs) {
  if (!h || (njobs > 0 && !jobs)) return fail(HOLO_ERR_ARG, "NULL argument");
  if (njobs < 0) return fail(HOLO_ERR_ARG, "njobs must be >= 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (int i = 0; i < njobs; ++i) {
    const holo_gs_job& j = jobs[i];
    if (!j.holo1 || !j.holo2 || !j.out) return fail(HOLO_ERR_ARG, "job " + std::to_string(i) + ": NULL argument");
    if (j.iters < 0) return fail(HOLO_ERR_ARG, "job " + std::to_string(i) + ": iterations must be >= 0");
    HOLO_TRY(validate_transfer(&j.fwd, "transfer"));
    HOLO_TRY(validate_transfer(&j.bwd, "invtransfer"));
  }
  std::vector<int> order;
  for (int i = 0; i < njobs; ++i) {
    if (jobs[i].iters == 0) HOLO_TRY(holo_gabor_wavefront(h, stream, jobs[i].nx, jobs[i].ny, jobs[i].holo1, jobs[i].out));
    else order.push_back(i);
  }
  auto key = [&](int i) { return std::make_pair(jobs[i].nx, jobs[i].ny); };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
  const size_t budget = h->batch_ws_limit ? h->batch_ws_limit : env_size("HOLO_BATCH_WS_BYTES", 8ull << 30);
  for (size_t g0 = 0; g0 < order.size();) {
    size_t g1 = g0 + 1;
    while (g1 < order.size() && key(order[g1]) == key(order[g0])) ++g1;
    const int nx = jobs[order[g0]].nx, ny = jobs[order[g0]].ny;
    Geom g;
    int Ex, Ey;
    // The batched entry point is the throughput path: its row passes take the length's default
    // elements per thread (at 1024: 16, one wavefront per line, three FFT stages instead of the
    // single-pair call's five), columns keep the single-plane tiles. 1024^2 x 16 pairs: 50.4k ->
    // 53.2k pair-iterations/s (all-default tiles 47.9k). The tiles depend on the size only, so a
    // pair's result does not depend on what it is batched with.
    Tuning gt = plane_tiles(h, nx, ny, true);
    gt.erow = 0;
    gt.rows = h->tune.rows;
    // 2048-tall columns take 2-column tiles here: 16 pairs x 20 iterations, 10.9k -> 11.25k
    // pair-iterations/s (at 1024^2 they cost the lane-swap columns 20%); a tile's width only
    // chooses which columns a workgroup transforms, so the results are the same
    if (nx == 2048 && ny == 2048 && h->tune.cols <= 0) gt.cols = 2;
    HOLO_TRY(make_geom(h, nx, ny, &g, &Ex, &Ey, 256, &gt));
    const size_t plane = static_cast<size_t>(nx) * ny;
    // per pair: field, forward and backward tables, the amplitudes sqrt(I1), sqrt(I2) (so that
    // captured iterations reference workspace addresses only)
    const size_t perpair = plane * (3 * sizeof(float2) + 2 * sizeof(float));
    size_t chunk = std::max<size_t>(1, std::min<size_t>(kMaxJobs, budget / perpair));
    // Pairs per launch: as many as keep the fields and amplitudes (16 B/px) of the chunk within
    // 256 MiB, so that each pass finds most of what the previous one wrote in the Infinity
    // Cache (1024^2, 32 pairs: 54.8k pair-iterations/s in one chunk of 32, 63.0k in chunks of
    // 16, 58.6k in chunks of 8); HOLO_GS_CHUNK sets the count instead.
    const size_t cap = env_size("HOLO_GS_CHUNK", 0);
    chunk = std::min(chunk, cap > 0 ? cap : std::max<size_t>(1, (256ull << 20) / (plane * 16)));
    const float sc = static_cast<float>(1.0 / (static_cast<double>(nx) * ny));
    for (size_t c0 = g0; c0 < g1; c0 += chunk) {
      const size_t c1 = std::min(g1, c0 + chunk);
      HOLO_TRY(ensure_ws(h, perpair * (c1 - c0)));
      std::vector<GsPair> pairs(c1 - c0);
      int maxit = 0, minit = std::numeric_limits<int>::max();
      unsigned long long conjmask = 0, hqmask = 0;
      for (size_t q = c0; q < c1; ++q) {
        const holo_gs_job& j = jobs[order[q]];
        GsPair& p = pairs[q - c0];
        p.B = reinterpret_cast<float2*>(static_cast<char*>(h->ws) + perpair * (q - c0));
        float2* hf = p.B + plane;
        HOLO_TRY(gs_transfers(g, &j.fwd, &j.bwd, &p, hf, hf + plane, st));
        float* a1 = reinterpret_cast<float*>(hf + 2 * plane);
        float* a2 = a1 + plane;
        HOLO_TRY(launch_sqrt_plane(static_cast<const float*>(j.holo1), a1, plane, st));
        HOLO_TRY(launch_sqrt_plane(static_cast<const float*>(j.holo2), a2, plane, st));
        p.A1 = a1;
        p.A2 = a2;
        p.out = static_cast<float2*>(j.out);
        p.iters = j.iters;
        HOLO_TRY(launch_rows_fft(g, Ex, -1, 1, j.holo1, 0, p.B, 0, 1.0f, 1, st));
        maxit = std::max(maxit, j.iters);
        minit = std::min(minit, j.iters);
        if (p.conj_b) conjmask |= 1ull << (q - c0);
        hqmask |= static_cast<unsigned long long>(p.hq_f | p.hq_b << 1) << (2 * (q - c0));
      }
      // iterations 0 .. minit-2 are identical for the chunk (every pair active, none storing its
      // result): graphs of up to kGsGraphIters of them, captured once per chunk layout; the rest
      // (the last iteration of each pair, pairs with more iterations) are launched directly
      const int body = h->gs_batch_graphs ? minit - 1 : 0;
      if (body > 0) {
        holo_context::GsKey key{nx, ny, h->ws, nullptr, 0, g.W, g.rows, Ex, Ey, 0,
                                static_cast<int>(pairs.size()), conjmask, hqmask};
        const int gchunk = std::min(body, kGsGraphIters);
        hipGraphExec_t full = nullptr, rest = nullptr;
        const int np = static_cast<int>(pairs.size());
        HOLO_TRY(gs_graph(h, key, gchunk, g, Ex, Ey, pairs.data(), sc, &full, np));
        if (body % gchunk) HOLO_TRY(gs_graph(h, key, body % gchunk, g, Ex, Ey, pairs.data(), sc, &rest, np));
        for (int c = 0; c < body / gchunk; ++c) HOLO_HIP(hipGraphLaunch(full, st));
        if (rest) HOLO_HIP(hipGraphLaunch(rest, st));
      }
      for (int it = std::max(body, 0); it < maxit; ++it)
        HOLO_TRY(gs_round(h, g, Ex, Ey, pairs.data(), static_cast<int>(pairs.size()), it, sc, st));
    }
    g0 = g1;
  }
  return HOLO_OK;
}

int holo_set_gs_graphs(holo_handle h, int enable) {
  if (!h) return fail(HOLO_ERR_ARG, "handle is NULL");
  h->gs_batch_graphs = enable != 0;
  return HOLO_OK;
}

int holo_set_profiling(holo_handle h, int enable) {
  if (!h) return fail(HOLO_ERR_ARG, "handle is NULL");
  h->prof = enable != 0;
  return HOLO_OK;
}

int holo_profile_read(holo_handle h, int pass, double* total_ms, long long* launches,
                      long long* units, int reset) {
  if (!h || pass < 0 || pass >= 8) return fail(HOLO_ERR_ARG, "bad handle or pass");
  prof_collect(h);
  if (total_ms) *total_ms = h->prof_ms[pass];
  if (launches) *launches = h->prof_launch[pass];
  if (units) *units = h->prof_units[pass];
  if (reset) {
    h->prof_ms[pass] = 0.0;
    h->prof_launch[pass] = 0;
    h->prof_units[pass] = 0;
  }
  return HOLO_OK;
}

int holo_set_tuning_for_size(holo_handle h, int nx, int ny, int batch, int spg_gen, int spg_out,
                             int cols_per_wg, int rows_per_wg) {
  if (!h) return fail(HOLO_ERR_ARG, "handle is NULL");
  if (cols_per_wg > 0 && (cols_per_wg & (cols_per_wg - 1)) != 0)
    return fail(HOLO_ERR_ARG, "cols_per_wg must be a power of two");
  // the five tile fields are replaced; the size's kernel variants (elements per thread, exchange
  // buffers, row prefetch) stay
  Tuning t;
  auto it = h->size_tune.find({nx, ny});
  if (it != h->size_tune.end()) t = it->second;
  t.batch = batch;
  t.spg_gen = spg_gen;
  t.spg_out = spg_out;
  t.cols = cols_per_wg;
  t.rows = rows_per_wg;
  h->size_tune[{nx, ny}] = t;
  return HOLO_OK;
}

int holo_size_variants(holo_handle h, int nx, int ny, int* col_db, int* row_pf, int set) {
  if (!h || !col_db || !row_pf) return fail(HOLO_ERR_ARG, "NULL argument");
  if (set) {
    if (*col_db < -1 || *col_db > 3) return fail(HOLO_ERR_ARG, "col_db must be -1, 0, 1, 2 or 3");
    if (*row_pf < -1 || *row_pf > 3) return fail(HOLO_ERR_ARG, "row_pf must be -1, 0, 1, 2 or 3");
    Tuning& t = h->size_tune[{nx, ny}];
    t.col_db = *col_db;
    t.row_pf = *row_pf;
    return HOLO_OK;
  }
  const Tuning t = eff_tuning(h, nx, ny);
  *col_db = t.col_db;
  *row_pf = t.row_pf;
  return HOLO_OK;
}

int holo_get_tuning_for_size(holo_handle h, int nx, int ny, int slices, int* out5) {
  if (!h || !out5) return fail(HOLO_ERR_ARG, "NULL argument");
  Geom g;
  int Ex, Ey;
  HOLO_TRY(make_geom(h, nx, ny, &g, &Ex, &Ey, 512));
  const ReconPlan rp = recon_plan(h, g, slices > 0 ? slices : 1 << 20);
  out5[0] = rp.m;
  out5[1] = rp.spg_gen;
  out5[2] = rp.spg_out;
  out5[3] = g.W;
  out5[4] = g.rows;
  return HOLO_OK;
}

int holo_reconst_batch(holo_handle h, void* stream, const holo_job* jobs, int njobs) {
  if (!h || (njobs > 0 && !jobs)) return fail(HOLO_ERR_ARG, "NULL argument");
  if (njobs < 0) return fail(HOLO_ERR_ARG, "njobs must be >= 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (int i = 0; i < njobs; ++i) {
    const holo_job& j = jobs[i];
    if (!j.src || !j.proj) return fail(HOLO_ERR_ARG, "job " + std::to_string(i) + ": NULL input or projection");
    if (j.slices < 1) return fail(HOLO_ERR_ARG, "job " + std::to_string(i) + ": slices must be >= 1");
    if (!j.front.generated || !j.dz.generated)
      return fail(HOLO_ERR_ARG, "job " + std::to_string(i) + ": batched reconstruction needs generated transfers");
    HOLO_TRY(validate_transfer(&j.front, "front"));
    HOLO_TRY(validate_transfer(&j.dz, "dz"));
    if (j.front.wavlen != j.dz.wavlen || j.front.dx != j.dz.dx)
      return fail(HOLO_ERR_ARG, "job " + std::to_string(i) + ": transfers must share wavlen and dx");
  }
  // group by (nx, ny, wavlen, dx), keeping submission order inside a group
  std::vector<int> order(njobs);
  for (int i = 0; i < njobs; ++i) order[i] = i;
  auto key = [&](int i) {
    return std::make_tuple(jobs[i].nx, jobs[i].ny, jobs[i].front.wavlen, jobs[i].front.dx);
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key(a) < key(b); });
  const size_t budget = h->batch_ws_limit ? h->batch_ws_limit : env_size("HOLO_BATCH_WS_BYTES", 8ull << 30);
  const size_t round_bytes = env_size("HOLO_BATCH_ROUND_BYTES", 320ull << 20);
  for (size_t g0 = 0; g0 < order.size();) {
    size_t g1 = g0 + 1;
    while (g1 < order.size() && key(order[g1]) == key(order[g0])) ++g1;
    const holo_job& j0 = jobs[order[g0]];
    Geom g;
    int Ex, Ey;
    HOLO_TRY(make_geom(h, j0.nx, j0.ny, &g, &Ex, &Ey, 512));
    int maxs = 0;
    for (size_t q = g0; q < g1; ++q) maxs = std::max(maxs, jobs[order[q]].slices);
    ReconPlan rp = recon_plan(h, g, maxs);
    const size_t plane = static_cast<size_t>(g.nx) * g.ny;
    // one job's spectrum and round of slices stay within the workspace budget where they can
    const size_t fit = budget / (plane * sizeof(float2));
    if (fit >= 2 && static_cast<size_t>(rp.m) > fit - 1) rp.m = static_cast<int>(fit - 1);
    const size_t perjob = plane * sizeof(float2) * (1 + static_cast<size_t>(rp.m));
    // Holograms per launch: as many as keep one round's spectra and column-transformed
    // slices near the 256 MB Infinity Cache, where pass B finds what pass A wrote (measured:
    // 768^2 batches 0.59 -> 0.94 of the HBM roofline against a launch over every hologram),
    // and within the workspace budget; at least one.
    const size_t cap = std::min(budget, round_bytes);
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(kMaxJobs, cap / perjob));
    for (size_t c0 = g0; c0 < g1; c0 += chunk) {
      const size_t c1 = std::min(g1, c0 + chunk);
      HOLO_TRY(ensure_ws(h, perjob * (c1 - c0)));
      std::vector<float2*> F(c1 - c0);
      int cmax = 0;
      for (size_t q = c0; q < c1; ++q) {
        const holo_job& j = jobs[order[q]];
        F[q - c0] = reinterpret_cast<float2*>(static_cast<char*>(h->ws) + perjob * (q - c0));
        HOLO_TRY(launch_fill(j.proj, plane, 0x7fffffffu, st));
        hipEvent_t pe = prof_begin(h, st);
        HOLO_TRY(forward_spectrum(g, Ex, Ey, j.src, j.src_real, F[q - c0], st));
        prof_end(h, st, pe, HOLO_PASS_FWD_FFT, 1);
        cmax = std::max(cmax, j.slices);
      }
      for (int b0 = 0; b0 < cmax; b0 += rp.m) {
        JobTable jt{};
        int maxnb = 0;
        for (size_t q = c0; q < c1; ++q) {
          const holo_job& j = jobs[order[q]];
          if (j.slices <= b0) continue;
          JobDesc& d = jt.d[jt.n++];
          d.F = F[q - c0];
          d.G = F[q - c0] + plane;
          d.proj = j.proj;
          d.zl0 = j.front.z / j.front.wavlen;
          d.kbase = b0;
          d.dzl = j.dz.z / j.front.wavlen;
          d.nbatch = std::min(rp.m, j.slices - b0);
          maxnb = std::max(maxnb, d.nbatch);
        }
        HOLO_TRY(launch_round(h, g, Ex, Ey, rp, j0.front.wavlen, j0.front.dx, jt, maxnb, st));
      }
    }
    g0 = g1;
  }
  return HOLO_OK;
}

int holo_autotune(holo_handle h, void* stream, int nx, int ny, double wavlen, double dx, int* out5) {
  if (!h) return fail(HOLO_ERR_ARG, "handle is NULL");
  if (!(wavlen > 0) || !(dx > 0)) return fail(HOLO_ERR_ARG, "wavlen and dx must be positive");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Tuning saved_global = h->tune;
  h->tune = Tuning();
  // the tiles in force before the search (shipped defaults or an earlier choice) compete with
  // the search's winner at the end
  Tuning incumbent;
  {
    auto it = h->size_tune.find({nx, ny});
    if (it != h->size_tune.end()) incumbent = it->second;
  }
  const size_t plane = static_cast<size_t>(nx) * ny;
  const size_t pbytes = plane * sizeof(float2);
  hipEvent_t e0 = ev_get(h), e1 = ev_get(h);
  if (!e0 || !e1) return fail(HOLO_ERR_HIP, "event creation failed");
  // time one round of one hologram with m slices under tuning t; returns ms per slice
  auto time_round = [&](const Tuning& t, bool a, bool b, double* ms_per_slice) -> int {
    h->size_tune[{nx, ny}] = t;
    Geom g;
    int Ex, Ey;
    int r = make_geom(h, nx, ny, &g, &Ex, &Ey, 512);
    if (r != HOLO_OK) return r;
    // batch 0 (the default round, e.g. an incumbent without a size entry) is timed at the
    // round size a long reconstruction would use, not capped at zero slices
    const ReconPlan rp = recon_plan(h, g, t.batch > 0 ? t.batch : 1 << 20);
    HOLO_TRY(ensure_ws(h, pbytes * (1 + static_cast<size_t>(rp.m)) + plane * sizeof(float), pbytes));
    float2* F = static_cast<float2*>(h->ws);
    float* proj = reinterpret_cast<float*>(F + plane * (1 + rp.m));
    JobTable jt{};
    jt.n = 1;
    jt.d[0] = JobDesc{F, F + plane, proj, -1e4 / wavlen, -200.0 / wavlen, rp.m, 0};
    HOLO_TRY(launch_round(h, g, Ex, Ey, rp, wavlen, dx, jt, rp.m, st, a, b));  // warm-up
    HOLO_HIP(hipEventRecord(e0, st));
    constexpr int kReps = 3;
    for (int i = 0; i < kReps; ++i) HOLO_TRY(launch_round(h, g, Ex, Ey, rp, wavlen, dx, jt, rp.m, st, a, b));
    HOLO_HIP(hipEventRecord(e1, st));
    HOLO_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    HOLO_HIP(hipEventElapsedTime(&ms, e0, e1));
    *ms_per_slice = ms / (kReps * rp.m);
    return HOLO_OK;
  };
  // a synthetic spectrum: any field of the right size (the kernels' cost does not depend on it)
  HOLO_TRY(ensure_ws(h, pbytes * 2));
  hipLaunchKernelGGL(k_transfer_natural, dim3((nx + 255) / 256, ny), dim3(256), 0, st, nx, ny, nx, 1.0e3 / wavlen,
                     wavlen, dx, static_cast<float2*>(h->ws));
  HOLO_TRY(check_launch());
  Tuning best;
  double best_total = 1e30;
  // rounds near half and near all of the Infinity Cache (pass B then reads pass A's slices
  // from it), and 2 GB rounds for sides where a cache-sized round holds fewer slices than a
  // column workgroup advances (16): there long rounds amortise each launch's tail instead. (At
  // 1024^2 long rounds can win a single-hologram race yet cost batched calls their
  // several-holograms-per-launch rounds, 1.06 -> 0.89 of the roofline on the 64-hologram mix.)
  const size_t mb[3] = {128ull << 20, 256ull << 20, 2048ull << 20};
  int tried[3] = {0, 0, 0};
  for (size_t bi = 0; bi < 3; ++bi) {
    if (bi == 2 && mb[1] / pbytes >= 16) continue;
    Tuning t;
    t.batch = static_cast<int>(std::max<size_t>(1, std::min<size_t>(64, mb[bi] / pbytes)));
    tried[bi] = t.batch;
    if ((bi >= 1 && t.batch == tried[0]) || (bi == 2 && t.batch == tried[1])) continue;
    double ba = 1e30, bb = 1e30;
    Tuning ta = t, tb = t;
    // pass A (columns, generates H_k): columns per workgroup, slices per workgroup and one or
    // two exchange buffers. The round is cut to a multiple of the group, so that no workgroup
    // runs a short remainder group (1280 x 1024: 25 slices in groups of 8 left one of 1).
    for (int W : {2, 4, 8, 16})
      for (int sg : {4, 8, 16})
        for (int db : {-1, 1, 2, 3}) {
          if (db >= 2 && !col_xs_form(ny)) continue;
          Tuning c = t;
          c.cols = W;
          if (t.batch > sg) c.batch = t.batch / sg * sg;
          c.spg_gen = std::min(sg, c.batch);
          c.spg_out = 8;
          c.col_db = db;
          double v;
          if (time_round(c, true, false, &v) != HOLO_OK) continue;
          if (v < ba) { ba = v; ta = c; }
        }
    // pass B (rows, forms the projection): rows per workgroup, slices per workgroup and rows
    // loaded one slice ahead or not (with the twiddles in registers or in LDS)
    // (half the round among the group sizes: the round in two equal groups)
    for (int rows : {1, 2, 4, 8})
      for (int so : {4, 8, 16, ta.batch / 2})
        for (int pf : {-1, 1, 2, 3}) {
          if (so < 1 || (so == ta.batch / 2 && (so == 4 || so == 8 || so == 16))) continue;
          Tuning c = ta;
          c.rows = rows;
          c.spg_out = std::min(so, ta.batch);
          c.row_pf = pf;
          double v;
          if (time_round(c, false, true, &v) != HOLO_OK) continue;
          if (v < bb) { bb = v; tb = c; }
        }
    if (ba + bb < best_total) {
      best_total = ba + bb;
      best = tb;
    }
  }
  // final: both passes together, best of three runs each; the incumbent stays unless the
  // winner is at least 2% faster
  if (best_total < 1e29) {
    auto full = [&](const Tuning& t) {
      double m = 1e30;
      for (int r = 0; r < 3; ++r) {
        double v;
        if (time_round(t, true, true, &v) != HOLO_OK) return 1e30;
        m = std::min(m, v);
      }
      return m;
    };
    const double vb = full(best), vi = full(incumbent);
    if (vi < 1e29 && vi <= vb * 1.02) best = incumbent;
  }
  g_err.clear();
  h->tune = saved_global;
  if (best_total >= 1e29) {
    h->size_tune.erase({nx, ny});
    return fail(HOLO_ERR_SIZE, "no tile configuration fits this size");
  }
  h->size_tune[{nx, ny}] = best;
  // the tiles in force, with defaults resolved (an incumbent without a size entry keeps none)
  return out5 ? holo_get_tuning_for_size(h, nx, ny, 0, out5) : HOLO_OK;
}

int holo_workspace_size(holo_handle h, size_t* bytes) {
  if (!h || !bytes) return fail(HOLO_ERR_ARG, "NULL argument");
  *bytes = h->ws_bytes;
  return HOLO_OK;
}

int holo_reserve_workspace(holo_handle h, size_t bytes) {
  if (!h) return fail(HOLO_ERR_ARG, "handle is NULL");
  return ensure_ws(h, bytes);
}

int holo_set_workspace_limit(holo_handle h, size_t bytes) {
  if (!h) return fail(HOLO_ERR_ARG, "handle is NULL");
  h->batch_ws_limit = bytes;
  return HOLO_OK;
}

}  // extern "C"
