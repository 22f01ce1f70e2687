// Host z-linking of per-slice bounding rectangles into particle neighbourhoods
// (ccl.jl:266-466, particle_detection.jl:58-96). Neighbourhoods are rows of one table in start
// order, marked dead instead of erased, so that the reference's dict order (insertion order
// here) and its sequential semantics hold without a hash map:
//  - a rectangle is compared with every neighbourhood present at that moment (those started by
//    earlier rectangles of the slice included), all overlapping ones grow, and without overlap
//    it starts a neighbourhood (ccl.jl:328-363); the 3-D variant compares only neighbourhoods
//    that ended on the previous slice when the slice began (ccl.jl:378-409);
//  - after each slice after the first, 1 x 1 x 1 neighbourhoods are deleted;
//  - finalize: each neighbourhood in turn absorbs the first other one it overlaps in 3-D,
//    passes repeat until none changes (ccl.jl:441-466), then the depth-1, aspect and area rules
//    (ccl.jl:424-439).
// The table is stored as columns of int32 (pixel and slice indices). The 2-D scan asks a
// uniform grid for candidates: a row is registered in every cell its rectangle covers (cells
// are added as it grows), and strict overlap implies a shared cell, so the candidates hold every
// overlapping row; the per-rectangle updates do not interact, so visiting order is free. Only
// rows grown or started in a slice can become 1 x 1 x 1, so only those are checked.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <limits>
#include <vector>

namespace holo {

// |a - b| without int32 overflow
inline int64_t link_extent(int32_t a, int32_t b) { return std::llabs(static_cast<int64_t>(a) - b); }

struct LinkTable {
  // the 2-D box of a row is stored packed: the overlap test of a candidate row, the hot loop of
  // the z-linking, then reads one cache line instead of four
  struct Box { int32_t x0, y0, x1, y1; };
  std::vector<Box> xy;
  std::vector<int32_t> z0, z1, born;
  std::vector<uint8_t> dead;
  std::vector<uint32_t> live;  // live rows in start order (valid after each rebuild)

  int32_t& x0(size_t i) { return xy[i].x0; }
  int32_t& y0(size_t i) { return xy[i].y0; }
  int32_t& x1(size_t i) { return xy[i].x1; }
  int32_t& y1(size_t i) { return xy[i].y1; }
  int32_t x0(size_t i) const { return xy[i].x0; }
  int32_t y0(size_t i) const { return xy[i].y0; }
  int32_t x1(size_t i) const { return xy[i].x1; }
  int32_t y1(size_t i) const { return xy[i].y1; }

  size_t size() const { return xy.size(); }
  void reserve(size_t n) {
    xy.reserve(n);
    for (auto* v : {&z0, &z1, &born}) v->reserve(n);
    dead.reserve(n);
    live.reserve(n);
  }
  void start(const int32_t* r, int32_t s) {
    xy.push_back(Box{r[0], r[1], r[2], r[3]});
    z0.push_back(s);
    z1.push_back(s);
    born.push_back(s);
    dead.push_back(0);
  }
  // judge_overlap2d(rect, (x_min, y_min, x_max, y_max) of row i) (ccl.jl:279-288)
  bool overlap2d(const int32_t* r, size_t i) const {
    const Box& b = xy[i];
    return r[0] < b.x1 && r[2] > b.x0 && r[1] < b.y1 && r[3] > b.y0;
  }
  // judge_overlap3d (ccl.jl:290-299)
  bool overlap3d(size_t a, size_t b) const {
    const Box &p = xy[a], &q = xy[b];
    return p.x0 < q.x1 && p.x1 > q.x0 && p.y0 < q.y1 && p.y1 > q.y0 && z0[a] < z1[b] && z1[a] > z0[b];
  }
  void grow(size_t i, const int32_t* r) {
    Box& b = xy[i];
    b.x0 = std::min(b.x0, r[0]);
    b.y0 = std::min(b.y0, r[1]);
    b.x1 = std::max(b.x1, r[2]);
    b.y1 = std::max(b.y1, r[3]);
  }
  void rebuild_live() {
    live.clear();
    for (size_t i = 0; i < size(); ++i)
      if (!dead[i]) live.push_back(static_cast<uint32_t>(i));
  }
};

// Cells of side `cell` over [gx0, gx1] x [gy0, gy1], each listing the rows registered in it.
struct LinkGrid {
  int64_t gx0 = 0, gy0 = 0, cell = 16, nx = 0, ny = 0;
  int shift = 4;  // cell = 2^shift: cell indices by shifts (coordinates are never below gx0, gy0)
  std::vector<std::vector<uint32_t>> cells;
  std::vector<uint32_t> stamp;
  uint32_t qid = 0;

  // bounds of n rectangles given by their four coordinates; false for none or an inverted one
  template <class F>
  bool init(int64_t n, F&& rect) {
    if (n == 0) return false;
    int64_t x0 = std::numeric_limits<int64_t>::max(), y0 = x0, x1 = std::numeric_limits<int64_t>::min(), y1 = x1;
    for (int64_t q = 0; q < n; ++q) {
      int32_t r[4];
      rect(q, r);
      if (r[0] > r[2] || r[1] > r[3]) return false;  // inverted rectangles: full scans
      x0 = std::min<int64_t>(x0, r[0]);
      y0 = std::min<int64_t>(y0, r[1]);
      x1 = std::max<int64_t>(x1, r[2]);
      y1 = std::max<int64_t>(y1, r[3]);
    }
    gx0 = x0;
    gy0 = y0;
    while (((x1 - x0) / cell + 1) * ((y1 - y0) / cell + 1) > (int64_t{1} << 20)) cell = int64_t{1} << ++shift;
    nx = (x1 - x0) / cell + 1;
    ny = (y1 - y0) / cell + 1;
    cells.assign(static_cast<size_t>(nx * ny), {});
    return true;
  }
  int64_t cx(int64_t x) const { return (x - gx0) >> shift; }
  int64_t cy(int64_t y) const { return (y - gy0) >> shift; }
  // registers row i in the cells of [x0, x1] x [y0, y1] outside the cell range `old` covered
  // (old, when given, lies inside the new range: rows only grow). Only the new cells are
  // walked: a growing neighbourhood spans many cells and gains few at a time.
  void add(uint32_t i, int32_t x0, int32_t y0, int32_t x1, int32_t y1, const int64_t* old) {
    const int64_t a0 = cx(x0), a1 = cx(x1), b0 = cy(y0), b1 = cy(y1);
    for (int64_t b = b0; b <= b1; ++b) {
      if (!old || b < old[1] || b > old[3]) {
        for (int64_t a = a0; a <= a1; ++a) cells[b * nx + a].push_back(i);
      } else {
        for (int64_t a = a0; a < old[0]; ++a) cells[b * nx + a].push_back(i);
        for (int64_t a = old[2] + 1; a <= a1; ++a) cells[b * nx + a].push_back(i);
      }
    }
  }
};

// Starts a row for rectangle r at slice s (registered in the grid when there is one).
inline void link_start(LinkTable& t, LinkGrid* grid, const int32_t* r, int32_t s) {
  const uint32_t i = static_cast<uint32_t>(t.size());
  t.start(r, s);
  if (grid) grid->add(i, r[0], r[1], r[2], r[3], nullptr);
}

// sub: on entry (3-D) the rows grown or started in slice s - 1, possibly repeated; a 3-D row
// ends on s - 1 exactly when it was grown or started there, so the live ones among them are the
// candidates. On exit: the rows grown or started in slice s. t.live is kept only without a grid
// (the grid path does not read it; finalize rebuilds it).
inline void link_update(LinkTable& t, LinkGrid* grid, const int32_t* rects, int64_t n, int32_t s, bool three_d,
                        std::vector<uint32_t>& sub, std::vector<uint32_t>& touched, std::vector<int32_t>& seen) {
  const size_t m0 = t.size();  // rows started before this slice
  touched.clear();
  if (three_d) {
    seen.resize(t.size(), 0);
    size_t w = 0;
    for (uint32_t i : sub)
      if (!t.dead[i] && seen[i] != s && t.z1[i] == s - 1) {
        seen[i] = s;
        sub[w++] = i;
      }
    sub.resize(w);
  }
  for (int64_t q = 0; q < n; ++q) {
    const int32_t* r = rects + 4 * q;
    bool hit = false;
    if (three_d) {
      for (uint32_t i : sub) {
        if (t.overlap2d(r, i)) {
          t.grow(i, r);
          t.z1[i] = s;
          touched.push_back(i);
          hit = true;
        }
      }
    } else {
      auto visit = [&](uint32_t i) {
        if (t.dead[i] || !t.overlap2d(r, i)) return;
        int64_t old[4];
        if (grid) old[0] = grid->cx(t.x0(i)), old[1] = grid->cy(t.y0(i)), old[2] = grid->cx(t.x1(i)), old[3] = grid->cy(t.y1(i));
        t.grow(i, r);
        if (grid) grid->add(i, t.x0(i), t.y0(i), t.x1(i), t.y1(i), old);
        if (s < t.z0[i]) t.z0[i] = s;
        else if (s > t.z1[i]) t.z1[i] = s;
        touched.push_back(i);
        hit = true;
      };
      if (grid) {
        const uint32_t qid = ++grid->qid;
        grid->stamp.resize(t.size(), 0);
        for (int64_t b = grid->cy(r[1]); b <= grid->cy(r[3]); ++b)
          for (int64_t a = grid->cx(r[0]); a <= grid->cx(r[2]); ++a) {
            const std::vector<uint32_t>& c = grid->cells[b * grid->nx + a];
            for (size_t k = 0; k < c.size(); ++k) {  // c may grow while visited (growth adds cells)
              const uint32_t i = c[k];
              if (grid->stamp[i] == qid) continue;
              grid->stamp[i] = qid;
              visit(i);
            }
          }
      } else {
        for (uint32_t i : t.live) visit(i);
        for (size_t i = m0; i < t.size(); ++i) visit(static_cast<uint32_t>(i));  // started earlier in this slice
      }
    }
    if (!hit) link_start(t, grid, r, s);
  }
  bool died = false;
  auto unit = [&](uint32_t i) {
    if (!t.dead[i] && link_extent(t.x0(i), t.x1(i)) == 1 && link_extent(t.y0(i), t.y1(i)) == 1 &&
        link_extent(t.z0[i], t.z1[i]) == 1) {
      t.dead[i] = 1;
      died = true;
    }
  };
  for (uint32_t i : touched) unit(i);
  for (size_t i = m0; i < t.size(); ++i) unit(static_cast<uint32_t>(i));
  if (three_d) {  // the next slice's candidates
    sub.assign(touched.begin(), touched.end());
    for (size_t i = m0; i < t.size(); ++i) sub.push_back(static_cast<uint32_t>(i));
  }
  if (grid) return;
  if (died) {
    size_t w = 0;
    for (uint32_t i : t.live)
      if (!t.dead[i]) t.live[w++] = i;
    t.live.resize(w);
  }
  for (size_t i = m0; i < t.size(); ++i)
    if (!t.dead[i]) t.live.push_back(static_cast<uint32_t>(i));
}

// The first live row after or before k, in start order, that overlaps row k in 3-D (-1: none):
// the smallest such index, found among the rows of k's grid cells or by a full scan.
inline int64_t link_first_overlap(const LinkTable& t, LinkGrid* grid, uint32_t k) {
  int64_t best = -1;
  if (!grid) {
    for (uint32_t k2 : t.live)
      if (k2 != k && !t.dead[k2] && t.overlap3d(k, k2)) return k2;
    return -1;
  }
  // merged rows die as the passes go; they leave the cells as the scan passes them (dead is
  // final, and nothing else walks the grid during the scan)
  for (int64_t b = grid->cy(t.y0(k)); b <= grid->cy(t.y1(k)); ++b)
    for (int64_t a = grid->cx(t.x0(k)); a <= grid->cx(t.x1(k)); ++a) {
      std::vector<uint32_t>& c = grid->cells[b * grid->nx + a];
      size_t w = 0;
      for (size_t q = 0; q < c.size(); ++q) {
        const uint32_t k2 = c[q];
        if (t.dead[k2]) continue;
        c[w++] = k2;
        if (k2 != k && (best < 0 || k2 < best) && t.overlap3d(k, k2)) best = k2;
      }
      c.resize(w);
    }
  return best;
}

inline void link_delete_duplicates(LinkTable& t) {
  t.rebuild_live();
  LinkGrid grid_store;
  const bool use_grid = grid_store.init(static_cast<int64_t>(t.live.size()), [&](int64_t q, int32_t* r) {
    const uint32_t i = t.live[q];
    r[0] = t.x0(i), r[1] = t.y0(i), r[2] = t.x1(i), r[3] = t.y1(i);
  });
  LinkGrid* grid = use_grid ? &grid_store : nullptr;
  if (grid)
    for (uint32_t i : t.live) grid->add(i, t.x0(i), t.y0(i), t.x1(i), t.y1(i), nullptr);
  bool changed = true;
  while (changed) {
    changed = false;
    t.rebuild_live();
    const std::vector<uint32_t> a = t.live;
    for (size_t p = 0; p < a.size(); ++p) {
      const uint32_t k = a[p];
      if (t.dead[k]) continue;
      const int64_t f = link_first_overlap(t, grid, k);
      if (f >= 0) {
        const uint32_t k2 = static_cast<uint32_t>(f);
        int64_t old[4];
        if (grid) old[0] = grid->cx(t.x0(k)), old[1] = grid->cy(t.y0(k)), old[2] = grid->cx(t.x1(k)), old[3] = grid->cy(t.y1(k));
        t.x0(k) = std::min(t.x0(k), t.x0(k2));
        t.y0(k) = std::min(t.y0(k), t.y0(k2));
        t.z0[k] = std::min(t.z0[k], t.z0[k2]);
        t.x1(k) = std::max(t.x1(k), t.x1(k2));
        t.y1(k) = std::max(t.y1(k), t.y1(k2));
        t.z1[k] = std::max(t.z1[k], t.z1[k2]);
        if (grid) grid->add(k, t.x0(k), t.y0(k), t.x1(k), t.y1(k), old);
        t.dead[k2] = 1;
        changed = true;
      }
    }
  }
}

// Julia Int / Int: x / 0 is Inf for x > 0 and NaN for 0 / 0 (extents are non-negative).
inline double link_ratio(int64_t a, int64_t b) {
  if (b != 0) return static_cast<double>(a) / static_cast<double>(b);
  return a > 0 ? std::numeric_limits<double>::infinity() : std::numeric_limits<double>::quiet_NaN();
}

inline void link_finalize(LinkTable& t) {
  link_delete_duplicates(t);
  for (size_t i = 0; i < t.size(); ++i) {
    if (t.dead[i]) continue;
    const int64_t dx = link_extent(t.x0(i), t.x1(i)), dy = link_extent(t.y0(i), t.y1(i));
    const double r = link_ratio(dx, dy);
    if (link_extent(t.z0[i], t.z1[i]) == 1 || r > 3.0 || r < 1.0 / 3.0 || dx * dy < 10) t.dead[i] = 1;
  }
}

}  // namespace holo
