// CPU oracle — TEST INFRASTRUCTURE ONLY (see iv.hpp).
// rectangle_size_bound drivers: the host-side initial grid of the reference
// (find_critical_intervals_in, rectangle_size_bound.cu:276-310, with
// Variables::from_la_r1_r2, variables.hpp:37-51), the per-cell enumeration of
// its kernel (rectangle_size_bound.cu:161-236), and the widest-dimension
// branch-and-bound over 7-D cells.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "criteria.hpp"
#include "rsb_chain.hpp"

namespace ccor {
namespace rsb {

static_assert(sizeof(Cell) == 112, "record layout");

// variables.hpp:53-66 — radii[i].lb raised to the largest lb of the smaller
// radii, for i >= 1 (radii[0] is left as is).
static Cell tighten_cell(Cell c) {
  double lb = 0.0;
  for (int i = 5; i > 0; --i) {
    if (lb < c.r[i].lb) lb = c.r[i].lb;
    else c.r[i].lb = lb;
  }
  return c;
}
// variables.hpp:68-80
static bool empty_cell(const Cell& c) {
  if (c.la.empty()) return true;
  for (int i = 0; i < 6; ++i)
    if (c.r[i].lb >= c.r[i].ub) return true;
  return false;
}

// Variables::from_la_r1_r2 (variables.hpp:37-51).
static Cell from_la_r1_r2(const Cell& v, const IV& la, const IV& r1, const IV& r2) {
  Cell r;
  r.la = v.la.intersect(la);
  r.r[0] = v.r[0].intersect(r1);
  r.r[1] = v.r[1].intersect(r2);
  for (int i = 2; i < 6; ++i) {
    r.r[i] = v.r[i];
    r.r[i].tighten_ub(r2.ub);
  }
  return tighten_cell(r);
}

// Widest dimension of (lambda, r1..r6), absolute width, ties to the lowest
// index, split into get_subinterval(x, 0, 2) and (x, 1, 2).
static void bisect(const Cell& c, Cell* out) {
  const IV* d[7] = {&c.la, &c.r[0], &c.r[1], &c.r[2], &c.r[3], &c.r[4], &c.r[5]};
  int best = 0;
  double bw = rn_sub(d[0]->ub, d[0]->lb);
  for (int k = 1; k < 7; ++k) {
    double w = rn_sub(d[k]->ub, d[k]->lb);
    if (w > bw) {
      bw = w;
      best = k;
    }
  }
  for (int h = 0; h < 2; ++h) {
    out[h] = c;
    IV* t = best == 0 ? &out[h].la : &out[h].r[best - 1];
    *t = get_subinterval(*d[best], h, 2);
  }
}

}  // namespace rsb
}  // namespace ccor

using namespace ccor;

extern "C" {

// Initial cells of the reference's host loop over (lambda, r1, r2) pieces of
// `outer` (14 doubles); nonempty cells are written in loop order to out
// (capacity nla*nr1*nr2 records); returns their number.
size_t or_rsb_grid(const double* outer, int nla, int nr1, int nr2, double* out) {
  rsb::Cell v;
  std::memcpy(&v, outer, sizeof v);
  size_t n = 0;
  for (int a = 0; a < nla; ++a) {
    IV la = get_subinterval(v.la, a, nla);
    for (int b = 0; b < nr1; ++b) {
      IV r1 = get_subinterval(v.r[0], b, nr1);
      for (int c = 0; c < nr2; ++c) {
        IV r2r = v.r[1];
        r2r.tighten_ub(r1.ub);
        IV r2 = get_subinterval(r2r, c, nr2);
        r1.tighten_lb(r2.lb);
        rsb::Cell s = rsb::from_la_r1_r2(v, la, r1, r2);
        if (rsb::empty_cell(s)) continue;
        std::memcpy(out + 14 * n, &s, sizeof s);
        ++n;
      }
    }
  }
  return n;
}

// Branch-and-bound over n cells: every cell gets the reference's leaf
// verdict; unproven cells below max_depth are split on their widest
// dimension, unproven cells at max_depth are reported in `out` (up to cap;
// *n_out counts all).  per_level[d] = number of cells evaluated at depth d
// (max_depth+1 entries); verdict_hist[d*21 + v+1] = histogram of verdicts;
// *trap = number of cells that reached a trap site.
int or_rsb_bnb(const double* cells, size_t n, int max_depth, double* out, size_t cap, size_t* n_out,
               uint64_t* per_level, uint64_t* verdict_hist, int* trap, int threads) {
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  std::vector<rsb::Cell> cur(n), next, fin;
  if (n) std::memcpy(cur.data(), cells, n * sizeof(rsb::Cell));
  int any_trap = 0;
  for (int d = 0; d <= max_depth && !cur.empty(); ++d) {
    per_level[d] = cur.size();
    std::vector<int32_t> verdict(cur.size());
    std::vector<uint8_t> tr(cur.size());
    std::atomic<size_t> pos{0};
    auto work = [&]() {
      for (;;) {
        size_t s = pos.fetch_add(16);
        if (s >= cur.size()) break;
        size_t e = std::min(s + 16, cur.size());
        for (size_t i = s; i < e; ++i) {
          g_trap = 0;
          verdict[i] = rsb::cell_verdict(cur[i]);
          tr[i] = g_trap ? 1 : 0;
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    next.clear();
    for (size_t i = 0; i < cur.size(); ++i) {
      any_trap += tr[i];
      verdict_hist[d * 21 + verdict[i] + 1]++;
      if (verdict[i] >= 0) continue;
      if (d == max_depth) {
        fin.push_back(cur[i]);
      } else {
        rsb::Cell ch[2];
        rsb::bisect(cur[i], ch);
        next.push_back(ch[0]);
        next.push_back(ch[1]);
      }
    }
    cur.swap(next);
  }
  *n_out = fin.size();
  *trap = any_trap;
  const size_t m = std::min(fin.size(), cap);
  if (m) std::memcpy(out, fin.data(), m * sizeof(rsb::Cell));
  return 0;
}

// The reference kernel's enumeration of one cell (rectangle_size_bound.cu:
// 161-236) with piece counts n[4] for r3..r6; unproven leaves (tightened, as
// the reference pushes them) to out (up to cap), *n_out counts all.
int or_rsb_find_critical(const double* cell, const int* n, double* out, size_t cap, size_t* n_out, int* trap,
                         int threads) {
  rsb::Cell in;
  std::memcpy(&in, cell, sizeof in);
  *n_out = 0;
  *trap = 0;
  g_trap = 0;
  bool pre = rsb::shortcut_even_split_recursion(in.la, in.r[0], in.r[1]) ||
             rsb::shortcut_uneven_split_recursion(in.la, in.r[0], in.r[1]);
  if (g_trap) *trap = 1;
  if (pre) return 0;
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  std::mutex mu;
  std::vector<rsb::Cell> all;
  std::atomic<size_t> next{0};
  int any_trap = 0;
  const size_t tasks = (size_t)n[0] * n[1];
  auto work = [&]() {
    std::vector<rsb::Cell> mine;
    int t = 0;
    for (;;) {
      size_t s = next.fetch_add(1);
      if (s >= tasks) break;
      const int i3 = (int)(s / n[1]), i4 = (int)(s % n[1]);
      g_trap = 0;
      IV r3 = get_subinterval(in.r[2], i3, n[0]);
      bool skip = rsb::shortcut_uneven_split_recursion(in.la, in.r[0], in.r[1], r3) ||
                  rsb::shortcut_r1_r2_opposite_corners_recursion(in.la, in.r[0], in.r[1], r3) ||
                  rsb::shortcut_r1_r2_vertical_strip(in.la, in.r[0], in.r[1], r3);
      if (g_trap) t = 1;
      if (skip) continue;
      IV r4 = get_subinterval(in.r[3], i4, n[1]);
      for (int i5 = 0; i5 < n[2]; ++i5) {
        IV r5 = get_subinterval(in.r[4], i5, n[2]);
        for (int i6 = 0; i6 < n[3]; ++i6) {
          IV r6 = get_subinterval(in.r[5], i6, n[3]);
          rsb::Cell v{in.la, {in.r[0], in.r[1], r3, r4, r5, r6}};
          v = rsb::tighten_cell(v);
          g_trap = 0;
          // the leaf chain of cell_verdict after its shortcut steps, on the
          // already tightened leaf (tighten is idempotent)
          int verdict = rsb::chain_verdict(v);
          if (g_trap) t = 1;
          if (verdict < 0) mine.push_back(v);
        }
      }
    }
    std::lock_guard<std::mutex> g(mu);
    all.insert(all.end(), mine.begin(), mine.end());
    any_trap |= t;
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  *n_out = all.size();
  *trap |= any_trap;
  const size_t m = std::min(all.size(), cap);
  if (m) std::memcpy(out, all.data(), m * sizeof(rsb::Cell));
  return 0;
}

// Two children per cell (widest dimension).
void or_rsb_bisect(const double* cells, size_t n, double* out) {
  for (size_t i = 0; i < n; ++i) {
    rsb::Cell c;
    std::memcpy(&c, cells + 14 * i, sizeof c);
    rsb::Cell ch[2];
    rsb::bisect(c, ch);
    std::memcpy(out + 28 * i, ch, sizeof ch);
  }
}

}  // extern "C"
