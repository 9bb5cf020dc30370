// CPU oracle — TEST INFRASTRUCTURE ONLY (see iv.hpp).
// The tight_rectangle enumeration of one kernel launch, its host-side merge
// and the small branch-and-bound built on it, restated as plain loops.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "criteria.hpp"

namespace ccor {
namespace tr {

// kernel_find_critical_intervals_in (tight_rectangle.cu:169-262) for the
// r1 piece `i1` of `n[0]`; unproven leaves are appended to `out`.  The
// reference's grid-stride loops visit every (i1, i2, i3, i4) exactly once, so
// a sequential loop nest yields the same multiset of leaves.
static void find_critical_r1(const Cell& in, const int n[4], int i1, std::vector<Cell>& out, int& trap) {
  g_trap = 0;
  const IV weight = required_weight_for(in.la);
  const IV r1r = r1_range(in.la).intersect(in.r[0]);
  if (r1r.empty()) return;
  const IV r1 = get_subinterval(r1r, i1, n[0]);
  const IV r2r = r2_range(in.la, r1).intersect(in.r[1]);
  if (r2r.empty()) return;
  if (r1_strategies(in.la, r1, r2r.ub)) return;
  for (int i2 = 0; i2 < n[1]; ++i2) {
    const IV r2 = get_subinterval(r2r, i2, n[1]);
    if (r1_strategies(in.la, r1, r2.ub)) continue;
    const IV r3r = r3_range(in.la, r1, r2).intersect(in.r[2]);
    if (r3r.empty()) continue;
    if (r1_r2_strategies(in.la, r1, r2, r3r.ub)) continue;
    for (int i3 = 0; i3 < n[2]; ++i3) {
      const IV r3 = get_subinterval(r3r, i3, n[2]);
      if (r1_r2_strategies(in.la, r1, r2, r3.ub)) continue;
      const IV r4r = r4_range(in.la, r1, r2, r3).intersect(in.r[3]);
      if (r4r.empty()) continue;
      if (r1_r2_r3_strategies(in.la, r1, r2, r3, r4r.ub)) continue;
      IV R4 = weight - r1 - r2 - r3;
      R4.tighten_lb(0.0);
      if (R4.empty()) continue;
      for (int i4 = 0; i4 < n[3]; ++i4) {
        const IV r4 = get_subinterval(r4r, i4, n[3]);
        if (r1_r2_r3_strategies(in.la, r1, r2, r3, r4.ub)) continue;
        Cell v{in.la, {r1, r2, r3, r4}};
        v.r[2].tighten_lb(v.r[3].lb);
        v.r[1].tighten_lb(v.r[2].lb);
        v.r[0].tighten_lb(v.r[1].lb);
        IV R = weight - v.r[0] - v.r[1] - v.r[2] - v.r[3];
        R.tighten_lb(0.0);
        if (R.empty()) continue;
        if (!two_by_two_strategies(v, R) && !r1_r2_r3_strategies(v, R4) && !r1_r2_large_r3_r4_gaps_strategy(v, R))
          out.push_back(v);
      }
    }
  }
  if (g_trap) trap = 1;
}

// Widest dimension (absolute width ub - lb, ties to the lowest index among
// lambda, r1..r4) split into the two halves get_subinterval(x, k, 2).
static void bisect(const Cell& c, Cell* out) {
  const IV* d[5] = {&c.la, &c.r[0], &c.r[1], &c.r[2], &c.r[3]};
  int best = 0;
  double bw = rn_sub(d[0]->ub, d[0]->lb);
  for (int k = 1; k < 5; ++k) {
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

// merge_adjacent_r1 (tight_rectangle.cu:121-159): sort by r1 (lb, ub), then
// join every record whose r1 touches the running record's r1.
static std::vector<Cell> merge_adjacent_r1(std::vector<Cell> v) {
  std::vector<Cell> res;
  if (v.empty()) return res;
  std::stable_sort(v.begin(), v.end(), [](const Cell& a, const Cell& b) {
    return a.r[0].lb < b.r[0].lb || (a.r[0].lb == b.r[0].lb && a.r[0].ub < b.r[0].ub);
  });
  res.push_back(v.front());
  for (const Cell& c : v) {
    Cell& p = res.back();
    bool touch = p.r[0].lb <= c.r[0].lb ? p.r[0].ub >= c.r[0].lb : c.r[0].ub >= p.r[0].lb;
    if (touch) {
      for (int k = 0; k < 4; ++k) p.r[k].do_join(c.r[k]);
      continue;
    }
    res.push_back(c);
  }
  return res;
}

}  // namespace tr
}  // namespace ccor

using namespace ccor;

extern "C" {

// Enumeration of ncells cells (80-byte records) with piece counts n[4];
// leaves are written to out (up to cap records), *n_out = number found,
// *trap = 1 if a reference trap site was reached.  Work is split over
// (cell, r1 piece) tasks with dynamic chunking.
int or_tr_find_critical(const double* cells, size_t ncells, const int* n, double* out, size_t cap, size_t* n_out,
                        int* trap, int threads) {
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  const size_t tasks = ncells * (size_t)n[0];
  std::atomic<size_t> next{0};
  std::mutex mu;
  std::vector<tr::Cell> all;
  int any_trap = 0;
  auto work = [&]() {
    std::vector<tr::Cell> mine;
    int t = 0;
    for (;;) {
      size_t s = next.fetch_add(1);
      if (s >= tasks) break;
      const double* p = cells + 10 * (s / n[0]);
      tr::Cell c;
      c.la = IV(p[0], p[1]);
      for (int k = 0; k < 4; ++k) c.r[k] = IV(p[2 + 2 * k], p[3 + 2 * k]);
      tr::find_critical_r1(c, n, (int)(s % n[0]), mine, t);
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
  *trap = any_trap;
  const size_t m = all.size() < cap ? all.size() : cap;
  if (m) std::memcpy(out, all.data(), m * sizeof(tr::Cell));
  return 0;
}

// The leaf criteria of n tight_rectangle leaves (80-byte records), as the
// enumeration forms them before its leaf test (tight_rectangle.cu:217-246):
// R4 = weight - r1 - r2 - r3 and R = weight - r1 - .. - r4 (lower bounds
// raised to 0) from the leaf's radii with their lower bounds tightened
// downwards from r4.  out[i] = tr::leaf_strategy_bits, with bit 31 set when R
// is empty (the enumeration skips such a leaf).
void or_tr_leaf_masks(const double* cells, size_t n, uint32_t* out) {
  for (size_t i = 0; i < n; ++i) {
    tr::Cell c;
    std::memcpy(&c, cells + 10 * i, sizeof c);
    IV weight = tr::required_weight_for(c.la);
    IV R4 = weight - c.r[0] - c.r[1] - c.r[2];
    R4.tighten_lb(0.0);
    c.r[2].tighten_lb(c.r[3].lb);
    c.r[1].tighten_lb(c.r[2].lb);
    c.r[0].tighten_lb(c.r[1].lb);
    IV R = weight - c.r[0] - c.r[1] - c.r[2] - c.r[3];
    R.tighten_lb(0.0);
    out[i] = R.empty() || R4.empty() ? 1u << 31 : tr::leaf_strategy_bits(c, R4, R);
  }
}

// The n pieces get_subinterval([lb, ub], i, n) as 2n doubles.
void or_subintervals(double lb, double ub, int n, double* out) {
  for (int i = 0; i < n; ++i) {
    IV p = get_subinterval(IV(lb, ub), i, n);
    out[2 * i] = p.lb;
    out[2 * i + 1] = p.ub;
  }
}

// Two children per cell, split on the widest dimension.
void or_tr_bisect(const double* cells, size_t n, double* out) {
  static_assert(sizeof(tr::Cell) == 80, "record layout");
  for (size_t i = 0; i < n; ++i) {
    tr::Cell c;
    std::memcpy(&c, cells + 10 * i, sizeof c);
    tr::Cell ch[2];
    tr::bisect(c, ch);
    std::memcpy(out + 20 * i, ch, sizeof ch);
  }
}

// merge_adjacent_r1 of n records; returns the merged count.
size_t or_tr_merge_adjacent_r1(const double* cells, size_t n, double* out) {
  std::vector<tr::Cell> v(n);
  if (n) std::memcpy(v.data(), cells, n * sizeof(tr::Cell));
  std::vector<tr::Cell> m = tr::merge_adjacent_r1(v);
  if (!m.empty()) std::memcpy(out, m.data(), m.size() * sizeof(tr::Cell));
  return m.size();
}

}  // extern "C"

// kernel_compute_distance_from_theoretical_worst_cases
// (tight_rectangle.cu:343-456), sequential over the cells in order.  Output
// record (same layout as the library's report block): ints have_2d, have_3d;
// then IV la_2d, la_3d, d_r1_2d, 2 cells, d_r2_2d, 2 cells, doubles lb_h1_2d,
// ub_la_by_S1, ub_R3_2d, IV d_r1_3d, 2 cells, d_r2_3d, 2 cells, d_r3_3d,
// 2 cells, double ub_R4_3d, IV S1, h2, S3, double pocket_3d, int trap.
namespace {
struct MaxDist {
  int have_2d, have_3d;
  ccor::IV la_2d, la_3d;
  ccor::IV d_r1_2d;
  ccor::tr::Cell d_r1_2d_crit[2];
  ccor::IV d_r2_2d;
  ccor::tr::Cell d_r2_2d_crit[2];
  double lb_h1_2d, ub_la_by_S1, ub_R3_2d;
  ccor::IV d_r1_3d;
  ccor::tr::Cell d_r1_3d_crit[2];
  ccor::IV d_r2_3d;
  ccor::tr::Cell d_r2_3d_crit[2];
  ccor::IV d_r3_3d;
  ccor::tr::Cell d_r3_3d_crit[2];
  double ub_R4_3d;
  ccor::IV S1, h2, S3;
  double pocket_3d;
  int trap;
};
}  // namespace

extern "C" size_t or_tr_max_distances_size() { return sizeof(MaxDist); }

extern "C" void or_tr_max_distances(const double* cells, size_t n, void* out) {
  using namespace ccor;
  MaxDist r;
  std::memset(&r, 0, sizeof r);
  g_trap = 0;
  r.la_3d = IV::pt(1.0);
  r.la_2d = IV::pt(2.0);
  r.d_r1_2d = r.d_r2_2d = IV::pt(0.0);
  r.d_r1_3d = r.d_r2_3d = r.d_r3_3d = IV::pt(0.0);
  r.pocket_3d = -1.0;
  r.lb_h1_2d = 1.0;
  r.ub_la_by_S1 = 1.0;
  for (size_t i = 0; i < n; ++i) {
    tr::Cell v;
    std::memcpy(&v, cells + 10 * i, sizeof v);
    if (v.r[2].ub < 0.1) {
      r.have_2d = 1;
      r.la_2d.do_join(v.la);
      IV r1t = 0.25 * (iv_square(v.la) + 1.0);
      IV d1 = v.r[0] - r1t;
      if (d1.lb < r.d_r1_2d.lb) r.d_r1_2d_crit[0] = v;
      if (d1.ub > r.d_r1_2d.ub) r.d_r1_2d_crit[1] = v;
      r.d_r1_2d.do_join(d1);
      IV d2 = v.r[1] - IV::pt(0.25);
      if (d2.lb < r.d_r2_2d.lb) r.d_r2_2d_crit[0] = v;
      if (d2.ub > r.d_r2_2d.ub) r.d_r2_2d_crit[1] = v;
      r.d_r2_2d.do_join(d2);
      IV R3 = tr::required_weight_for(v.la) - v.r[0] - v.r[1];
      if (R3.ub > r.ub_R3_2d) r.ub_R3_2d = R3.ub;
      double S1 = mul_rd(2.0, sqrt_rd(add_rd(v.r[0].lb, -0.25)));
      double q = div_ru(v.la.ub, S1);
      if (q > r.ub_la_by_S1) r.ub_la_by_S1 = q;
      double half = add_ru(v.la.ub, -mul_rd(0.5, S1));
      double h1 = mul_rd(2.0, sqrt_rd(add_rd(v.r[0].lb, -mul_ru(half, half))));
      if (h1 < r.lb_h1_2d) r.lb_h1_2d = h1;
    } else {
      r.have_3d = 1;
      r.la_3d.do_join(v.la);
      IV la2 = iv_square(v.la);
      IV la4 = iv_square(la2);
      IV th = (16.0 * la4 + 40.0 * la2 + 9.0) / (256.0 * la2);
      IV d1 = v.r[0] - th;
      if (d1.lb < r.d_r1_3d.lb) r.d_r1_3d_crit[0] = v;
      if (d1.ub > r.d_r1_3d.ub) r.d_r1_3d_crit[1] = v;
      IV d2 = v.r[1] - th;
      if (d2.lb < r.d_r2_3d.lb) r.d_r2_3d_crit[0] = v;
      if (d2.ub > r.d_r2_3d.ub) r.d_r2_3d_crit[1] = v;
      IV d3 = v.r[2] - th;
      if (d3.lb < r.d_r3_3d.lb) r.d_r3_3d_crit[0] = v;
      if (d3.ub > r.d_r3_3d.ub) r.d_r3_3d_crit[1] = v;
      r.d_r1_3d.do_join(d1);
      r.d_r2_3d.do_join(d2);
      r.d_r3_3d.do_join(d3);
      IV R4 = tr::required_weight_for(v.la) - v.r[0] - v.r[1] - v.r[2];
      if (R4.ub > r.ub_R4_3d) r.ub_R4_3d = R4.ub;
      IV S1 = iv_sqrt(4.0 * v.r[0] - 1.0);
      IV h2 = iv_sqrt(4.0 * v.r[1] - iv_square(v.la - S1));
      IV S3 = iv_sqrt(4.0 * v.r[2] - iv_square(1.0 - h2));
      double pw = (v.la - S1 - S3).ub;
      if (pw > r.pocket_3d) {
        if (r.pocket_3d < 0.0) {
          r.S1 = S1;
          r.h2 = h2;
          r.S3 = S3;
        }
        r.pocket_3d = pw;
      }
      r.S1.do_join(S1);
      r.h2.do_join(h2);
      r.S3.do_join(S3);
    }
  }
  r.trap = g_trap;
  std::memcpy(out, &r, sizeof r);
}
