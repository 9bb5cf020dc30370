// CPU oracle — TEST INFRASTRUCTURE ONLY (see iv.hpp).
// extern "C" surface used by tests/, the smoke check and the benchmark's CPU
// baseline.  Never linked into the product library.
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <thread>
#include <vector>

#include "criteria.hpp"
#include "rsb_chain.hpp"

using namespace ccor;

namespace {

// The synthetic batched-criterion workload: ten criteria on a 4-D box
// (lambda, r1, r2, r3), evaluated without short-circuit.  Bit k of the
// verdict mask is criterion k; bit 16+k flags a trap inside criterion k.
//   0 TR r1_strategies(la, r1, r2.ub)                  r1_strategies.cu:27
//   1 TR r1_r2_strategies_simple(la, r1, r2, r3.ub)     r1_r2_strategies.cu:46
//   2 TR r1_r2_strategies_stacked(la, r1, r2, r3.ub)    r1_r2_strategies.cu:74
//   3 TR r1_r2_r3_strategy_horizontal(.., ub_r4)        r1_r2_r3_strategies.cu:28
//   4 TR r1_r2_r3_strategy_cover_strip(.., ub_r4)       r1_r2_r3_strategies.cu:69
//   5 RSB shortcut_even_split_recursion(la, r1, r2)     even_split_recursion.cu:42
//   6 RSB shortcut_uneven_split_recursion(la, r1, r2)   uneven_split_recursion.cu:78
//   7 RSB shortcut_uneven_split_recursion(la,r1,r2,r3)  uneven_split_recursion.cu:102
//   8 RSB shortcut_r1_r2_opposite_corners_recursion     r1_r2_opposite_corners_recursion.cu:28
//   9 RSB shortcut_r1_r2_vertical_strip                 r1_r2_vertical_strip.cu:28
// ub_r4 = min(r3.ub, (required_weight_for(la)-r1-r2-r3).ub) as in
// tight_rectangle.cu:97-102.
uint32_t eval_box(const IV& la, const IV& r1, const IV& r2, const IV& r3, uint64_t* ops_per_crit) {
  double a = r3.ub, b = (tr::required_weight_for(la) - r1 - r2 - r3).ub;
  const double ub_r4 = a < b ? a : b;
  uint32_t mask = 0;
  for (int k = 0; k < 10; ++k) {
    g_trap = 0;
#ifdef CC_ORACLE_COUNT
    uint64_t before = g_op_count;
#endif
    bool ok = false;
    switch (k) {
      case 0: ok = tr::r1_strategies(la, r1, r2.ub); break;
      case 1: ok = tr::r1_r2_strategies_simple(la, r1, r2, r3.ub); break;
      case 2: ok = tr::r1_r2_strategies_stacked(la, r1, r2, r3.ub); break;
      case 3: ok = tr::r1_r2_r3_strategy_horizontal(la, r1, r2, r3, ub_r4); break;
      case 4: ok = tr::r1_r2_r3_strategy_cover_strip(la, r1, r2, r3, ub_r4); break;
      case 5: ok = rsb::shortcut_even_split_recursion(la, r1, r2); break;
      case 6: ok = rsb::shortcut_uneven_split_recursion(la, r1, r2); break;
      case 7: ok = rsb::shortcut_uneven_split_recursion(la, r1, r2, r3); break;
      case 8: ok = rsb::shortcut_r1_r2_opposite_corners_recursion(la, r1, r2, r3); break;
      case 9: ok = rsb::shortcut_r1_r2_vertical_strip(la, r1, r2, r3); break;
    }
#ifdef CC_ORACLE_COUNT
    if (ops_per_crit) ops_per_crit[k] += g_op_count - before;
#else
    (void)ops_per_crit;
#endif
    if (ok) mask |= 1u << k;
    if (g_trap) mask |= 1u << (16 + k);
  }
  return mask;
}

inline IV box_iv(const double* soa, size_t ld, size_t dim, size_t i) {
  return IV(soa[(2 * dim) * ld + i], soa[(2 * dim + 1) * ld + i]);
}

}  // namespace

extern "C" {

// Evaluates the ten-criterion workload on n boxes stored as 8 SoA arrays of
// stride ld (la.lb, la.ub, r1.lb, r1.ub, r2.lb, r2.ub, r3.lb, r3.ub).
// Uses `threads` worker threads with dynamic chunking (threads <= 0: all).
int or_c2_eval(const double* soa, size_t n, size_t ld, uint32_t* mask, uint8_t* first_ok, int threads) {
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  std::atomic<size_t> next{0};
  const size_t chunk = 256;
  auto work = [&]() {
    for (;;) {
      size_t s = next.fetch_add(chunk);
      if (s >= n) break;
      size_t e = s + chunk < n ? s + chunk : n;
      for (size_t i = s; i < e; ++i) {
        uint32_t m = eval_box(box_iv(soa, ld, 0, i), box_iv(soa, ld, 1, i), box_iv(soa, ld, 2, i),
                              box_iv(soa, ld, 3, i), nullptr);
        mask[i] = m;
        uint8_t f = 255;
        for (int k = 0; k < 10; ++k)
          if (m & (1u << k)) {
            f = (uint8_t)k;
            break;
          }
        first_ok[i] = f;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  return 0;
}

// Counting build only: per-criterion totals of directed FP64 operations over
// the n boxes (ops[10]); returns -1 when the library was built without counting.
int or_c2_count_ops(const double* soa, size_t n, size_t ld, uint64_t* ops) {
#ifdef CC_ORACLE_COUNT
  for (int k = 0; k < 10; ++k) ops[k] = 0;
  for (size_t i = 0; i < n; ++i)
    eval_box(box_iv(soa, ld, 0, i), box_iv(soa, ld, 1, i), box_iv(soa, ld, 2, i), box_iv(soa, ld, 3, i), ops);
  return 0;
#else
  (void)soa;
  (void)n;
  (void)ld;
  (void)ops;
  return -1;
#endif
}

// Directed-op count of the rectangle_size_bound verdict of each cell
// (counting build only; -1 otherwise).  ops[i] = operations of cell i,
// verdict[i] = its verdict.
int or_rsb_count_ops(const double* cells, size_t n, uint64_t* ops, int32_t* verdict) {
#ifdef CC_ORACLE_COUNT
  for (size_t i = 0; i < n; ++i) {
    const double* p = cells + 14 * i;
    rsb::Cell c;
    c.la = IV(p[0], p[1]);
    for (int k = 0; k < 6; ++k) c.r[k] = IV(p[2 + 2 * k], p[3 + 2 * k]);
    uint64_t before = g_op_count;
    verdict[i] = rsb::cell_verdict(c);
    ops[i] = g_op_count - before;
  }
  return 0;
#else
  (void)cells;
  (void)n;
  (void)ops;
  (void)verdict;
  return -1;
#endif
}

// Per-step directed-op counts of every step evaluated on its own (counting
// build only): ops[20*i + k] for cell i, step k; mask[i] as or_rsb_masks.
int or_rsb_count_steps(const double* cells, size_t n, uint64_t* ops, uint64_t* mask) {
#ifdef CC_ORACLE_COUNT
  for (size_t i = 0; i < n; ++i) {
    const double* p = cells + 14 * i;
    rsb::Cell c;
    c.la = IV(p[0], p[1]);
    for (int k = 0; k < 6; ++k) c.r[k] = IV(p[2 + 2 * k], p[3 + 2 * k]);
    mask[i] = rsb::cell_mask(c, ops + 20 * i);
  }
  return 0;
#else
  (void)cells;
  (void)n;
  (void)ops;
  (void)mask;
  return -1;
#endif
}

// Primitive access for the golden-vector tests: op 0..5 = add, mul, div
// (binary) and sqrt (unary, b ignored), in round-down (dir=0) or round-up
// (dir=1).
double or_prim(int op, int dir, double a, double b) {
  switch (op) {
    case 0: return dir ? add_ru(a, b) : add_rd(a, b);
    case 1: return dir ? mul_ru(a, b) : mul_rd(a, b);
    case 2: return dir ? div_ru(a, b) : div_rd(a, b);
    case 3: return dir ? sqrt_ru(a) : sqrt_rd(a);
    default: return 0.0;
  }
}

// Interval-level helpers exposed for closed-form tests.
void or_tr_critical_ratio(double lb, double ub, double* out) {
  IV r = tr::critical_ratio(IV(lb, ub));
  out[0] = r.lb;
  out[1] = r.ub;
}
double or_two_disks_maximize_height(double r1lb, double r2lb, double ub_w) {
  return rsb::two_disks_maximize_height(IV(r1lb, r1lb), IV(r2lb, r2lb), ub_w);
}
int or_disk_satisfies_size_bound(double r_sq_ub, double h_lb) { return rsb::disk_satisfies_size_bound(r_sq_ub, h_lb); }
double or_bound_required_height(double r_sq_ub) { return rsb::bound_required_height(r_sq_ub); }

// Circle-circle intersection of two point circles (centres and squared
// radii given as points); writes the two points as 8 doubles (x.lb, x.ub,
// y.lb, y.ub per point) and returns definitely_intersecting.
int or_intersection(const double* c1, const double* c2, double* out) {
  Circle a{{IV::pt(c1[0]), IV::pt(c1[1])}, IV::pt(c1[2])};
  Circle b{{IV::pt(c2[0]), IV::pt(c2[1])}, IV::pt(c2[2])};
  Intersection x = intersection(a, b);
  for (int p = 0; p < 2; ++p) {
    out[4 * p + 0] = x.p[p].x.lb;
    out[4 * p + 1] = x.p[p].x.ub;
    out[4 * p + 2] = x.p[p].y.lb;
    out[4 * p + 3] = x.p[p].y.ub;
  }
  return x.definitely_intersecting;
}

// Interval binary operation (0 add, 1 sub, 2 mul, 3 div, 4 scalar-mul with b
// as the point b.lb, 5 square of a, 6 sqrt of a, 7 reciprocal of a).
void or_iv_op(int op, const double* a, const double* b, double* out) {
  IV x(a[0], a[1]), y(b[0], b[1]), r(0, 0);
  g_trap = 0;
  switch (op) {
    case 0: r = x + y; break;
    case 1: r = x - y; break;
    case 2: r = x * y; break;
    case 3: r = x / y; break;
    case 4: r = x * b[0]; break;
    case 5: r = iv_square(x); break;
    case 6: r = iv_sqrt(x); break;
    case 7: r = iv_reciprocal(x); break;
  }
  out[0] = r.lb;
  out[1] = r.ub;
  out[2] = g_trap;
}

// Round-to-nearest helpers: out = {0.5*(a+b), fma(a,b,c), (a-b)/c}.
void or_rn(double a, double b, double c, double* out) {
  out[0] = rn_mid(a, b);
  out[1] = rn_fma(a, b, c);
  out[2] = rn_div(rn_sub(a, b), c);
}

// Verdict of n 7-D rectangle_size_bound cells (14 doubles each: la, r1..r6
// as lb,ub pairs): out[i] = first succeeding step (see rsb_chain.hpp) or -1;
// trap[i] = 1 when a reference trap site was reached.
int or_rsb_masks(const double* cells, size_t n, uint64_t* out, int threads) {
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      size_t s = next.fetch_add(16);
      if (s >= n) break;
      size_t e = s + 16 < n ? s + 16 : n;
      for (size_t i = s; i < e; ++i) {
        const double* p = cells + 14 * i;
        rsb::Cell c;
        c.la = IV(p[0], p[1]);
        for (int k = 0; k < 6; ++k) c.r[k] = IV(p[2 + 2 * k], p[3 + 2 * k]);
        out[i] = rsb::cell_mask(c);
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  return 0;
}

// C14's stages on n point-or-interval cells (14 doubles each) with np r7
// pieces per cell (pieces: n * np (lb, ub) pairs): out[7 * i ..] and
// bits[np * i ..] as rsb::seven_disk_probe.
void or_rsb_c14_probe(const double* cells, size_t n, const double* pieces, int np, double* out, uint32_t* bits) {
  for (size_t i = 0; i < n; ++i) {
    rsb::Cell c;
    c.la = IV(cells[14 * i], cells[14 * i + 1]);
    for (int k = 0; k < 6; ++k) c.r[k] = IV(cells[14 * i + 2 + 2 * k], cells[14 * i + 3 + 2 * k]);
    rsb::seven_disk_probe(c, pieces + 2 * np * i, np, out + 7 * i, bits + np * i);
  }
}

int or_rsb_verdicts(const double* cells, size_t n, int32_t* out, uint8_t* trap, int threads) {
  if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
  if (threads < 1) threads = 1;
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      size_t s = next.fetch_add(16);
      if (s >= n) break;
      size_t e = s + 16 < n ? s + 16 : n;
      for (size_t i = s; i < e; ++i) {
        const double* p = cells + 14 * i;
        rsb::Cell c;
        c.la = IV(p[0], p[1]);
        for (int k = 0; k < 6; ++k) c.r[k] = IV(p[2 + 2 * k], p[3 + 2 * k]);
        g_trap = 0;
        out[i] = rsb::cell_verdict(c);
        trap[i] = (uint8_t)g_trap;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  return 0;
}

// Counting build: the calling thread's directed-operation counter (the
// drivers run on the calling thread when given one thread), and its reset.
uint64_t or_op_count() {
#ifdef CC_ORACLE_COUNT
  return g_op_count;
#else
  return 0;
#endif
}
void or_op_count_reset() {
#ifdef CC_ORACLE_COUNT
  g_op_count = 0;
#endif
}

}  // extern "C"

