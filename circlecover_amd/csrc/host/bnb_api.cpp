// cc_run_bnb: the branch-and-bound boundary in one call, from a
// configuration to the sorted unproven cells (no reference counterpart: the
// reference subdivides with fixed piece counts, rectangle_size_bound.cu:
// 161-236, 276-310; this drives the widest-dimension bisection workload over
// the same cells).  It selects one of the drivers and owns the output array,
// which the caller releases with cc_free.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "runtime.hpp"

using namespace cchost;

namespace {

bool iv_less(const cc_iv& a, const cc_iv& b) { return a.lb < b.lb || (a.lb == b.lb && a.ub < b.ub); }

// Lexicographic order of (lambda, r1..r6) by (lb, ub), as the reference's
// Variables::operator< (rectangle_size_bound/variables.hpp:82-94).
bool cell_less(const cc_rsb_cell& a, const cc_rsb_cell& b) {
  if (iv_less(a.la, b.la)) return true;
  if (iv_less(b.la, a.la)) return false;
  for (int i = 0; i < 6; ++i) {
    if (iv_less(a.r[i], b.r[i])) return true;
    if (iv_less(b.r[i], a.r[i])) return false;
  }
  return false;
}

}  // namespace

extern "C" {

int cc_run_bnb(cc_ctx* ctx, const cc_bnb_config* cfg, cc_rsb_cell** unproven, size_t* n, cc_bnb_stats* stats) {
  if (!ctx || !cfg || !unproven || !n) return CC_E_ARG;
  *unproven = nullptr;
  *n = 0;
  if (cfg->max_depth < 0 || cfg->max_depth >= CC_BNB_MAX_DEPTH) return fail(ctx, CC_E_ARG, "max_depth out of range");
  if (cfg->form != CC_BNB_PLAIN && cfg->form != CC_BNB_BATCHED && cfg->form != CC_BNB_LEARNED &&
      cfg->form != CC_BNB_PLAIN_EXACT)
    return fail(ctx, CC_E_ARG, "unknown form");
  if (cfg->form == CC_BNB_LEARNED && !cfg->ranker) return fail(ctx, CC_E_ARG, "the learned form needs a ranker");
  clear_interrupt(ctx);
  // initial cells: the caller's, or the grid of the configuration
  std::vector<cc_rsb_cell> grid;
  const cc_rsb_cell* cells = cfg->cells;
  size_t ncells = cfg->ncells;
  if (!cells) {
    if (cfg->nla < 1 || cfg->nr1 < 1 || cfg->nr2 < 1 || !(cfg->la_lb <= cfg->la_ub) || cfg->la_lb <= 0.0)
      return fail(ctx, CC_E_ARG, "grid needs la_lb <= la_ub, la_lb > 0 and nla, nr1, nr2 >= 1");
    const double w = 9.0 / 64.0;  // rectangle_size_bound/values.hpp:33-34: squared radius bound (3/8)^2
    const cc_rsb_cell outer{{cfg->la_lb, cfg->la_ub}, {{0.0, w}, {0.0, w}, {0.0, w}, {0.0, w}, {0.0, w}, {0.0, w}}};
    const size_t full = (size_t)cfg->nla * cfg->nr1 * cfg->nr2;
    if (full > ((size_t)1 << 31)) return fail(ctx, CC_E_ARG, "grid too large");
    try {
      grid.resize(full);
    } catch (const std::bad_alloc&) {
      return fail(ctx, CC_E_ARG, "grid too large for host memory");
    }
    grid.resize(cc_rsb_grid(&outer, cfg->nla, cfg->nr1, cfg->nr2, grid.data(), grid.size()));
    cells = grid.data();
    ncells = grid.size();
  } else if (ncells == 0) {
    if (stats) std::memset(stats, 0, sizeof *stats);
    return CC_OK;
  }
  // The drivers write into a caller buffer and report the size they needed;
  // start from a guess and retry once with the exact size.
  size_t cap = 4096, got = 0;
  cc_rsb_cell* out = nullptr;
  int rc = CC_E_OVERFLOW;
  for (int attempt = 0; attempt < 2 && rc == CC_E_OVERFLOW; ++attempt) {
    std::free(out);
    out = static_cast<cc_rsb_cell*>(std::malloc((cap ? cap : 1) * sizeof(cc_rsb_cell)));
    if (!out) return fail(ctx, CC_E_ARG, "cannot allocate the output array");
    switch (cfg->form) {
      case CC_BNB_PLAIN: rc = cc_rsb_bnb(ctx, cells, ncells, cfg->max_depth, out, cap, &got, stats); break;
      case CC_BNB_PLAIN_EXACT:
        rc = cc_rsb_bnb_exact(ctx, cells, ncells, cfg->max_depth, out, cap, &got, stats);
        break;
      case CC_BNB_BATCHED:
        rc = cc_rsb_bnb_batched(ctx, cells, ncells, cfg->max_depth, cfg->params, out, cap, &got, stats);
        break;
      default:
        rc = cc_rsb_bnb_learned(ctx, cells, ncells, cfg->max_depth, cfg->ranker, cfg->params, out, cap, &got,
                                stats);
    }
    if (rc == CC_E_OVERFLOW) cap = got;
  }
  if (rc != CC_OK) {
    std::free(out);
    return rc;
  }
  std::sort(out, out + got, cell_less);
  *unproven = out;
  *n = got;
  return CC_OK;
}

void cc_free(void* p) { std::free(p); }

}  // extern "C"
