// rectangle_size_bound drivers of the C ABI: the reference's initial grid
// (host loop of find_critical_intervals_in, Prover/rectangle_size_bound/
// rectangle_size_bound.cu:276-310), its per-cell kernel enumeration (:161-236,
// host wrapper :238-274) over a batch of cells, and the widest-dimension
// branch-and-bound over 7-D cells (level-synchronous form).
#include <chrono>
#include <cmath>
#include <cstring>

#include "runtime.hpp"

using namespace cchost;

namespace {

constexpr size_t kRec = 112, kS3 = 24;
// Reference piece counts for r3..r6 (rectangle_size_bound.cu:39-57).
constexpr int kInitCounts[4] = {64, 64, 64, 64};
constexpr int kSubdivCounts[4] = {16, 16, 16, 32};

struct Iv {
  double lb, ub;
};
struct Cell7 {
  Iv la;
  Iv r[6];
};
static_assert(sizeof(Cell7) == kRec, "record layout");

unsigned blocks_for(uint64_t threads, unsigned block) { return (unsigned)((threads + block - 1) / block); }

// get_subinterval (Prover/subintervals.hpp:32-43), lb + diff*i fused.
Iv piece(Iv x, int i, int n) {
  const double diff = (x.ub - x.lb) / (double)n;
  return Iv{std::fma(diff, (double)i, x.lb), i + 1 == n ? x.ub : std::fma(diff, (double)(i + 1), x.lb)};
}
Iv intersect(Iv a, Iv b) { return Iv{a.lb < b.lb ? b.lb : a.lb, a.ub > b.ub ? b.ub : a.ub}; }

// Variables::tighten / empty / from_la_r1_r2 (variables.hpp:37-80).
Cell7 tighten(Cell7 c) {
  double lb = 0.0;
  for (int i = 5; i > 0; --i) {
    if (lb < c.r[i].lb) lb = c.r[i].lb;
    else c.r[i].lb = lb;
  }
  return c;
}
bool empty(const Cell7& c) {
  if (c.la.lb > c.la.ub) return true;
  for (int i = 0; i < 6; ++i)
    if (c.r[i].lb >= c.r[i].ub) return true;
  return false;
}
Cell7 from_la_r1_r2(const Cell7& v, Iv la, Iv r1, Iv r2) {
  Cell7 r;
  r.la = intersect(v.la, la);
  r.r[0] = intersect(v.r[0], r1);
  r.r[1] = intersect(v.r[1], r2);
  for (int i = 2; i < 6; ++i) {
    r.r[i] = v.r[i];
    if (r2.ub < r.r[i].ub) r.r[i].ub = r2.ub;
  }
  return tighten(r);
}

int read_counts(cc_ctx* ctx, int k) {
  CC_HIP(ctx, hipMemcpyAsync(ctx->h_counts, ctx->d_counts, k * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CC_OK;
}

// The reference's kernel shape (rectangle_size_bound.cu:161-251): a (64, 64)
// grid of 16x16 blocks per cell, the cells of the batch along the grid's
// third dimension; leaves in scratch().leaves as enumerate_dev leaves them
// (counter slots 1 leaves, 2 trap), relaunched once if the buffer was short.
int enumerate_plain(cc_ctx* ctx, const void* d_cells, size_t ncells, const int n[4], size_t* n_leaves, int* trap) {
  Scratch& S = scratch(ctx);
  unsigned long long* cnt = static_cast<unsigned long long*>(ctx->d_counts);
  int* d_trap = reinterpret_cast<int*>(cnt + 2);
  hipFunction_t f;
  CC_HIP(ctx, kernel(ctx, ctx->enum_form == CC_ENUM_PLAIN ? "ccrd_rsb_plain" : "ccrd_rsb_plain_x", &f));
  if (ncells > 65535) return fail(ctx, CC_E_ARG, "plain enumeration: at most 65535 cells a call");
  if (!S.leaves.p) CC_HIP(ctx, S.leaves.ensure((size_t)1 << 20));
  uint64_t cb = 0;
  int rc;
  for (int pass = 0; pass < 2; ++pass) {
    const uint64_t cap = S.leaves.cap / kRec;
    CC_HIP(ctx, hipMemsetAsync(cnt + 1, 0, sizeof(unsigned long long), ctx->stream));
    CC_HIP(ctx, launch(f, dim3(64, 64, (unsigned)ncells), dim3(16, 16), 0, ctx->stream, d_cells, n[0], n[1], n[2],
                       n[3], ctx->d_konst, S.leaves.p, cap, cnt + 1, d_trap));
    if ((rc = read_counts(ctx, 3))) return rc;
    cb = ctx->h_counts[1];
    if (cb <= cap) break;
    if (pass == 1) return fail(ctx, CC_E_OVERFLOW, "leaf count grew between identical launches");
    CC_HIP(ctx, S.leaves.ensure(cb * kRec));
  }
  *trap = (int)(ctx->h_counts[2] & 0xffffffffu);
  *n_leaves = cb;
  return CC_OK;
}

// Enumeration of the device cells with counts n[4] (r3..r6); leaves land in
// scratch().leaves.  Counter slots: 0 level A, 1 leaves, 2 trap (int), 3
// listed non-empty leaves, 4 leaves listed for C14 (both lists 8 bytes per
// leaf, in scratch().enum_list / enum_b).  Level B runs over the level-A
// records in chunks whose leaves fit the lists even if none were empty (up
// to 2^28 leaves a chunk), so the lists cannot overflow and the counts stay
// on the device until the end.
int enumerate_dev(cc_ctx* ctx, const void* d_cells, size_t ncells, const int n[4], size_t* n_leaves, int* trap) {
  Scratch& S = scratch(ctx);
  unsigned long long* cnt = static_cast<unsigned long long*>(ctx->d_counts);
  int* d_trap = reinterpret_cast<int*>(cnt + 2);
  hipFunction_t fa, frows, fla, flb;
  CC_HIP(ctx, kernel(ctx, "ccrd_rsb_enum_r3", &fa));
  CC_HIP(ctx, kernel(ctx, "ccrd_rsb_enum_rows", &frows));
  CC_HIP(ctx, kernel(ctx, "ccrd_rsb_enum_leaf_a", &fla));
  CC_HIP(ctx, kernel(ctx, "ccrd_rsb_enum_leaf_b", &flb));
  *n_leaves = 0;
  *trap = 0;
  CC_HIP(ctx, hipMemsetAsync(cnt, 0, 5 * sizeof(unsigned long long), ctx->stream));
  if (ncells == 0) return CC_OK;
  if (ctx->enum_form != CC_ENUM_BATCHED) return enumerate_plain(ctx, d_cells, ncells, n, n_leaves, trap);
  for (int k = 1; k < 4; ++k)
    if (n[k] < 1 || n[k] > 256) return fail(ctx, CC_E_ARG, "r4..r6 piece counts must be 1..256");
  const uint64_t nc = ncells, ta = nc * (uint64_t)n[0];
  CC_HIP(ctx, S.s1.ensure(ta * kS3));
  CC_HIP(ctx, launch(fa, dim3(blocks_for(ta, 128)), dim3(128), 0, ctx->stream, d_cells, nc, n[0], ctx->d_konst,
                     S.s1.p, cnt + 0, d_trap));
  int rc = read_counts(ctx, 1);
  if (rc) return rc;
  const uint64_t ca = ctx->h_counts[0];
  if (ca > 0xffffffffull) return fail(ctx, CC_E_ARG, "enumeration too large");
  const uint64_t per = (uint64_t)n[1] * (uint64_t)n[2] * (uint64_t)n[3];  // leaves per level-A record
  constexpr uint64_t kListLeaves = 1ull << 28;                           // list room: up to 2 GiB each
  constexpr size_t kRef = 8;                                              // bytes per listed leaf
  const uint64_t chunk = per >= kListLeaves ? 1 : kListLeaves / per;     // records per chunk
  const uint64_t room = (ca < chunk ? ca : chunk) * per;
  uint64_t cb = 0;
  if (ca) {
    CC_HIP(ctx, S.enum_list.ensure(room * kRef));
    CC_HIP(ctx, S.enum_b.ensure(room * kRef));
    if (!S.leaves.p) CC_HIP(ctx, S.leaves.ensure((size_t)1 << 20));
    const unsigned grid = 8192;  // stages A and B: grid-stride over the lists
    for (int pass = 0; pass < 2; ++pass) {  // again with room for every unproven leaf if they overflowed
      const uint64_t cap = S.leaves.cap / kRec;
      CC_HIP(ctx, hipMemsetAsync(cnt + 1, 0, sizeof(unsigned long long), ctx->stream));
      for (uint64_t r0 = 0; r0 < ca; r0 += chunk) {
        const uint64_t nr = ca - r0 < chunk ? ca - r0 : chunk;
        const uint64_t rows = nr * (uint64_t)n[1] * (uint64_t)n[2];
        CC_HIP(ctx, hipMemsetAsync(cnt + 3, 0, 2 * sizeof(unsigned long long), ctx->stream));
        CC_HIP(ctx, launch(frows, dim3(blocks_for(rows, 1024)), dim3(1024), 0, ctx->stream, d_cells, S.s1.p, r0, nr,
                           n[1], n[2], n[3], S.enum_list.p, cnt + 3));
        CC_HIP(ctx, launch(fla, dim3(grid), dim3(128), 0, ctx->stream, d_cells, S.s1.p, S.enum_list.p,
                           (const unsigned long long*)(cnt + 3), n[1], n[2], n[3], ctx->d_konst, S.enum_b.p, cnt + 4,
                           d_trap));
        CC_HIP(ctx, launch(flb, dim3(grid), dim3(128), 0, ctx->stream, d_cells, S.s1.p, S.enum_b.p,
                           (const unsigned long long*)(cnt + 4), n[1], n[2], n[3], ctx->d_konst, S.leaves.p, cap,
                           cnt + 1, d_trap));
      }
      if ((rc = read_counts(ctx, 3))) return rc;
      cb = ctx->h_counts[1];
      if (cb <= cap) break;
      if (pass == 1) return fail(ctx, CC_E_OVERFLOW, "leaf count grew between identical launches");
      CC_HIP(ctx, S.leaves.ensure(cb * kRec));
    }
  } else if ((rc = read_counts(ctx, 3))) {
    return rc;
  }
  *trap = (int)(ctx->h_counts[2] & 0xffffffffu);
  *n_leaves = cb;
  return CC_OK;
}

int upload(cc_ctx* ctx, DevBuf& b, const void* src, size_t bytes) {
  CC_HIP(ctx, b.ensure(bytes));
  if (bytes) CC_HIP(ctx, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return CC_OK;
}

int copy_out(cc_ctx* ctx, const void* d_src, size_t n, cc_rsb_cell* out, size_t cap, size_t* n_out) {
  *n_out = n;
  const size_t m = n < cap ? n : cap;
  if (m) {
    CC_HIP(ctx, hipMemcpyAsync(out, d_src, m * kRec, hipMemcpyDeviceToHost, ctx->stream));
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (n > cap) return fail(ctx, CC_E_OVERFLOW, "output capacity too small");
  return CC_OK;
}

}  // namespace

extern "C" {

size_t cc_rsb_grid(const cc_rsb_cell* outer, int nla, int nr1, int nr2, cc_rsb_cell* out, size_t cap) {
  Cell7 v;
  std::memcpy(&v, outer, sizeof v);
  size_t n = 0;
  for (int a = 0; a < nla; ++a) {
    const Iv la = piece(v.la, a, nla);
    for (int b = 0; b < nr1; ++b) {
      Iv r1 = piece(v.r[0], b, nr1);
      for (int c = 0; c < nr2; ++c) {
        Iv r2r = v.r[1];
        if (r1.ub < r2r.ub) r2r.ub = r1.ub;
        const Iv r2 = piece(r2r, c, nr2);
        if (r1.lb < r2.lb) r1.lb = r2.lb;
        const Cell7 s = from_la_r1_r2(v, la, r1, r2);
        if (empty(s)) continue;
        if (n < cap) std::memcpy(out + n, &s, sizeof s);
        ++n;
      }
    }
  }
  return n;
}

int cc_rsb_enumerate(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, const int counts[4], cc_rsb_cell* out,
                     size_t cap, size_t* n_out) {
  if (!ctx || !n_out || !counts || (ncells && !cells) || (cap && !out)) return CC_E_ARG;
  for (int k = 0; k < 4; ++k)
    if (counts[k] < 1) return fail(ctx, CC_E_ARG, "piece counts must be >= 1");
  if (ncells > 0xffffffffull) return fail(ctx, CC_E_ARG, "too many cells");
  CC_HIP(ctx, hipSetDevice(ctx->device));
  Scratch& S = scratch(ctx);
  int rc = upload(ctx, S.cells, cells, ncells * kRec);
  if (rc) return rc;
  size_t n = 0;
  int trap = 0;
  if ((rc = enumerate_dev(ctx, S.cells.p, ncells, counts, &n, &trap))) return rc;
  if ((rc = copy_out(ctx, S.leaves.p, n, out, cap, n_out))) return rc;
  if (trap) return fail(ctx, CC_E_TRAP, "a leaf reached a reference trap site");
  return CC_OK;
}

int cc_rsb_find_critical_in(cc_ctx* ctx, const cc_rsb_cell* cell, int is_init, cc_rsb_cell* out, size_t cap,
                            size_t* n_out) {
  if (!cell) return CC_E_ARG;
  return cc_rsb_enumerate(ctx, cell, 1, is_init ? kInitCounts : kSubdivCounts, out, cap, n_out);
}

}  // extern "C"

namespace {

// cc_rsb_bnb (exact: cc_rsb_bnb_exact), and with log != NULL the records
// of every evaluated cell (cc_rsb_bnb_log).
int rsb_bnb_run(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, bool exact, cc_rsb_cell* out,
                size_t cap, size_t* n_out, cc_bnb_stats* stats, cc_bnb_record* log, size_t log_cap, size_t* n_log) {
  if (!ctx || !n_out || (ncells && !cells) || (cap && !out) || max_depth < 0 || max_depth >= CC_BNB_MAX_DEPTH)
    return CC_E_ARG;
  if (n_log && !log && log_cap) return CC_E_ARG;
  CC_HIP(ctx, hipSetDevice(ctx->device));
  clear_interrupt(ctx);
  const auto t0 = std::chrono::steady_clock::now();
  Scratch& S = scratch(ctx);
  hipFunction_t f;
  CC_HIP(ctx, kernel(ctx, exact ? "ccrd_rsb_bnb_level_x" : "ccrd_rsb_bnb_level", &f));
  cc_bnb_stats st;
  std::memset(&st, 0, sizeof st);
  int rc = upload(ctx, S.cells, cells, ncells * kRec);
  if (rc) return rc;
  unsigned long long* cnt = static_cast<unsigned long long*>(ctx->d_counts);
  unsigned long long* hist = cnt + 8;  // 21 bins in slots 8..28
  CC_HIP(ctx, hipMemsetAsync(cnt, 0, 32 * sizeof(unsigned long long), ctx->stream));
  void* d_log = nullptr;
  const uint64_t log_cap64 = n_log ? log_cap : 0;
  if (n_log) {
    CC_HIP(ctx, S.log.ensure(log_cap64 * sizeof(cc_bnb_record) + 8));
    d_log = S.log.p;
  }
  CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  uint64_t n = ncells, unproven = 0;
  int d = 0;
  for (; d <= max_depth && n; ++d) {
    if (interrupted(ctx)) return fail(ctx, CC_E_INTERRUPTED, "interrupted");
    st.cells_per_depth[d] = n;
    const int at_limit = d == max_depth;
    CC_HIP(ctx, S.next.ensure(at_limit ? kRec : 2 * n * kRec));
    if (at_limit) CC_HIP(ctx, S.leaves.ensure(n * kRec));  // every cell of the last level may stay unproven
    const uint64_t cap_u = S.leaves.cap / kRec;
    CC_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(unsigned long long), ctx->stream));
    CC_HIP(ctx, launch(f, dim3(blocks_for(n, 128)), dim3(128), 0, ctx->stream, S.cells.p, n, at_limit, ctx->d_konst,
                       S.next.p, S.leaves.p, cap_u, cnt, hist, d_log, log_cap64, d));
    if ((rc = read_counts(ctx, 3))) return rc;
    n = 2 * ctx->h_counts[0];
    unproven = ctx->h_counts[1];
    S.next.swap(S.cells);
  }
  CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  CC_HIP(ctx, hipMemcpyAsync(ctx->h_counts, cnt, 32 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0;
  CC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  st.trap_cells = ctx->h_counts[2];
  for (int v = 0; v < 21; ++v) {
    st.verdict_hist[v] = ctx->h_counts[8 + v];
    st.cell_evals += st.verdict_hist[v] * (uint64_t)(v == 0 ? 20 : v);
  }
  st.depth_reached = d - 1;
  st.device_ms = ms;
  if (n_log) {
    const uint64_t logged = ctx->h_counts[3];
    *n_log = logged;
    const size_t m = logged < log_cap64 ? logged : log_cap64;
    if (m) {
      CC_HIP(ctx, hipMemcpyAsync(log, d_log, m * sizeof(cc_bnb_record), hipMemcpyDeviceToHost, ctx->stream));
      CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
  }
  if ((rc = copy_out(ctx, S.leaves.p, unproven, out, cap, n_out))) return rc;
  st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = st;
  return CC_OK;
}

}  // namespace

extern "C" {

int cc_rsb_bnb(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, cc_rsb_cell* out, size_t cap,
               size_t* n_out, cc_bnb_stats* stats) {
  return rsb_bnb_run(ctx, cells, ncells, max_depth, false, out, cap, n_out, stats, nullptr, 0, nullptr);
}

int cc_rsb_bnb_exact(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, cc_rsb_cell* out,
                     size_t cap, size_t* n_out, cc_bnb_stats* stats) {
  return rsb_bnb_run(ctx, cells, ncells, max_depth, true, out, cap, n_out, stats, nullptr, 0, nullptr);
}

int cc_rsb_bnb_log(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, cc_bnb_record* log,
                   size_t log_cap, size_t* n_log, cc_rsb_cell* out, size_t cap, size_t* n_out, cc_bnb_stats* stats) {
  if (!n_log) return CC_E_ARG;
  return rsb_bnb_run(ctx, cells, ncells, max_depth, false, out, cap, n_out, stats, log, log_cap, n_log);
}

}  // extern "C"
