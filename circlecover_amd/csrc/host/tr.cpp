// tight_rectangle drivers of the C ABI: the reference's per-cell enumeration
// (find_critical_intervals_in_call_kernel, Prover/tight_rectangle/
// tight_rectangle.cu:264-298) generalised to a batch of cells, and the small
// level-synchronous branch-and-bound on top of it.
#include <algorithm>
#include <cstring>
#include <vector>

#include "runtime.hpp"

using namespace cchost;

namespace {

// Record sizes of the device level structs (csrc/tr_enum.hip).
constexpr size_t kCell = 80, kAux = 32, kS1 = 40, kS2 = 56, kS3 = 88;
// Reference piece counts (tight_rectangle.cu:45-62).
constexpr int kInitCounts[4] = {128, 128, 128, 128};
constexpr int kSubdivCounts[4] = {128, 128, 128, 64};

unsigned blocks_for(uint64_t threads, unsigned block) { return (unsigned)((threads + block - 1) / block); }

// Reads the level counters [0, k) into the pinned mirror.
int read_counts(cc_ctx* ctx, int k) {
  CC_HIP(ctx, hipMemcpyAsync(ctx->h_counts, ctx->d_counts, k * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CC_OK;
}

// The reference's kernel shape (tight_rectangle.cu:169-277): a (128, 128)
// grid of 16x16 blocks per cell, the cells of the batch along the grid's
// third dimension; same outputs and counter slots as enumerate_dev (3
// leaves, 4 trap), the same relaunch when the leaf buffer was short.
int enumerate_plain(cc_ctx* ctx, const void* d_cells, size_t ncells, const int n[4], size_t* n_leaves, int* trap) {
  Scratch& S = scratch(ctx);
  unsigned long long* cnt = static_cast<unsigned long long*>(ctx->d_counts);
  int* d_trap = reinterpret_cast<int*>(cnt + 4);
  hipFunction_t f;
  CC_HIP(ctx, kernel(ctx, ctx->enum_form == CC_ENUM_PLAIN ? "ccrd_tr_plain" : "ccrd_tr_plain_x", &f));
  if (ncells > 65535) return fail(ctx, CC_E_ARG, "plain enumeration: at most 65535 cells a call");
  if (!S.leaves.p) CC_HIP(ctx, S.leaves.ensure((size_t)1 << 20));
  uint64_t c4 = 0;
  int rc;
  for (int pass = 0; pass < 2; ++pass) {
    const uint64_t cap = S.leaves.cap / kCell;
    CC_HIP(ctx, hipMemsetAsync(cnt + 3, 0, sizeof(unsigned long long), ctx->stream));
    CC_HIP(ctx, S.verdict.ensure(cap * sizeof(uint32_t)));
    CC_HIP(ctx, S.tr_piece.ensure(cap * sizeof(uint32_t)));
    CC_HIP(ctx, launch(f, dim3(128, 128, (unsigned)ncells), dim3(16, 16), 0, ctx->stream, d_cells, n[0], n[1], n[2],
                       n[3], ctx->d_konst, S.leaves.p, cap, S.verdict.p, S.tr_piece.p, cnt + 3, d_trap));
    if ((rc = read_counts(ctx, 5))) return rc;
    c4 = ctx->h_counts[3];
    if (c4 <= cap) break;
    if (pass == 1) return fail(ctx, CC_E_OVERFLOW, "leaf count grew between identical launches");
    CC_HIP(ctx, S.leaves.ensure(c4 * kCell));
  }
  *trap = (int)(ctx->h_counts[4] & 0xffffffffu);
  *n_leaves = c4;
  return CC_OK;
}

// Enumerates the device cells d_cells[0..ncells) with piece counts n[4];
// the unproven leaves land in scratch().leaves, their number in *n_leaves.
// Counter slots: 0..3 level counts, 4 trap flag (as int).
int enumerate_dev(cc_ctx* ctx, const void* d_cells, size_t ncells, const int n[4], size_t* n_leaves, int* trap) {
  Scratch& S = scratch(ctx);
  unsigned long long* cnt = static_cast<unsigned long long*>(ctx->d_counts);
  int* d_trap = reinterpret_cast<int*>(cnt + 4);
  hipFunction_t f0, f1, f2, f3, f4;
  CC_HIP(ctx, kernel(ctx, "ccrd_tr_level0", &f0));
  CC_HIP(ctx, kernel(ctx, "ccrd_tr_level1", &f1));
  CC_HIP(ctx, kernel(ctx, "ccrd_tr_level2", &f2));
  CC_HIP(ctx, kernel(ctx, "ccrd_tr_level3", &f3));
  CC_HIP(ctx, kernel(ctx, "ccrd_tr_level4", &f4));
  *n_leaves = 0;
  *trap = 0;
  CC_HIP(ctx, hipMemsetAsync(cnt, 0, 8 * sizeof(unsigned long long), ctx->stream));
  if (ncells == 0) return CC_OK;
  if (ctx->enum_form != CC_ENUM_BATCHED) return enumerate_plain(ctx, d_cells, ncells, n, n_leaves, trap);
  const uint64_t nc = ncells;
  CC_HIP(ctx, S.aux.ensure(ncells * kAux));
  CC_HIP(ctx, launch(f0, dim3(blocks_for(nc, 256)), dim3(256), 0, ctx->stream, d_cells, nc, ctx->d_konst, S.aux.p));
  // level 1: every (cell, r1 piece)
  const uint64_t t1 = nc * (uint64_t)n[0];
  CC_HIP(ctx, S.s1.ensure(t1 * kS1));
  CC_HIP(ctx, launch(f1, dim3(blocks_for(t1, 256)), dim3(256), 0, ctx->stream, d_cells, nc, n[0], ctx->d_konst,
                     S.aux.p, S.s1.p, cnt + 0));
  int rc = read_counts(ctx, 1);
  if (rc) return rc;
  const uint64_t c1 = ctx->h_counts[0];
  // level 2
  const uint64_t t2 = c1 * (uint64_t)n[1];
  if (t2) {
    CC_HIP(ctx, S.s2.ensure(t2 * kS2));
    CC_HIP(ctx, launch(f2, dim3(blocks_for(t2, 256)), dim3(256), 0, ctx->stream, d_cells, S.s1.p, c1, n[1],
                       ctx->d_konst, S.aux.p, S.s2.p, cnt + 1, d_trap));
  }
  if ((rc = read_counts(ctx, 2))) return rc;
  const uint64_t c2 = t2 ? ctx->h_counts[1] : 0;
  // level 3
  const uint64_t t3 = c2 * (uint64_t)n[2];
  if (t3) {
    CC_HIP(ctx, S.s3.ensure(t3 * kS3));
    CC_HIP(ctx, launch(f3, dim3(blocks_for(t3, 256)), dim3(256), 0, ctx->stream, d_cells, S.s2.p, c2, n[2],
                       ctx->d_konst, S.aux.p, S.s3.p, cnt + 2, d_trap));
  }
  if ((rc = read_counts(ctx, 3))) return rc;
  const uint64_t c3 = t3 ? ctx->h_counts[2] : 0;
  // level 4: leaves; re-run once with the exact size if the buffer was short
  const uint64_t t4 = c3 * (uint64_t)n[3];
  uint64_t c4 = 0;
  if (t4) {
    if (!S.leaves.p) CC_HIP(ctx, S.leaves.ensure((size_t)1 << 20));
    for (int pass = 0; pass < 2; ++pass) {
      const uint64_t cap = S.leaves.cap / kCell;
      CC_HIP(ctx, hipMemsetAsync(cnt + 3, 0, sizeof(unsigned long long), ctx->stream));
      CC_HIP(ctx, S.verdict.ensure((S.leaves.cap / kCell) * sizeof(uint32_t)));
      CC_HIP(ctx, S.tr_piece.ensure((S.leaves.cap / kCell) * sizeof(uint32_t)));
      CC_HIP(ctx, launch(f4, dim3(blocks_for(t4, 128)), dim3(128), 0, ctx->stream, d_cells, S.s3.p, c3, n[3],
                         ctx->d_konst, S.aux.p, S.leaves.p, cap, S.verdict.p, S.tr_piece.p, cnt + 3, d_trap));
      if ((rc = read_counts(ctx, 5))) return rc;
      c4 = ctx->h_counts[3];
      if (c4 <= cap) break;
      if (pass == 1) return fail(ctx, CC_E_OVERFLOW, "leaf count grew between identical launches");
      CC_HIP(ctx, S.leaves.ensure(c4 * kCell));
    }
  } else if ((rc = read_counts(ctx, 5))) {
    return rc;
  }
  *trap = (int)(ctx->h_counts[4] & 0xffffffffu);
  *n_leaves = c4;
  return CC_OK;
}

int copy_out(cc_ctx* ctx, const void* d_src, size_t n, cc_tr_cell* out, size_t cap, size_t* n_out) {
  *n_out = n;
  const size_t m = n < cap ? n : cap;
  if (m) {
    CC_HIP(ctx, hipMemcpyAsync(out, d_src, m * kCell, hipMemcpyDeviceToHost, ctx->stream));
    CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  if (n > cap) return fail(ctx, CC_E_OVERFLOW, "output capacity too small");
  return CC_OK;
}

int upload_cells(cc_ctx* ctx, const cc_tr_cell* cells, size_t ncells) {
  Scratch& S = scratch(ctx);
  CC_HIP(ctx, S.cells.ensure(ncells * kCell));
  if (ncells) CC_HIP(ctx, hipMemcpyAsync(S.cells.p, cells, ncells * kCell, hipMemcpyHostToDevice, ctx->stream));
  return CC_OK;
}

}  // namespace

namespace cchost {

// The leaves of a batch of cells, merged per cell as merge_adjacent_r1
// (tight_rectangle.cu:129-159) merges one launch's leaves, appended in cell
// order.  The merge result is the connected components of the union of the
// leaves' r1 intervals, each with the hull of all four radii and the cell's
// lambda, in order of their least r1 lower bound: whatever the order of
// equal keys in the reference's sort, since the radii are never negative
// zero or NaN and the hull is then the same bits in any order.  Every leaf
// of a (cell, r1 piece) bucket holds the piece's r1 upper bound, so the
// device forms each bucket's hull (cc_tr_bucket_hull) and only the buckets
// come back: per cell, they are sorted by their least r1 lower bound and
// joined where that bound is at most the running r1 upper bound, as the
// reference's sweep joins leaves.
int tr_enumerate_merged(cc_ctx* ctx, const cc_tr_cell* cells, size_t ncells, const int counts[4],
                        std::vector<cc_tr_cell>& append_to, int* trap) {
  CC_HIP(ctx, hipSetDevice(ctx->device));
  int rc = upload_cells(ctx, cells, ncells);
  if (rc) return rc;
  size_t n = 0;
  if ((rc = enumerate_dev(ctx, scratch(ctx).cells.p, ncells, counts, &n, trap))) return rc;
  if (n == 0) return CC_OK;
  Scratch& S = scratch(ctx);
  const size_t n1 = (size_t)counts[0], nb = ncells * n1;
  constexpr size_t kWords = 10;
  std::vector<unsigned long long> h(nb * kWords + 1);
  for (size_t k = 0; k < nb; ++k) {
    unsigned long long* b = h.data() + kWords * k;
    for (int q = 0; q < 4; ++q) b[q] = ~0ull, b[4 + q] = 0ull;
    b[8] = ~0ull;
    b[9] = 0ull;
  }
  h[nb * kWords] = 0;  // the flag of a leaf with r1.lb > r1.ub
  CC_HIP(ctx, S.tr_buckets.ensure(h.size() * sizeof(unsigned long long)));
  auto* d_b = static_cast<unsigned long long*>(S.tr_buckets.p);
  CC_HIP(ctx, hipMemcpyAsync(d_b, h.data(), h.size() * sizeof(unsigned long long), hipMemcpyHostToDevice,
                             ctx->stream));
  hipFunction_t fh;
  CC_HIP(ctx, kernel(ctx, "cc_tr_bucket_hull", &fh));
  CC_HIP(ctx, launch(fh, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, S.leaves.p, S.verdict.p, S.tr_piece.p,
                     (uint64_t)n, (int)n1, d_b, reinterpret_cast<int*>(d_b + nb * kWords)));
  CC_HIP(ctx, hipMemcpyAsync(h.data(), d_b, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h[nb * kWords] & 0xffffffffull) return fail(ctx, CC_E_TRAP, "a leaf's r1 lower bound exceeds its upper bound");
  auto value = [](unsigned long long k) {
    const unsigned long long b = (k >> 63) ? k ^ (1ull << 63) : ~k;
    double x;
    std::memcpy(&x, &b, sizeof x);
    return x;
  };
  std::vector<cc_tr_cell> hulls;
  for (size_t c = 0; c < ncells; ++c) {
    hulls.clear();
    for (size_t p = 0; p < n1; ++p) {
      const unsigned long long* b = h.data() + kWords * (c * n1 + p);
      if (b[9] == 0) continue;
      if (b[8] != b[4]) return fail(ctx, CC_E_TRAP, "leaves of one r1 piece differ in r1's upper bound");
      cc_tr_cell v;
      v.la = cells[c].la;
      for (int q = 0; q < 4; ++q) v.r[q] = cc_iv{value(b[q]), value(b[4 + q])};
      hulls.push_back(v);
    }
    std::stable_sort(hulls.begin(), hulls.end(),
                     [](const cc_tr_cell& x, const cc_tr_cell& y) { return x.r[0].lb < y.r[0].lb; });
    for (size_t k = 0; k < hulls.size(); ++k) {
      const cc_tr_cell& v = hulls[k];
      if (k > 0 && v.r[0].lb <= append_to.back().r[0].ub) {
        cc_tr_cell& p = append_to.back();
        for (int q = 0; q < 4; ++q) {
          if (v.r[q].lb < p.r[q].lb) p.r[q].lb = v.r[q].lb;
          if (v.r[q].ub > p.r[q].ub) p.r[q].ub = v.r[q].ub;
        }
      } else {
        append_to.push_back(v);
      }
    }
  }
  return CC_OK;
}

}  // namespace cchost

extern "C" {

int cc_tr_enumerate(cc_ctx* ctx, const cc_tr_cell* cells, size_t ncells, const int counts[4], cc_tr_cell* out,
                    size_t cap, size_t* n_out) {
  if (!ctx || !n_out || !counts || (ncells && !cells) || (cap && !out)) return CC_E_ARG;
  for (int k = 0; k < 4; ++k)
    if (counts[k] < 1) return fail(ctx, CC_E_ARG, "piece counts must be >= 1");
  if (ncells > 0xffffffffull) return fail(ctx, CC_E_ARG, "too many cells");
  CC_HIP(ctx, hipSetDevice(ctx->device));
  int rc = upload_cells(ctx, cells, ncells);
  if (rc) return rc;
  size_t n = 0;
  int trap = 0;
  if ((rc = enumerate_dev(ctx, scratch(ctx).cells.p, ncells, counts, &n, &trap))) return rc;
  if ((rc = copy_out(ctx, scratch(ctx).leaves.p, n, out, cap, n_out))) return rc;
  if (trap) return fail(ctx, CC_E_TRAP, "a cell reached a sqrt of a possibly negative interval");
  return CC_OK;
}

int cc_tr_find_critical_in(cc_ctx* ctx, const cc_tr_cell* cell, int is_init, cc_tr_cell* out, size_t cap,
                           size_t* n_out) {
  if (!cell) return CC_E_ARG;
  return cc_tr_enumerate(ctx, cell, 1, is_init ? kInitCounts : kSubdivCounts, out, cap, n_out);
}

int cc_tr_bnb(cc_ctx* ctx, const cc_tr_cell* cells, size_t ncells, const int counts[4], int levels,
              cc_tr_cell* out, size_t cap, size_t* n_out, uint64_t* per_level) {
  if (!ctx || !n_out || !counts || (ncells && !cells) || (cap && !out) || levels < 0 || levels > 30)
    return CC_E_ARG;
  for (int k = 0; k < 4; ++k)
    if (counts[k] < 1) return fail(ctx, CC_E_ARG, "piece counts must be >= 1");
  CC_HIP(ctx, hipSetDevice(ctx->device));
  clear_interrupt(ctx);
  Scratch& S = scratch(ctx);
  hipFunction_t fb;
  CC_HIP(ctx, kernel(ctx, "ccrd_tr_bisect", &fb));
  int rc = upload_cells(ctx, cells, ncells);
  if (rc) return rc;
  size_t n = 0;
  int trap = 0, any_trap = 0;
  if ((rc = enumerate_dev(ctx, S.cells.p, ncells, counts, &n, &trap))) return rc;
  any_trap |= trap;
  if (per_level) per_level[0] = n;
  static const int kOne[4] = {1, 1, 1, 1};
  for (int l = 1; l <= levels && n; ++l) {
    if (interrupted(ctx)) return fail(ctx, CC_E_INTERRUPTED, "interrupted");
    const uint64_t nn = n;
    CC_HIP(ctx, S.next.ensure(2 * nn * kCell));
    CC_HIP(ctx, launch(fb, dim3(blocks_for(2 * nn, 256)), dim3(256), 0, ctx->stream, S.leaves.p, nn, S.next.p));
    S.next.swap(S.cells);  // children become the next level's input cells
    if ((rc = enumerate_dev(ctx, S.cells.p, 2 * nn, kOne, &n, &trap))) return rc;
    any_trap |= trap;
    if (per_level) per_level[l] = n;
  }
  if ((rc = copy_out(ctx, S.leaves.p, n, out, cap, n_out))) return rc;
  if (any_trap) return fail(ctx, CC_E_TRAP, "a cell reached a sqrt of a possibly negative interval");
  return CC_OK;
}

}  // extern "C"
