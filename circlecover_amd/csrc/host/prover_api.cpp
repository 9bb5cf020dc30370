// The reference's two prover entry points on top of the C ABI:
//   circlecover::tight_rectangle::find_critical_intervals()
//     (Prover/tight_rectangle/tight_rectangle.cu:458-537)
//   circlecover::rectangle_size_bound::find_critical_intervals()
//     (Prover/rectangle_size_bound/rectangle_size_bound.cu:312-382)
// with the same generation schedules, per-launch merges, report format
// (Prover/output_case.hpp:35-54, the Interval and Variables operator<<)
// and SIGINT behaviour (Prover/interrupt.cpp:30-46).  Launches of one
// generation are batched into few device calls; merges are applied per
// launch exactly as the reference does, so the records are the same.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <new>
#include <sstream>
#include <stdexcept>
#include <functional>
#include <string>
#include <system_error>
#include <vector>

#include "runtime.hpp"

using namespace cchost;

namespace cchost {
int tr_enumerate_merged(cc_ctx* ctx, const cc_tr_cell* cells, size_t ncells, const int counts[4],
                        std::vector<cc_tr_cell>& append_to, int* trap);
}

namespace {

// ------------------------------------------------------------------ output
constexpr size_t kOutputWidth = 120;

void begin_case(std::ostream& o, const char* message) {  // output_case.hpp:35-49
  const size_t mlen = std::strlen(message);
  if (mlen >= kOutputWidth - 2) {
    o << message << std::endl;
    return;
  }
  const size_t pre = (kOutputWidth - mlen - 2) / 2, post = kOutputWidth - mlen - 2 - pre;
  o << std::string(pre, '-') << ' ' << message << ' ' << std::string(post, '-') << std::endl;
}
void end_case(std::ostream& o) { o << std::string(kOutputWidth, '-') << std::endl; }  // :51-54

// Interval operator<< (algcuda/interval.hpp:294-303): right-aligned, width
// 26, precision 19.
void put_iv(std::ostream& o, const cc_iv& x) {
  std::ios_base::fmtflags f = o.flags(std::ios::right);
  std::streamsize p = o.precision(19);
  o << '[' << std::setw(26) << x.lb << ", " << std::setw(26) << x.ub << ']';
  o.flags(f);
  o.precision(p);
}
void put_cell(std::ostream& o, const cc_tr_cell& v) {  // tight_rectangle.cu:539-545
  o << "λ = ";
  put_iv(o, v.la);
  for (int i = 0; i < 4; ++i) {
    o << ", r²_" << (i + 1) << " = ";
    put_iv(o, v.r[i]);
  }
}
void put_cell(std::ostream& o, const cc_rsb_cell& v) {  // variables.hpp:112-118
  o << "λ = ";
  put_iv(o, v.la);
  for (int i = 0; i < 6; ++i) {
    o << ", r²_" << (i + 1) << " = ";
    put_iv(o, v.r[i]);
  }
}

// get_subinterval (subintervals.hpp:32-43), lb + diff*i fused.
cc_iv piece(cc_iv x, int i, int n) {
  const double diff = (x.ub - x.lb) / (double)n;
  return cc_iv{std::fma(diff, (double)i, x.lb), i + 1 == n ? x.ub : std::fma(diff, (double)(i + 1), x.lb)};
}
void do_join(cc_iv& a, const cc_iv& b) {
  if (b.lb < a.lb) a.lb = b.lb;
  if (b.ub > a.ub) a.ub = b.ub;
}
bool iv_less(const cc_iv& a, const cc_iv& b) { return a.lb < b.lb || (a.lb == b.lb && a.ub < b.ub); }
bool iv_eq(const cc_iv& a, const cc_iv& b) { return a.lb == b.lb && a.ub == b.ub; }

// merge_critical_intervals<N> (rectangle_size_bound.cu:103-130): sort by
// (lambda, radii) lexicographically, then join radii[N..6) of consecutive
// records whose lambda and radii[0..N) are equal.
std::vector<cc_rsb_cell> merge_critical(std::vector<cc_rsb_cell> v, int N) {
  std::vector<cc_rsb_cell> res;
  if (v.empty()) return res;
  std::sort(v.begin(), v.end(), [](const cc_rsb_cell& a, const cc_rsb_cell& b) {
    if (iv_less(a.la, b.la)) return true;
    if (iv_less(b.la, a.la)) return false;
    for (int i = 0; i < 6; ++i) {
      if (iv_less(a.r[i], b.r[i])) return true;
      if (iv_less(b.r[i], a.r[i])) return false;
    }
    return false;
  });
  res.push_back(v.front());
  for (const cc_rsb_cell& c : v) {
    cc_rsb_cell& p = res.back();
    bool same = iv_eq(c.la, p.la);
    for (int i = 0; same && i < N; ++i) same = iv_eq(p.r[i], c.r[i]);
    if (same) {
      for (int i = N; i < 6; ++i) do_join(p.r[i], c.r[i]);
    } else {
      res.push_back(c);
    }
  }
  return res;
}

// -------------------------------------------------------------- interrupt
volatile std::sig_atomic_t g_interrupted = 0;
void (*g_prev_handler)(int) = nullptr;
void on_sigint(int) {
  if (g_interrupted) std::_Exit(SIGINT);
  g_interrupted = 1;
}
struct Interruptible {  // start/stop_interruptible_computation (interrupt.cpp:38-46)
  Interruptible() {
    g_interrupted = 0;
    g_prev_handler = std::signal(SIGINT, &on_sigint);
  }
  ~Interruptible() {
    std::signal(SIGINT, g_prev_handler);
    g_interrupted = 0;
  }
};

// ------------------------------------------------------------- checkpoint
// A proof in progress (no counterpart in the reference, whose only resume
// point is a lambda offset): the proof's parameters, the phase (0 the
// initial launches, g >= 1 generation g), how many launches or cells of the
// phase are done, the records the phase started from and those it has found
// so far.  Written to <path>.tmp, then renamed over <path>.
template <class Rec>
struct Checkpoint {
  std::string path;
  char kind[8] = {0};
  int32_t params[24] = {0};
  int32_t phase = 0;
  uint64_t pos = 0;
  std::vector<Rec> current, next;

  // CC_OK and loaded = false if there is no file; CC_E_ARG if the file is
  // not a checkpoint of this proof with these parameters.
  int load(cc_ctx* ctx, bool& loaded) {
    loaded = false;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return CC_OK;
    char magic[8], k[8];
    int32_t pr[24], ph = 0, pad = 0;
    uint64_t ps = 0, nc = 0, nn = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && !std::memcmp(magic, "CCCKPT02", 8) && std::fread(k, 1, 8, f) == 8 &&
              !std::memcmp(k, kind, 8) && std::fread(pr, sizeof pr, 1, f) == 1 && !std::memcmp(pr, params, sizeof pr) &&
              std::fread(&ph, 4, 1, f) == 1 && std::fread(&pad, 4, 1, f) == 1 && std::fread(&ps, 8, 1, f) == 1 &&
              std::fread(&nc, 8, 1, f) == 1 && std::fread(&nn, 8, 1, f) == 1;
    // The record counts must account for exactly the rest of the file, so
    // that a truncated or corrupt file is refused before anything is sized
    // from it.
    if (ok) {
      const long here = std::ftell(f);
      ok = here >= 0 && std::fseek(f, 0, SEEK_END) == 0;
      const long end = ok ? std::ftell(f) : -1;
      ok = ok && end >= here && std::fseek(f, here, SEEK_SET) == 0;
      const uint64_t rest = ok ? (uint64_t)(end - here) : 0, per = sizeof(Rec);
      ok = ok && nc <= rest / per && nn <= rest / per && (nc + nn) * per == rest;
    }
    if (ok) {
      try {
        current.resize(nc);
        next.resize(nn);
      } catch (const std::bad_alloc&) {
        ok = false;
      }
      ok = ok && (nc == 0 || std::fread(current.data(), sizeof(Rec), nc, f) == nc) &&
           (nn == 0 || std::fread(next.data(), sizeof(Rec), nn, f) == nn);
    }
    std::fclose(f);
    if (!ok) {
      current.clear();
      next.clear();
    }
    if (!ok) return fail(ctx, CC_E_ARG, "checkpoint file does not match this proof");
    phase = ph;
    pos = ps;
    loaded = true;
    return CC_OK;
  }
  int save(cc_ctx* ctx) const {
    const std::string tmp = path + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return fail(ctx, CC_E_ARG, "cannot write the checkpoint file");
    const int32_t pad = 0;
    const uint64_t nc = current.size(), nn = next.size();
    bool ok = std::fwrite("CCCKPT02", 1, 8, f) == 8 && std::fwrite(kind, 1, 8, f) == 8 &&
              std::fwrite(params, sizeof params, 1, f) == 1 && std::fwrite(&phase, 4, 1, f) == 1 &&
              std::fwrite(&pad, 4, 1, f) == 1 && std::fwrite(&pos, 8, 1, f) == 1 && std::fwrite(&nc, 8, 1, f) == 1 &&
              std::fwrite(&nn, 8, 1, f) == 1 &&
              (nc == 0 || std::fwrite(current.data(), sizeof(Rec), nc, f) == nc) &&
              (nn == 0 || std::fwrite(next.data(), sizeof(Rec), nn, f) == nn);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) return fail(ctx, CC_E_ARG, "cannot write the checkpoint file");
    return CC_OK;
  }
};

// Saves a proof's progress after its batches (every 2 s at most, and
// always before it stops) and stops it after max_batches batches of this
// call (0: no limit).
template <class Rec>
struct Progress {
  Checkpoint<Rec>* ck = nullptr;
  int max_batches = 0, batches = 0;
  bool stopped = false;  // max_batches reached
  std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();

  // The phase's state after `done` of its launches or cells: CC_OK to go
  // on, CC_E_INTERRUPTED to stop (the checkpoint then holds this state).
  int after_batch(cc_ctx* ctx, int phase, uint64_t done, const std::vector<Rec>& current,
                  const std::vector<Rec>& next) {
    ++batches;
    const bool stop = max_batches > 0 && batches >= max_batches;
    stopped = stop;
    const auto now = std::chrono::steady_clock::now();
    if (ck && (stop || now - last >= std::chrono::seconds(2))) {
      if (int rc = save(ctx, phase, done, current, next)) return rc;
      last = now;
    }
    return stop ? CC_E_INTERRUPTED : CC_OK;
  }
  int save(cc_ctx* ctx, int phase, uint64_t done, const std::vector<Rec>& current, const std::vector<Rec>& next) {
    if (!ck) return CC_OK;
    ck->phase = phase;
    ck->pos = done;
    ck->current = current;
    ck->next = next;
    return ck->save(ctx);
  }
};

// Host mirror of the device tr_maxdist block (tr_enum.hip).
struct MaxDist {
  int have_2d, have_3d;
  cc_iv la_2d, la_3d;
  cc_iv d_r1_2d;
  cc_tr_cell d_r1_2d_crit[2];
  cc_iv d_r2_2d;
  cc_tr_cell d_r2_2d_crit[2];
  double lb_h1_2d, ub_la_by_S1, ub_R3_2d;
  cc_iv d_r1_3d;
  cc_tr_cell d_r1_3d_crit[2];
  cc_iv d_r2_3d;
  cc_tr_cell d_r2_3d_crit[2];
  cc_iv d_r3_3d;
  cc_tr_cell d_r3_3d_crit[2];
  double ub_R4_3d;
  cc_iv S1, h2, S3;
  double pocket_3d;
  int trap;
};

int max_distances(cc_ctx* ctx, const std::vector<cc_tr_cell>& crit, MaxDist* out) {
  hipFunction_t f;
  CC_HIP(ctx, kernel(ctx, "ccrd_tr_max_distances", &f));
  DevBuf d_cells, d_out;
  CC_HIP(ctx, d_cells.ensure(crit.size() * sizeof(cc_tr_cell) + 8));
  CC_HIP(ctx, d_out.ensure(sizeof(MaxDist)));
  if (!crit.empty())
    CC_HIP(ctx, hipMemcpyAsync(d_cells.p, crit.data(), crit.size() * sizeof(cc_tr_cell), hipMemcpyHostToDevice,
                               ctx->stream));
  const uint64_t n = crit.size();
  // one workgroup of 256 lanes: the kernel's reduction assumes exactly that
  CC_HIP(ctx, launch(f, dim3(1), dim3(256), 0, ctx->stream, d_cells.p, n, ctx->d_konst, d_out.p));
  CC_HIP(ctx, hipMemcpyAsync(out, d_out.p, sizeof(MaxDist), hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // the reference's interval sqrt traps on a possibly negative argument
  // (ACU/interval.cuh:339-350)
  if (out->trap) return fail(ctx, CC_E_TRAP, "a critical cell reached a sqrt of a possibly negative interval");
  return CC_OK;
}

// The Max_distances lines of the final report (tight_rectangle.cu:498-536)
// for the critical cells `current`.
int write_tr_criticals(cc_ctx* ctx, const std::vector<cc_tr_cell>& current, std::ostream& o) {
  if (current.empty()) return CC_OK;
  MaxDist r;
  if (const int rc = max_distances(ctx, current, &r)) return rc;
  auto crit_line = [&](const cc_tr_cell& c) {
    o << "\t\t";
    put_cell(o, c);
    o << std::endl;
  };
  if (r.have_2d) {  // tight_rectangle.cu:508-519
    o << "Two disks: Range for lambda that contains 2-disk critical intervals: ";
    put_iv(o, r.la_2d);
    o << "; maximal distances from theoretical values: " << std::endl;
    o << "\tr_1^2: ";
    put_iv(o, r.d_r1_2d);
    o << ", r_2^2: ";
    put_iv(o, r.d_r2_2d);
    o << ", R_3 <= " << r.ub_R3_2d << ", h_1 >= " << r.lb_h1_2d << ", lambda / S_1 <= " << r.ub_la_by_S1
      << std::endl;
    o << "\tDominating criticals for r_1: " << std::endl;
    crit_line(r.d_r1_2d_crit[0]);
    crit_line(r.d_r1_2d_crit[1]);
    o << "\tDominating criticals for r_2: " << std::endl;
    crit_line(r.d_r2_2d_crit[0]);
    crit_line(r.d_r2_2d_crit[1]);
  }
  if (r.have_3d) {  // :521-533
    o << "Three disks: Range for lambda that contains 3-disk critical intervals: ";
    put_iv(o, r.la_3d);
    o << "; maximal distances from theoretical values: " << std::endl;
    o << "\tr_1^2: ";
    put_iv(o, r.d_r1_3d);
    o << ", r_2^2: ";
    put_iv(o, r.d_r2_3d);
    o << ", r_3^2: ";
    put_iv(o, r.d_r3_3d);
    o << ", R_4 <= " << r.ub_R4_3d << std::endl;
    o << "\tS_1: ";
    put_iv(o, r.S1);
    o << ", h_2: ";
    put_iv(o, r.h2);
    o << ", S_3: ";
    put_iv(o, r.S3);
    o << ", pocket width <= " << r.pocket_3d << std::endl;
    o << "\tDominating criticals for r_1: " << std::endl;
    crit_line(r.d_r1_3d_crit[0]);
    crit_line(r.d_r1_3d_crit[1]);
    o << "\tDominating criticals for r_2: " << std::endl;
    crit_line(r.d_r2_3d_crit[0]);
    crit_line(r.d_r2_3d_crit[1]);
    o << "\tDominating criticals for r_3: " << std::endl;
    crit_line(r.d_r3_3d_crit[0]);
    crit_line(r.d_r3_3d_crit[1]);
  }
  return CC_OK;
}

// Batches of at most this many launches share one enumeration call.
constexpr size_t kTrBatch = 16;

// find_critical_intervals_in (tight_rectangle.cu:300-320) for the cells
// `launches` (one reference launch each), appending each launch's merged
// leaves in launch order.
// Launches [start, ...) of `launches`; after each batch, after(done) may
// stop the run (CC_E_INTERRUPTED) once the batch's records are appended.
int tr_launches(cc_ctx* ctx, const std::vector<cc_tr_cell>& launches, bool init, std::vector<cc_tr_cell>& append_to,
                std::ostream& err, size_t start, const std::function<int(size_t)>& after) {
  for (size_t b = start; b < launches.size(); b += kTrBatch) {
    if (g_interrupted || interrupted(ctx)) return CC_E_INTERRUPTED;
    const size_t m = std::min(kTrBatch, launches.size() - b);
    int trap = 0;
    const size_t before = append_to.size();
    int rc = tr_enumerate_merged(ctx, launches.data() + b, m, init ? ctx->tr_init_counts : ctx->tr_gen_counts,
                                 append_to, &trap);
    if (rc || trap) append_to.resize(before);
    if (rc) return rc;
    if (trap) return fail(ctx, CC_E_TRAP, "a cell reached a sqrt of a possibly negative interval");
    err << "launch " << (b + m) << " of " << launches.size() << " done (" << append_to.size()
        << " critical intervals found)                \r" << std::flush;
    if (int rc = after(b + m)) return rc;
  }
  return CC_OK;
}

}  // namespace

extern "C" {

int cc_set_enumeration_form(cc_ctx* ctx, int form) {
  if (!ctx) return CC_E_ARG;
  if (form != CC_ENUM_BATCHED && form != CC_ENUM_PLAIN && form != CC_ENUM_PLAIN_EXACT)
    return fail(ctx, CC_E_ARG, "unknown enumeration form");
  ctx->enum_form = form;
  return CC_OK;
}

int cc_tr_prove_shape(cc_ctx* ctx, const int init_counts[4], const int gen_counts[4]) {
  if (!ctx) return CC_E_ARG;
  static const int kInit[4] = {128, 128, 128, 128}, kGen[4] = {128, 128, 128, 64};
  const int* a = init_counts ? init_counts : kInit;
  const int* b = gen_counts ? gen_counts : kGen;
  for (int k = 0; k < 4; ++k)
    if (a[k] < 1 || b[k] < 1 || a[k] > 4096 || b[k] > 4096) return fail(ctx, CC_E_ARG, "piece counts must be 1..4096");
  std::memcpy(ctx->tr_init_counts, a, sizeof ctx->tr_init_counts);
  std::memcpy(ctx->tr_gen_counts, b, sizeof ctx->tr_gen_counts);
  return CC_OK;
}

int cc_rsb_prove_shape(cc_ctx* ctx, const int init_counts[4], const int gen_grid[3], const int gen_counts[4]) {
  if (!ctx) return CC_E_ARG;
  static const int kInit[4] = {64, 64, 64, 64}, kGen[4] = {16, 16, 16, 32}, kGrid[3] = {64, 32, 32};
  const int* a = init_counts ? init_counts : kInit;
  const int* b = gen_counts ? gen_counts : kGen;
  const int* g = gen_grid ? gen_grid : kGrid;
  for (int k = 0; k < 4; ++k)  // r4..r6 pieces of the enumeration: 1..256
    if (a[k] < 1 || b[k] < 1 || (k > 0 && (a[k] > 256 || b[k] > 256)) || a[k] > 4096 || b[k] > 4096)
      return fail(ctx, CC_E_ARG, "piece counts must be 1..4096 for r3 and 1..256 for r4..r6");
  for (int k = 0; k < 3; ++k)
    if (g[k] < 1 || g[k] > 4096) return fail(ctx, CC_E_ARG, "generation grid sizes must be 1..4096");
  std::memcpy(ctx->rsb_init_counts, a, sizeof ctx->rsb_init_counts);
  std::memcpy(ctx->rsb_gen_counts, b, sizeof ctx->rsb_gen_counts);
  std::memcpy(ctx->rsb_gen_grid, g, sizeof ctx->rsb_gen_grid);
  return CC_OK;
}

int cc_interrupt(cc_ctx* ctx) {
  if (!ctx) return CC_E_ARG;
  ctx->interrupt.store(1, std::memory_order_relaxed);
  return CC_OK;
}

int cc_tr_prove(cc_ctx* ctx, int la_init, int la_first, int la_count, int generations, FILE* out) {
  return cc_tr_prove_ex(ctx, la_init, la_first, la_count, generations, nullptr, 0, out);
}

int cc_tr_prove_ex(cc_ctx* ctx, int la_init, int la_first, int la_count, int generations, const char* checkpoint,
                   int max_batches, FILE* out) {
  if (!ctx || max_batches < 0) return CC_E_ARG;
  clear_interrupt(ctx);
  static const int kSubdivSchedule[11] = {1, 2, 2, 2, 2, 1, 2, 1, 2, 1, 2};  // tight_rectangle.cu:50-53
  if (la_init <= 0) la_init = 256;
  if (la_first < 0 || la_first >= la_init) la_first = 0;
  if (la_count <= 0 || la_first + la_count > la_init) la_count = la_init - la_first;
  if (generations < 0 || generations > 11) generations = 11;
  std::ostringstream o;
  std::ostream& err = std::cerr;
  begin_case(o, "Critical intervals for rectangles (tight)");
  const cc_tr_cell initial{{1.0, 2.0898841580413818342}, {{0.0, DBL_MAX}, {0.0, DBL_MAX}, {0.0, DBL_MAX}, {0.0, DBL_MAX}}};
  std::vector<cc_tr_cell> current, next, launches;
  Checkpoint<cc_tr_cell> ck;
  Progress<cc_tr_cell> pr;
  bool resumed = false;
  int rc = CC_OK;
  if (checkpoint) {
    ck.path = checkpoint;
    std::memcpy(ck.kind, "TR", 2);
    const int32_t p[4] = {la_init, la_first, la_count, generations};
    std::memcpy(ck.params, p, sizeof p);
    std::memcpy(ck.params + 4, ctx->tr_init_counts, sizeof ctx->tr_init_counts);
    std::memcpy(ck.params + 8, ctx->tr_gen_counts, sizeof ctx->tr_gen_counts);
    if ((rc = ck.load(ctx, resumed))) return rc;
    pr.ck = &ck;
  }
  pr.max_batches = max_batches;
  const int phase0 = resumed ? ck.phase : 0;
  size_t pos0 = resumed ? (size_t)ck.pos : 0, last_done = 0;
  if (resumed) {
    current = ck.current;
    next = ck.next;
    err << "Resuming from phase " << phase0 << " after " << pos0 << " launches" << std::endl;
  }
  if (phase0 == 0) {
    err << "Initializing..." << std::endl;
    for (int a = la_first; a < la_first + la_count; ++a) {
      cc_tr_cell s = initial;
      s.la = piece(initial.la, a, la_init);
      launches.push_back(s);
    }
    rc = tr_launches(ctx, launches, true, next, err, pos0, [&](size_t done) {
      last_done = done;
      return pr.after_batch(ctx, 0, done, current, next);
    });
    if (rc == CC_E_INTERRUPTED && !pr.stopped) {
      const int r2 = pr.save(ctx, 0, last_done > pos0 ? last_done : pos0, current, next);
      if (r2) return r2;
    }
    if (rc) return rc;
    current.swap(next);
    next.clear();
  }
  err << "Starting refinement - press Ctrl-C to interrupt and use best result achieved so far..." << std::endl;
  bool completed = true;
  {
    Interruptible guard;
    for (int g = phase0 > 1 ? phase0 : 1; g <= generations; ++g) {
      err << "Running generation " << g << " on " << current.size() << " intervals..." << std::endl;
      launches.clear();
      const int nla = kSubdivSchedule[g - 1];
      for (const cc_tr_cell& s : current)
        for (int a = 0; a < nla; ++a) {
          cc_tr_cell c = s;
          c.la = piece(s.la, a, nla);
          launches.push_back(c);
        }
      const size_t start = g == phase0 ? pos0 : 0;
      last_done = start;
      rc = tr_launches(ctx, launches, false, next, err, start, [&](size_t done) {
        last_done = done;
        return pr.after_batch(ctx, g, done, current, next);
      });
      if (rc == CC_E_INTERRUPTED && pr.stopped) return rc;  // to be resumed: no report yet
      if (rc == CC_E_INTERRUPTED) {                        // keep the best complete generation
        err << "Interrupted during generation " << g << ": reporting generation " << (g - 1) << std::endl;
        if ((rc = pr.save(ctx, g, last_done, current, next))) return rc;
        completed = false;
        break;
      }
      if (rc) return rc;
      current.swap(next);
      next.clear();
      if (current.empty()) break;
    }
  }
  if ((rc = write_tr_criticals(ctx, current, o))) return rc;
  end_case(o);
  const std::string s = o.str();
  std::fwrite(s.data(), 1, s.size(), out ? out : stdout);
  std::fflush(out ? out : stdout);
  if (checkpoint && completed) std::remove(checkpoint);  // the proof is done
  return CC_OK;
}

int cc_tr_report_criticals(cc_ctx* ctx, const cc_tr_cell* cells, size_t n, char* buf, size_t cap, size_t* len) {
  if (!ctx || !len || (n && !cells)) return CC_E_ARG;
  std::ostringstream o;
  begin_case(o, "Critical intervals for rectangles (tight)");
  const std::vector<cc_tr_cell> crit(cells, cells + n);
  if (const int rc = write_tr_criticals(ctx, crit, o)) return rc;
  end_case(o);
  const std::string t = o.str();
  *len = t.size();
  if (!buf || cap <= t.size()) return CC_E_OVERFLOW;
  std::memcpy(buf, t.c_str(), t.size() + 1);
  return CC_OK;
}

int cc_rsb_prove(cc_ctx* ctx, int la_init, int la_first, int la_count, int r1_init, int r2_init, int generations,
                 FILE* out) {
  return cc_rsb_prove_ex(ctx, la_init, la_first, la_count, r1_init, r2_init, generations, nullptr, 0, out);
}

int cc_rsb_prove_ex(cc_ctx* ctx, int la_init, int la_first, int la_count, int r1_init, int r2_init, int generations,
                    const char* checkpoint, int max_batches, FILE* out) {
  if (!ctx || max_batches < 0) return CC_E_ARG;
  clear_interrupt(ctx);
  // rectangle_size_bound.cu:36-62 and values.hpp:28-35
  if (la_init <= 0) la_init = 256;
  if (la_first < 0 || la_first >= la_init) la_first = 0;
  if (la_count <= 0 || la_first + la_count > la_init) la_count = la_init - la_first;
  if (r1_init <= 0) r1_init = 128;
  if (r2_init <= 0) r2_init = 128;
  if (generations < 0) generations = 1;
  const int* kInit = ctx->rsb_init_counts;
  const int* kSub = ctx->rsb_gen_counts;
  const int* kGrid = ctx->rsb_gen_grid;
  const double w = 9.0 / 64.0;
  std::ostringstream o;
  std::ostream& err = std::cerr;
  begin_case(o, "Critical intervals for rectangles with small disks");
  err << "Size bound: r <= " << 3.0 / 8.0 << ", critical ratio: " << 0.61 << std::endl;
  const cc_rsb_cell initial{{1.0, 2.5}, {{0.0, w}, {0.0, w}, {0.0, w}, {0.0, w}, {0.0, w}, {0.0, w}}};
  // find_critical_intervals_in (rectangle_size_bound.cu:276-310) over the
  // lambda pieces [first, first + count) of `outer`, from cell `start` of
  // that grid on; after(done) may stop the run after a batch of cells
  auto run = [&](const cc_rsb_cell& outer, int nla, int first, int count, int nr1, int nr2, const int* counts,
                 std::vector<cc_rsb_cell>& append_to, size_t start,
                 const std::function<int(size_t)>& after) -> int {
    std::vector<cc_rsb_cell> cells;
    for (int a = first; a < first + count; ++a) {
      cc_rsb_cell sub = outer;
      sub.la = piece(outer.la, a, nla);
      std::vector<cc_rsb_cell> part((size_t)nr1 * nr2);
      // one lambda piece: the grid's lambda loop with a single piece equal
      // to `sub.la` (from_la_r1_r2 intersects it with itself)
      const size_t m = cc_rsb_grid(&sub, 1, nr1, nr2, part.data(), part.size());
      cells.insert(cells.end(), part.begin(), part.begin() + m);
    }
    const size_t m = cells.size();
    const size_t chunk = 4096;
    for (size_t b = start; b < m; b += chunk) {
      if (g_interrupted || interrupted(ctx)) return CC_E_INTERRUPTED;
      const size_t k = std::min(chunk, m - b);
      std::vector<cc_rsb_cell> leaves(1 << 16);
      size_t n = 0;
      int rc = cc_rsb_enumerate(ctx, cells.data() + b, k, counts, leaves.data(), leaves.size(), &n);
      if (rc == CC_E_OVERFLOW) {
        leaves.resize(n);
        rc = cc_rsb_enumerate(ctx, cells.data() + b, k, counts, leaves.data(), leaves.size(), &n);
      }
      if (rc) return rc;
      append_to.insert(append_to.end(), leaves.begin(), leaves.begin() + n);
      err << "cells " << (b + k) << " of " << m << " done (" << append_to.size()
          << " critical intervals found)                \r" << std::flush;
      if ((rc = after(b + k))) return rc;
    }
    return CC_OK;
  };
  std::vector<cc_rsb_cell> current, next;
  Checkpoint<cc_rsb_cell> ck;
  Progress<cc_rsb_cell> pr;
  bool resumed = false;
  int rc = CC_OK;
  if (checkpoint) {
    ck.path = checkpoint;
    std::memcpy(ck.kind, "RSB", 3);
    const int32_t p[6] = {la_init, la_first, la_count, r1_init, r2_init, generations};
    std::memcpy(ck.params, p, sizeof p);
    std::memcpy(ck.params + 6, ctx->rsb_init_counts, sizeof ctx->rsb_init_counts);
    std::memcpy(ck.params + 10, ctx->rsb_gen_counts, sizeof ctx->rsb_gen_counts);
    std::memcpy(ck.params + 14, ctx->rsb_gen_grid, sizeof ctx->rsb_gen_grid);
    if ((rc = ck.load(ctx, resumed))) return rc;
    pr.ck = &ck;
  }
  pr.max_batches = max_batches;
  const int phase0 = resumed ? ck.phase : 0;
  const size_t pos0 = resumed ? (size_t)ck.pos : 0;
  size_t last_done = pos0;
  if (resumed) {
    current = ck.current;
    next = ck.next;
    err << "Resuming from phase " << phase0 << " at position " << pos0 << std::endl;
  }
  if (phase0 == 0) {
    err << "Initializing..." << std::endl;
    rc = run(initial, la_init, la_first, la_count, r1_init, r2_init, kInit, next, pos0, [&](size_t done) {
      last_done = done;
      return pr.after_batch(ctx, 0, done, current, next);
    });
    if (rc == CC_E_INTERRUPTED && !pr.stopped) {
      const int r2 = pr.save(ctx, 0, last_done, current, next);
      if (r2) return r2;
    }
    if (rc) return rc;
    current = merge_critical(next, 0);
    next.clear();
  }
  if (current.empty()) {
    err << "No critical regions after initialization - done!" << std::endl;
    end_case(o);
    const std::string s = o.str();
    std::fwrite(s.data(), 1, s.size(), out ? out : stdout);
    if (checkpoint) std::remove(checkpoint);  // the proof is done
    return CC_OK;
  }
  bool completed = true;
  {
    Interruptible guard;
    const auto none = [](size_t) { return CC_OK; };
    for (int g = phase0 > 1 ? phase0 : 1; g <= generations; ++g) {
      err << "Running generation " << g << " on " << current.size() << " intervals..." << std::endl;
      bool stop = false;
      // position in a generation: the cells of `current` done
      for (size_t j = g == phase0 ? pos0 : 0; j < current.size(); ++j) {
        const size_t mark = next.size();  // the leaves of cells 0..j-1
        rc = run(current[j], kGrid[0], 0, kGrid[0], kGrid[1], kGrid[2], kSub, next, 0,
                 none);  // rectangle_size_bound.cu:50-52
        if (rc == CC_E_INTERRUPTED) {
          // cell j is redone from its start on resume: drop its partial leaves
          next.resize(mark);
          err << "Interrupted during generation " << g << ": reporting generation " << (g - 1) << std::endl;
          if ((rc = pr.save(ctx, g, j, current, next))) return rc;
          stop = true;
          break;
        }
        if (rc) return rc;
        if ((rc = pr.after_batch(ctx, g, j + 1, current, next)) == CC_E_INTERRUPTED) return rc;  // to be resumed
        if (rc) return rc;
      }
      if (stop) {
        completed = false;
        break;
      }
      current = merge_critical(next, g % 3 == 0 ? 3 : g % 3 == 1 ? 2 : 1);
      next.clear();
      if (current.empty()) break;
    }
  }
  for (const cc_rsb_cell& c : current) {  // rectangle_size_bound.cu:374-379
    o << "Critical: ";
    put_cell(o, c);
    o << std::endl;
  }
  end_case(o);
  const std::string s = o.str();
  std::fwrite(s.data(), 1, s.size(), out ? out : stdout);
  std::fflush(out ? out : stdout);
  if (checkpoint && completed) std::remove(checkpoint);  // the proof is done
  return CC_OK;
}

}  // extern "C"

// ------------------------------------------------------ C++ entry points
namespace {
struct Ctx {
  cc_ctx* p = nullptr;
  Ctx() {
    const int rc = cc_init(0, &p);
    if (rc != CC_OK) throw std::system_error(rc, std::generic_category(), "no usable gfx950 device");
  }
  ~Ctx() { cc_destroy(p); }
};
void check(cc_ctx* ctx, int rc) {
  if (rc == CC_OK || rc == CC_E_INTERRUPTED) return;
  if (rc == CC_E_OVERFLOW) throw std::logic_error(cc_last_error(ctx));
  throw std::system_error(rc, std::generic_category(), cc_last_error(ctx));
}
}  // namespace

namespace circlecover {
namespace tight_rectangle {
__attribute__((visibility("default"))) void find_critical_intervals() {
  Ctx c;
  check(c.p, cc_tr_prove(c.p, 0, 0, 0, -1, stdout));
}
}  // namespace tight_rectangle
namespace rectangle_size_bound {
__attribute__((visibility("default"))) void find_critical_intervals() {
  Ctx c;
  check(c.p, cc_rsb_prove(c.p, 0, 0, 0, 0, 0, -1, stdout));
}
}  // namespace rectangle_size_bound
}  // namespace circlecover
