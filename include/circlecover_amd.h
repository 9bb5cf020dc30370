/* circlecover_amd — C ABI of the MI355X (gfx950) interval branch-and-bound prover.
 *
 * Drop-in boundary for the reference's CUDA path
 * (phillip-keldenich/circlecover, directory Prover/).  Plain C types only:
 * no exceptions, no C++ or torch types cross this boundary.  Every function
 * returning int returns CC_OK (0) or one of the CC_E_* codes; the message of
 * the last failure on a context is available through cc_last_error().
 *
 * Threading: a context is used by one host thread at a time (cc_interrupt
 * excepted, which any thread may call).  Functions
 * ending in _async enqueue work on the context's HIP stream and take DEVICE
 * pointers; all other functions are synchronous and take HOST pointers.
 */
#ifndef CIRCLECOVER_AMD_H_
#define CIRCLECOVER_AMD_H_

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CC_OK 0
#define CC_E_HIP 1         /* HIP runtime error (reference: std::system_error, ACW/error.hpp:84-98) */
#define CC_E_OVERFLOW 2    /* output capacity too small; *n_out holds the size needed
                              (reference: relaunch with exact size, TR/tight_rectangle.cu:281-294) */
#define CC_E_TRAP 3        /* a geometric invariant failed in a cell (reference: device trap, ACU/exit.cuh:28-31) */
#define CC_E_INTERRUPTED 4 /* cc_interrupt() was called (reference: SIGINT flag, P/interrupt.cpp:30-46) */
#define CC_E_ARG 5         /* invalid argument */
#define CC_E_NODEVICE 6    /* no usable gfx950 device */

/* An FP64 interval; layout of algcuda::Interval<double> (ACU/interval.hpp:84-242). */
typedef struct cc_iv {
  double lb, ub;
} cc_iv;

/* tight_rectangle::Variables, 80 bytes (Prover/tight_rectangle/strategies.cuh:33-50):
 * lambda and the four largest SQUARED radii. */
typedef struct cc_tr_cell {
  cc_iv la;
  cc_iv r[4];
} cc_tr_cell;

/* rectangle_size_bound::Variables, 112 bytes (Prover/rectangle_size_bound/variables.hpp:33-110):
 * lambda and the six largest SQUARED radii. */
typedef struct cc_rsb_cell {
  cc_iv la;
  cc_iv r[6];
} cc_rsb_cell;

typedef struct cc_ctx cc_ctx;

/* ---- context (replaces the implicit CUDA context and algcuda::device wrappers,
 *      Prover/algcuda/cuda-cpp-wrapper/src/{device,memory,error}.cpp) ---- */
int cc_init(int device, cc_ctx** out);
void cc_destroy(cc_ctx* ctx);
const char* cc_last_error(const cc_ctx* ctx);
int cc_synchronize(cc_ctx* ctx);
/* Device memory owned by the caller through the context's device. */
int cc_dev_alloc(cc_ctx* ctx, size_t bytes, void** dptr);
int cc_dev_free(cc_ctx* ctx, void* dptr);
int cc_h2d(cc_ctx* ctx, void* dst, const void* src, size_t bytes);
int cc_d2h(cc_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy with a plain streaming kernel (bytes a multiple of
 * 16); its known traffic calibrates the HBM counters. */
int cc_copy_async(cc_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
/* Stream-ordered timing: cc_timer_start records an event on the context
 * stream, cc_timer_stop records a second one, waits for it and returns the
 * elapsed milliseconds between the two. */
int cc_timer_start(cc_ctx* ctx);
int cc_timer_stop(cc_ctx* ctx, float* ms);

/* ---- batched criterion evaluation (no reference counterpart: the reference
 *      evaluates criteria only inside its enumeration kernels,
 *      TR/tight_rectangle.cu:169-262, RSB/rectangle_size_bound.cu:161-236) ---- */
#define CC_BATCH_C2 0      /* the ten-criterion 4-D box set (see cc_eval_batch) */
#define CC_PREC_FP64 0     /* outward-rounded FP64 intervals */
#define CC_PREC_FP32_PRE 1 /* FP32 pre-filter: every criterion but 4 is first tried as the FP32
                              shadow of its FP64 computation, which proves a box only when the
                              FP64 criterion does; the FP64 criterion runs where the shadow
                              leaves the box undecided.  Outputs equal CC_PREC_FP64's. */

/* Evaluates criterion set `prover` on n boxes stored as 8 SoA FP64 arrays of
 * stride ld >= n: (la.lb, la.ub, r1.lb, r1.ub, r2.lb, r2.ub, r3.lb, r3.ub).
 * verdict_mask[i] bit k = criterion k proved box i; bit 16+k = criterion k
 * hit a sqrt-domain trap.  first_ok[i] = lowest proving criterion or 255.
 * Criteria (reference files under Prover/):
 *   0 tight_rectangle::r1_strategies(la, r1, r2.ub)            r1_strategies.cu:27
 *   1 r1_r2_strategies_simple(la, r1, r2, r3.ub)                r1_r2_strategies.cu:46
 *   2 r1_r2_strategies_stacked(la, r1, r2, r3.ub)               r1_r2_strategies.cu:74
 *   3 r1_r2_r3_strategy_horizontal(la, r1, r2, r3, ub_r4)       r1_r2_r3_strategies.cu:28
 *   4 r1_r2_r3_strategy_cover_strip(la, r1, r2, r3, ub_r4)      r1_r2_r3_strategies.cu:69
 *   5 rectangle_size_bound::shortcut_even_split_recursion       even_split_recursion.cu:42
 *   6 shortcut_uneven_split_recursion(la, r1, r2)               uneven_split_recursion.cu:78
 *   7 shortcut_uneven_split_recursion(la, r1, r2, r3)           uneven_split_recursion.cu:102
 *   8 shortcut_r1_r2_opposite_corners_recursion                 r1_r2_opposite_corners_recursion.cu:28
 *   9 shortcut_r1_r2_vertical_strip                             r1_r2_vertical_strip.cu:28
 * with ub_r4 = min(r3.ub, (required_weight_for(la)-r1-r2-r3).ub), TR/tight_rectangle.cu:97-102. */
int cc_eval_batch(cc_ctx* ctx, int prover, int precision, const double* soa, size_t n, size_t ld,
                  uint32_t* verdict_mask, uint8_t* first_ok);
int cc_eval_batch_async(cc_ctx* ctx, int prover, int precision, const double* d_soa, size_t n, size_t ld,
                        uint32_t* d_verdict_mask, uint8_t* d_first_ok);
/* No counterpart in the reference: cc_eval_batch_async for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree
 * or allocating library routine: the batch's scratch (deferred-box list and
 * rows, criterion 0's staging rows, a counter) is one static device array of
 * CC_NOALLOC_SCRATCH_BYTES per GPU, set up by cc_init and held until the
 * process ends.  n <= CC_NOALLOC_MAX_BOXES (CC_E_ARG beyond).  Same outputs
 * as cc_eval_batch_async.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams or two contexts) race, and the
 * caller must order them. */
#define CC_NOALLOC_MAX_BOXES 262144u
#define CC_NOALLOC_SCRATCH_BYTES 29376768u
int cc_eval_batch_noalloc_async(cc_ctx* ctx, int prover, int precision, const double* d_soa, size_t n, size_t ld,
                                uint32_t* d_verdict_mask, uint8_t* d_first_ok);

/* rectangle_size_bound verdict of n 7-D cells (cc_rsb_cell layout), each
 * evaluated exactly as the reference kernel treats a leaf cell
 * (RSB/rectangle_size_bound.cu:162-229): shortcuts S1..S5, tighten(),
 * Intermediate_values, criteria C1..C14 in order.  verdict[i] = index of the
 * first succeeding step (S1..S5 = 0..4, empty/vacuous cell = 5, C1..C14 =
 * 6..19) or -1 if unproven; trap[i] = 1 if a reference trap site was hit. */
int cc_rsb_eval_cells(cc_ctx* ctx, const cc_rsb_cell* cells, size_t n, int32_t* verdict, uint8_t* trap);
/* Every step of cc_rsb_eval_cells evaluated independently (no
 * short-circuit): bit k of mask[i] = step k succeeds, bit 20+k = step k hit a
 * reference trap site.  Used to pin each criterion separately. */
int cc_rsb_eval_masks(cc_ctx* ctx, const cc_rsb_cell* cells, size_t n, uint64_t* mask);
int cc_rsb_eval_cells_async(cc_ctx* ctx, const cc_rsb_cell* d_cells, size_t n, int32_t* d_verdict, uint8_t* d_trap);

/* ---- tight_rectangle enumeration (replaces kernel_find_critical_intervals_in
 *      and its host wrapper, Prover/tight_rectangle/tight_rectangle.cu:169-298) ---- */

/* One launch of the reference: the unproven leaves of `cell` with the
 * reference's piece counts (is_init: 128 for r1..r4; otherwise 128,128,128,64;
 * tight_rectangle.cu:45-62).  Leaves are written in unspecified order (the
 * reference's order is that of its atomicAdd); *n_out = number of leaves.
 * Returns CC_E_OVERFLOW with *n_out set when cap is too small (the caller
 * enlarges its buffer and calls again, as tight_rectangle.cu:283-293 does),
 * CC_E_TRAP if a cell reached a reference trap site. */
int cc_tr_find_critical_in(cc_ctx* ctx, const cc_tr_cell* cell, int is_init, cc_tr_cell* out, size_t cap,
                           size_t* n_out);
/* The same enumeration over a batch of cells with counts[4] pieces for
 * r1..r4 (each >= 1); the leaves of all cells are returned together. */
int cc_tr_enumerate(cc_ctx* ctx, const cc_tr_cell* cells, size_t ncells, const int counts[4], cc_tr_cell* out,
                    size_t cap, size_t* n_out);
/* Level-synchronous branch-and-bound (no reference counterpart; the small
 * tight_rectangle workload): enumerate `cells` with counts[4], then `levels`
 * times split every unproven leaf on its widest dimension (absolute width,
 * ties to the lowest of lambda, r1..r4; halves get_subinterval(x, k, 2)) and
 * re-evaluate both halves with piece counts 1.  Returns the final unproven
 * leaves; per_level (levels+1 entries, may be NULL) receives the number of
 * unproven leaves after each level. */
int cc_tr_bnb(cc_ctx* ctx, const cc_tr_cell* cells, size_t ncells, const int counts[4], int levels,
              cc_tr_cell* out, size_t cap, size_t* n_out, uint64_t* per_level);

/* Synthetic boxes (host): splitmix64(seed); la.lb ~ U[1,2.5]; three radius
 * lower bounds ~ U[0.02,0.7] sorted descending; each width 2^U[-20,-6];
 * ub = lb + width rounded upward.  Writes 8*n doubles in SoA order (ld = n). */
void cc_gen_c2_boxes(uint64_t seed, size_t n, double* soa);

/* ---- rectangle_size_bound enumeration and branch-and-bound ---- */

/* The reference's initial cells (host loop of find_critical_intervals_in,
 * Prover/rectangle_size_bound/rectangle_size_bound.cu:276-310, with
 * Variables::from_la_r1_r2, variables.hpp:37-51): lambda, r1 and r2 of
 * `outer` split into nla, nr1 and nr2 pieces, r2 limited by r1.ub, r3..r6
 * limited by r2.ub, tightened, empty cells skipped.  Writes up to cap cells in
 * loop order and returns the number of nonempty cells. */
size_t cc_rsb_grid(const cc_rsb_cell* outer, int nla, int nr1, int nr2, cc_rsb_cell* out, size_t cap);
/* One launch of the reference's kernel (rectangle_size_bound.cu:161-236):
 * shortcuts S1/S2 on the cell, S3..S5 per r3 piece, the leaf chain on every
 * (r3, r4, r5, r6) piece; is_init selects the piece counts (64,64,64,64 or
 * 16,16,16,32, :39-57).  Unproven leaves (tightened) in unspecified order;
 * CC_E_OVERFLOW with *n_out set when cap is too small, as the reference's
 * relaunch protocol (:256-269) expects. */
int cc_rsb_find_critical_in(cc_ctx* ctx, const cc_rsb_cell* cell, int is_init, cc_rsb_cell* out, size_t cap,
                            size_t* n_out);
/* The same enumeration over a batch of cells with counts[4] pieces for r3..r6
 * (r4..r6: 1..256 each; CC_E_ARG otherwise). */
int cc_rsb_enumerate(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, const int counts[4], cc_rsb_cell* out,
                     size_t cap, size_t* n_out);

#define CC_BNB_MAX_DEPTH 64
typedef struct cc_bnb_stats {
  uint64_t cells_per_depth[CC_BNB_MAX_DEPTH]; /* cells evaluated at each depth */
  uint64_t verdict_hist[21];                  /* [v+1] = cells with verdict v (cc_rsb_eval_cells codes) */
  uint64_t trap_cells;                        /* cells that reached a reference trap site */
  uint64_t cell_evals;                        /* criterion evaluations: v+1 per cell proven by step v, 20 per unproven */
  int depth_reached;
  double device_ms;                           /* first launch to last kernel end (HIP events) */
  double wall_ms;                             /* entry to return, host clock */
} cc_bnb_stats;

/* Widest-dimension branch-and-bound over 7-D cells (no reference
 * counterpart; the reference subdivides with fixed piece counts instead).
 * Every cell gets the cc_rsb_eval_cells verdict; an unproven cell below
 * max_depth is split on its widest dimension (absolute width, ties to the
 * lowest of lambda, r1..r6; halves get_subinterval(x, k, 2)), an unproven
 * cell at max_depth is returned.  Level-synchronous: one launch per depth
 * with wave-aggregated compaction of the children. */
int cc_rsb_bnb(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, cc_rsb_cell* out, size_t cap,
               size_t* n_out, cc_bnb_stats* stats);
/* cc_rsb_bnb with the exact shortcuts of cc_rsb_bnb_batched (bisections whose
 * test fails at their smallest argument exit at once, returning what the
 * full loop returns) in the level kernel: the same per-cell arithmetic as
 * the batched form, level-synchronous.  Same outputs as cc_rsb_bnb. */
int cc_rsb_bnb_exact(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, cc_rsb_cell* out,
                     size_t cap, size_t* n_out, cc_bnb_stats* stats);

/* Launch shape of the batched branch-and-bound. */
typedef struct cc_fast_params {
  int cpw_a;         /* cells a wave takes per work-queue fetch, stage A (multiple of 64) */
  int cpw_b;         /* the same for stage B (C14) */
  int blocks_per_cu; /* persistent workgroups (256 threads) per compute unit */
  int use_graph;     /* capture the depth loop as a hipGraph */
  int waves_per_simd; /* occupancy target of the stage kernels: 1, 2 or 4 (register budget 512/W) */
} cc_fast_params;

/* The branch-and-bound of cc_rsb_bnb, batched (no reference counterpart):
 * persistent stage kernels pull cells from device work queues (stage A:
 * every step before C14; stage B: C14 on the cells stage A leaves), all
 * counts stay in device memory, the loop over depths is one hipGraph, and
 * bisections whose test fails at their smallest argument exit at once (an
 * exact shortcut: they return what the full loop returns).  Same unproven
 * cells, per-depth counts and verdict histogram as cc_rsb_bnb; if a level
 * outgrows the queues the call falls back to cc_rsb_bnb.  params == NULL
 * uses the autotuned shape (or a default). */
int cc_rsb_bnb_batched(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth,
                       const cc_fast_params* params, cc_rsb_cell* out, size_t cap, size_t* n_out,
                       cc_bnb_stats* stats);
/* Times cc_rsb_bnb_batched over a small grid of launch shapes on `cells`
 * and keeps the fastest as the context's default. */
int cc_rsb_autotune(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, cc_fast_params* best,
                    double* best_ms);


/* ---- learned criterion ranker (no reference counterpart) ----
 * An 8 -> 64 -> 64 -> 11 MLP (ReLU; BF16 weights, FP32 accumulate) that
 * predicts from a box (centre and log2 width of lambda, r1, r2, r3) which of
 * the ten cc_eval_batch criteria proves it first (classes 0..9) or that none
 * does (class 10).  It only reorders work: every verdict is still decided by
 * the interval criteria. */
#define CC_RANKER_PARAMS 5451 /* W1 8x64, b1 64, W2 64x64, b2 64, W3 64x11, b3 11 (row-major k, n) */
typedef struct cc_ranker {
  float params[CC_RANKER_PARAMS]; /* FP32 master weights */
  float norm[16];                 /* feature mean[8], std[8] */
} cc_ranker;

/* Trains on the boxes with label[i] <= 10 (the first proving criterion, 10:
 * none proves the box; other labels leave the box out): softmax cross-entropy, Adam (beta 0.9/0.999, eps 1e-8), `batch` boxes per
 * step, He-uniform initial weights from `seed`.  epoch_loss (epochs entries,
 * may be NULL) receives the mean training loss of each epoch. */
int cc_ranker_train(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const uint8_t* label, int epochs,
                    int batch, float lr, uint64_t seed, cc_ranker* model, float* epoch_loss);
/* Gradient of the mean loss over the B boxes sample[0..B) with respect to
 * every parameter (one forward and backward pass, no update). */
int cc_ranker_gradient(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const uint8_t* label,
                       const uint32_t* sample, int B, const cc_ranker* model, float* grad, float* loss);
/* Batched MFMA inference: top1[i] = predicted class of box i; logits
 * (n*11, may be NULL) receives the raw outputs. */
int cc_ranker_rank(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const cc_ranker* model, uint8_t* top1,
                   float* logits);

/* ---- learned step order in the batched branch-and-bound (no reference
 *      counterpart; the reference tries its steps in one fixed order,
 *      RSB/rectangle_size_bound.cu:162-229) ---- */

/* One cell of a branch-and-bound run as the prover logs it: the cell, the
 * first step that proved it (cc_rsb_eval_cells numbering, -1: unproven at
 * its depth) and its depth. */
typedef struct cc_bnb_record {
  cc_rsb_cell cell;
  int32_t first_step;
  int32_t depth;
} cc_bnb_record;
/* cc_rsb_bnb that also logs every evaluated cell: log[0 .. min(*n_log,
 * log_cap)) receives the records (unspecified order), *n_log the number of
 * cells evaluated. */
int cc_rsb_bnb_log(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth, cc_bnb_record* log,
                   size_t log_cap, size_t* n_log, cc_rsb_cell* out, size_t cap, size_t* n_out, cc_bnb_stats* stats);

/* A ranker of rectangle_size_bound steps: the 8 -> 64 -> 64 -> 11 network of
 * cc_ranker on a cell's lambda, r1, r2, r3 (centre and log2 width), whose
 * class c stands for step class_step[c] (0..18, the steps before C14), or
 * for the reference's order when class_step[c] < 0. */
typedef struct cc_rsb_ranker {
  cc_ranker net;
  int32_t class_step[11];
} cc_rsb_ranker;
/* Trains a ranker on n logged records: classes 0..9 are the ten steps
 * before C14 that prove most of the records first, class 10 everything else
 * (C14, unproven cells, rarer steps); then cc_ranker_train with these
 * labels (epochs, batch, lr, seed as there). */
int cc_rsb_ranker_train(cc_ctx* ctx, const cc_bnb_record* records, size_t n, int epochs, int batch, float lr,
                        uint64_t seed, cc_rsb_ranker* model, float* epoch_loss);
/* cc_rsb_bnb_batched with the learned order: at every depth the ranker
 * classifies the level's cells on the GPU (MFMA inference), the cells are
 * grouped by predicted step (counting sort), and stage A tries each cell's
 * predicted step first, then the others in the reference's order.  The
 * interval criteria stay the only judge: the same cells are proven and
 * split at every depth, so cells_per_depth and the unproven cells equal
 * cc_rsb_bnb's; verdict_hist counts the first step that succeeded in the
 * order tried, and trap_cells may count trap sites of steps tried out of
 * order. */
int cc_rsb_bnb_learned(cc_ctx* ctx, const cc_rsb_cell* cells, size_t ncells, int max_depth,
                       const cc_rsb_ranker* ranker, const cc_fast_params* params, cc_rsb_cell* out, size_t cap,
                       size_t* n_out, cc_bnb_stats* stats);

/* ---- the branch-and-bound in one call (no reference counterpart; the
 *      reference subdivides with fixed piece counts instead) ---- */
#define CC_BNB_PLAIN 0   /* level-synchronous: cc_rsb_bnb */
#define CC_BNB_BATCHED 1 /* persistent stage kernels and one hipGraph: cc_rsb_bnb_batched */
#define CC_BNB_LEARNED 2 /* batched with the learned step order: cc_rsb_bnb_learned */
#define CC_BNB_PLAIN_EXACT 3 /* level-synchronous with the exact shortcuts: cc_rsb_bnb_exact */
typedef struct cc_bnb_config {
  const cc_rsb_cell* cells; /* initial cells (host memory); NULL: the grid below */
  size_t ncells;
  double la_lb, la_ub;      /* grid: lambda in [la_lb, la_ub] split into nla pieces, squared radii in */
  int nla, nr1, nr2;        /* [0, 9/64] with r1 in nr1 and r2 (<= r1) in nr2 pieces, built as cc_rsb_grid */
  int max_depth;            /* 0 .. CC_BNB_MAX_DEPTH-1 */
  int form;                 /* CC_BNB_PLAIN, CC_BNB_BATCHED, CC_BNB_LEARNED or CC_BNB_PLAIN_EXACT */
  const cc_fast_params* params; /* batched forms' launch shape; NULL: autotuned or default */
  const cc_rsb_ranker* ranker;  /* CC_BNB_LEARNED */
} cc_bnb_config;
/* Runs the widest-dimension branch-and-bound of cc_rsb_bnb in the chosen
 * form and returns the unproven cells in lexicographic (lambda, r1..r6)
 * order, the order of the reference's Variables::operator<
 * (rectangle_size_bound/variables.hpp:82-94), in an array the library
 * allocates: *unproven (NULL when *n == 0), released with cc_free.  All
 * forms give the same unproven cells and cells per depth (the learned
 * form's verdict histogram and trap count follow its order, see
 * cc_rsb_bnb_learned).  Interruptible (cc_interrupt) in the plain forms. */
int cc_run_bnb(cc_ctx* ctx, const cc_bnb_config* cfg, cc_rsb_cell** unproven, size_t* n, cc_bnb_stats* stats);
/* Releases an array returned by cc_run_bnb. */
void cc_free(void* p);

typedef struct cc_first_stats {
  uint64_t proven;          /* boxes some criterion proves */
  uint64_t criterion_evals; /* criteria evaluated over all boxes */
  uint64_t evals_proven;    /* criteria evaluated over the proven boxes */
  double rank_ms;           /* inference and grouping (HIP events) */
  double eval_ms;           /* criterion evaluation (HIP events) */
  uint64_t second_stage;    /* boxes past criterion 0 in the cascade form (else all) */
} cc_first_stats;

/* First-success evaluation of the ten-criterion set: criteria are tried one
 * at a time until one proves the box.  model == NULL: plain order 0..9.
 * Otherwise boxes are grouped by the ranker's predicted class: a predicted
 * criterion is tried first, then the others in plain order; the boxes
 * predicted to be unprovable (class 10) form the last group, in plain order,
 * so that they do not hold up the waves of boxes that are proven early.  Whether a box is proven
 * does not depend on the order; first[i] = the criterion that proved it
 * (255: none), evals[i] = criteria evaluated, trap[i] = 1 if an evaluated
 * criterion reached a trap site. */
int cc_c2_eval_first(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const cc_ranker* model, uint8_t* first,
                     uint8_t* evals, uint8_t* trap, cc_first_stats* stats);
/* As cc_c2_eval_first; cascade != 0 first tries criterion 0 (which proves
 * most boxes) on every box, then runs the remaining criteria, in plain or
 * ranked order, only on the boxes it leaves open, ranking only those.  Plain
 * cascade gives exactly plain order's first/evals/trap; ranked cascade
 * proves the same boxes. */
int cc_c2_eval_first_ex(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const cc_ranker* model, int cascade,
                        uint8_t* first, uint8_t* evals, uint8_t* trap, cc_first_stats* stats);

/* ---- interval text (host only, no context) ----
 * The reference's interval writer and reader, so that a report can be read
 * back and compared bit for bit. */
/* Writes "[lb, ub]" as the reference's operator<< does
 * (Prover/algcuda/cuda-utils/cpp/include/algcuda/interval.hpp:294-303:
 * precision 19, each bound right-aligned in 26 columns) with a terminating
 * NUL; *len = its length.  CC_E_OVERFLOW (with *len set) if cap <= *len. */
int cc_format_interval(const cc_iv* x, char* buf, size_t cap, size_t* len);
/* Reads one interval from text[0..n) as the reference's operator>> does
 * (interval.hpp:305-351): after white space, "[a, b]" or "[a; b]", or a
 * scalar x read as [x, x]; numbers in the decimal form an iostream reads
 * (no inf, nan or hex).  *consumed (may be NULL) = characters used.
 * CC_E_ARG where the reference's stream would fail. */
int cc_parse_interval(const char* text, size_t n, cc_iv* out, size_t* consumed);

/* ---- whole proofs (replace find_critical_intervals(), see also the C++
 *      declarations at the end of this header) ---- */

/* The tight_rectangle proof of Prover/tight_rectangle/tight_rectangle.cu:
 * 458-537: la_init initial lambda launches over [1, 2.0898841580413818342]
 * (<= 0: the reference's 256), of which the pieces [la_first, la_first +
 * la_count) are run (la_count <= 0: all), then `generations` refinement generations of
 * the reference schedule {1,2,2,2,2,1,2,1,2,1,2} (< 0: all 11), each launch's
 * leaves merged by merge_adjacent_r1, and the Max_distances report.  Writes
 * the reference's stdout report to `out` (NULL: stdout), progress to
 * stderr.  SIGINT during the generations ends the run with the last
 * complete generation, as the reference does. */
int cc_tr_prove(cc_ctx* ctx, int la_init, int la_first, int la_count, int generations, FILE* out);
/* The report that cc_tr_prove writes for a final critical set `cells`
 * (tight_rectangle.cu:498-536: the case banner, the Max_distances lines of
 * kernel_compute_distance_from_theoretical_worst_cases, :343-456, computed
 * on the device, and the closing line), into buf as a NUL-terminated
 * string; *len receives its length.  CC_E_OVERFLOW if buf is NULL or cap <=
 * *len; CC_E_TRAP if a three-disk cell takes the square root of a possibly
 * negative interval (the reference's kernel traps there). */
int cc_tr_report_criticals(cc_ctx* ctx, const cc_tr_cell* cells, size_t n, char* buf, size_t cap, size_t* len);
/* The rectangle_size_bound proof of rectangle_size_bound.cu:312-382:
 * initial grid la_init x r1_init x r2_init (<= 0: 256, 128, 128) over the
 * lambda pieces [la_first, la_first + la_count) (the reference's
 * start_from_lambda_offset, :276; la_count <= 0: to the end), merge<0>,
 * `generations` refinement generations (< 0: the reference's 1) with the
 * 64x32x32 host grid per critical cell and merge<2>/<1>/<3>, then one
 * "Critical: " line per remaining cell. */
int cc_rsb_prove(cc_ctx* ctx, int la_init, int la_first, int la_count, int r1_init, int r2_init, int generations,
                 FILE* out);
/* Checkpointed forms of the two proofs (no counterpart in the reference,
 * whose only resume point is a lambda offset; for runs longer than a job's
 * time limit).  checkpoint (may be NULL) names a file: a run starts from it
 * if it exists (CC_E_ARG if it belongs to another proof or other
 * parameters), saves its progress there after its batches (every 2 s at
 * most: 16 launches of tight_rectangle, 4096 cells of rectangle_size_bound,
 * or one critical cell of a generation) and when it is interrupted, and
 * removes it when the proof completes.  max_batches > 0 stops the run after
 * that many batches with CC_E_INTERRUPTED and no report: a later call with
 * the same checkpoint continues it, and the final report is the one an
 * uninterrupted run writes. */
int cc_tr_prove_ex(cc_ctx* ctx, int la_init, int la_first, int la_count, int generations, const char* checkpoint,
                   int max_batches, FILE* out);
int cc_rsb_prove_ex(cc_ctx* ctx, int la_init, int la_first, int la_count, int r1_init, int r2_init, int generations,
                    const char* checkpoint, int max_batches, FILE* out);
/* Piece counts of the two proofs on this context, for shortened runs of the
 * same procedure (NULL: the reference's constants, which are also the
 * defaults): tight_rectangle's r1..r4 counts of the initial launches
 * (128^4) and of the generation launches (128,128,128,64;
 * TR/tight_rectangle.cu:44-62); rectangle_size_bound's r3..r6 counts of the
 * initial launches (64^4) and of the generation launches (16,16,16,32), and
 * the lambda x r1 x r2 host grid of a generation (64x32x32;
 * RSB/rectangle_size_bound.cu:36-62).  A checkpoint records the shape it was
 * made with. */
int cc_tr_prove_shape(cc_ctx* ctx, const int init_counts[4], const int gen_counts[4]);
/* The form of both proofs' per-cell enumerations (cc_tr_enumerate,
 * cc_rsb_enumerate, the find_critical_in calls and cc_tr_prove /
 * cc_rsb_prove on top of them).  No counterpart in the reference: an option
 * for measuring the batched form against the reference's own kernel shape.
 *   CC_ENUM_BATCHED      level kernels with compaction between levels and the
 *                        exact shortcuts (the default);
 *   CC_ENUM_PLAIN        the reference's kernel shape (TR/tight_rectangle.cu:
 *                        169-277, RSB/rectangle_size_bound.cu:161-251): one
 *                        block per outer piece pair, the inner pieces in a 16x16
 *                        block, no compaction, one atomic per unproven leaf,
 *                        the reference's operation sequence;
 *   CC_ENUM_PLAIN_EXACT  that shape with the batched form's exact shortcuts.
 * Every form gives the same leaves (as a set) and reports. */
#define CC_ENUM_BATCHED 0
#define CC_ENUM_PLAIN 1
#define CC_ENUM_PLAIN_EXACT 2
int cc_set_enumeration_form(cc_ctx* ctx, int form);
int cc_rsb_prove_shape(cc_ctx* ctx, const int init_counts[4], const int gen_grid[3], const int gen_counts[4]);
/* Asks the call running on ctx to stop at its next launch or level boundary
 * (reference: was_interrupted, Prover/interrupt.cpp:30-46, polled at
 * TR/tight_rectangle.cu:481-489 and RSB/rectangle_size_bound.cu:342-349).
 * The one function that may be called from another thread while a call runs
 * on ctx.  Polled by cc_tr_prove[_ex], cc_rsb_prove[_ex], cc_tr_bnb,
 * cc_rsb_bnb, cc_run_bnb (level-synchronous forms) and cc_ranker_train, each
 * of which clears the flag when it starts, as the reference's
 * start_interruptible_computation does (interrupt.cpp:38-41): an interrupt
 * never stops a later call.  Results:
 *   - a proof interrupted during its refinement generations writes the
 *     report of the last complete generation and returns CC_OK (the
 *     reference's "best result achieved so far");
 *   - a proof interrupted during its initial launches returns
 *     CC_E_INTERRUPTED without a report (the reference cannot be
 *     interrupted there; its SIGINT ends the process), after saving its
 *     checkpoint if it has one;
 *   - the other calls return CC_E_INTERRUPTED and their outputs are unset. */
int cc_interrupt(cc_ctx* ctx);

/* FP32 pre-filter in plain first-success order: each box is converted
 * outward to FP32 and criterion 0 (r1_strategies, the usual first proof) is
 * tried in outward-rounded FP32 intervals; a box it proves is decided there
 * (first = 0, f32_decided = 1), every other box is evaluated in FP64 as
 * cc_c2_eval_first does (the FP32 attempt counts as one evaluation).
 * widths (4*n floats, may be NULL): per box the widths of
 * critical_ratio(lambda) and of required_weight_for(lambda) - r1 in FP32 and
 * in FP64 (cr32, cr64, rem32, rem64). */
int cc_c2_eval_first_f32(cc_ctx* ctx, const double* soa, size_t n, size_t ld, uint8_t* first, uint8_t* evals,
                         uint8_t* trap, uint8_t* f32_decided, float* widths, cc_first_stats* stats);
/* As cc_c2_eval_first_f32; cascade != 0 runs it in the cascade form of
 * cc_c2_eval_first_ex: the FP32 shadow of criterion 0 on every box, FP64
 * criterion 0 on the boxes it leaves undecided, the other criteria on the
 * boxes still open.  Same outputs as cascade = 0. */
int cc_c2_eval_first_f32_ex(cc_ctx* ctx, const double* soa, size_t n, size_t ld, int cascade, uint8_t* first,
                            uint8_t* evals, uint8_t* trap, uint8_t* f32_decided, float* widths,
                            cc_first_stats* stats);

/* Per-criterion report of the FP32 shadows used by CC_PREC_FP32_PRE (no
 * reference counterpart: the reference computes in FP64 only).  f32_mask[i]
 * bit k = the FP32 shadow of criterion k (the cc_eval_batch numbering)
 * proves box i; a set bit implies bit k of cc_eval_batch's verdict mask.
 * Criterion 4 has no shadow (bit 4 never set).  inflation (10*n floats, may
 * be NULL): per box and criterion, the width of the criterion's weight
 * budget in FP32 over its width in FP64 (-1 where undefined); the budget is
 * the weight a criterion starts from minus the disks it places:
 *   0: W-r1  1,2: W-r1-r2  3: W-r1-r2-r3  (W = required_weight_for(la))
 *   5,6,9: R-r1-r2  7: R-r1-r3  8: R-r1-r2-r3  (R = 0.61*la). */
int cc_c2_f32_report(cc_ctx* ctx, const double* soa, size_t n, size_t ld, uint32_t* f32_mask, float* inflation);

/* get_subinterval (Prover/subintervals.hpp:32-43) on the host: the n pieces
 * of [lb, ub] as 2n doubles (lb + diff*i with one rounding, the last piece
 * ending at ub).  Used to split lambda ranges into initial cells. */
void cc_subintervals(double lb, double ub, int n, double* out);

/* Directed-rounding primitive self-test: out[8*i+k] for k = add_rd, add_ru,
 * mul_rd, mul_ru, div_rd, div_ru, sqrt_rd(a), sqrt_ru(a). */
int cc_prim_selftest(cc_ctx* ctx, const double* a, const double* b, size_t n, double* out);
/* Round-to-nearest helpers self-test: out[3*i+k] = 0.5*(a+b), fma(a,b,c), (a-b)/c. */
int cc_prim_rn_selftest(cc_ctx* ctx, const double* a, const double* b, const double* c, size_t n, double* out);

#ifdef __cplusplus
}
#endif
#ifdef __cplusplus
/* The reference's C++ entry points (Prover/prover.cpp:23-31), exported by the
 * library: each creates a context on device 0, runs cc_tr_prove /
 * cc_rsb_prove with the reference's constants and throws std::system_error
 * (HIP errors) or std::logic_error (capacity) as the reference does. */
namespace circlecover {
namespace tight_rectangle {
void find_critical_intervals();
}
namespace rectangle_size_bound {
void find_critical_intervals();
}
}  // namespace circlecover
#endif

#endif
