// Context management, device memory, timing and the batched-criterion entry
// points of the C ABI.
#include "runtime.hpp"

#include <cstring>
#include <new>

extern "C" const unsigned char cc_prover_rd_hsaco[];
extern "C" const unsigned char cc_prover_rd_hsaco_end[];

namespace cchost {

int fail(cc_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}
int hip_fail(cc_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, CC_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

hipError_t kernel(cc_ctx* ctx, const char* name, hipFunction_t* f) {
  auto it = ctx->fns.find(name);
  if (it != ctx->fns.end()) {
    *f = it->second;
    return hipSuccess;
  }
  hipError_t e = hipModuleGetFunction(f, ctx->mod, name);
  if (e == hipSuccess) ctx->fns.emplace(name, *f);
  return e;
}

Scratch& scratch(cc_ctx* ctx) {
  if (!ctx->scratch) ctx->scratch = new Scratch();
  return *ctx->scratch;
}
void release_scratch(cc_ctx* ctx) {
  delete ctx->scratch;
  ctx->scratch = nullptr;
}

int c2_batch_check(cc_ctx* ctx, int prover, int precision, const double* d_soa, size_t n, size_t ld,
                   const uint32_t* d_mask, const uint8_t* d_first) {
  if (!ctx) return CC_E_ARG;
  if (prover != CC_BATCH_C2) return fail(ctx, CC_E_ARG, "unknown criterion set");
  if (precision != CC_PREC_FP64 && precision != CC_PREC_FP32_PRE) return fail(ctx, CC_E_ARG, "unknown precision");
  if (ld < n) return fail(ctx, CC_E_ARG, "ld < n");
  if (n == 0) return CC_OK;
  if (!d_soa || !d_mask || !d_first) return fail(ctx, CC_E_ARG, "null buffer");
  if (n > 0xffffffffull) return fail(ctx, CC_E_ARG, "batch too large");
  return CC_OK;
}

int c2_batch_launch(cc_ctx* ctx, int prover, int precision, const double* d_soa, size_t n, size_t ld,
                    uint32_t* d_mask, uint8_t* d_first, const C2Scratch& s) {
  if (int rc = c2_batch_check(ctx, prover, precision, d_soa, n, ld, d_mask, d_first)) return rc;
  if (n == 0) return CC_OK;
  const unsigned block = 256;
  const uint64_t ns = c2_lanes(n), blocks = ns / block, nwaves = ns / 64;
  const uint64_t n64 = n, ld64 = ld, one = 1;
  hipFunction_t f_eval = precision == CC_PREC_FP32_PRE ? ctx->f_c2_eval_f32pre : ctx->f_c2_eval;
  CC_HIP(ctx, launch(ctx->f_zero_u64, dim3(1), dim3(64), 0, ctx->stream, s.cnt, one));
  CC_HIP(ctx, launch(f_eval, dim3((unsigned)blocks), dim3(block), 0, ctx->stream, d_soa, n64, ld64, ctx->d_konst,
                     d_mask, d_first, s.list, s.cnt, s.dstage, s.r1_stage, s.r1_id, s.r1_cnt, ns));
  // The deferred boxes are few (well under 1% of a typical batch); 1024
  // blocks of 16 groups of 16 lanes cover 16384 of them in one sweep and
  // stride over more.
  CC_HIP(ctx, launch(ctx->f_c2_deferred, dim3(1024), dim3(block), 0, ctx->stream, (const double*)s.dstage, n64,
                     ctx->d_konst, (const uint32_t*)s.list, (const unsigned long long*)s.cnt, d_mask, d_first));
  // Criterion 0's remaining strategies: the staging slots of 32 waves per block.
  const uint64_t r1_blocks = (nwaves + 31) / 32;
  CC_HIP(ctx, launch(ctx->f_c2_r1, dim3((unsigned)r1_blocks), dim3(block), 0, ctx->stream,
                     (const double*)s.r1_stage, (const uint32_t*)s.r1_id, (const uint32_t*)s.r1_cnt, nwaves, ns,
                     ctx->d_konst, d_mask, d_first));
  return CC_OK;
}

}  // namespace cchost

using namespace cchost;

extern "C" {

int cc_init(int device, cc_ctx** out) {
  if (!out) return CC_E_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return CC_E_NODEVICE;
  cc_ctx* ctx = new (std::nothrow) cc_ctx();
  if (!ctx) return CC_E_ARG;
  ctx->device = device;
  auto bail = [&](hipError_t e, const char* what) {
    hip_fail(ctx, e, what);
    cc_destroy(ctx);
    return CC_E_HIP;
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return bail(e, "hipSetDevice");
  hipDeviceProp_t prop;
  if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) return bail(e, "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    cc_destroy(ctx);
    return CC_E_NODEVICE;
  }
  if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
    return bail(e, "hipStreamCreate");
  if ((e = hipModuleLoadData(&ctx->mod, cc_prover_rd_hsaco)) != hipSuccess) return bail(e, "hipModuleLoadData");
  struct {
    hipFunction_t* f;
    const char* name;
  } fns[] = {{&ctx->f_init_konst, "ccrd_init_konst"}, {&ctx->f_c2_eval, "ccrd_c2_eval"}, {&ctx->f_prim, "ccrd_prim"},
                {&ctx->f_prim_rn, "ccrd_prim_rn"},
                {&ctx->f_rsb_verdict, "ccrd_rsb_verdict"},
                {&ctx->f_rsb_masks, "ccrd_rsb_masks"},
                {&ctx->f_c2_eval_f32pre, "ccrd_c2_eval_f32pre"},
                {&ctx->f_zero_u64, "cc_zero_u64"},
                {&ctx->f_c2_deferred, "ccrd_c2_strip_deferred"},
                {&ctx->f_c2_r1, "ccrd_c2_r1_pass"}};
  for (auto& fn : fns)
    if ((e = hipModuleGetFunction(fn.f, ctx->mod, fn.name)) != hipSuccess) return bail(e, fn.name);
  if ((e = hipEventCreate(&ctx->ev0)) != hipSuccess) return bail(e, "hipEventCreate");
  if ((e = hipEventCreate(&ctx->ev1)) != hipSuccess) return bail(e, "hipEventCreate");
  if ((e = hipMalloc(&ctx->d_konst, kKonstBytes)) != hipSuccess) return bail(e, "hipMalloc(konst)");
  if ((e = hipMalloc(&ctx->d_counts, 64 * sizeof(unsigned long long))) != hipSuccess) return bail(e, "hipMalloc(counts)");
  if ((e = hipHostMalloc((void**)&ctx->h_counts, 64 * sizeof(unsigned long long), hipHostMallocDefault)) !=
      hipSuccess)
    return bail(e, "hipHostMalloc(counts)");
  (void)noalloc_setup(ctx);  // a failure disables cc_eval_batch_noalloc_async only
  if ((e = launch(ctx->f_init_konst, dim3(1), dim3(64), 0, ctx->stream, ctx->d_konst)) != hipSuccess)
    return bail(e, "ccrd_init_konst");
  if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
  *out = ctx;
  return CC_OK;
}

void cc_destroy(cc_ctx* ctx) {
  if (!ctx) return;
  if (ctx->device >= 0) (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_konst) (void)hipFree(ctx->d_konst);
  if (ctx->d_counts) (void)hipFree(ctx->d_counts);
  if (ctx->h_counts) (void)hipHostFree(ctx->h_counts);
  release_scratch(ctx);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->mod) (void)hipModuleUnload(ctx->mod);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char* cc_last_error(const cc_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int cc_synchronize(cc_ctx* ctx) {
  if (!ctx) return CC_E_ARG;
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CC_OK;
}

int cc_dev_alloc(cc_ctx* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return CC_E_ARG;
  CC_HIP(ctx, hipSetDevice(ctx->device));
  CC_HIP(ctx, hipMalloc(dptr, bytes ? bytes : 1));
  return CC_OK;
}
int cc_dev_free(cc_ctx* ctx, void* dptr) {
  if (!ctx) return CC_E_ARG;
  if (dptr) CC_HIP(ctx, hipFree(dptr));
  return CC_OK;
}
int cc_h2d(cc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return CC_E_ARG;
  CC_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CC_OK;
}
int cc_d2h(cc_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return CC_E_ARG;
  CC_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CC_OK;
}
int cc_copy_async(cc_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
  if (!ctx || (bytes && (!d_dst || !d_src)) || bytes % 16) return CC_E_ARG;
  if (!bytes) return CC_OK;
  hipFunction_t f;
  CC_HIP(ctx, kernel(ctx, "cc_copy_u4", &f));
  const uint64_t n = bytes / 16;
  CC_HIP(ctx, launch(f, dim3(2048), dim3(256), 0, ctx->stream, d_src, d_dst, n));
  return CC_OK;
}

int cc_timer_start(cc_ctx* ctx) {
  if (!ctx) return CC_E_ARG;
  CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  return CC_OK;
}
int cc_timer_stop(cc_ctx* ctx, float* ms) {
  if (!ctx || !ms) return CC_E_ARG;
  CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  CC_HIP(ctx, hipEventSynchronize(ctx->ev1));
  CC_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  return CC_OK;
}

int cc_eval_batch_async(cc_ctx* ctx, int prover, int precision, const double* d_soa, size_t n, size_t ld,
                        uint32_t* d_mask, uint8_t* d_first) {
  if (int rc = c2_batch_check(ctx, prover, precision, d_soa, n, ld, d_mask, d_first)) return rc;
  if (n == 0) return CC_OK;
  DevBuf& db = scratch(ctx).c2_defer;
  CC_HIP(ctx, db.ensure(c2_scratch_bytes(n)));
  return c2_batch_launch(ctx, prover, precision, d_soa, n, ld, d_mask, d_first, c2_scratch_at(db.p, n));
}

int cc_eval_batch(cc_ctx* ctx, int prover, int precision, const double* soa, size_t n, size_t ld,
                  uint32_t* mask, uint8_t* first) {
  if (int rc = c2_batch_check(ctx, prover, precision, soa, n, ld, mask, first)) return rc;
  if (n == 0) return CC_OK;
  CC_HIP(ctx, hipSetDevice(ctx->device));
  double* d_soa = nullptr;
  uint32_t* d_mask = nullptr;
  uint8_t* d_first = nullptr;
  int rc = CC_OK;
  hipError_t e = hipMalloc(&d_soa, 8 * n * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&d_mask, n * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&d_first, n);
  for (int k = 0; e == hipSuccess && k < 8; ++k)
    e = hipMemcpyAsync(d_soa + k * n, soa + k * ld, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) rc = hip_fail(ctx, e, "eval_batch staging");
  if (rc == CC_OK) rc = cc_eval_batch_async(ctx, prover, precision, d_soa, n, n, d_mask, d_first);
  if (rc == CC_OK) {
    e = hipMemcpyAsync(mask, d_mask, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(first, d_first, n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "eval_batch readback");
  }
  (void)hipStreamSynchronize(ctx->stream);
  if (d_soa) (void)hipFree(d_soa);
  if (d_mask) (void)hipFree(d_mask);
  if (d_first) (void)hipFree(d_first);
  return rc;
}

int cc_rsb_eval_cells_async(cc_ctx* ctx, const cc_rsb_cell* d_cells, size_t n, int32_t* d_verdict,
                            uint8_t* d_trap) {
  if (!ctx) return CC_E_ARG;
  if (n == 0) return CC_OK;
  if (!d_cells || !d_verdict || !d_trap) return fail(ctx, CC_E_ARG, "null buffer");
  const unsigned block = 128;
  const uint64_t blocks = (n + block - 1) / block;
  if (blocks > 0x7fffffffull) return fail(ctx, CC_E_ARG, "batch too large");
  uint64_t n64 = n;
  CC_HIP(ctx, launch(ctx->f_rsb_verdict, dim3((unsigned)blocks), dim3(block), 0, ctx->stream, (const double*)d_cells,
                     n64, ctx->d_konst, d_verdict, d_trap));
  return CC_OK;
}

int cc_rsb_eval_cells(cc_ctx* ctx, const cc_rsb_cell* cells, size_t n, int32_t* verdict, uint8_t* trap) {
  if (!ctx) return CC_E_ARG;
  if (n == 0) return CC_OK;
  if (!cells || !verdict || !trap) return fail(ctx, CC_E_ARG, "null buffer");
  CC_HIP(ctx, hipSetDevice(ctx->device));
  cc_rsb_cell* dc = nullptr;
  int32_t* dv = nullptr;
  uint8_t* dt = nullptr;
  int rc = CC_OK;
  hipError_t e = hipMalloc(&dc, n * sizeof(cc_rsb_cell));
  if (e == hipSuccess) e = hipMalloc(&dv, n * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&dt, n);
  if (e == hipSuccess) e = hipMemcpyAsync(dc, cells, n * sizeof(cc_rsb_cell), hipMemcpyHostToDevice, ctx->stream);
  if (e != hipSuccess) rc = hip_fail(ctx, e, "rsb_eval staging");
  if (rc == CC_OK) rc = cc_rsb_eval_cells_async(ctx, dc, n, dv, dt);
  if (rc == CC_OK) {
    e = hipMemcpyAsync(verdict, dv, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(trap, dt, n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "rsb_eval readback");
  }
  (void)hipStreamSynchronize(ctx->stream);
  if (dc) (void)hipFree(dc);
  if (dv) (void)hipFree(dv);
  if (dt) (void)hipFree(dt);
  return rc;
}

int cc_rsb_eval_masks(cc_ctx* ctx, const cc_rsb_cell* cells, size_t n, uint64_t* mask) {
  if (!ctx) return CC_E_ARG;
  if (n == 0) return CC_OK;
  if (!cells || !mask) return fail(ctx, CC_E_ARG, "null buffer");
  CC_HIP(ctx, hipSetDevice(ctx->device));
  cc_rsb_cell* dc = nullptr;
  uint64_t* dm = nullptr;
  int rc = CC_OK;
  hipError_t e = hipMalloc(&dc, n * sizeof(cc_rsb_cell));
  if (e == hipSuccess) e = hipMalloc(&dm, n * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMemcpyAsync(dc, cells, n * sizeof(cc_rsb_cell), hipMemcpyHostToDevice, ctx->stream);
  uint64_t n64 = n;
  if (e == hipSuccess)
    e = launch(ctx->f_rsb_masks, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, ctx->stream, (const double*)dc, n64,
               ctx->d_konst, dm);
  if (e == hipSuccess) e = hipMemcpyAsync(mask, dm, n * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) rc = hip_fail(ctx, e, "rsb_eval_masks");
  (void)hipStreamSynchronize(ctx->stream);
  if (dc) (void)hipFree(dc);
  if (dm) (void)hipFree(dm);
  return rc;
}

int cc_prim_selftest(cc_ctx* ctx, const double* a, const double* b, size_t n, double* out) {
  if (!ctx || !a || !b || !out) return CC_E_ARG;
  if (n == 0) return CC_OK;
  CC_HIP(ctx, hipSetDevice(ctx->device));
  double *da = nullptr, *db = nullptr, *dout = nullptr;
  int rc = CC_OK;
  hipError_t e = hipMalloc(&da, n * 8);
  if (e == hipSuccess) e = hipMalloc(&db, n * 8);
  if (e == hipSuccess) e = hipMalloc(&dout, n * 64);
  if (e == hipSuccess) e = hipMemcpyAsync(da, a, n * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(db, b, n * 8, hipMemcpyHostToDevice, ctx->stream);
  uint64_t n64 = n;
  if (e == hipSuccess)
    e = launch(ctx->f_prim, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (const double*)da,
               (const double*)db, n64, dout);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, n * 64, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) rc = hip_fail(ctx, e, "prim_selftest");
  (void)hipStreamSynchronize(ctx->stream);
  if (da) (void)hipFree(da);
  if (db) (void)hipFree(db);
  if (dout) (void)hipFree(dout);
  return rc;
}

int cc_prim_rn_selftest(cc_ctx* ctx, const double* a, const double* b, const double* c, size_t n, double* out) {
  if (!ctx || !a || !b || !c || !out) return CC_E_ARG;
  if (n == 0) return CC_OK;
  CC_HIP(ctx, hipSetDevice(ctx->device));
  double *da = nullptr, *db = nullptr, *dc = nullptr, *dout = nullptr;
  int rc = CC_OK;
  hipError_t e = hipMalloc(&da, n * 8);
  if (e == hipSuccess) e = hipMalloc(&db, n * 8);
  if (e == hipSuccess) e = hipMalloc(&dc, n * 8);
  if (e == hipSuccess) e = hipMalloc(&dout, n * 24);
  if (e == hipSuccess) e = hipMemcpyAsync(da, a, n * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(db, b, n * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dc, c, n * 8, hipMemcpyHostToDevice, ctx->stream);
  uint64_t n64 = n;
  if (e == hipSuccess)
    e = launch(ctx->f_prim_rn, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, (const double*)da,
               (const double*)db, (const double*)dc, n64, dout);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, n * 24, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) rc = hip_fail(ctx, e, "prim_rn_selftest");
  (void)hipStreamSynchronize(ctx->stream);
  for (double* p : {da, db, dc, dout})
    if (p) (void)hipFree(p);
  return rc;
}

}  // extern "C"
