// Learned criterion ranker (ml.hip) and first-success evaluation
// (c2_first.hip): training loop, gradient access, batched ranking and the
// ranked-versus-plain evaluation of the ten-criterion box set.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "runtime.hpp"

using namespace cchost;

namespace {

// classes 0..9: the criterion that proves a box first; class 10: none does
constexpr int kIn = 8, kH = 64, kOut = 11, kCrit = 10, kRec = kIn + 4 * kH + kOut;
constexpr int kW1 = 0, kB1 = kIn * kH, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH, kB3 = kW3 + kH * kOut;
static_assert(kB3 + kOut == CC_RANKER_PARAMS, "parameter count");

struct MlState {
  DevBuf soa, label, perm, rec, loss, params, m, v, grad, norm, wb, bias, top1, idx, counts, first, evals, trap, bstart,
      partial, pend, list, stage, stage_id, packed, dec, widths;
};
constexpr int kSlice = 64;  // samples per partial gradient sum (ml.hip)

// Gradient of the batch in `rec` (B samples), then Adam if update != 0.
int grad_step(cc_ctx* ctx, MlState& M, int B, int t, float lr, int update, void* grad_out) {
  hipFunction_t fp, fa;
  CC_HIP(ctx, kernel(ctx, "ccml_grad_partial", &fp));
  CC_HIP(ctx, kernel(ctx, "ccml_grad_adam", &fa));
  const int slices = (B + kSlice - 1) / kSlice;
  CC_HIP(ctx, M.partial.ensure((size_t)slices * CC_RANKER_PARAMS * sizeof(float)));
  CC_HIP(ctx, launch(fp, dim3((CC_RANKER_PARAMS + 255) / 256, (unsigned)slices), dim3(256), 0, ctx->stream, M.rec.p,
                     B, M.partial.p));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  CC_HIP(ctx, launch(fa, dim3((CC_RANKER_PARAMS + 255) / 256), dim3(256), 0, ctx->stream, M.partial.p, slices,
                     M.params.p, M.m.p, M.v.p, t, lr, b1, b2, eps, update, grad_out));
  return CC_OK;
}

MlState& ml(cc_ctx* ctx) {
  Scratch& S = scratch(ctx);
  if (!S.ml) {
    S.ml = new MlState();
    S.ml_free = [](void* p) { delete static_cast<MlState*>(p); };
  }
  return *static_cast<MlState*>(S.ml);
}

struct Rng {  // splitmix64
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next() >> 11) * 0x1.0p-53; }
};

float feature_host(const double* soa, size_t ld, size_t i, int k) {
  const double lo = soa[(2 * (k >> 1)) * ld + i], hi = soa[(2 * (k >> 1) + 1) * ld + i];
  if (k & 1) {
    const double w = hi - lo;
    return (float)(w > 0x1p-60 ? std::log2(w) : -60.0);
  }
  return (float)(0.5 * (lo + hi));
}

uint16_t bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

int upload(cc_ctx* ctx, DevBuf& b, const void* src, size_t bytes) {
  CC_HIP(ctx, b.ensure(bytes));
  if (bytes) CC_HIP(ctx, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return CC_OK;
}

// Uploads the SoA boxes as a dense (8, n) block.
int upload_soa(cc_ctx* ctx, DevBuf& b, const double* soa, size_t n, size_t ld) {
  CC_HIP(ctx, b.ensure(8 * n * sizeof(double) + 8));
  for (int r = 0; r < 8; ++r)
    if (n)
      CC_HIP(ctx, hipMemcpyAsync(static_cast<double*>(b.p) + r * n, soa + r * ld, n * sizeof(double),
                                 hipMemcpyHostToDevice, ctx->stream));
  return CC_OK;
}

// BF16 weight block [W1 | W2 | W3] and FP32 biases [b1 | b2 | b3] of a model.
int upload_inference(cc_ctx* ctx, MlState& M, const cc_ranker* model) {
  std::vector<uint16_t> wb(kIn * kH + kH * kH + kH * kOut);
  std::vector<float> bias(2 * kH + kOut);
  size_t o = 0;
  for (int p = kW1; p < kB1; ++p) wb[o++] = bf16_bits(model->params[p]);
  for (int p = kW2; p < kB2; ++p) wb[o++] = bf16_bits(model->params[p]);
  for (int p = kW3; p < kB3; ++p) wb[o++] = bf16_bits(model->params[p]);
  for (int j = 0; j < kH; ++j) bias[j] = model->params[kB1 + j];
  for (int j = 0; j < kH; ++j) bias[kH + j] = model->params[kB2 + j];
  for (int c = 0; c < kOut; ++c) bias[2 * kH + c] = model->params[kB3 + c];
  int rc = upload(ctx, M.wb, wb.data(), wb.size() * sizeof(uint16_t));
  if (!rc) rc = upload(ctx, M.bias, bias.data(), bias.size() * sizeof(float));
  if (!rc) rc = upload(ctx, M.norm, model->norm, sizeof model->norm);
  return rc;
}

// Ranks the positions [0, n) (n_dev not null: [0, *n_dev), at most n) of
// the device SoA (stride ld): box i, or box list[i] when list is not null;
// top1 (and logits) by position.  The kernel is persistent: at most 512
// workgroups of four waves, the two waves per SIMD its registers allow, so
// that each wave loads the weights once for several 32-box tiles.
int rank_dev(cc_ctx* ctx, MlState& M, const void* d_soa, size_t ld, const uint32_t* list, size_t n,
             const uint32_t* n_dev, float* d_logits) {
  hipFunction_t f;
  CC_HIP(ctx, kernel(ctx, "ccml_rank", &f));
  CC_HIP(ctx, M.top1.ensure(n + 1));
  const uint64_t n64 = n, ld64 = ld;
  const uint64_t tiles = (n + 31) / 32;
  unsigned blocks = (unsigned)((tiles + 3) / 4);
  if (blocks > 512) blocks = 512;
  if (blocks == 0) blocks = 1;
  const uint64_t bs = 1;
  CC_HIP(ctx, launch(f, dim3(blocks), dim3(256), 0, ctx->stream, d_soa, n64, ld64, bs, list, n_dev, M.norm.p, M.wb.p,
                     M.bias.p, M.top1.p, d_logits));
  return CC_OK;
}

}  // namespace

namespace cchost {

int ml_upload_inference(cc_ctx* ctx, const cc_ranker* model, const void** norm, const void** wb, const void** bias) {
  MlState& M = ml(ctx);
  if (int rc = upload_inference(ctx, M, model)) return rc;
  *norm = M.norm.p;
  *wb = M.wb.p;
  *bias = M.bias.p;
  return CC_OK;
}

}  // namespace cchost

extern "C" {

int cc_ranker_train(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const uint8_t* label, int epochs,
                    int batch, float lr, uint64_t seed, cc_ranker* model, float* epoch_loss) {
  if (!ctx || !soa || !label || !model || ld < n || epochs < 0 || batch < 1 || batch > (1 << 20)) return CC_E_ARG;
  CC_HIP(ctx, hipSetDevice(ctx->device));
  clear_interrupt(ctx);
  MlState& M = ml(ctx);
  std::vector<uint32_t> lab;  // boxes some criterion proves
  for (size_t i = 0; i < n; ++i)
    if (label[i] < kOut) lab.push_back((uint32_t)i);
  if (lab.empty()) return fail(ctx, CC_E_ARG, "no proven boxes to train on");
  // feature normalisation over the training boxes
  for (int k = 0; k < kIn; ++k) {
    double s = 0, s2 = 0;
    for (uint32_t i : lab) {
      const double f = feature_host(soa, ld, i, k);
      s += f;
      s2 += f * f;
    }
    const double mean = s / lab.size(), var = s2 / lab.size() - mean * mean;
    model->norm[k] = (float)mean;
    model->norm[kIn + k] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
  }
  // He-uniform initial weights, zero biases
  Rng rng{seed};
  std::memset(model->params, 0, sizeof model->params);
  auto init = [&](int off, int fan_in, int count) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int p = 0; p < count; ++p) model->params[off + p] = (float)((2.0 * rng.unit() - 1.0) * a);
  };
  init(kW1, kIn, kIn * kH);
  init(kW2, kH, kH * kH);
  init(kW3, kH, kH * kOut);
  int rc = upload_soa(ctx, M.soa, soa, n, ld);
  if (!rc) rc = upload(ctx, M.label, label, n);
  if (!rc) rc = upload(ctx, M.params, model->params, sizeof model->params);
  if (!rc) rc = upload(ctx, M.norm, model->norm, sizeof model->norm);
  if (rc) return rc;
  CC_HIP(ctx, M.m.ensure(sizeof model->params));
  CC_HIP(ctx, M.v.ensure(sizeof model->params));
  CC_HIP(ctx, hipMemsetAsync(M.m.p, 0, sizeof model->params, ctx->stream));
  CC_HIP(ctx, hipMemsetAsync(M.v.p, 0, sizeof model->params, ctx->stream));
  CC_HIP(ctx, M.rec.ensure((size_t)batch * kRec * sizeof(float)));
  CC_HIP(ctx, M.loss.ensure(lab.size() * sizeof(float)));
  CC_HIP(ctx, M.perm.ensure(lab.size() * sizeof(uint32_t)));
  hipFunction_t ffb;
  CC_HIP(ctx, kernel(ctx, "ccml_fwd_bwd", &ffb));
  const uint64_t ld_dev = n;
  int t = 0;
  std::vector<uint32_t> perm = lab;
  std::vector<float> losses(lab.size());
  for (int e = 0; e < epochs; ++e) {
    if (interrupted(ctx)) return fail(ctx, CC_E_INTERRUPTED, "interrupted");
    for (size_t i = perm.size() - 1; i > 0; --i) {  // Fisher-Yates
      const size_t j = (size_t)(rng.next() % (i + 1));
      const uint32_t tmp = perm[i];
      perm[i] = perm[j];
      perm[j] = tmp;
    }
    CC_HIP(ctx, hipMemcpyAsync(M.perm.p, perm.data(), perm.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                               ctx->stream));
    for (size_t off = 0; off < perm.size(); off += (size_t)batch) {
      const int B = (int)(perm.size() - off < (size_t)batch ? perm.size() - off : (size_t)batch);
      ++t;
      const uint32_t* sample = static_cast<const uint32_t*>(M.perm.p) + off;
      float* loss = static_cast<float*>(M.loss.p) + off;
      CC_HIP(ctx, launch(ffb, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, ctx->stream, M.soa.p, ld_dev, sample,
                         M.label.p, B, M.norm.p, M.params.p, M.rec.p, loss));
      if ((rc = grad_step(ctx, M, B, t, lr, 1, nullptr))) return rc;
    }
    if (epoch_loss) {
      CC_HIP(ctx, hipMemcpyAsync(losses.data(), M.loss.p, losses.size() * sizeof(float), hipMemcpyDeviceToHost,
                                 ctx->stream));
      CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      double s = 0;
      for (float x : losses) s += x;
      epoch_loss[e] = (float)(s / losses.size());
    }
  }
  CC_HIP(ctx, hipMemcpyAsync(model->params, M.params.p, sizeof model->params, hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CC_OK;
}

int cc_ranker_gradient(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const uint8_t* label,
                       const uint32_t* sample, int B, const cc_ranker* model, float* grad, float* loss) {
  if (!ctx || !soa || !label || !sample || !model || !grad || B < 1 || ld < n) return CC_E_ARG;
  for (int s = 0; s < B; ++s)
    if (sample[s] >= n || label[sample[s]] >= kOut) return fail(ctx, CC_E_ARG, "sample without a label");
  CC_HIP(ctx, hipSetDevice(ctx->device));
  MlState& M = ml(ctx);
  int rc = upload_soa(ctx, M.soa, soa, n, ld);
  if (!rc) rc = upload(ctx, M.label, label, n);
  if (!rc) rc = upload(ctx, M.params, model->params, sizeof model->params);
  if (!rc) rc = upload(ctx, M.norm, model->norm, sizeof model->norm);
  if (!rc) rc = upload(ctx, M.perm, sample, (size_t)B * sizeof(uint32_t));
  if (rc) return rc;
  CC_HIP(ctx, M.rec.ensure((size_t)B * kRec * sizeof(float)));
  CC_HIP(ctx, M.loss.ensure((size_t)B * sizeof(float)));
  CC_HIP(ctx, M.grad.ensure(sizeof model->params));
  hipFunction_t ffb;
  CC_HIP(ctx, kernel(ctx, "ccml_fwd_bwd", &ffb));
  const uint64_t ld_dev = n;
  CC_HIP(ctx, launch(ffb, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, ctx->stream, M.soa.p, ld_dev, M.perm.p,
                     M.label.p, B, M.norm.p, M.params.p, M.rec.p, M.loss.p));
  if ((rc = grad_step(ctx, M, B, 1, 0.f, 0, M.grad.p))) return rc;
  CC_HIP(ctx, hipMemcpyAsync(grad, M.grad.p, sizeof model->params, hipMemcpyDeviceToHost, ctx->stream));
  if (loss) CC_HIP(ctx, hipMemcpyAsync(loss, M.loss.p, (size_t)B * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CC_OK;
}

int cc_ranker_rank(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const cc_ranker* model, uint8_t* top1,
                   float* logits) {
  if (!ctx || !soa || !model || !top1 || ld < n) return CC_E_ARG;
  if (n == 0) return CC_OK;
  CC_HIP(ctx, hipSetDevice(ctx->device));
  MlState& M = ml(ctx);
  int rc = upload_soa(ctx, M.soa, soa, n, ld);
  if (!rc) rc = upload_inference(ctx, M, model);
  if (rc) return rc;
  float* d_logits = nullptr;
  if (logits) {
    CC_HIP(ctx, M.grad.ensure(n * kOut * sizeof(float)));
    d_logits = static_cast<float*>(M.grad.p);
  }
  if ((rc = rank_dev(ctx, M, M.soa.p, n, nullptr, n, nullptr, d_logits))) return rc;
  CC_HIP(ctx, hipMemcpyAsync(top1, M.top1.p, n, hipMemcpyDeviceToHost, ctx->stream));
  if (logits)
    CC_HIP(ctx, hipMemcpyAsync(logits, d_logits, n * kOut * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CC_OK;
}

int cc_c2_eval_first(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const cc_ranker* model, uint8_t* first,
                     uint8_t* evals, uint8_t* trap, cc_first_stats* stats) {
  return cc_c2_eval_first_ex(ctx, soa, n, ld, model, 0, first, evals, trap, stats);
}

// First-success evaluation.  stage1 = 0: none (every box to the second
// stage); 1: FP64 criterion 0 on every box (cascade); 2: FP32 shadow of
// criterion 0, then FP64 criterion 0 on the boxes it leaves undecided
// (cascade; f32_decided and widths as cc_c2_eval_first_f32 writes them).
static int first_success(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const cc_ranker* model, int stage1,
                         uint8_t* first, uint8_t* evals, uint8_t* trap, uint8_t* f32_decided, float* widths,
                         cc_first_stats* stats) {
  if (!ctx || !soa || !first || !evals || !trap || ld < n) return CC_E_ARG;
  if (stage1 == 2 && (!f32_decided || model)) return CC_E_ARG;
  const bool cascade = stage1 != 0;
  if (n > 0xffffffffull) return fail(ctx, CC_E_ARG, "too many boxes");
  CC_HIP(ctx, hipSetDevice(ctx->device));
  MlState& M = ml(ctx);
  int rc = upload_soa(ctx, M.soa, soa, n, ld);
  if (!rc && model) rc = upload_inference(ctx, M, model);
  if (rc) return rc;
  CC_HIP(ctx, M.idx.ensure(n * sizeof(uint32_t) + 4));
  CC_HIP(ctx, M.first.ensure(n + 1));
  CC_HIP(ctx, M.evals.ensure(n + 1));
  CC_HIP(ctx, M.trap.ensure(n + 1));
  CC_HIP(ctx, M.bstart.ensure(16 * sizeof(uint64_t)));
  // every buffer is in place before the timed region starts
  constexpr uint32_t kChunk = 2048;  // boxes per ccrd_c2_first0 block
  const uint64_t nb = (n + kChunk - 1) / kChunk;
  if (cascade && n) {
    CC_HIP(ctx, M.pend.ensure(2 * (nb + 1) * sizeof(uint32_t)));
    CC_HIP(ctx, M.list.ensure(n * sizeof(uint32_t) + 4));
    CC_HIP(ctx, M.stage.ensure(8 * n * sizeof(double)));
    CC_HIP(ctx, M.stage_id.ensure(n * sizeof(uint32_t)));
    CC_HIP(ctx, M.packed.ensure(8 * n * sizeof(double)));
    if (stage1 == 2) {
      CC_HIP(ctx, M.dec.ensure(n + 1));
      if (widths) CC_HIP(ctx, M.widths.ensure(4 * n * sizeof(float)));
    }
  }
  if (model && n) {
    CC_HIP(ctx, M.top1.ensure(n + 1));
    CC_HIP(ctx, M.counts.ensure(2 * (size_t)1024 * kOut * sizeof(unsigned long long)));
  }
  hipEvent_t e[4];
  for (auto& x : e) CC_HIP(ctx, hipEventCreate(&x));
  struct Cleanup {
    hipEvent_t* e;
    ~Cleanup() {
      for (int k = 0; k < 4; ++k) (void)hipEventDestroy(e[k]);
    }
  } cleanup{e};
  const uint64_t ld_dev = n, n64 = n;
  // stage 1 (cascade only): criterion 0 on every box, then the boxes it
  // leaves open packed densely (M.packed, stride n, with their indices in
  // M.list) and their number, all in device memory: the kernels after it
  // read the count there.  Without the cascade the positions are the boxes.
  CC_HIP(ctx, hipEventRecord(e[0], ctx->stream));
  const void* data = M.soa.p;       // SoA the second stage reads
  int by_pos = 0;                   // data indexed by position (packed) or by box
  const uint32_t* list = nullptr;   // box of each position (null: the position)
  const uint32_t* m_dev = nullptr;  // null: n positions
  if (cascade && n) {
    uint32_t* pend = static_cast<uint32_t*>(M.pend.p);
    uint32_t* offs = pend + nb + 1;
    hipFunction_t f0, fs, fl;
    CC_HIP(ctx, kernel(ctx, stage1 == 2 ? "ccrd_c2_first0_f32" : "ccrd_c2_first0", &f0));
    CC_HIP(ctx, kernel(ctx, "cc_scan_u32", &fs));
    CC_HIP(ctx, kernel(ctx, "ccrd_c2_first0_list", &fl));
    const uint32_t nb32 = (uint32_t)nb;
    if (stage1 == 2) {
      CC_HIP(ctx, launch(f0, dim3(nb32), dim3(256), 0, ctx->stream, M.soa.p, ld_dev, n64, ctx->d_konst, M.first.p,
                         M.evals.p, M.trap.p, M.dec.p, widths ? M.widths.p : nullptr, pend, M.stage.p, M.stage_id.p));
    } else {
      CC_HIP(ctx, launch(f0, dim3(nb32), dim3(256), 0, ctx->stream, M.soa.p, ld_dev, n64, ctx->d_konst, M.first.p,
                         M.evals.p, M.trap.p, pend, M.stage.p, M.stage_id.p));
    }
    CC_HIP(ctx, launch(fs, dim3(1), dim3(1024), 0, ctx->stream, pend, nb32, offs));
    CC_HIP(ctx, launch(fl, dim3(nb32), dim3(256), 0, ctx->stream, M.stage.p, M.stage_id.p, n64, pend, offs,
                       M.packed.p, M.list.p));
    data = M.packed.p;
    by_pos = 1;
    list = static_cast<const uint32_t*>(M.list.p);
    m_dev = offs + nb;
  }
  CC_HIP(ctx, hipEventRecord(e[1], ctx->stream));
  // ranking: the positions grouped into buckets by predicted class (box
  // indices in M.idx); the cascade's packed boxes move with them (M.stage),
  // while the full batch is read in place (moving it costs more than the
  // scattered reads save)
  const uint32_t* idx = list;
  const uint64_t* bstart = nullptr;  // null: every position in bucket 0 (plain order)
  if (model && n) {
    if ((rc = rank_dev(ctx, M, data, ld_dev, by_pos ? nullptr : list, n, m_dev, nullptr))) return rc;
    hipFunction_t fc, fp, fs;
    CC_HIP(ctx, kernel(ctx, "ccml_bucket_count", &fc));
    CC_HIP(ctx, kernel(ctx, "ccml_bucket_prefix", &fp));
    CC_HIP(ctx, kernel(ctx, "ccml_bucket_scatter", &fs));
    // about 4096 positions per block, at most 1024 blocks (one prefix group)
    uint64_t nb = (n + 4095) / 4096;
    if (cascade) nb = (nb + 7) / 8;  // the cascade leaves a small fraction
    const uint32_t nblocks = (uint32_t)(nb < 1 ? 1 : nb > 1024 ? 1024 : nb);
    CC_HIP(ctx, M.counts.ensure(2 * (size_t)nblocks * kOut * sizeof(unsigned long long)));
    unsigned long long* cnt = static_cast<unsigned long long*>(M.counts.p);
    unsigned long long* off = cnt + (size_t)nblocks * kOut;
    CC_HIP(ctx, launch(fc, dim3(nblocks), dim3(256), 0, ctx->stream, M.top1.p, n64, m_dev, cnt));
    CC_HIP(ctx, launch(fp, dim3(1), dim3(1024), 0, ctx->stream, cnt, nblocks, off, M.bstart.p));
    CC_HIP(ctx, launch(fs, dim3(nblocks), dim3(256), 0, ctx->stream, M.top1.p, n64, m_dev, off, list, M.idx.p,
                       by_pos ? data : nullptr, ld_dev, by_pos ? M.stage.p : nullptr));
    if (by_pos) data = M.stage.p;
    idx = static_cast<const uint32_t*>(M.idx.p);
    bstart = static_cast<const uint64_t*>(M.bstart.p);
  }
  CC_HIP(ctx, hipEventRecord(e[2], ctx->stream));
  if (n) {
    hipFunction_t f;
    CC_HIP(ctx, kernel(ctx, "ccrd_c2_first", &f));
    const int done0 = stage1;  // evaluations spent on criterion 0 in stage 1
    // one position per lane; the cascade's count is known on the device
    // only, so its grid covers a quarter of the boxes, grid-stride
    uint64_t blocks = (n + 255) / 256;
    if (cascade) blocks = (blocks + 3) / 4;
    CC_HIP(ctx, launch(f, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, data, ld_dev, idx, n64, m_dev, bstart,
                       ctx->d_konst, M.first.p, M.evals.p, M.trap.p, done0, by_pos));
  }
  CC_HIP(ctx, hipEventRecord(e[3], ctx->stream));
  uint32_t m = (uint32_t)n;
  if (m_dev) CC_HIP(ctx, hipMemcpyAsync(&m, m_dev, sizeof m, hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipMemcpyAsync(first, M.first.p, n, hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipMemcpyAsync(evals, M.evals.p, n, hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipMemcpyAsync(trap, M.trap.p, n, hipMemcpyDeviceToHost, ctx->stream));
  if (stage1 == 2 && n) {
    CC_HIP(ctx, hipMemcpyAsync(f32_decided, M.dec.p, n, hipMemcpyDeviceToHost, ctx->stream));
    if (widths)
      CC_HIP(ctx, hipMemcpyAsync(widths, M.widths.p, 4 * n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  }
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (stats) {
    float s1 = 0, rk = 0, s2 = 0;
    CC_HIP(ctx, hipEventElapsedTime(&s1, e[0], e[1]));
    CC_HIP(ctx, hipEventElapsedTime(&rk, e[1], e[2]));
    CC_HIP(ctx, hipEventElapsedTime(&s2, e[2], e[3]));
    std::memset(stats, 0, sizeof *stats);
    stats->rank_ms = model ? rk : 0.0;
    stats->eval_ms = (double)s1 + s2;
    stats->second_stage = m;
    for (size_t i = 0; i < n; ++i) {
      stats->criterion_evals += evals[i];
      if (first[i] < kCrit) {
        stats->proven += 1;
        stats->evals_proven += evals[i];
      }
    }
  }
  return CC_OK;
}

int cc_c2_eval_first_ex(cc_ctx* ctx, const double* soa, size_t n, size_t ld, const cc_ranker* model, int cascade,
                        uint8_t* first, uint8_t* evals, uint8_t* trap, cc_first_stats* stats) {
  return first_success(ctx, soa, n, ld, model, cascade ? 1 : 0, first, evals, trap, nullptr, nullptr, stats);
}

int cc_c2_eval_first_f32_ex(cc_ctx* ctx, const double* soa, size_t n, size_t ld, int cascade, uint8_t* first,
                            uint8_t* evals, uint8_t* trap, uint8_t* f32_decided, float* widths,
                            cc_first_stats* stats) {
  if (cascade) return first_success(ctx, soa, n, ld, nullptr, 2, first, evals, trap, f32_decided, widths, stats);
  return cc_c2_eval_first_f32(ctx, soa, n, ld, first, evals, trap, f32_decided, widths, stats);
}

int cc_c2_eval_first_f32(cc_ctx* ctx, const double* soa, size_t n, size_t ld, uint8_t* first, uint8_t* evals,
                         uint8_t* trap, uint8_t* f32_decided, float* widths, cc_first_stats* stats) {
  if (!ctx || !soa || !first || !evals || !trap || !f32_decided || ld < n) return CC_E_ARG;
  CC_HIP(ctx, hipSetDevice(ctx->device));
  MlState& M = ml(ctx);
  int rc = upload_soa(ctx, M.soa, soa, n, ld);
  if (rc) return rc;
  CC_HIP(ctx, M.first.ensure(n + 1));
  CC_HIP(ctx, M.evals.ensure(n + 1));
  CC_HIP(ctx, M.trap.ensure(n + 1));
  CC_HIP(ctx, M.top1.ensure(n + 1));
  float* d_w = nullptr;
  if (widths) {
    CC_HIP(ctx, M.grad.ensure(4 * n * sizeof(float) + 16));
    d_w = static_cast<float*>(M.grad.p);
  }
  hipFunction_t f;
  CC_HIP(ctx, kernel(ctx, "ccrd_c2_first_f32", &f));
  CC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if (n) {
    const uint64_t n64 = n, ld_dev = n;
    CC_HIP(ctx, launch(f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, M.soa.p, ld_dev, n64,
                       ctx->d_konst, M.first.p, M.evals.p, M.trap.p, M.top1.p, d_w));
  }
  CC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  CC_HIP(ctx, hipMemcpyAsync(first, M.first.p, n, hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipMemcpyAsync(evals, M.evals.p, n, hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipMemcpyAsync(trap, M.trap.p, n, hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipMemcpyAsync(f32_decided, M.top1.p, n, hipMemcpyDeviceToHost, ctx->stream));
  if (widths) CC_HIP(ctx, hipMemcpyAsync(widths, d_w, 4 * n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (stats) {
    float ms = 0;
    CC_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    std::memset(stats, 0, sizeof *stats);
    stats->eval_ms = ms;
    for (size_t i = 0; i < n; ++i) {
      stats->criterion_evals += evals[i];
      if (first[i] < kCrit) {
        stats->proven += 1;
        stats->evals_proven += evals[i];
      }
    }
  }
  return CC_OK;
}

int cc_c2_f32_report(cc_ctx* ctx, const double* soa, size_t n, size_t ld, uint32_t* f32_mask, float* inflation) {
  if (!ctx || !soa || !f32_mask || ld < n) return CC_E_ARG;
  CC_HIP(ctx, hipSetDevice(ctx->device));
  MlState& M = ml(ctx);
  int rc = upload_soa(ctx, M.soa, soa, n, ld);
  if (rc) return rc;
  CC_HIP(ctx, M.grad.ensure((n + 1) * 10 * sizeof(float) + (n + 1) * sizeof(uint32_t)));
  uint32_t* d_mask = static_cast<uint32_t*>(M.grad.p);
  float* d_infl = inflation ? reinterpret_cast<float*>(d_mask + n + 1) : nullptr;
  hipFunction_t f;
  CC_HIP(ctx, kernel(ctx, "ccrd_c2_f32_report", &f));
  if (n) {
    const uint64_t n64 = n, ld_dev = n;
    CC_HIP(ctx, launch(f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, M.soa.p, ld_dev, n64,
                       ctx->d_konst, d_mask, d_infl));
  }
  CC_HIP(ctx, hipMemcpyAsync(f32_mask, d_mask, n * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
  if (inflation)
    CC_HIP(ctx, hipMemcpyAsync(inflation, d_infl, 10 * n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  CC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return CC_OK;
}

}  // extern "C"

extern "C" {

int cc_rsb_ranker_train(cc_ctx* ctx, const cc_bnb_record* records, size_t n, int epochs, int batch, float lr,
                        uint64_t seed, cc_rsb_ranker* model, float* epoch_loss) {
  if (!ctx || !model || (n && !records)) return CC_E_ARG;
  // the ten steps before C14 (0..18) that prove the most records first
  uint64_t count[19] = {0};
  for (size_t i = 0; i < n; ++i)
    if (records[i].first_step >= 0 && records[i].first_step < 19) ++count[records[i].first_step];
  int order[19];
  for (int k = 0; k < 19; ++k) order[k] = k;
  std::sort(order, order + 19, [&](int a, int b) { return count[a] > count[b] || (count[a] == count[b] && a < b); });
  int cls_of_step[20];
  for (int k = 0; k < 20; ++k) cls_of_step[k] = kCrit;  // class 10: the reference's order
  for (int c = 0; c < kCrit; ++c) {
    model->class_step[c] = count[order[c]] ? order[c] : -1;
    if (count[order[c]]) cls_of_step[order[c]] = c;
  }
  model->class_step[kCrit] = -1;
  // features: lambda, r1, r2, r3 of each record (8 SoA rows), labels by class
  std::vector<double> soa(8 * n);
  std::vector<uint8_t> label(n);
  for (size_t i = 0; i < n; ++i) {
    const cc_rsb_cell& c = records[i].cell;
    const double v[8] = {c.la.lb, c.la.ub, c.r[0].lb, c.r[0].ub, c.r[1].lb, c.r[1].ub, c.r[2].lb, c.r[2].ub};
    for (int r = 0; r < 8; ++r) soa[r * n + i] = v[r];
    const int s = records[i].first_step;
    label[i] = (uint8_t)(s >= 0 && s < 19 ? cls_of_step[s] : kCrit);
  }
  return cc_ranker_train(ctx, soa.data(), n, n, label.data(), epochs, batch, lr, seed, &model->net, epoch_loss);
}

}  // extern "C"
