// circlecover_prover: the reference's executable (Prover/prover.cpp:33-37)
// over libcirclecover_amd.  Without arguments it runs the tight_rectangle
// proof and then the rectangle_size_bound proof with the reference's
// constants.  Optional flags shorten a run of the same procedure:
//   --only tr|rsb        run one of the two proofs
//   --tr-la N            initial lambda launches of tight_rectangle (256)
//   --tr-la-range F,C    run only the lambda pieces F..F+C-1
//   --rsb-la-range F,C   the same for rectangle_size_bound
//   --tr-generations G   refinement generations of tight_rectangle (11)
//   --rsb-grid A,B,C     initial lambda x r1 x r2 grid (256,128,128)
//   --rsb-generations G  refinement generations of rectangle_size_bound (1)
//   --tr-counts A,B,C,D      r1..r4 pieces of the initial launches (128^4)
//   --tr-gen-counts A,B,C,D  r1..r4 pieces of the generation launches (128,128,128,64)
//   --rsb-counts A,B,C,D     r3..r6 pieces of the initial launches (64^4)
//   --rsb-gen-counts A,B,C,D r3..r6 pieces of the generation launches (16,16,16,32)
//   --rsb-gen-grid A,B,C     lambda x r1 x r2 host grid of a generation (64,32,32)
//   --checkpoint P       save progress to P.tr / P.rsb and resume from them
//   --max-batches N      stop each proof after N batches, exit status 3
//                        (resume later with the same --checkpoint)
//   --enum-form F        batched (default), plain or plain-exact: the form of
//                        the per-cell enumerations (cc_set_enumeration_form)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

#include "circlecover_amd.h"

int main(int argc, const char* argv[]) {
  bool tr = true, rsb = true, custom = false;
  int tr_la = 0, tr_gen = -1, rsb_la = 0, rsb_r1 = 0, rsb_r2 = 0, rsb_gen = -1, enum_form = CC_ENUM_BATCHED;
  int tr_first = 0, tr_count = 0, rsb_first = 0, rsb_count = 0, max_batches = 0;
  const char* ckpt = nullptr;
  int tr_c[4], tr_gc[4], rsb_c[4], rsb_gc[4], rsb_gg[3];
  bool have_tr_c = false, have_tr_gc = false, have_rsb_c = false, have_rsb_gc = false, have_rsb_gg = false;
  auto four = [](const char* v, int* o) { return std::sscanf(v, "%d,%d,%d,%d", o, o + 1, o + 2, o + 3) == 4; };
  for (int i = 1; i < argc; ++i) {
    const char* a = argv[i];
    const char* v = i + 1 < argc ? argv[i + 1] : nullptr;
    if (!std::strcmp(a, "--only") && v) {
      tr = !std::strcmp(v, "tr");
      rsb = !std::strcmp(v, "rsb");
      ++i;
    } else if (!std::strcmp(a, "--tr-la") && v) {
      tr_la = std::atoi(v), custom = true, ++i;
    } else if (!std::strcmp(a, "--tr-la-range") && v) {
      if (std::sscanf(v, "%d,%d", &tr_first, &tr_count) != 2) return 2;
      custom = true, ++i;
    } else if (!std::strcmp(a, "--rsb-la-range") && v) {
      if (std::sscanf(v, "%d,%d", &rsb_first, &rsb_count) != 2) return 2;
      custom = true, ++i;
    } else if (!std::strcmp(a, "--tr-generations") && v) {
      tr_gen = std::atoi(v), custom = true, ++i;
    } else if (!std::strcmp(a, "--rsb-grid") && v) {
      if (std::sscanf(v, "%d,%d,%d", &rsb_la, &rsb_r1, &rsb_r2) != 3) {
        std::cerr << "--rsb-grid expects A,B,C" << std::endl;
        return 2;
      }
      custom = true, ++i;
    } else if (!std::strcmp(a, "--rsb-generations") && v) {
      rsb_gen = std::atoi(v), custom = true, ++i;
    } else if (!std::strcmp(a, "--tr-counts") && v) {
      if (!four(v, tr_c)) return 2;
      have_tr_c = custom = true, ++i;
    } else if (!std::strcmp(a, "--tr-gen-counts") && v) {
      if (!four(v, tr_gc)) return 2;
      have_tr_gc = custom = true, ++i;
    } else if (!std::strcmp(a, "--rsb-counts") && v) {
      if (!four(v, rsb_c)) return 2;
      have_rsb_c = custom = true, ++i;
    } else if (!std::strcmp(a, "--rsb-gen-counts") && v) {
      if (!four(v, rsb_gc)) return 2;
      have_rsb_gc = custom = true, ++i;
    } else if (!std::strcmp(a, "--rsb-gen-grid") && v) {
      if (std::sscanf(v, "%d,%d,%d", rsb_gg, rsb_gg + 1, rsb_gg + 2) != 3) return 2;
      have_rsb_gg = custom = true, ++i;
    } else if (!std::strcmp(a, "--checkpoint") && v) {
      ckpt = v, custom = true, ++i;
    } else if (!std::strcmp(a, "--max-batches") && v) {
      max_batches = std::atoi(v), custom = true, ++i;
    } else if (!std::strcmp(a, "--enum-form") && v) {
      if (!std::strcmp(v, "batched")) enum_form = CC_ENUM_BATCHED;
      else if (!std::strcmp(v, "plain")) enum_form = CC_ENUM_PLAIN;
      else if (!std::strcmp(v, "plain-exact")) enum_form = CC_ENUM_PLAIN_EXACT;
      else {
        std::cerr << "--enum-form expects batched, plain or plain-exact" << std::endl;
        return 2;
      }
      custom = true, ++i;
    } else {
      std::cerr << "unknown argument " << a << std::endl;
      return 2;
    }
  }
  if (!custom) {  // the reference's main
    if (tr) circlecover::tight_rectangle::find_critical_intervals();
    if (rsb) circlecover::rectangle_size_bound::find_critical_intervals();
    return 0;
  }
  cc_ctx* ctx = nullptr;
  if (cc_init(0, &ctx) != CC_OK) {
    std::cerr << "no usable gfx950 device" << std::endl;
    return 1;
  }
  int rc = cc_set_enumeration_form(ctx, enum_form);
  if (rc == CC_OK) rc = cc_tr_prove_shape(ctx, have_tr_c ? tr_c : nullptr, have_tr_gc ? tr_gc : nullptr);
  if (rc == CC_OK)
    rc = cc_rsb_prove_shape(ctx, have_rsb_c ? rsb_c : nullptr, have_rsb_gg ? rsb_gg : nullptr,
                            have_rsb_gc ? rsb_gc : nullptr);
  if (rc != CC_OK) {
    std::cerr << "error " << rc << ": " << cc_last_error(ctx) << std::endl;
    cc_destroy(ctx);
    return 2;
  }
  const std::string ck_tr = ckpt ? std::string(ckpt) + ".tr" : "", ck_rsb = ckpt ? std::string(ckpt) + ".rsb" : "";
  if (tr) rc = cc_tr_prove_ex(ctx, tr_la, tr_first, tr_count, tr_gen, ckpt ? ck_tr.c_str() : nullptr, max_batches, stdout);
  if (rc == CC_OK && rsb)
    rc = cc_rsb_prove_ex(ctx, rsb_la, rsb_first, rsb_count, rsb_r1, rsb_r2, rsb_gen, ckpt ? ck_rsb.c_str() : nullptr,
                         max_batches, stdout);
  if (rc != CC_OK && rc != CC_E_INTERRUPTED) std::cerr << "error " << rc << ": " << cc_last_error(ctx) << std::endl;
  cc_destroy(ctx);
  if (rc == CC_E_INTERRUPTED && max_batches > 0) return 3;  // stopped after --max-batches: resume later
  return rc == CC_OK || rc == CC_E_INTERRUPTED ? 0 : 1;
}
