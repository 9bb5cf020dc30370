"""cc_eval_batch_noalloc_async, the batch entry point that allocates nothing
(its scratch is the library's static per-GPU array), against
cc_eval_batch_async on the same boxes: bit-identical masks and first-success
bytes on two calls in a row, the second with fewer boxes (the array keeps
the first call's contents), and sizes beyond the maximum refused."""
import numpy as np
import pytest

from circlecover_amd import gen_c2_boxes
from circlecover_amd.native import CC_E_ARG, CC_PREC_FP32_PRE, CC_PREC_FP64, NOALLOC_MAX_BOXES, CCError


def _run(prover, soa, n, ld, noalloc, precision):
    assert soa.shape == (8, ld)
    d_soa = prover.alloc(soa.nbytes)
    d_mask, d_first = prover.alloc(4 * n), prover.alloc(n)
    try:
        d_soa.upload(soa)
        # outputs start as garbage: every box must be written
        d_mask.upload(np.full(n, 0xA5A5A5A5, dtype=np.uint32))
        d_first.upload(np.full(n, 0x5A, dtype=np.uint8))
        call = prover.eval_batch_noalloc_async if noalloc else prover.eval_batch_async
        call(d_soa, n, ld, d_mask, d_first, precision)
        prover.synchronize()
        return d_mask.download(np.empty(n, dtype=np.uint32)), d_first.download(np.empty(n, dtype=np.uint8))
    finally:
        for b in (d_soa, d_mask, d_first):
            b.free()


@pytest.mark.gpu
def test_gpu_noalloc_batch_matches_allocating_entry(prover):
    print("noalloc: two calls, 200000 then 70001 boxes", flush=True)
    for seed, n, ld, precision in ((0xC1C1EC0000000002, 200000, 200017, CC_PREC_FP64),
                                   (0xC1C1EC0000000005, 70001, 70001, CC_PREC_FP64)):
        soa = np.ascontiguousarray(gen_c2_boxes(seed, ld))
        want_m, want_f = _run(prover, soa, n, ld, False, precision)
        got_m, got_f = _run(prover, soa, n, ld, True, precision)
        assert np.array_equal(got_m, want_m) and np.array_equal(got_f, want_f)
        assert (want_f != 0x5A).all() and (want_m & 0x3FF).any()


@pytest.mark.gpu
def test_gpu_noalloc_batch_fp32_prefilter_and_limits(prover):
    soa = np.ascontiguousarray(gen_c2_boxes(0xC1C1EC0000000007, 4096))
    want_m, want_f = _run(prover, soa, 4096, 4096, False, CC_PREC_FP64)
    got_m, got_f = _run(prover, soa, 4096, 4096, True, CC_PREC_FP32_PRE)
    assert np.array_equal(got_m, want_m) and np.array_equal(got_f, want_f)
    # beyond the static array's capacity: refused before any launch
    d = prover.alloc(64)
    try:
        with pytest.raises(CCError) as e:
            prover.eval_batch_noalloc_async(d, NOALLOC_MAX_BOXES + 1, NOALLOC_MAX_BOXES + 1, d, d)
        assert e.value.code == CC_E_ARG
    finally:
        d.free()
