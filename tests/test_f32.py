"""FP32 outward pre-filter on the ten-criterion box set: boxes it accepts
must be boxes the FP64 path proves (FP32-accept and FP64-reject must not
happen), and the final first-success results equal the plain FP64 ones."""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from circlecover_amd import gen_c2_boxes

SEED = 0xC1C1EC0000000002


def test_f32_ratio_constants_bracket_061():
    up = np.float32(0.610000014305114746)
    down = np.float32(0.609999954700469971)
    assert Fraction(float(down)) < Fraction(61, 100) < Fraction(float(up))
    assert np.nextafter(down, np.float32(1)) == up  # adjacent FP32 numbers
    assert float(np.float32(9 / 64)) == 9 / 64 and float(np.float32(195 / 256)) == 195 / 256


@pytest.mark.gpu
def test_gpu_f32_prefilter_identity(prover):
    soa = np.concatenate([gen_c2_boxes(SEED, 1 << 16), gen_c2_boxes(SEED + 5, 1 << 14)], axis=1)
    first, evals, trap, dec, w, st = prover.eval_first_f32(soa, with_widths=True)
    mask, of = oracle.c2_eval(soa)
    # FP32 accepts only boxes FP64 criterion 0 proves
    accepted = dec == 1
    assert accepted.mean() > 0.5
    assert ((mask[accepted] & 1) == 1).all()
    # same first proving criterion as the plain FP64 order everywhere
    assert (first == of).all()
    # FP32 enclosures are wider than the FP64 ones (the FP64 path keeps the
    # reference's inward scalar-product quirk, so this is checked on the bulk)
    assert (w[:, 0] >= w[:, 1]).mean() > 0.999 and (w[:, 2] >= w[:, 3]).mean() > 0.999
    assert np.median(w[:, 2] / np.maximum(w[:, 3], 1e-300)) > 1.0


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1, 2049, 1 << 15])
def test_gpu_f32_cascade_matches_per_box_form(prover, n):
    """The cascade form of the pre-filter (FP32 shadow of criterion 0 on every
    box, FP64 criterion 0 on the undecided ones, the rest on the boxes still
    open) gives the per-box form's outputs exactly."""
    soa = gen_c2_boxes(SEED + 7, n)
    a = prover.eval_first_f32(soa, with_widths=True)
    b = prover.eval_first_f32(soa, with_widths=True, cascade=True)
    for x, y in zip(a[:5], b[:5]):
        assert (x == y).all()
    assert a[5].proven == b[5].proven and a[5].evals_proven == b[5].evals_proven
