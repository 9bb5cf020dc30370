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


def _f32_sets():
    from test_c2_batch import cover_strip_boxes, edge_boxes, small_radius_boxes
    return {"seeded": gen_c2_boxes(SEED + 3, 1 << 16), "small_radii": small_radius_boxes(1 << 15),
            "cover_strip": cover_strip_boxes(1 << 14), "edge": edge_boxes()}


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["seeded", "small_radii", "cover_strip", "edge"])
def test_gpu_f32_shadows_imply_fp64(prover, name):
    """Each criterion's FP32 shadow proves only boxes its FP64 form proves
    (per criterion, FP32-accept and FP64-reject never happen), and the
    pre-filtered batch (every shadow first, FP64 where undecided) returns the
    FP64 batch's masks and first criteria exactly."""
    soa = _f32_sets()[name]
    m64, f64 = prover.eval_batch(soa)
    m32, infl = prover.f32_report(soa)
    bad = m32 & ~(m64 & 0x3FF)
    assert not bad.any(), f"{int((bad != 0).sum())} boxes: FP32 proves a criterion FP64 rejects"
    assert not (m32 & (1 << 4)).any()  # criterion 4 has no shadow
    mp, fp = prover.eval_batch(soa, precision=1)
    assert (mp == m64).all() and (fp == f64).all()
    assert infl.shape == (soa.shape[1], 10) and (infl[:, 4] == -1).all()


@pytest.mark.gpu
def test_gpu_f32_shadows_decide_most_fp64_proofs(prover):
    """The shadows are not vacuous: on the seeded and size-bounded boxes each
    criterion's shadow proves most of the boxes its FP64 form proves, and
    the FP32 budgets are wider than the FP64 ones (outward FP32 rounding)."""
    sets = _f32_sets()
    soa = np.concatenate([sets["seeded"], sets["small_radii"]], axis=1)
    m64, _ = prover.eval_batch(soa)
    m32, infl = prover.f32_report(soa)
    for k in (0, 1, 2, 3, 5, 6, 7, 8, 9):
        p64 = ((m64 >> k) & 1).astype(bool)
        p32 = ((m32 >> k) & 1).astype(bool)
        assert p64.sum() > 100, (k, int(p64.sum()))
        assert p32.sum() >= 0.5 * p64.sum(), (k, int(p32.sum()), int(p64.sum()))
        r = infl[:, k][infl[:, k] >= 0]
        assert r.size > 0.9 * soa.shape[1] and np.median(r) > 1.0, (k, r.size)
