"""Learned criterion ranker: the numpy restatement is pinned by finite
differences; the GPU's gradients (one forward/backward pass) match it; the
MFMA inference matches its forward; training lowers the loss; and ranked
first-success evaluation decides exactly the boxes plain order decides,
with fewer criterion evaluations per proven box."""
import numpy as np
import pytest

import oracle
from circlecover_amd import gen_c2_boxes
from oracle import mlp

SEED = 0xC1C1EC0000000002


def test_numpy_mlp_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = rng.standard_normal(mlp.N_PARAMS) * 0.2
    x = rng.standard_normal((32, 8))
    y = rng.integers(0, mlp.K_OUT, 32)
    _, g, _ = mlp.loss_and_grad(params, x, y, quantize=False)
    eps = 1e-6
    for i in rng.choice(mlp.N_PARAMS, 60, replace=False):
        a, b = params.copy(), params.copy()
        a[i] += eps
        b[i] -= eps
        num = (mlp.loss_and_grad(a, x, y, quantize=False)[0] - mlp.loss_and_grad(b, x, y, quantize=False)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-7 + 1e-6 * abs(g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5e-3, 3.0e38], dtype=np.float32)
    r = mlp.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa wins (1.0)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.0078125) and r[3] == np.float32(1.0078125)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)


def labelled_boxes(seed, n):
    soa = gen_c2_boxes(seed, n)
    _, first = oracle.c2_eval(soa)
    return soa, first


def class_labels(first):
    """Ranker classes: the first proving criterion, 10 where none proves."""
    return np.where(first == 255, 10, first).astype(np.uint8)


@pytest.mark.gpu
def test_gpu_gradient_matches_numpy(prover):
    from circlecover_amd.native import Ranker

    soa, first = labelled_boxes(SEED + 1, 1 << 14)
    label = class_labels(first)
    # every class present, the unprovable class 10 included
    sample = np.concatenate([np.flatnonzero(first < 10)[:1000], np.flatnonzero(first == 255)[:200]]).astype(np.uint32)
    model, _ = prover.ranker_train(soa, label, epochs=1, batch=4096, seed=3)  # a trained, non-trivial point
    g, loss = prover.ranker_gradient(soa, label, sample, model)
    x = mlp.features(soa, sample, np.ctypeslib.as_array(model.norm))
    ref_loss, ref_g, ref_losses = mlp.loss_and_grad(model.params_array(), x, label[sample].astype(int))
    assert np.allclose(loss, ref_losses, rtol=1e-4, atol=1e-5)
    scale = np.abs(ref_g).max()
    assert np.abs(g - ref_g).max() <= 1e-4 * scale, (np.abs(g - ref_g).max(), scale)
    assert isinstance(model, Ranker)


@pytest.mark.gpu
def test_gpu_mfma_inference_matches_numpy(prover):
    soa, first = labelled_boxes(SEED + 1, 1 << 14)
    model, _ = prover.ranker_train(soa, class_labels(first), epochs=2, batch=4096, seed=4)
    test = gen_c2_boxes(SEED + 9, 5000)
    top1, lg = prover.ranker_rank(test, model, with_logits=True)
    x = mlp.features(test, np.arange(5000), np.ctypeslib.as_array(model.norm))
    ref = mlp.logits(model.params_array(), x)
    assert np.abs(lg - ref).max() <= 2e-2 * max(1.0, np.abs(ref).max())
    # top-1 agrees wherever the margin exceeds the rounding noise
    srt = np.sort(ref, axis=1)
    clear = srt[:, -1] - srt[:, -2] > 5e-2
    assert (top1[clear] == np.argmax(ref, axis=1)[clear]).all()


@pytest.mark.gpu
def test_gpu_plain_first_success_matches_oracle(prover):
    soa = gen_c2_boxes(SEED, 1 << 15)
    first, evals, trap, st = prover.eval_first(soa)
    _, of = oracle.c2_eval(soa)
    assert (first == of).all()
    assert (evals == np.where(of < 10, of.astype(int) + 1, 10)).all()


@pytest.mark.gpu
def test_gpu_training_and_ranked_evaluation(prover):
    soa, first = labelled_boxes(SEED + 1, 1 << 17)
    model, losses = prover.ranker_train(soa, class_labels(first), epochs=4, batch=4096, seed=1)
    assert losses[-1] < losses[0]
    test = gen_c2_boxes(SEED, 1 << 16)
    mask, of = oracle.c2_eval(test)
    pf, pe, _, ps = prover.eval_first(test)
    rf, re, _, rs = prover.eval_first(test, model)
    # same boxes proven; every reported criterion really proves its box
    assert ((rf < 10) == (of < 10)).all()
    ok = rf < 10
    assert (((mask[ok] >> rf[ok].astype(np.uint32)) & 1) == 1).all()
    assert rs.proven == ps.proven
    # the ranker never needs more evaluations per proven box than plain order
    assert rs.evals_proven <= ps.evals_proven


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1, 2047, 2049, 1 << 15])
def test_gpu_plain_cascade_matches_plain_order(prover, n):
    """Criterion 0 on every box first, the rest on the boxes it leaves open:
    the same first/evals/trap as plain order, whatever the chunk borders."""
    soa = gen_c2_boxes(SEED + 5, n)
    pf, pe, pt, ps = prover.eval_first(soa)
    cf, ce, ct, cs = prover.eval_first(soa, cascade=True)
    assert (cf == pf).all() and (ce == pe).all() and (ct == pt).all()
    assert cs.proven == ps.proven and cs.evals_proven == ps.evals_proven
    # the second stage holds exactly the boxes criterion 0 does not prove
    assert cs.second_stage == int((pf != 0).sum())
    if n:
        _, of = oracle.c2_eval(soa)
        assert (cf == of).all()


@pytest.mark.gpu
def test_gpu_ranked_cascade(prover):
    soa, first = labelled_boxes(SEED + 1, 1 << 18)
    # the cascade's ranker sees only the boxes criterion 0 leaves open (about
    # 8 %, so more epochs than above for a comparable number of steps)
    label = class_labels(first)
    model, _ = prover.ranker_train(soa, np.where(label == 0, 255, label).astype(np.uint8), epochs=30, batch=4096,
                                   seed=1)
    test = gen_c2_boxes(SEED, 1 << 16)
    mask, of = oracle.c2_eval(test)
    pf, pe, _, ps = prover.eval_first(test)
    rf, re, _, rs = prover.eval_first(test, model, cascade=True)
    assert ((rf < 10) == (of < 10)).all()
    ok = rf < 10
    assert (((mask[ok] >> rf[ok].astype(np.uint32)) & 1) == 1).all()
    # criterion 0 still decides first wherever it proves the box
    assert (rf[of == 0] == 0).all() and (re[of == 0] == 1).all()
    assert rs.proven == ps.proven
    assert rs.evals_proven <= ps.evals_proven
