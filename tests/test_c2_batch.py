"""Ten-criterion batch (4-D boxes): GPU verdicts against the CPU oracle,
bit-exact.  Boxes come from the seeded generator plus hand-built edge cases
(point intervals, degenerate radii, boxes straddling the criteria's
thresholds, lambda at the break-even point)."""
import numpy as np
import pytest

import oracle
from circlecover_amd import gen_c2_boxes

SEED = 0xC1C1EC0000000002


def edge_boxes() -> np.ndarray:
    rows = []
    lam2 = 1.0357971111816713  # break-even lambda_2 of critical_ratio
    for la in (1.0, lam2, 1.05, 1.5, 2.0898841580413818, 2.5):
        for r in ((0.49, 0.245, 0.1), (0.55, 0.3, 0.27), (0.7, 0.7, 0.7), (0.3, 0.2, 0.02), (0.25, 0.25, 0.25),
                  (0.140625, 0.140625, 0.140625), (0.6, 0.5, 0.4), (0.5625, 0.2, 0.0)):
            for w in (0.0, 2.0 ** -30, 2.0 ** -8):
                box = []
                for lb in (la, *r):
                    box += [lb, lb + w]
                rows.append(box)
    return np.array(rows, dtype=np.float64).T.copy()


def small_radius_boxes(n: int) -> np.ndarray:
    """The generator's boxes with radii scaled into the size-bounded regime
    (squared radii <= 9/64), where the rectangle_size_bound shortcuts fire."""
    soa = gen_c2_boxes(SEED + 7, n)
    soa[2:] *= 0.2
    return soa


def cover_strip_boxes(n: int) -> np.ndarray:
    """Boxes aimed at the cover-strip criterion's width search: equal
    squared radii r2 = r3 (both height tests bind together), r1 near 1/4
    (the search starts where its test is not yet monotone), lambda just above
    the starting width 0.3 (the test holds up to lambda), radii a hair apart,
    and random boxes."""
    rng = np.random.default_rng(11)
    la = 1.0 + 7.0 * rng.random(n)
    r1 = rng.random(n) * 0.6 * la
    r2 = rng.random(n) * r1
    r3 = rng.random(n) * r2
    kind = np.arange(n) % 6
    m = kind == 0
    r3[m] = r2[m]
    m = kind == 1
    r1[m] = 0.25 + 0.05 * rng.random(m.sum())
    r2[m] = 0.25 - 1e-3 * rng.random(m.sum())
    r3[m] = r2[m]
    m = kind == 2
    la[m] = 0.3 + 0.1 * rng.random(m.sum())
    m = kind == 3
    r2[m] = 0.25 * (1 + 1e-12 * rng.random(m.sum()))
    r3[m] = 0.25 * (1 - 1e-9 * rng.random(m.sum()))
    m = kind == 4
    la[m] = 1.0 + 0.5 * rng.random(m.sum())
    r1[m] = 0.2 + 0.3 * rng.random(m.sum())
    r2[m] = r1[m] * (0.5 + 0.5 * rng.random(m.sum()))
    r3[m] = r2[m] * (0.5 + 0.5 * rng.random(m.sum()))
    w = np.where(rng.random(n) < 0.5, 0.0, 2.0 ** -20)
    soa = np.empty((8, n))
    for k, v in enumerate((la, r1, r2, r3)):
        soa[2 * k] = v
        soa[2 * k + 1] = v + w
    return soa


def test_oracle_c2_deterministic_and_threads_agree():
    soa = np.concatenate([gen_c2_boxes(SEED, 3000), small_radius_boxes(3000)], axis=1)
    m1, f1 = oracle.c2_eval(soa, threads=1)
    m4, f4 = oracle.c2_eval(soa, threads=4)
    assert (m1 == m4).all() and (f1 == f4).all()
    # every criterion fires on some boxes and fails on others
    bits = (m1[:, None] >> np.arange(10)) & 1
    assert bits.any(axis=0).all(), bits.any(axis=0)
    assert (~bits.astype(bool)).any(axis=0).all()
    # first_ok is the lowest set bit
    low = np.where((m1 & 0x3FF) == 0, 255, np.argmax(bits, axis=1))
    assert (low == f1).all()


@pytest.mark.gpu
def test_gpu_c2_matches_oracle_seeded(prover):
    soa = gen_c2_boxes(SEED, 1 << 16)
    gm, gf = prover.eval_batch(soa)
    om, of = oracle.c2_eval(soa)
    bad = np.flatnonzero((gm != om) | (gf != of))
    assert bad.size == 0, f"{bad.size} boxes differ; first {bad[:5]}: gpu {gm[bad[:5]]} oracle {om[bad[:5]]}"


@pytest.mark.gpu
def test_gpu_c2_matches_oracle_small_radii(prover):
    soa = small_radius_boxes(1 << 15)
    gm, gf = prover.eval_batch(soa)
    om, of = oracle.c2_eval(soa)
    assert (gm == om).all() and (gf == of).all()
    assert ((om >> 6) & 7).any()  # the uneven-split / opposite-corner shortcuts are exercised


@pytest.mark.gpu
def test_gpu_c2_matches_oracle_cover_strip_regimes(prover):
    soa = cover_strip_boxes(1 << 15)
    gm, gf = prover.eval_batch(soa)
    om, of = oracle.c2_eval(soa)
    bad = np.flatnonzero((gm != om) | (gf != of))
    assert bad.size == 0, f"{bad.size} boxes differ; first {bad[:5]}: gpu {gm[bad[:5]]} oracle {om[bad[:5]]}"
    strip = (om >> 4) & 1
    assert strip.any() and not strip.all()  # the criterion both proves and fails boxes here


@pytest.mark.gpu
def test_gpu_c2_matches_oracle_edge_cases(prover):
    soa = edge_boxes()
    gm, gf = prover.eval_batch(soa)
    om, of = oracle.c2_eval(soa)
    assert (gm == om).all() and (gf == of).all()


@pytest.mark.gpu
def test_gpu_c2_full_size_sampled(prover):
    """2^22 boxes on the GPU; a seeded 1/64 sample re-checked on the CPU, and
    the batch result must not depend on the batch split (ragged tail)."""
    n = 1 << 22
    soa = gen_c2_boxes(SEED, n)
    gm, gf = prover.eval_batch(soa)
    rng = np.random.default_rng(3)
    idx = np.sort(rng.choice(n, n // 64, replace=False))
    om, of = oracle.c2_eval(soa[:, idx])
    assert (gm[idx] == om).all() and (gf[idx] == of).all()
    k = 1000003  # ragged split
    am, af = prover.eval_batch(soa[:, :k])
    bm, bf = prover.eval_batch(soa[:, k:])
    assert (np.concatenate([am, bm]) == gm).all() and (np.concatenate([af, bf]) == gf).all()


@pytest.mark.gpu
def test_gpu_c2_empty_batch(prover):
    m, f = prover.eval_batch(np.empty((8, 0)))
    assert m.size == 0 and f.size == 0
