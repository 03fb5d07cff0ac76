"""Pin the float64 MLP oracle: backward formulas vs finite differences, bf16 rounding
vs known bit patterns, feature layout vs a hand-built example."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_bf16_rounding_known_values():
    assert mo.bf16(1.0) == 1.0
    assert mo.bf16(1.0 + 2 ** -8) == 1.0                 # halfway, ties to even
    assert mo.bf16(1.0 + 3 * 2 ** -9) == 1.0 + 2 ** -7    # above halfway rounds up
    assert mo.bf16(-0.5) == -0.5


def test_gradients_match_finite_differences():
    mo.EMULATE_BF16 = False
    try:
        rng = np.random.default_rng(0)
        fp, hd, B = 8, 6, 5
        n_params = fp * hd + hd + hd * hd + hd + hd + 1
        p = rng.normal(0, 0.5, n_params)
        x = rng.normal(0, 1, (B, fp))
        t = rng.normal(0, 1, B)
        g = mo.gradients(x, t, p, fp, hd)

        def loss(pp):
            y, _, _ = mo.forward(x, pp, fp, hd)
            return np.mean((y - t) ** 2)
        eps = 1e-6
        for i in rng.choice(n_params, 40, replace=False):
            e = np.zeros(n_params)
            e[i] = eps
            num = (loss(p + e) - loss(p - e)) / (2 * eps)
            assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), i
    finally:
        mo.EMULATE_BF16 = True


def test_features_layout():
    keys = np.array([[0.5, 0.1, 0.3, 0.9, 0.2, 0.4]], np.float32)   # n = 3
    area = np.array([1.0, 2.0, 3.0], np.float32)
    height = np.array([0.25, 0.5, 0.75], np.float32)
    x = mo.features(keys, area, height, 16, thresholds=[0.25, 0.5, 0.75])
    # sequence order by SV = [0.5, 0.1, 0.3] -> slots 1, 2, 0: positions 2, 0, 1
    assert x[0, 0:3].tolist() == [mo.bf16(np.float32(2) / np.float32(3)), 0.0,
                                  mo.bf16(np.float32(1) / np.float32(3))]
    # machine keys 0.9, 0.2, 0.4 -> buckets 3, 0, 1 of 4 machines, divided by 3
    assert x[0, 3:6].tolist() == [1.0, 0.0, mo.bf16(np.float32(1) / np.float32(3))]
    assert x[0, 6:9].tolist() == [2.0, 3.0, 1.0]
    assert x[0, 9:12].tolist() == [0.5, 0.75, 0.25]
    assert np.all(x[0, 12:] == 0)
    # one machine: every bucket is 0
    assert np.all(mo.features(keys, area, height, 16)[0, 3:6] == 0)


def test_machine_aggregate_features_layout():
    """Hand-worked aggregates: slots 0..3 on machines (by key) 1, 0, 1, 1 of two."""
    keys = np.array([[0.4, 0.3, 0.2, 0.1, 0.9, 0.5, 0.6, 0.75]], np.float32)   # n = 4; 0.5 is machine 0
    area = np.array([1.0, 2.0, 3.0, 4.0], np.float32)
    height = np.array([0.5, 0.25, 1.0, 0.75], np.float32)
    t = np.zeros((2, 3, 4), np.float32)
    t[0, 0] = [1, 2, 4, 8]          # only slot 1 is on machine 0
    t[0, 1] = [0.5, 0.25, 0.5, 0.5]
    t[0, 2] = [9, 3, 9, 9]
    t[1, 0] = [1, 2, 4, 8]          # slots 0, 2, 3: 1 + 4 + 8
    t[1, 1] = [0.5, 0.25, 0.125, 1.0]
    t[1, 2] = [0.5, 9, 0.25, 0.75]
    x = mo.features(keys, area, height, 32, thresholds=[0.5], machine_terms=t)
    # slot columns scaled by 1/8; sequence order slots 3, 2, 1, 0
    assert x[0, 0:4].tolist() == [0.75 / 8, 0.5 / 8, 0.25 / 8, 0.0]
    assert x[0, 4:8].tolist() == [1 / 8, 0.0, 1 / 8, 1 / 8]
    assert x[0, 8:12].tolist() == [4 / 8, 3 / 8, 2 / 8, 1 / 8]
    assert x[0, 12:16].tolist() == [0.75 / 8, 1 / 8, 0.25 / 8, 0.5 / 8]
    # machine 0: sums 2 and 0.25, count 1 * 2 / 4, tallest 3; machine 1: 13, 1.625, 3 * 2 / 4, 0.75
    assert x[0, 16:20].tolist() == [2.0, 0.25, 0.5, 3.0]
    assert x[0, 20:24].tolist() == [13.0, 1.625, 1.5, 0.75]
    assert np.all(x[0, 24:] == 0)
    # a machine with no slots: zeros
    keys[0, 4:] = 0.9
    x = mo.features(keys, area, height, 32, thresholds=[0.5], machine_terms=t)
    assert x[0, 16:20].tolist() == [0.0, 0.0, 0.0, 0.0] and x[0, 22] == 2.0


def test_machine_terms_scale_to_even_split():
    """machine_terms scales processing time and plate-relative area so that the
    parts spread evenly over identical machines sum to 1 per machine."""
    from brkga_fft2_amd.problem import synthetic_problem
    from brkga_fft2_amd.surrogate import machine_terms
    prob, n, M = synthetic_problem("C2", seed=1)
    t = machine_terms(prob, prob.instance_parts, M).astype(np.float64)
    assert t.shape == (M, 3, n)
    assert np.allclose(t[:, 0].sum(axis=1), M) and np.allclose(t[:, 1].sum(axis=1), M)
    assert t[:, 2].max() == 1.0


def test_machine_bucket_matches_decoder_assignment():
    """Buckets agree with the machine the C oracle places every part on,
    including keys exactly on a threshold (they belong to the lower bucket)."""
    from collections import Counter
    from oracle import oracle as orc
    from tests.helpers import syn_instance
    prob, n, M, packed = syn_instance("C3", seed=2)
    thr = np.asarray(packed["thresholds"], np.float64)
    rng = np.random.default_rng(5)
    keys = rng.random((6, 2 * n))
    keys[:, n:n + 3] = thr[:3]           # on the thresholds
    keys[:, n + 3] = 0.0
    _, log, _ = orc.decode(packed, keys, with_log=True)
    slot_type = np.asarray(packed["slot_type"])
    type_id = np.asarray(packed["type_id"])
    for b in range(keys.shape[0]):
        mach = mo.machine_bucket(keys[b, n:], thr)
        want = Counter(zip(mach.tolist(), type_id[slot_type].tolist()))
        got = Counter((int(r[0]), int(r[1])) for r in log[b])
        assert got == want


def _net_params(fp, hidden, rng):
    p = np.concatenate([rng.normal(0, 1 / np.sqrt(fp), fp * hidden), rng.normal(0, 0.1, hidden),
                        rng.normal(0, 1 / np.sqrt(hidden), hidden * hidden), rng.normal(0, 0.1, hidden),
                        np.abs(rng.normal(0, 1 / np.sqrt(hidden), hidden)), [0.1]])
    return p


def test_gradient_tolerance_accepts_float32_and_rejects_a_dropped_chunk():
    """The GPU gradient test's element-wise tolerance: a float32 evaluation of
    the same network (products and sums rounded to float32) stays inside it,
    while a gradient that loses one 64-row chunk of the batch (as a split-k
    partial left out would) falls outside it for most elements of every
    tensor (targets below every prediction and w3 > 0, so the products of the
    output-side gradients share one sign)."""
    rng = np.random.default_rng(5)
    n, hidden, B = 50, 128, 1024
    fp = 4 * n + (-4 * n) % 32
    x = mo.features(rng.random((B, 2 * n)).astype(np.float32), rng.random(n), rng.random(n), fp,
                    [1 / 3, 2 / 3])
    p = _net_params(fp, hidden, rng)
    y, _, _ = mo.forward(x, p, fp, hidden)
    t = y - 1.0 - rng.random(B)              # every prediction above its target
    g = mo.gradients(x, t, p, fp, hidden)
    tol = mo.gradient_tolerance(x, t, p, fp, hidden)
    # float32 evaluation
    p32 = p.astype(np.float32).astype(np.float64)
    g32 = mo.gradients(x.astype(np.float32).astype(np.float64), t.astype(np.float32).astype(np.float64),
                       p32, fp, hidden).astype(np.float32).astype(np.float64)
    assert np.all(np.abs(g32 - g) <= tol)
    # one 64-row chunk of the batch missing from the weight gradients
    keep = np.ones(B, bool)
    keep[128:192] = False
    gd = mo.gradients(x[keep], t[keep], p, fp, hidden) * (keep.sum() / B)   # same 2/B scale as the full batch
    sizes = [fp * hidden, hidden, hidden * hidden, hidden, hidden, 1]
    offs = np.cumsum([0] + sizes)
    for k, name in enumerate(("W1", "b1", "W2", "b2", "w3", "b3")):
        sl = slice(offs[k], offs[k + 1])
        frac = np.mean(np.abs(gd[sl] - g[sl]) > tol[sl])
        assert frac > 0.5, (name, frac)   # elements with no product in the dropped rows stay put


def test_lane_tree_sum_order():
    """The kernel's summation order: 64 strided lane sums, then a butterfly.
    Exactly representable terms give the plain sum; a case where float32
    rounding depends on the order tells it apart from the sequential sum."""
    v = np.arange(200, dtype=np.float32) * np.float32(0.25)
    assert mo.lane_tree_sum(v) == np.float32(v.astype(np.float64).sum())
    big = np.zeros(65, np.float32)
    big[0], big[64], big[1] = 2.0 ** 24, 1.0, 1.0   # lane 0 adds 2^24 + 1 (rounds away), lane 1 holds 1
    # sequential: ((2^24 + 1) + 1) rounds twice to 2^24; lanes: (2^24 + 1) -> 2^24, then + 1 -> 2^24 too
    assert mo.lane_tree_sum(big) == np.float32(2.0 ** 24)
    small = np.zeros(66, np.float32)
    small[0], small[1], small[64], small[65] = 2.0 ** 24, 1.0, 1.0, 1.0
    # lanes: lane 0 = 2^24 + 1 -> 2^24; lane 1 = 1 + 1 = 2; butterfly 2^24 + 2 exact
    assert mo.lane_tree_sum(small) == np.float32(2.0 ** 24 + 2)
    seq = np.float32(0)
    for x in small:
        seq = np.float32(seq + x)
    assert seq == np.float32(2.0 ** 24)   # the slot-order sum loses both ones
