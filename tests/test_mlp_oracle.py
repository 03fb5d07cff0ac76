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
