"""Surrogate MLP on the GPU: features, forward and gradients vs the float64 oracle,
training, and what screening gains over plain selection."""
import sys

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from tests.helpers import syn_instance

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


THR = np.array([1.0 / 3.0, 2.0 / 3.0])   # three machines


def _net(n=50, hidden=512, seed=0, terms=False):
    from brkga_fft2_amd.surrogate import MLPSurrogate
    rng = np.random.default_rng(seed)
    area = rng.random(n).astype(np.float32)
    height = rng.random(n).astype(np.float32)
    mt = (rng.random((THR.size + 1, 3, n)) * 3.0 / n).astype(np.float32) if terms else None
    net = MLPSurrogate(n, area, height, THR, hidden=hidden, seed=seed, machine_terms=mt)
    return net, area, height


@pytest.mark.parametrize("terms", [False, True])
def test_features_bit_exact(terms):
    _say("mlp features terms=%s" % terms)
    import torch
    net, area, height = _net(terms=terms)
    assert net.fp == 256 and (net.machine_terms is not None) == terms   # 4n (+ 4M) rounded up to 64
    keys = np.random.default_rng(1).random((37, 100)).astype(np.float32)
    keys[:, 50:52] = THR.astype(np.float32)   # near the thresholds
    keys[3, :50] = keys[3, 7]                  # all sequence keys tied
    keys[5, 50:] = 0.9                         # every slot on the last machine: the others get zeros
    x = net.features(torch.from_numpy(keys).cuda()).cpu().numpy().view(np.uint16)
    got = (x.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    want = mo.features(keys, area, height, net.fp, THR, machine_terms=net.machine_terms)
    np.testing.assert_array_equal(got, want)
    if terms:
        assert np.all(got[5, 200:208] == 0) and got[5, 210] == 3.0   # count 50 * 3 / 50
        assert np.all(got[:, 212:] == 0)
    net.close()


@pytest.mark.parametrize("B,same_sign,terms", [(64, False, False), (200, False, True), (333, False, False),
                                               (1024, False, True), (1024, True, False)])
def test_forward_and_gradients_match_oracle(B, same_sign, terms):
    """Element-wise tolerance (mlp_oracle.gradient_tolerance): |GPU - float64| <=
    (K 2^-23 + 2^-6) S + T per gradient element, S the sum of |products| the
    element adds and T one product (float32 accumulation of K terms, one-ulp
    bf16 flips of the two stored operands of a product, one flipped ReLU
    mask); also cosine >= 0.999 per tensor.  B = 333 and 1024 take the split-k
    weight gradients; with same_sign the targets lie below every prediction, so
    the output-side gradients add products of one sign and a lost split-k
    partial exceeds the bound (test_mlp_oracle pins that on the CPU)."""
    _say("mlp gradients B=%d same_sign=%s terms=%s" % (B, same_sign, terms))
    import torch
    net, area, height = _net(seed=B, terms=terms)
    rng = np.random.default_rng(B)
    p = net.get_params().astype(np.float64)
    o_b1 = net.fp * net.hidden
    p[o_b1:o_b1 + net.hidden] = rng.normal(0, 0.1, net.hidden)   # non-zero b1
    net.set_params(p.astype(np.float32))
    p = net.get_params().astype(np.float64)
    keys = rng.random((B, 100)).astype(np.float32)
    dk = torch.from_numpy(keys).cuda()
    y = net.predict(dk).cpu().numpy().astype(np.float64)
    x = mo.features(keys, area, height, net.fp, THR, machine_terms=net.machine_terms)
    y_ref, _, _ = mo.forward(x, p, net.fp, net.hidden)
    assert np.max(np.abs(y - y_ref)) <= 2e-2 * max(1.0, np.max(np.abs(y_ref)))
    if same_sign:
        t = (y_ref - 1.0 - rng.random(B)).astype(np.float32)
    else:
        t = rng.normal(0, 1, B).astype(np.float32)
    g = net.gradients(dk, torch.from_numpy(t).cuda()).cpu().numpy().astype(np.float64)
    t64 = t.astype(np.float64)
    g_ref = mo.gradients(x, t64, p, net.fp, net.hidden)
    tol = mo.gradient_tolerance(x, t64, p, net.fp, net.hidden)
    fp, hd = net.fp, net.hidden
    sizes = [fp * hd, hd, hd * hd, hd, hd, 1]
    o = 0
    for name, sz in zip(["W1", "b1", "W2", "b2", "w3", "b3"], sizes):
        a, b, tl = g[o:o + sz], g_ref[o:o + sz], tol[o:o + sz]
        o += sz
        bad = np.flatnonzero(np.abs(a - b) > tl)
        assert bad.size == 0, (name, bad[:5], (a - b)[bad[:5]], tl[bad[:5]])
        if sz > 1:
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300)
            assert cos >= 0.999, name
    # the loss a training step reports is the mean squared error before its update
    loss = net.train_step(dk, torch.from_numpy(t).cuda())
    loss_ref = float(np.mean((y_ref - t64) ** 2))
    assert abs(loss - loss_ref) <= 2e-2 * max(1.0, loss_ref)
    net.close()


def test_training_fits_a_batch():
    _say("mlp training")
    import torch
    net, _, _ = _net(seed=3)
    rng = np.random.default_rng(3)
    keys = torch.from_numpy(rng.random((512, 100)).astype(np.float32)).cuda()
    t = torch.from_numpy(rng.normal(0, 1, 512).astype(np.float32)).cuda()
    first = net.train_step(keys, t)
    for _ in range(60):
        last = net.train_step(keys, t)
    assert last < 0.5 * first
    net.close()


@pytest.mark.parametrize("config", ["C2", "C3"])
def test_screening_gains_over_plain_selection(config):
    """After online training on exact decodes, the network's picks from a fresh pool decode
    to better makespans than the pool's first rows (what plain BRKGA decodes); C3 has four
    machines with different plates, where the per-machine aggregates carry the ranking."""
    _say("surrogate gain %s" % config)
    import torch
    from brkga_fft2_amd.native import Decoder
    from brkga_fft2_amd.surrogate import SurrogateBRKGA, spearman
    prob, n, M, packed = syn_instance(config, seed=4)
    dec = Decoder(packed, max_bins=n)
    ga = SurrogateBRKGA(dec, prob, prob.instance_parts, num_individuals=512, pool=2048, warmup=3,
                        seed=7, train_steps=32, batch=1024)
    ga.run(10)
    _say("spearman per generation %s" % np.round(ga.log["spearman"], 3).tolist())
    assert all(np.isfinite(ga.log["loss"][1:]))
    # fresh pool from the current population
    lib = ga._lib
    s = ga.stream.cuda_stream
    lib.brkga_sort_fitness(ga.fit.data_ptr(), ga.N, ga.order.data_ptr(), s)
    lib.brkga_breed_pool(ga.pop.data_ptr(), ga.order.data_ptr(), ga.N, ga.G2, ga.ne, 0, 2048,
                         ga.rho, 12345, 999, ga.poolbuf.data_ptr(), s)
    pool = ga.poolbuf[:2048].contiguous()
    pred = ga.net.predict(pool).cpu().numpy()
    exact = torch.empty(2048, dtype=torch.float64, device=pool.device)
    dec.evaluate_device(pool, exact)
    f = exact.cpu().numpy()
    k = 512
    picked = f[np.argsort(pred, kind="stable")[:k]]
    plain = f[:k]
    rho = spearman(pred, f)
    _say("rho %.3f picked mean %.4f plain mean %.4f" % (rho, picked.mean(), plain.mean()))
    assert rho > 0.3
    assert picked.mean() < plain.mean()
    assert picked.min() <= plain.min()
    ga.close()
    dec.close()
