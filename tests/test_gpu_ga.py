"""BRKGA generation loops on the GPU: the host drop-in and the device graph loop."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle
from tests.helpers import oracle_threads, ref_instance, syn_instance

pytestmark = pytest.mark.gpu


class _OracleEvaluator:
    """Same evaluate_batch contract, decoded by the CPU oracle (test checker)."""

    def __init__(self, packed):
        self.packed = packed

    def evaluate_batch(self, chromosomes):
        ms, _, _ = oracle.decode(self.packed, np.asarray(chromosomes), n_threads=oracle_threads())
        return ms.tolist()


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_host_loop_matches_oracle_driven_loop():
    """The drop-in BRKGA loop driven by the GPU evaluator follows exactly the trajectory of
    the same loop driven by the CPU oracle (same np.random seed, bit-exact fitness)."""
    _say("host BRKGA loop")
    from brkga_fft2_amd.brkga import BRKGA
    from brkga_fft2_amd.problem import thresholds_for
    prob, n, M, packed = syn_instance("C1", seed=5)
    kw = dict(num_generations=6, num_individuals=64, num_elites=13, num_mutants=10, eliteCProb=0.7)
    np.random.seed(11)
    gpu = BRKGA(prob, n, M, thresholds_for(M), prob.instance_parts, None, **kw)
    gpu.fit()
    np.random.seed(11)
    cpu = BRKGA(prob, n, M, thresholds_for(M), prob.instance_parts, None,
                evaluator=_OracleEvaluator(packed), **kw)
    cpu.fit()
    assert gpu.history["min"] == cpu.history["min"]
    assert gpu.history["mean"] == cpu.history["mean"]
    assert gpu.best_fitness == cpu.best_fitness
    assert np.array_equal(gpu.solution, cpu.solution)
    mins = np.minimum.accumulate(gpu.history["min"])
    assert np.all(np.diff(mins) <= 0)


@pytest.mark.parametrize("cfg", ["C1", "C2", "C3"])   # one, two and four machines
def test_device_loop_graph_equals_eager_and_oracle(cfg):
    _say("device GA %s" % cfg)
    from brkga_fft2_amd.native import Decoder, DeviceGA
    prob, n, M, packed = syn_instance(cfg, seed=2)
    N, ne, nm, gens = 128, 26, 19, 5
    outs = []
    for use_graph in (True, False):
        dec = Decoder(packed, max_bins=n)
        ga = DeviceGA(dec, N, ne, nm, elite_prob=0.7, seed=99, history_capacity=gens)
        ga.run(gens, use_graph=use_graph)
        best, keys, hist, done = ga.best()
        pop, fit = ga.population()
        outs.append((best, keys, hist, pop, fit))
        assert done == gens and len(hist) == gens + 1
        assert np.all(np.diff(np.minimum.accumulate(hist)) <= 0)
        assert best == np.minimum.accumulate(hist)[-1]
        st = dec.stats()
        assert st["overflows"] == 0
        ga.close()
        dec.close()
    (b1, k1, h1, p1, f1), (b2, k2, h2, p2, f2) = outs
    assert b1 == b2 and np.array_equal(k1, k2) and np.array_equal(h1, h2)
    assert np.array_equal(p1, p2) and np.array_equal(f1, f2)
    # the stored fitness is the exact decode of the stored keys
    want, _, _ = oracle.decode(packed, p1, n_threads=oracle_threads())
    np.testing.assert_array_equal(f1, want)
    bw, _, _ = oracle.decode(packed, k1[None])
    assert bw[0] == b1


def test_device_loop_keeps_elites():
    _say("device GA elites")
    from brkga_fft2_amd.native import Decoder, DeviceGA
    prob, n, M, packed = ref_instance("P25M2-4")
    N, ne, nm = 100, 10, 15
    dec = Decoder(packed, max_bins=n)
    init = np.random.default_rng(3).random((N, 2 * n)).astype(np.float32)
    ga = DeviceGA(dec, N, ne, nm, seed=5, history_capacity=4, init_population=init)
    pop0, fit0 = ga.population()
    assert np.array_equal(pop0, init)
    ga.run(1, use_graph=True)
    pop1, fit1 = ga.population()
    order = np.lexsort((np.arange(N), fit0))[:ne]
    np.testing.assert_array_equal(pop1[:ne], pop0[order])
    np.testing.assert_array_equal(fit1[:ne], fit0[order])
    ga.close()
    dec.close()


def test_graph_recaptured_after_decoder_buffers_move():
    """A larger population decoded through the GA's decoder reallocates the
    decoder's per-call device buffers; the next graph replay must use the new
    ones (the graph is captured again), so the trajectory stays the eager one."""
    _say("device GA graph after a larger decode")
    from brkga_fft2_amd.native import Decoder, DeviceGA
    prob, n, M, packed = syn_instance("C2", seed=2)
    N, ne, nm = 128, 26, 19
    dec = Decoder(packed, max_bins=n)
    ga = DeviceGA(dec, N, ne, nm, elite_prob=0.7, seed=17, history_capacity=8)
    ga.run(2, use_graph=True)
    big = np.random.default_rng(5).random((3 * N, 2 * n)).astype(np.float32)
    ms, _ = dec.evaluate(big)
    want, _, _ = oracle.decode(packed, big, n_threads=oracle_threads())
    np.testing.assert_array_equal(ms, want)
    ga.run(3, use_graph=True)
    pop, fit = ga.population()
    best, _, hist, done = ga.best()
    ga.close()
    dec.close()
    dec2 = Decoder(packed, max_bins=n)
    ge = DeviceGA(dec2, N, ne, nm, elite_prob=0.7, seed=17, history_capacity=8)
    ge.run(5, use_graph=False)
    pop2, fit2 = ge.population()
    best2, _, hist2, _ = ge.best()
    ge.close()
    dec2.close()
    assert done == 5
    np.testing.assert_array_equal(pop, pop2)
    np.testing.assert_array_equal(fit, fit2)
    np.testing.assert_array_equal(hist, hist2)
    assert best == best2
    want, _, _ = oracle.decode(packed, pop, n_threads=oracle_threads())
    np.testing.assert_array_equal(fit, want)


def test_device_loop_reports_bin_overflow():
    """With too few bins per context the device loop cannot grow the workspace
    (no allocation inside a generation): it refuses with BRKGA_EOVERFLOW instead
    of carrying NaN fitness into the population."""
    _say("device GA overflow")
    from brkga_fft2_amd.native import Decoder, DecoderError, DeviceGA
    prob, n, M, packed = syn_instance("C2", seed=2)
    init = np.random.default_rng(4).random((64, 2 * n)).astype(np.float32)
    _, _, st = oracle.decode(packed, init, n_threads=oracle_threads())
    assert st["bins_opened"] > 64 * M   # some context needs a second bin
    dec = Decoder(packed, max_bins=1)
    with pytest.raises(DecoderError) as e:
        DeviceGA(dec, 64, 13, 10, seed=1, init_population=init)
    assert e.value.code == -4
    dec.close()


def _parents_of(child, pop, elites, non_elites):
    """(elite row, non-elite row, mask of genes taken from the elite) that explain
    ``child`` gene by gene, or None."""
    for e in elites:
        from_e = child == pop[e]
        rest = ~from_e
        if not rest.any():
            return e, None, from_e
        for o in non_elites:
            if np.array_equal(child[rest], pop[o][rest]):
                return e, o, from_e
    return None


def _check_offspring(children, pop, elites, non_elites, rho):
    taken = total = 0
    for c in children:
        got = _parents_of(c, pop, elites, non_elites)
        assert got is not None, "an offspring gene comes from neither of one elite and one non-elite parent"
        _, _, from_e = got
        taken += int(from_e.sum())
        total += from_e.size
    share = taken / total
    # binomial bound: 6 standard deviations of the share of elite genes
    assert abs(share - rho) <= 6 * np.sqrt(rho * (1 - rho) / total), (share, rho, total)


def test_breed_kernel_semantics():
    """BRKGA_alg3.py:188-208 on the device: the elites are copied, the mutants are
    fresh keys in [0, 1) on the float32 grid, and every offspring gene comes from
    that row's one elite parent or its one non-elite parent, the elite's with
    probability rho (share within binomial bounds)."""
    _say("device GA breed semantics")
    from brkga_fft2_amd.native import Decoder, DeviceGA
    prob, n, M, packed = ref_instance("P25M2-4")
    N, ne, nm, rho = 120, 24, 18, 0.7
    dec = Decoder(packed, max_bins=n)
    init = np.random.default_rng(8).random((N, 2 * n)).astype(np.float32)
    ga = DeviceGA(dec, N, ne, nm, elite_prob=rho, seed=23, init_population=init)
    pop0, fit0 = ga.population()
    ga.run(1, use_graph=False)
    pop1, _ = ga.population()
    ga.close()
    dec.close()
    order = np.lexsort((np.arange(N), fit0))
    elites, non_elites = order[:ne], order[ne:]
    np.testing.assert_array_equal(pop1[:ne], pop0[elites])
    mut = pop1[ne:ne + nm]
    assert mut.min() >= 0.0 and mut.max() < 1.0
    assert np.all(mut * 16777216.0 == np.floor(mut * 16777216.0))
    assert (mut[:, None, :] == pop0[None, :, :]).sum(-1).max() <= 2   # fresh rows, not copies
    _check_offspring(pop1[ne + nm:], pop0, elites, non_elites, rho)


def test_breed_pool_semantics():
    """brkga_breed_pool (the screened generation's oversized pool): same rules
    as the generation's breeding, on a pool larger than the population."""
    _say("breed pool semantics")
    import ctypes
    import torch
    from brkga_fft2_amd.surrogate import _lib
    lib = _lib()
    N, G2, ne, pm, po, rho = 100, 60, 20, 40, 160, 0.6
    rng = np.random.default_rng(2)
    pop = rng.random((N, G2)).astype(np.float32)
    fit = rng.random(N)
    fit[7] = np.nan   # a NaN fitness sorts last
    d_pop = torch.from_numpy(pop).cuda()
    d_fit = torch.from_numpy(fit).cuda()
    d_ord = torch.empty(N, dtype=torch.int32, device="cuda")
    d_out = torch.empty((pm + po, G2), dtype=torch.float32, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.brkga_sort_fitness(ctypes.c_void_p(d_fit.data_ptr()), N, ctypes.c_void_p(d_ord.data_ptr()), s) == 0
    assert lib.brkga_breed_pool(ctypes.c_void_p(d_pop.data_ptr()), ctypes.c_void_p(d_ord.data_ptr()), N, G2, ne,
                                pm, po, ctypes.c_double(rho), ctypes.c_uint64(9), ctypes.c_uint32(3),
                                ctypes.c_void_p(d_out.data_ptr()), s) == 0
    torch.cuda.synchronize()
    order = d_ord.cpu().numpy()
    key = np.where(np.isnan(fit), np.inf, fit)
    np.testing.assert_array_equal(order, np.lexsort((np.arange(N), key)))
    assert order[-1] == 7
    out = d_out.cpu().numpy()
    mut = out[:pm]
    assert mut.min() >= 0.0 and mut.max() < 1.0
    assert (mut[:, None, :] == pop[None, :, :]).sum(-1).max() <= 2
    _check_offspring(out[pm:], pop, order[:ne], order[ne:], rho)
