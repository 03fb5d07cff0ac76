"""BRKGA generation loops on the GPU: the host drop-in and the device graph loop."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle
from tests.helpers import syn_instance, ref_instance

pytestmark = pytest.mark.gpu


class _OracleEvaluator:
    """Same evaluate_batch contract, decoded by the CPU oracle (test checker)."""

    def __init__(self, packed):
        self.packed = packed

    def evaluate_batch(self, chromosomes):
        ms, _, _ = oracle.decode(self.packed, np.asarray(chromosomes), n_threads=os.cpu_count() or 4)
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
    want, _, _ = oracle.decode(packed, p1, n_threads=os.cpu_count() or 4)
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
