"""Every eval_mode of the reference's BRKGA (BRKGA_alg3.py:40-46,119-163) runs
on the one decoder: the batch modes decode whole populations, the cached modes
keep the reference's fitness cache (chromosomes rounded to 4 decimals)."""
import numpy as np
import pytest

from brkga_fft2_amd.brkga import BRKGA
from brkga_fft2_amd.problem import thresholds_for
from oracle import oracle
from tests.helpers import oracle_threads, syn_instance


class _Evaluator:
    """evaluate_batch contract, decoded by the oracle; counts what it decodes."""

    def __init__(self, packed):
        self.packed = packed
        self.decoded = 0

    def evaluate_batch(self, chromosomes):
        self.decoded += len(chromosomes)
        ms, _, _ = oracle.decode(self.packed, np.asarray(chromosomes), n_threads=oracle_threads())
        return ms.tolist()


class _Backend:
    def __init__(self, name):
        self.name = name


def _run(mode, backend=None, gens=4):
    prob, n, M, packed = syn_instance("C1", seed=5)
    ev = _Evaluator(packed)
    np.random.seed(3)
    g = BRKGA(prob, n, M, thresholds_for(M), prob.instance_parts, None, collision_backend=backend,
              eval_mode=mode, num_generations=gens, num_individuals=32, num_elites=7, num_mutants=5,
              evaluator=ev)
    g.fit()
    return g, ev


@pytest.mark.parametrize("mode", ["serial", "thread", "process", "wave_batch"])
def test_modes_follow_one_trajectory(mode):
    ref, _ = _run("native_full")
    got, _ = _run(mode)
    assert got.history["min"] == ref.history["min"] and got.history["mean"] == ref.history["mean"]
    assert np.array_equal(got.solution, ref.solution)


def test_auto_mode_and_cache():
    g, ev = _run("auto", backend=_Backend("torch_gpu"))
    assert g.eval_mode == "serial"
    assert _run("auto", backend=_Backend("numpy_cpu"))[0].eval_mode == "thread"
    assert g._cache_misses == ev.decoded and g._cache_hits == 0
    # a chromosome equal to a cached one after rounding to 4 decimals takes the cached makespan
    pop = np.random.default_rng(0).random((4, 2 * g.N)).astype(np.float32)
    first = g.cal_fitness(pop)
    near = pop.copy()
    near[1] += 1e-6
    d0 = ev.decoded
    again = g.cal_fitness(near)
    assert again[1] == first[1] and g._cache_hits >= 1
    assert ev.decoded - d0 == sum(1 for i in range(4)
                                  if g._hash_solution(near[i]) != g._hash_solution(pop[i]))


def test_unknown_and_unsupported_modes():
    prob, n, M, packed = syn_instance("C1", seed=5)
    g = BRKGA(prob, n, M, thresholds_for(M), prob.instance_parts, None, eval_mode="bogus",
              evaluator=_Evaluator(packed))
    with pytest.raises(ValueError, match="Unsupported eval_mode"):
        g.cal_fitness(np.zeros((2, 2 * n), np.float32))
    g = BRKGA(prob, n, M, thresholds_for(M), prob.instance_parts, None, collision_backend=_Backend("torch_cuda"),
              eval_mode="process", evaluator=_Evaluator(packed))
    with pytest.raises(ValueError, match="process"):
        g.cal_fitness(np.zeros((2, 2 * n), np.float32))
