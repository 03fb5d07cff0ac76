"""Greedy initial solution on the GPU (brkga_initial_solution) against the
oracle: the same seed chromosome bit for bit, the same makespan and the same
placements, on reference and synthetic instances."""
import sys

import numpy as np
import pytest

from brkga_fft2_amd.problem import (initial_machine_keys, initial_solution_order, thresholds_for,
                                    type_heights)
from oracle import oracle
from tests.helpers import ref_instance, syn_instance

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [("ref", "P25M2-0"), ("ref", "P50M2-0"), ("ref", "P75M2-3"), ("ref", "P100M4-0"),
         ("ref", "P150M4-1"), ("ref", "P200M4-0"), ("syn", "C2"), ("syn", "C3")]


@pytest.mark.parametrize("kind,name", CASES)
def test_gpu_greedy_matches_oracle(kind, name):
    _say("initial solution %s" % name)
    from brkga_fft2_amd.native import Decoder
    prob, n, M, packed = ref_instance(name) if kind == "ref" else syn_instance(name)
    th = type_heights(prob, packed)
    mk = initial_machine_keys(thresholds_for(M), M, np.random.default_rng(1))
    want_keys, want_ms, want_log = oracle.initial_solution(packed, initial_solution_order(packed, th), mk,
                                                           with_log=True)
    dec = Decoder(packed)
    keys, ms, log = dec.initial_solution(th, mk, with_log=True)
    np.testing.assert_array_equal(log, want_log)
    np.testing.assert_array_equal(keys, want_keys)
    assert ms == want_ms
    if name == "P50M2-0":
        assert ms == 246042.18342431998   # wave_batch_results.txt:6
    got, _ = dec.evaluate(keys[None, :])
    ref, _, _ = oracle.decode(packed, keys[None, :].astype(np.float64))
    np.testing.assert_array_equal(got, ref)
    dec.close()


def test_drop_in_helper_seeds_the_host_loop():
    """brkga.initial_solution returns the reference's (array, makespan); as
    BRKGA's initialSol it is population[0] and its decode is the first fitness."""
    _say("initial solution drop-in")
    from brkga_fft2_amd.brkga import BRKGA, initial_solution
    prob, n, M, packed = ref_instance("P50M2-0")
    arr, ms = initial_solution(prob, n, M, thresholds_for(M), prob.instance_parts, seed=5)
    assert arr.dtype == np.float64 and arr.shape == (2 * n,) and ms == 246042.18342431998
    np.random.seed(2)
    model = BRKGA(prob, n, M, thresholds_for(M), prob.instance_parts, arr, num_generations=1,
                  num_individuals=40, num_elites=4, num_mutants=6, eliteCProb=0.7)
    model.fit()
    assert model.history["min"][0] <= 246042.18342432
