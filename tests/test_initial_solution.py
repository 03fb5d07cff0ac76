"""Greedy initial solution (BRKGA_alg3.py:445-603): the oracle against the
reference's own log, and the chromosome's structure."""
import numpy as np
import pytest

from brkga_fft2_amd.problem import (initial_machine_keys, initial_solution_order, thresholds_for,
                                    type_heights)
from oracle import mlp_oracle, oracle
from tests.helpers import ref_instance, syn_instance


def _greedy(prob, M, packed, seed=0):
    order = initial_solution_order(packed, type_heights(prob, packed))
    mk = initial_machine_keys(thresholds_for(M), M, np.random.default_rng(seed))
    return oracle.initial_solution(packed, order, mk, with_log=True) + (mk,)


def test_oracle_reproduces_the_reference_log():
    """wave_batch_results.txt:6-10 (P50M2-0, population 500 of 100 keys): the
    greedy prints "Makespan of initial solution: 246042.18342431998" and the
    initial population's best fitness -- the decoded seed chromosome -- is
    246042.18342432.  Both values are exact doubles here."""
    prob, n, M, packed = ref_instance("P50M2-0")
    keys, ms, log, _ = _greedy(prob, M, packed)
    assert ms == 246042.18342431998
    dec, _, _ = oracle.decode(packed, keys[None, :].astype(np.float64))
    assert dec[0] == 246042.18342432


def test_order_is_height_then_area_descending_and_stable():
    packed = {"slot_type": np.array([0, 1, 2, 1, 3]), "area": np.array([5.0, 7.0, 7.0, 9.0])}
    height = np.array([10.0, 20.0, 20.0, 10.0])
    # height 20 and area 7 for slots 1, 2, 3 (a tie: instance order), then slot 4
    # (height 10, area 9), then slot 0 (height 10, area 5)
    assert initial_solution_order(packed, height).tolist() == [1, 2, 3, 4, 0]


@pytest.mark.parametrize("name", ["P25M2-1", "P100M4-2"])
def test_chromosome_structure(name):
    """Every part is placed; machine keys decode to the greedy's machine; the
    sequence keys of a machine are i/k in batch order; decoding the seed
    chromosome reproduces the greedy's batches on every machine."""
    prob, n, M, packed = ref_instance(name)
    keys, ms, log, mk = _greedy(prob, M, packed, seed=3)
    thr = np.asarray(packed["thresholds"])
    mach = mlp_oracle.machine_bucket(keys[n:].astype(np.float64), thr)
    assert np.all(mlp_oracle.machine_bucket(mk, thr) == np.arange(M))
    assert np.all(keys[n:] == mk.astype(np.float32)[mach])
    for m in range(M):
        sv = np.sort(keys[:n][mach == m])
        k = sv.size
        assert k == np.sum(log[:, 0] == m)
        np.testing.assert_array_equal(sv, (np.arange(k) * (1.0 / k)).astype(np.float32))
    assert np.all(log[:, 0] >= 0)
    dec, dlog, _ = oracle.decode(packed, keys[None, :].astype(np.float64), with_log=True)
    assert dec[0] < 1e15
    # the decode places every part (first fit in batch order); same part count per machine
    assert sorted(dlog[0][:, 0].tolist()) == sorted(log[:, 0].tolist())


def test_synthetic_instances_run():
    for cfg in ("C1", "C2"):
        prob, n, M, packed = syn_instance(cfg)
        keys, ms, log, _ = _greedy(prob, M, packed)
        assert np.isfinite(ms) and keys.shape == (2 * n,)
