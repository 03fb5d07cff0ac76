"""Run report as JSON (the reference exports model.history to Excel,
BRKGA_alg3.py:611-629): a host BRKGA run driven by the CPU oracle evaluator
is written and read back unchanged."""
import os

import numpy as np

from brkga_fft2_amd.brkga import BRKGA, read_history, write_history
from brkga_fft2_amd.problem import thresholds_for
from oracle import oracle
from tests.helpers import oracle_threads, syn_instance


class _OracleEvaluator:
    def __init__(self, packed):
        self.packed = packed

    def evaluate_batch(self, chromosomes):
        ms, _, _ = oracle.decode(self.packed, np.asarray(chromosomes), n_threads=oracle_threads())
        return ms.tolist()


def test_history_round_trip(tmp_path):
    prob, n, M, packed = syn_instance("C1", seed=5)
    np.random.seed(3)
    model = BRKGA(prob, n, M, thresholds_for(M), prob.instance_parts, None,
                  evaluator=_OracleEvaluator(packed), num_generations=3, num_individuals=32,
                  num_elites=7, num_mutants=5, eliteCProb=0.7)
    model.fit()
    path = tmp_path / "run.json"
    doc = write_history(path, [model, {"mean": [1.0], "min": [0.5], "time": [0.1]}],
                        instance="C1", seed=3)
    back = read_history(path)
    assert back == doc
    run = back["runs"][0]
    assert run["history"]["min"] == [float(v) for v in model.history["min"]]
    assert len(run["history"]["mean"]) == 4 and run["best_fitness"] == float(model.best_fitness)
    assert np.array_equal(np.asarray(run["best_chromosome"], np.float32), model.solution)
    assert back["meta"] == {"instance": "C1", "seed": 3}
    assert back["runs"][1] == {"history": {"mean": [1.0], "min": [0.5], "time": [0.1]}}
