"""BRKGA generation loop: drop-in for the reference's ``BRKGA`` class.

Host loop (`BRKGA.fit`) restates ``BRKGA_alg3.py:23-281`` with the same NumPy
call order, so a run seeded with ``np.random.seed`` draws the same random
numbers as the reference's:

* initial population ``np.random.uniform(0, 1, (N, 2n)).astype(float32)``,
  ``population[0] = initialSol`` (``:212-214``);
* ``partition``: ``np.argpartition(fitness, num_elites)`` then the elites
  sorted by fitness (``:178-186``);
* ``mating``: ``randint`` elite parents, ``randint`` non-elite parents,
  ``uniform`` crossover mask ``< eliteCProb`` (``:188-205``);
* ``mutants``: ``uniform(0, 1, (num_mutants, 2n))`` (``:207-208``);
* next population ``elites + mutants + offspring``, fitness of the elites
  reused (``:244-250``).

Fitness comes from the MI355X decoder through ``cal_fitness``.  Every
``eval_mode`` of the reference decodes on the GPU (``BRKGA_alg3.py:40-46,119-163``):

* ``native_full`` / ``wave_batch``: the whole population in one batch;
* ``serial`` / ``thread`` / ``process`` (and ``auto``, which the reference
  maps to ``serial`` for GPU collision backends and to ``thread`` otherwise):
  the reference's fitness cache keyed by the chromosome rounded to 4
  decimals -- a chromosome whose key was seen before takes the cached
  makespan -- and the misses decoded together in one batch instead of one
  per worker.  The cache changes which makespans a run sees, so it is kept
  for a run to follow the reference's trajectory in these modes.
"""
import json
import math
import os
import time

import numpy as np

from .evaluator import FullNativeDecoderEvaluator


class BRKGA:
    """Biased random-key GA over the population decoder."""

    def __init__(self, problem_data, nbParts, nbMachines, thresholds, instanceParts, initialSol,
                 collision_backend=None, eval_mode="auto", eval_workers=4, eval_chunksize=1,
                 num_generations=200, num_individuals=100, num_elites=12, num_mutants=18,
                 eliteCProb=0.7, evaluator=None):
        self.problem_data = problem_data
        self.nbMachines = nbMachines
        self.thresholds = thresholds
        self.N = nbParts
        self.instanceParts = instanceParts
        self.initialSol = initialSol
        self.collision_backend = collision_backend
        if eval_mode == "auto":   # BRKGA_alg3.py:40-46
            name = getattr(collision_backend, "name", None)
            gpu = name is None or "gpu" in name or "cuda" in name   # this decoder is a GPU one
            self.eval_mode = "serial" if gpu else "thread"
        else:
            self.eval_mode = eval_mode
        if self.eval_mode == "wave_batch" and os.getenv("ABRKGA_FULL_NATIVE_DECODER", "0").strip() not in \
                {"0", "false", "False"}:   # BRKGA_alg3.py:77-79
            self.eval_mode = "native_full"
        self.eval_workers = int(eval_workers) if eval_workers else 0
        self.eval_chunksize = int(eval_chunksize) if eval_chunksize else 1
        if self.eval_mode not in {"wave_batch", "native_full"}:   # BRKGA_alg3.py:96-101
            self._fitness_cache = {}
            self._cache_hits = 0
            self._cache_misses = 0
        self.num_generations = num_generations
        self.num_individuals = int(num_individuals)
        self.num_gene = 2 * self.N
        self.num_elites = int(num_elites)
        self.num_mutants = int(num_mutants)
        self.eliteCProb = eliteCProb
        self.used_bins = -1
        self.solution = None
        self.best_fitness = -1
        self.history = {"mean": [], "min": [], "time": []}
        self._native_decoder = evaluator or FullNativeDecoderEvaluator(
            problem_data, nbParts, nbMachines, thresholds, instanceParts, collision_backend)

    def _hash_solution(self, solution):
        """``BRKGA_alg3.py:114-116``: the chromosome rounded to 4 decimals."""
        return tuple(np.round(solution, 4))

    def evaluate_solution(self, solution):
        """Makespan of one chromosome (``BRKGA_alg3.py:103-112``)."""
        return self._native_decoder.evaluate_batch(np.asarray(solution)[None])[0]

    def cal_fitness(self, population):
        """``BRKGA_alg3.py:119-163``: a Python list of makespans."""
        if self.eval_mode in ("native_full", "wave_batch"):
            return self._native_decoder.evaluate_batch(np.array(population))
        if self.eval_mode not in ("serial", "thread", "process"):
            raise ValueError(f"Unsupported eval_mode: {self.eval_mode}")
        if self.eval_mode == "process" and "cuda" in getattr(self.collision_backend, "name", ""):
            raise ValueError("eval_mode=process is not supported with CUDA backends. Use thread or serial.")
        results = [None] * len(population)
        todo = []
        for i, sol in enumerate(population):
            key = self._hash_solution(sol)
            if key in self._fitness_cache:
                results[i] = self._fitness_cache[key]
                self._cache_hits += 1
            else:
                todo.append((i, key))
                self._cache_misses += 1
        if todo:
            batch = np.array([population[i] for i, _ in todo])
            new = self._native_decoder.evaluate_batch(batch)
            for (i, key), f in zip(todo, new):
                results[i] = f
                self._fitness_cache[key] = f
        return results

    def shutdown(self):
        """``BRKGA_alg3.py:164-175``: report the fitness cache's effectiveness."""
        if self.eval_mode not in {"wave_batch", "native_full"}:
            total = self._cache_hits + self._cache_misses
            if total > 0:
                hit_rate = 100.0 * self._cache_hits / total
                print(f"Fitness cache: {self._cache_hits}/{total} hits ({hit_rate:.1f}%), "
                      f"{len(self._fitness_cache)} unique solutions")

    def partition(self, population, fitness_list):
        fitness_arr = np.asarray(fitness_list)
        partition_indices = np.argpartition(fitness_arr, self.num_elites)
        elite_indices = partition_indices[:self.num_elites]
        non_elite_indices = partition_indices[self.num_elites:]
        elite_indices = elite_indices[np.argsort(fitness_arr[elite_indices])]
        return population[elite_indices], population[non_elite_indices], fitness_arr[elite_indices]

    def mating(self, elites, non_elites):
        num_offspring = self.num_individuals - self.num_elites - self.num_mutants
        elite_idx = np.random.randint(0, len(elites), size=num_offspring)
        non_elite_idx = np.random.randint(0, len(non_elites), size=num_offspring)
        elite_parents = elites[elite_idx]
        non_elite_parents = non_elites[non_elite_idx]
        mask = np.random.uniform(0.0, 1.0, size=(num_offspring, self.num_gene)) < self.eliteCProb
        return np.where(mask, elite_parents, non_elite_parents).astype(np.float32)

    def mutants(self):
        return np.random.uniform(low=0.0, high=1.0,
                                 size=(self.num_mutants, self.num_gene)).astype(np.float32)

    def fit(self, verbose=False, target=None):
        """Run the generations; stops early once ``target`` makespan is reached."""
        startfit = time.time()
        population = np.random.uniform(low=0.0, high=1.0,
                                       size=(self.num_individuals, self.num_gene)).astype(np.float32)
        if self.initialSol is not None:
            population[0] = self.initialSol
        fitness_list = self.cal_fitness(population)
        best_fitness = np.min(fitness_list)
        best_solution = population[np.argmin(fitness_list)]
        self.history["min"].append(np.min(fitness_list))
        self.history["mean"].append(np.mean(fitness_list))
        self.history["time"].append(time.time() - startfit)
        for g in range(self.num_generations):
            if target is not None and best_fitness <= target:
                break
            t0 = time.time()
            elites, non_elites, elite_fitness_list = self.partition(population, fitness_list)
            offsprings = self.mating(elites, non_elites)
            mutants = self.mutants()
            offspring = np.concatenate((mutants, offsprings), axis=0)
            offspring_fitness_list = self.cal_fitness(offspring)
            population = np.concatenate((elites, mutants, offsprings), axis=0)
            fitness_list = list(elite_fitness_list) + offspring_fitness_list
            min_idx = np.argmin(fitness_list)
            min_fitness = fitness_list[min_idx]
            if min_fitness < best_fitness:
                best_fitness = min_fitness
                best_solution = population[min_idx]
            self.history["min"].append(min_fitness)
            self.history["mean"].append(np.mean(fitness_list))
            self.history["time"].append(time.time() - startfit)
            if verbose:
                print("Generation %d: Best=%.4f, Time=%.2fs" % (g, best_fitness, time.time() - t0))
        self.used_bins = math.floor(best_fitness)
        self.best_fitness = best_fitness
        self.solution = best_solution
        return "feasible"


def initial_solution(problem_data, nbParts, nbMachines, thresholds, instanceParts, decoder=None,
                     seed=None, with_log=False):
    """Greedy initial solution of ``BRKGA_alg3.py:445-603`` on the GPU.

    Returns ``(array, makespan)`` like the reference's ``array`` (float64
    ``[2n]``, the ``initialSol`` argument of ``BRKGA``) and its printed
    "Makespan of initial solution"; with ``with_log`` also the placements.
    The reference draws each machine's key with ``random.uniform``; here a
    ``numpy`` generator seeded with ``seed`` draws it inside the same bucket
    (any such key decodes identically).  ``decoder``: an open ``Decoder`` of
    this problem, else one is created for the call.
    """
    from .native import Decoder
    from .problem import initial_machine_keys, pack_problem, type_heights
    packed = pack_problem(problem_data, nbParts, nbMachines, thresholds, instanceParts)
    own = decoder is None
    dec = Decoder(packed) if own else decoder
    try:
        mk = initial_machine_keys(thresholds, nbMachines, np.random.default_rng(seed))
        keys, ms, log = dec.initial_solution(type_heights(problem_data, packed), mk, with_log=with_log)
    finally:
        if own:
            dec.close()
    out = (keys.astype(np.float64), ms)
    return out + (log,) if with_log else out


def brkga_sizes(nb_parts, mult=10):
    """Population sizes of the reference's entry point (``BRKGA_alg3.py:608-618``)."""
    n = mult * nb_parts
    return dict(num_individuals=n, num_elites=math.ceil(n * 0.1),
                num_mutants=math.ceil(n * 0.15), eliteCProb=0.70)


def write_history(path, runs, **meta):
    """Write a run report as JSON (replaces the reference's Excel export of
    ``model.history``, one sheet per repetition, ``BRKGA_alg3.py:611-629``).

    ``runs``: one entry per repetition -- a fitted ``BRKGA`` (its ``history``
    columns ``mean``, ``min``, ``time`` and its best fitness and chromosome)
    or a plain ``{"mean": [...], "min": [...], "time": [...]}`` mapping.
    ``meta``: anything else to record (instance name, sizes, seed ...).
    The file holds ``{"meta": meta, "runs": [{"history": {...},
    "best_fitness": f, "best_chromosome": [...]}, ...]}``, values as plain
    JSON numbers (the 1e16 infeasibility penalty included).
    """
    out = []
    for r in runs:
        hist = r.history if hasattr(r, "history") else r
        rec = {"history": {k: [float(v) for v in hist[k]] for k in hist}}
        if hasattr(r, "best_fitness"):
            rec["best_fitness"] = float(r.best_fitness)
        sol = getattr(r, "solution", None)
        if sol is not None:
            rec["best_chromosome"] = [float(v) for v in np.asarray(sol).ravel()]
        out.append(rec)
    doc = {"meta": meta, "runs": out}
    tmp = str(path) + ".tmp"
    with open(tmp, "w") as f:
        json.dump(doc, f)
    os.replace(tmp, path)
    return doc


def read_history(path):
    """Load a report written by ``write_history``."""
    with open(path) as f:
        return json.load(f)
