"""Shared test helpers: reference instances, populations, log comparison."""
import json
import os

import numpy as np

from brkga_fft2_amd.problem import (SYN_HT, SYN_ST, SYN_VT, SYNTHETIC_CONFIGS, build_problem,
                                    load_reference_instance, pack_problem, random_polygon_mask,
                                    synthetic_population, synthetic_problem, thresholds_for)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def oracle_threads():
    """OpenMP threads for the oracle: this process's CPU share, at most 16 (a GPU
    box's share for one GPU, whatever the machine's core count)."""
    return max(1, min(16, len(os.sched_getaffinity(0))))


def ref_instance(name):
    prob, n, m = load_reference_instance(GOLDEN, name)
    packed = pack_problem(prob, n, m, thresholds_for(m), prob.instance_parts)
    return prob, n, m, packed


def syn_instance(cfg, seed=1):
    prob, n, m = synthetic_problem(cfg, seed=seed)
    packed = pack_problem(prob, n, m, thresholds_for(m), prob.instance_parts)
    return prob, n, m, packed


def syn_keys(cfg, n_chrom, n_parts, seed=7):
    dt = np.float64 if SYNTHETIC_CONFIGS[cfg]["key_dtype"] == "float64" else np.float32
    return synthetic_population(n_chrom, n_parts, seed=seed, dtype=dt)


def fingerprint_population(fp, n_parts):
    """Population of benchmark_harness.py:180-187 (default_rng(seed).random -> f32)."""
    rng = np.random.default_rng(fp["seed"])
    return rng.random((fp["num_individuals"], 2 * n_parts), dtype=np.float64).astype(np.float32)


def load_fingerprints():
    with open(os.path.join(GOLDEN, "fingerprints.json")) as f:
        return json.load(f)


def fingerprint_of(makespans):
    feas = [float(m) for m in makespans if m < 1e15]
    return ["%.6f" % m for m in feas[:5]], len(feas)


def feasible_logs_equal(ms, log_a, log_b):
    """Compare placement logs of feasible chromosomes; returns first mismatch or None."""
    for c in range(len(ms)):
        if ms[c] >= 1e15:
            continue
        if not np.array_equal(log_a[c], log_b[c]):
            rows = np.flatnonzero((log_a[c] != log_b[c]).any(axis=1))
            return c, int(rows[0]), log_a[c][rows[0]].tolist(), log_b[c][rows[0]].tolist()
    return None


def custom_instance(n, plates, bbox, nrot=4, seed=0, height=(5.0, 80.0)):
    """Seeded synthetic instance on arbitrary plates: ``plates`` lists (rows,
    cols) per machine; part i draws its polygon mask (bounding box in
    ``bbox`` cells per side) and height from ``default_rng(seed * 100003 + i)``
    as the benchmark configurations do."""
    masks, areas, vol, sup, hgt = {}, {}, {}, {}, {}
    for i in range(n):
        rng = np.random.default_rng(seed * 100003 + i)
        m = random_polygon_mask(rng, *bbox)
        masks[i] = m
        areas[i] = float(m.sum())
        hgt[i] = float(rng.uniform(*height))
        vol[i] = areas[i] * hgt[i]
        sup[i] = 0.0
    machines = [dict(L=r, W=c, ST=SYN_ST, VT=SYN_VT, SPT=0.0, HT=SYN_HT) for r, c in plates]
    prob = build_problem(np.arange(n), masks, areas, vol, sup, hgt, machines,
                         nrot={i: nrot for i in range(n)})
    M = len(plates)
    packed = pack_problem(prob, n, M, thresholds_for(M), prob.instance_parts)
    return prob, n, M, packed
