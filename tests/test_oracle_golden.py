"""The CPU oracle against every golden vector the reference holds for this path."""
import json
import os

import numpy as np
import pytest

from oracle import oracle
from tests.helpers import (GOLDEN, ref_instance, load_fingerprints, fingerprint_population,
                           fingerprint_of, oracle_threads)


def test_p75m2_0_placement_log():
    """results/golden/P75M2-0: chromosome 688 -> per-machine (part, bin, col, row, rot) log."""
    prob, n, M, packed = ref_instance("P75M2-0")
    with open(os.path.join(GOLDEN, "golden_P75M2-0.json")) as f:
        g = json.load(f)
    chrom = np.array(g["chromosome"], dtype=np.float32)[None]
    ms, log, _ = oracle.decode(packed, chrom, with_log=True)
    assert round(float(ms[0]), 6) == g["makespan"]
    for m in range(M):
        want = [(p["part_id"], p["bin_idx"], p["col"], p["row"], p["rot"])
                for p in g["machines"][str(m)]["placements"]]
        got = [tuple(int(v) for v in e[1:]) for e in log[0] if e[0] == m]
        assert got == want, "machine %d" % m
        n_bins = max(e[2] for e in log[0] if e[0] == m) + 1
        assert n_bins == g["machines"][str(m)]["n_bins"]


def test_p50m2_0_chromosome_462_grids():
    """results/inspection_output/pp_bin_m*_b*.txt: every final bin grid of chromosome 462
    (seed-42 population, verify_native_vs_pp.py:31-32,96-100) and makespan 448166.8336."""
    prob, n, M, packed = ref_instance("P50M2-0")
    np.random.seed(42)
    pop = np.random.uniform(0.0, 1.0, size=(500, 100)).astype(np.float32)
    ms, log, _ = oracle.decode(packed, pop[462:463], with_log=True)
    z = np.load(os.path.join(GOLDEN, "golden_P50M2-0_c462.npz"))
    assert round(float(ms[0]), 4) == float(z["makespan"])
    seen = set()
    for m in range(M):
        H, W = prob.machines[m].bin_length, prob.machines[m].bin_width
        grids, orders = {}, {}
        for e in log[0]:
            if e[0] != m:
                continue
            _, pid, b, col, row, rot = (int(v) for v in e)
            mask = prob.parts[pid].rotations[rot]
            h, w = mask.shape
            g = grids.setdefault(b, np.zeros((H, W), np.int64))
            g[row - h + 1:row + 1, col:col + w] += mask
            orders.setdefault(b, []).append(pid)
        for b, g in grids.items():
            key = "m%d_b%d" % (m, b)
            seen.add(key)
            shp = tuple(z["shape_" + key])
            want = np.unpackbits(z["grid_" + key])[:shp[0] * shp[1]].reshape(shp)
            assert g.max() == 1
            assert np.array_equal(g, want), key
            assert orders[b] == list(z["order_" + key]), key
    assert seen == {k[5:] for k in z.files if k.startswith("grid_")}


@pytest.mark.parametrize("name", sorted(load_fingerprints()))
def test_makespan_fingerprints(name):
    """First five feasible makespans + feasible count of the default_rng(127) population
    (benchmarks/20260422_..._baseline.json; PARALLEL_MACHINES_PLAN.md:468-473)."""
    fp = load_fingerprints()[name]
    prob, n, M, packed = ref_instance(name)
    pop = fingerprint_population(fp, n)
    ms, _, _ = oracle.decode(packed, pop, n_threads=oracle_threads())
    got, n_feas = fingerprint_of(ms)
    assert n_feas == fp["n_feasible"]
    assert got == fp["fingerprint"]
