"""Host-side rebuild of the reference's per-context bin state (``_last_contexts``)
from a placement log, checked with the oracle's logs against the reference's
golden files and against what create_golden.py / inspect_chromosome.py extract."""
import json
import os

import numpy as np

from brkga_fft2_amd.contexts import build_contexts, free_runs, machine_makespan
from brkga_fft2_amd.problem import thresholds_for
from oracle import oracle
from tests.helpers import GOLDEN, oracle_threads, ref_instance


def _contexts(prob, n, M, packed, keys):
    ms, log, _ = oracle.decode(packed, keys, with_log=True, n_threads=oracle_threads())
    return ms, build_contexts(prob, n, M, thresholds_for(M), prob.instance_parts, keys, log)


def test_free_runs_closed_form():
    g = np.array([[0, 1, 1, 0, 0, 0, 1],
                  [0, 0, 0, 0, 0, 0, 0],
                  [1, 1, 1, 1, 1, 1, 1],
                  [1, 0, 1, 0, 0, 1, 0]], np.uint8)
    assert free_runs(g).tolist() == [3, 7, 0, 2]


def test_create_golden_extraction_p75m2_0():
    """create_golden.py:166-190 run on the rebuilt contexts reproduces
    golden_P75M2-0.json: per machine the bins with area > 0, their count, and
    the machine makespan summed over them."""
    prob, n, M, packed = ref_instance("P75M2-0")
    with open(os.path.join(GOLDEN, "golden_P75M2-0.json")) as f:
        g = json.load(f)
    keys = np.array(g["chromosome"], dtype=np.float32)[None]
    ms, ctxs = _contexts(prob, n, M, packed, keys)
    worst = 0.0
    for m in range(M):
        ctx = ctxs[m][0]
        assert ctx.solution_idx == 0 and ctx.machine_idx == m and ctx.is_feasible and ctx.is_done
        active = [b for b in ctx.open_bins if b.area > 0]
        mk = machine_makespan(ctx, prob.machines[m].setup_time)
        assert round(mk, 6) == g["machines"][str(m)]["makespan"]
        assert len(active) == g["machines"][str(m)]["n_bins"]
        order = [p["part_id"] for p in g["machines"][str(m)]["placements"]]
        assert [int(p) for p in ctx.parts_sequence] == order
        assert [pid for b in ctx.open_bins for pid in b.parts_assigned] == \
            [p["part_id"] for p in sorted(g["machines"][str(m)]["placements"], key=lambda p: p["bin_idx"])]
        worst = max(worst, mk)
    assert worst == ms[0]   # same float64 sums, same order


def test_inspection_grids_p50m2_0_chromosome_462():
    """inspect_chromosome.py:538-566 writes each bin's grid; the rebuilt grids are
    the reference's pp_bin_m*_b*.txt grids of chromosome 462, with the parts
    in order, free runs and enclosure bounds consistent with the grid."""
    prob, n, M, packed = ref_instance("P50M2-0")
    np.random.seed(42)
    pop = np.random.uniform(0.0, 1.0, size=(500, 100)).astype(np.float32)
    ms, ctxs = _contexts(prob, n, M, packed, pop[462:463])
    z = np.load(os.path.join(GOLDEN, "golden_P50M2-0_c462.npz"))
    seen = set()
    for m in range(M):
        for b in ctxs[m][0].open_bins:
            key = "m%d_b%d" % (m, b.bin_idx)
            seen.add(key)
            shp = tuple(z["shape_" + key])
            want = np.unpackbits(z["grid_" + key])[:shp[0] * shp[1]].reshape(shp)
            assert np.array_equal(b.grid, want), key
            assert b.parts_assigned == list(z["order_" + key]), key
            rows = np.flatnonzero(b.grid.any(axis=1))
            assert b.min_occupied_row == rows[0] and b.max_occupied_row == rows[-1]
            assert b.enclosure_box_length == b.bin_length - rows[0]
            assert np.array_equal(b.vacancy_vector, free_runs(b.grid))
            area = 0.0
            for pid in b.parts_assigned:
                area += prob.parts[pid].area
            assert b.area == area
    assert seen == {k[5:] for k in z.files if k.startswith("grid_")}
    assert round(float(ms[0]), 4) == float(z["makespan"])


def test_infeasible_context_has_no_bins():
    """placement.py:46-53: the context of a machine that receives a part fitting
    it in neither orientation stops there, with no bins."""
    prob, n, M, packed = ref_instance("P50M2-0")
    np.random.seed(42)
    pop = np.random.uniform(0.0, 1.0, size=(500, 100)).astype(np.float32)
    ms, log, _ = oracle.decode(packed, pop[:40], with_log=True, n_threads=oracle_threads())
    bad = np.flatnonzero(ms >= 1e15)
    assert bad.size > 0
    ctxs = build_contexts(prob, n, M, thresholds_for(M), prob.instance_parts, pop[:40], log)
    for c in bad:
        infeasible = [ctxs[m][c] for m in range(M) if not ctxs[m][c].is_feasible]
        assert infeasible
        for ctx in infeasible:
            assert ctx.open_bins == []
            pid = int(ctx.parts_sequence[ctx.current_part_idx])
            H, W = ctx.bin_length, ctx.bin_width
            s0 = prob.parts[pid].shapes[0]
            assert (s0[0] > H or s0[1] > W) and (s0[1] > H or s0[0] > W)
    for c in np.flatnonzero(ms < 1e15):
        assert all(ctxs[m][c].is_feasible and ctxs[m][c].is_done for m in range(M))
        assert max(machine_makespan(ctxs[m][c], prob.machines[m].setup_time) for m in range(M)) == ms[c]
