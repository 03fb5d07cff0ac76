"""The drop-in evaluator on the GPU: the introspection surface the reference's
tools read (``_last_contexts``, ``_placement_log``, ``get_bin_stats``)."""
import json
import os
import sys

import numpy as np
import pytest

from brkga_fft2_amd.contexts import machine_makespan
from brkga_fft2_amd.problem import thresholds_for
from oracle import oracle
from tests.helpers import GOLDEN, oracle_threads, ref_instance, syn_instance, syn_keys

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _evaluator(prob, n, M):
    from brkga_fft2_amd.evaluator import WaveBatchEvaluator
    return WaveBatchEvaluator(prob, n, M, thresholds_for(M), prob.instance_parts, None, device="cuda")


def test_create_golden_on_last_contexts():
    """create_golden.py:152-190 against the drop-in: the single re-run with
    context collection gives the golden per-machine bins, makespans and
    placements of golden_P75M2-0.json."""
    _say("create_golden extraction P75M2-0")
    prob, n, M, packed = ref_instance("P75M2-0")
    with open(os.path.join(GOLDEN, "golden_P75M2-0.json")) as f:
        g = json.load(f)
    ev = _evaluator(prob, n, M)
    ev._collect_contexts = True
    ev._last_contexts = {}
    ev._placement_log = {}
    fit = ev.evaluate_batch(np.array([np.array(g["chromosome"], dtype=np.float32)]))
    ev._collect_contexts = False
    assert isinstance(fit, list) and round(fit[0], 6) == g["makespan"]
    for m in range(M):
        ctx = ev._last_contexts[m][0]
        active = [b for b in ctx.open_bins if b.area > 0]
        mk = sum(b.proc_time + b.proc_time_height + prob.machines[m].setup_time for b in active)
        assert round(mk, 6) == g["machines"][str(m)]["makespan"]
        assert len(active) == g["machines"][str(m)]["n_bins"]
        placements = [{"part_id": int(pid), "bin_idx": int(b), "col": int(c), "row": int(r), "rot": int(rot)}
                      for pid, b, c, r, rot, _ in ev._placement_log.get(m, [])]
        assert placements == g["machines"][str(m)]["placements"]
    ev.close()


def test_last_contexts_grids_chromosome_462():
    """inspect_chromosome.py:538-566 on the drop-in: every bin grid of P50M2-0
    chromosome 462 equals the reference's pp_bin_m*_b*.txt grid."""
    _say("inspection grids P50M2-0 #462")
    prob, n, M, packed = ref_instance("P50M2-0")
    np.random.seed(42)
    pop = np.random.uniform(0.0, 1.0, size=(500, 100)).astype(np.float32)
    ev = _evaluator(prob, n, M)
    ev._collect_contexts = True
    fit = ev.evaluate_batch(np.array([pop[462]]))
    z = np.load(os.path.join(GOLDEN, "golden_P50M2-0_c462.npz"))
    assert round(fit[0], 4) == float(z["makespan"])
    seen = set()
    for m in range(M):
        ctx = ev._last_contexts[m][0]
        assert machine_makespan(ctx, prob.machines[m].setup_time) <= fit[0]
        for b in ctx.open_bins:
            key = "m%d_b%d" % (m, b.bin_idx)
            seen.add(key)
            shp = tuple(z["shape_" + key])
            want = np.unpackbits(z["grid_" + key])[:shp[0] * shp[1]].reshape(shp)
            assert np.array_equal(b.grid, want), key
            assert b.parts_assigned == list(z["order_" + key]), key
    assert seen == {k[5:] for k in z.files if k.startswith("grid_")}
    ev.close()


def test_bin_stats_per_machine():
    """get_bin_stats (full_native_decoder.py:2286-2312): per machine the most
    bins one context opened, the bins opened and the contexts, counted against
    the oracle's placements; growth events when the workspace had to grow."""
    _say("bin stats")
    from brkga_fft2_amd.native import Decoder
    prob, n, M, packed = syn_instance("C2")
    keys = syn_keys("C2", 64, n, seed=41)
    ref_ms, ref_log, _ = oracle.decode(packed, keys, n_threads=oracle_threads(), with_log=True)
    assert np.all(ref_ms < 1e15)
    peak, total = [0] * M, [0] * M
    for c in range(len(keys)):
        for m in range(M):
            rows = ref_log[c][ref_log[c][:, 0] == m]
            nb = int(rows[:, 2].max()) + 1 if len(rows) else 0
            peak[m] = max(peak[m], nb)
            total[m] += nb
    for max_bins, grows in ((0, False), (1, True)):
        dec = Decoder(packed, max_bins=max_bins)
        ms, _ = dec.evaluate(keys)
        np.testing.assert_array_equal(ms, ref_ms)
        st = dec.bin_stats()
        assert len(st) == M
        for m in range(M):
            p, t, sols, batches, slots, growths = st[m]
            if not grows:
                assert (p, t, sols, batches) == (peak[m], total[m], len(keys), 1)
            assert slots >= peak[m]
            assert (growths > 0) == grows
        dec.reset_stats()
        assert all(s[:4] == (0, 0, 0, 0) and s[5] == 0 for s in dec.bin_stats())
        dec.close()
