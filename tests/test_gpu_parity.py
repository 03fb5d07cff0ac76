"""GPU decoder vs the CPU oracle: bit-exact makespans and placement logs."""
import json
import os
import sys

import numpy as np
import pytest

from brkga_fft2_amd.problem import synthetic_population
from oracle import oracle
from tests.helpers import (GOLDEN, custom_instance, feasible_logs_equal, fingerprint_of,
                           fingerprint_population, load_fingerprints, oracle_threads, ref_instance,
                           syn_instance, syn_keys)

pytestmark = pytest.mark.gpu


def _decoder(packed, **kw):
    from brkga_fft2_amd.native import Decoder
    return Decoder(packed, device=0, **kw)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("cfg,n_chrom", [("C1", 64), ("C2", 48), ("C3", 16), ("C5", 8)])
def test_synthetic_parity(cfg, n_chrom):
    _say("synthetic parity %s N=%d" % (cfg, n_chrom))
    prob, n, M, packed = syn_instance(cfg)
    keys = syn_keys(cfg, n_chrom, n)
    ref_ms, ref_log, _ = oracle.decode(packed, keys, n_threads=oracle_threads(), with_log=True)
    dec = _decoder(packed)
    ms, log = dec.evaluate(keys, with_log=True)
    np.testing.assert_array_equal(ms, ref_ms)
    assert feasible_logs_equal(ref_ms, log, ref_log) is None
    st = dec.stats()
    assert st["chromosomes"] == n_chrom and st["overflows"] == 0
    dec.close()


@pytest.mark.parametrize("cfg,n_chrom", [("C3", 1024), ("C5", 512)])
def test_full_population_parity(cfg, n_chrom):
    """Populations of the configurations' real sizes (C3: 4096 contexts on 512
    persistent workgroups, C5: 2048 contexts on 256), so the work queue hands
    out later contexts in the heaviest-first order: every makespan and every
    placement against the oracle, element-wise
    (verify_full_population_vs_pp.py:88-194 does the same for P50M2-0)."""
    _say("full population parity %s N=%d" % (cfg, n_chrom))
    prob, n, M, packed = syn_instance(cfg)
    keys = syn_keys(cfg, n_chrom, n, seed=23)
    ref_ms, ref_log, _ = oracle.decode(packed, keys, n_threads=oracle_threads(), with_log=True)
    dec = _decoder(packed)
    ms, log = dec.evaluate(keys, with_log=True)
    np.testing.assert_array_equal(ms, ref_ms)
    assert feasible_logs_equal(ref_ms, log, ref_log) is None
    st = dec.stats()
    assert st["contexts"] == n_chrom * M and st["overflows"] == 0
    dec.close()


@pytest.mark.parametrize("plates", [((600, 600), (520, 700)), ((1100, 640),)])
def test_tall_plates_vs_oracle(plates):
    """Plates with more than 512 valid top rows: the candidate rows of a
    rotation span several 512-row blocks of the vacancy filter, and the
    bottom-most ones (the preferred placements) must be found
    (collision_backend.py:115-128 searches the whole valid region)."""
    _say("tall plates %s" % (plates,))
    prob, n, M, packed = custom_instance(24, plates, (30, 160), seed=4)
    keys = synthetic_population(12, n, seed=31)
    ref_ms, ref_log, _ = oracle.decode(packed, keys, n_threads=oracle_threads(), with_log=True)
    assert np.sum(ref_ms < 1e15) >= 6
    # placements with top rows >= 512 exist, so the lower 512-row blocks are exercised
    tindex = {int(p): t for t, p in enumerate(packed["type_id"])}
    e = ref_log[ref_ms < 1e15].reshape(-1, 6)
    h = np.array([packed["rot_h"][packed["rot_base"][tindex[int(r[1])]] + r[5]] for r in e])
    assert np.sum(e[:, 4] - h + 1 >= 512) >= 10
    for dense, mfma in ((False, "fp4"), (True, "fp4"), (False, "int8")):
        dec = _decoder(packed, dense=dense, mfma=mfma)
        ms, log = dec.evaluate(keys, with_log=True)
        np.testing.assert_array_equal(ms, ref_ms)
        assert feasible_logs_equal(ref_ms, log, ref_log) is None
        dec.close()


@pytest.mark.parametrize("cfg,n_chrom", [("C1", 32), ("C2", 48), ("C3", 24), ("C5", 16)])
def test_int8_operands_agree(cfg, n_chrom):
    """The int8 MFMA operand format (brkga_options.mfma_format = BRKGA_MFMA_INT8),
    pruned and at full shape: the same placements and makespans as the oracle."""
    _say("int8 operands %s N=%d" % (cfg, n_chrom))
    prob, n, M, packed = syn_instance(cfg)
    keys = syn_keys(cfg, n_chrom, n, seed=17)
    ref_ms, ref_log, _ = oracle.decode(packed, keys, n_threads=oracle_threads(), with_log=True)
    for dense in (False, True):
        dec = _decoder(packed, mfma="int8", dense=dense)
        ms, log = dec.evaluate(keys, with_log=True)
        np.testing.assert_array_equal(ms, ref_ms)
        assert feasible_logs_equal(ref_ms, log, ref_log) is None
        st = dec.stats()
        assert st["mfma_ops"] % (2 * 32 * 32 * 32) == 0
        dec.close()


@pytest.mark.parametrize("dense,tiles", [(True, 8), (True, 1), (False, 2)])
def test_dense_mode_and_tile_shapes_agree(dense, tiles):
    """Full-shape correlation (no row pruning, no early exit) and every pass
    size give identical results."""
    _say("dense=%s tiles=%d" % (dense, tiles))
    prob, n, M, packed = syn_instance("C2")
    keys = syn_keys("C2", 16, n, seed=11)
    ref_ms, ref_log, _ = oracle.decode(packed, keys, with_log=True)
    dec = _decoder(packed, dense=dense, tiles_per_pass=tiles)
    t, w = dec.get_tiles()
    assert t.max() <= tiles and w.max() == 64
    ms, log = dec.evaluate(keys, with_log=True)
    np.testing.assert_array_equal(ms, ref_ms)
    assert feasible_logs_equal(ref_ms, log, ref_log) is None
    st = dec.stats()
    assert st["mfma_ops"] > 0 and st["algo_ops"] > 0
    dec.close()


@pytest.mark.parametrize("cfg", ["C2", "C3"])
def test_random_tile_tables_agree(cfg):
    """Random per-(machine, width class) pass and window sizes: same placements."""
    _say("random tile tables %s" % cfg)
    from brkga_fft2_amd.native import TILE_CLASSES
    prob, n, M, packed = syn_instance(cfg)
    keys = syn_keys(cfg, 12, n, seed=5)
    ref_ms, ref_log, _ = oracle.decode(packed, keys, n_threads=oracle_threads(), with_log=True)
    dec = _decoder(packed)
    rng = np.random.default_rng(3)
    for trial in range(3):
        t = rng.choice([0, 1, 2, 4, 8], size=(M, TILE_CLASSES))
        w = rng.choice([0, 1, 3, 4, 7, 16, 64], size=(M, TILE_CLASSES))
        dec.set_tiles(t, w)
        dec.set_screen(rng.choice([0, 2, 6, 9], size=(M, TILE_CLASSES)))
        ms, log = dec.evaluate(keys, with_log=True)
        np.testing.assert_array_equal(ms, ref_ms)
        assert feasible_logs_equal(ref_ms, log, ref_log) is None
    dec.close()


@pytest.mark.parametrize("cfg,n_chrom,mfma", [("C2", 24, "fp4"), ("C3", 12, "fp4"), ("C5", 6, "fp4"),
                                               ("C5", 4, "int8")])
def test_screen_settings_agree(cfg, n_chrom, mfma):
    """The screening pass only skips tiles that cannot hold a placement: no
    screening, every row count up to 16, and random per-class tables all give
    the oracle's placements, in both operand formats."""
    _say("screen settings %s N=%d %s" % (cfg, n_chrom, mfma))
    from brkga_fft2_amd.native import TILE_CLASSES
    prob, n, M, packed = syn_instance(cfg)
    keys = syn_keys(cfg, n_chrom, n, seed=17)
    ref_ms, ref_log, _ = oracle.decode(packed, keys, n_threads=oracle_threads(), with_log=True)
    mfma_counts = {}
    for rows in (-1, 1, 2, 6, 16):
        dec = _decoder(packed, mfma=mfma, screen_rows=rows)
        want = 0 if rows < 0 else rows
        assert (dec.get_screen() == want).all()
        ms, log = dec.evaluate(keys, with_log=True)
        np.testing.assert_array_equal(ms, ref_ms)
        assert feasible_logs_equal(ref_ms, log, ref_log) is None
        mfma_counts[rows] = dec.stats()["mfma_ops"]
        if rows == 6:
            rng = np.random.default_rng(9)
            for trial in range(2):
                dec.set_screen(rng.integers(0, 17, size=(M, TILE_CLASSES)))
                ms, log = dec.evaluate(keys, with_log=True)
                np.testing.assert_array_equal(ms, ref_ms)
                assert feasible_logs_equal(ref_ms, log, ref_log) is None
        dec.close()
    # the screen removes MFMA work on the large masks of C5
    if cfg == "C5":
        assert mfma_counts[6] < 0.5 * mfma_counts[-1]


@pytest.mark.parametrize("cfg,n_chrom", [("C2", 24), ("C3", 600), ("C5", 6)])
def test_occupancy_settings_agree(cfg, n_chrom):
    """The decode kernels compiled for two and for three workgroups per CU
    (their own register budgets, pass sizes and LDS layouts), and one
    workgroup per CU, give the oracle's placements; C3 at 600 chromosomes
    (2400 contexts) is large enough for the automatic choice to take three."""
    _say("occupancy settings %s N=%d" % (cfg, n_chrom))
    prob, n, M, packed = syn_instance(cfg)
    keys = syn_keys(cfg, n_chrom, n, seed=29)
    ref_ms, ref_log, _ = oracle.decode(packed, keys, n_threads=oracle_threads(), with_log=True)
    grids = {}
    for wpc in (0, 1, 2, 3):
        dec = _decoder(packed, wgs_per_cu=wpc)
        ms, log = dec.evaluate(keys, with_log=True)
        np.testing.assert_array_equal(ms, ref_ms)
        assert feasible_logs_equal(ref_ms, log, ref_log) is None
        grids[wpc] = dec.stats()["n_workgroups"]
        dec.close()
    import torch
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    n_ctx = n_chrom * M
    for k in (1, 2, 3):
        assert grids[k] == min(n_ctx, k * n_cu), (k, grids)
    assert grids[0] == (grids[3] if n_ctx >= 2 * 3 * n_cu else grids[2]), grids
    if cfg == "C3":
        assert grids[0] == 3 * n_cu


def test_autotune_keeps_results():
    _say("autotune C2")
    import torch
    from brkga_fft2_amd import autotune
    prob, n, M, packed = syn_instance("C2")
    keys = syn_keys("C2", 64, n, seed=8)
    ref_ms, _, _ = oracle.decode(packed, keys, n_threads=oracle_threads())
    dec = _decoder(packed)
    dk = torch.from_numpy(keys).cuda()
    rec = autotune.autotune(dec, packed, dk, reps=1, windows=(8, 64), cache=None)
    assert rec["trials"] > 0 and rec["tuned_ms"] > 0
    ms, _ = dec.evaluate(keys)
    np.testing.assert_array_equal(ms, ref_ms)
    dec.close()


def test_autotune_recovers_the_screen():
    """From a table with screening off, the plate-wide round turns it back on:
    on C5's large masks that is the difference between correlating every tile
    in full and skipping most of them."""
    _say("autotune C5 from screen off")
    import torch
    from brkga_fft2_amd import autotune
    prob, n, M, packed = syn_instance("C5")
    keys = syn_keys("C5", 256, n, seed=4)
    dec = _decoder(packed, screen_rows=-1)
    assert (dec.get_screen() == 0).all()
    dk = torch.from_numpy(keys).cuda()
    rec = autotune.autotune(dec, packed, dk, reps=1, tiles=(8,), windows=(64,), screens=(0, 6))
    used = sorted({autotune.chunk_class(w) for w in packed["rot_w"]})
    assert (np.asarray(rec["screen_rows"])[:, used] > 0).any()
    assert rec["tuned_ms"] < 0.8 * rec["default_ms"], rec
    dec.close()


def test_reference_golden_p75m2_0():
    _say("golden P75M2-0")
    prob, n, M, packed = ref_instance("P75M2-0")
    with open(os.path.join(GOLDEN, "golden_P75M2-0.json")) as f:
        g = json.load(f)
    dec = _decoder(packed)
    ms, log = dec.evaluate(np.array(g["chromosome"], dtype=np.float32)[None], with_log=True)
    assert round(float(ms[0]), 6) == g["makespan"]
    for m in range(M):
        want = [(p["part_id"], p["bin_idx"], p["col"], p["row"], p["rot"])
                for p in g["machines"][str(m)]["placements"]]
        got = [tuple(int(v) for v in e[1:]) for e in log[0] if e[0] == m]
        assert got == want
    dec.close()


@pytest.mark.parametrize("name", ["P25M2-2", "P50M2-0", "P75M2-4", "P100M4-0"])
def test_reference_fingerprints(name):
    _say("fingerprint %s" % name)
    fp = load_fingerprints()[name]
    prob, n, M, packed = ref_instance(name)
    pop = fingerprint_population(fp, n)
    dec = _decoder(packed)
    ms, _ = dec.evaluate(pop)
    got, n_feas = fingerprint_of(ms)
    assert n_feas == fp["n_feasible"]
    assert got == fp["fingerprint"]
    dec.close()


with open(os.path.join(GOLDEN, "ref_instances.json")) as _f:
    ALL_INSTANCES = list(json.load(_f))


def _routed_keys(packed, n_chrom, rng):
    """Half random keys; half with every part's machine key inside the bucket
    of a machine whose plate fits the part (so most chromosomes are feasible)."""
    n, M = int(packed["n_parts"]), int(packed["n_machines"])
    keys = rng.random((n_chrom, 2 * n))
    edges = np.concatenate([[0.0], np.asarray(packed["thresholds"], np.float64)[:M - 1], [1.0]])
    for c in range(n_chrom // 2, n_chrom):
        for j in range(n):
            t = int(packed["slot_type"][j])
            r0 = int(packed["rot_base"][t])
            h0, w0 = int(packed["rot_h"][r0]), int(packed["rot_w"][r0])
            ok = [m for m in range(M)
                  if (h0 <= packed["plate_h"][m] and w0 <= packed["plate_w"][m]) or
                     (w0 <= packed["plate_h"][m] and h0 <= packed["plate_w"][m])]
            if ok:
                m = ok[rng.integers(len(ok))]
                lo, hi = edges[m], edges[m + 1]
                keys[c, n + j] = lo + (hi - lo) * (0.05 + 0.9 * rng.random())
    return keys.astype(np.float32)


@pytest.mark.parametrize("name", ALL_INSTANCES)
def test_every_reference_instance_vs_oracle(name):
    """All 30 reference instances: placement logs and makespans vs the oracle."""
    _say("instance %s" % name)
    prob, n, M, packed = ref_instance(name)
    keys = _routed_keys(packed, 6, np.random.default_rng(ALL_INSTANCES.index(name)))
    ref_ms, ref_log, _ = oracle.decode(packed, keys, n_threads=oracle_threads(), with_log=True)
    assert np.sum(ref_ms < 1e15) >= 1
    dec = _decoder(packed)
    ms, log = dec.evaluate(keys, with_log=True)
    np.testing.assert_array_equal(ms, ref_ms)
    assert feasible_logs_equal(ref_ms, log, ref_log) is None
    dec.close()


def test_full_population_vs_oracle_p50m2_0():
    """verify_full_population_vs_pp.py:88-194: seed-42 population, element-wise."""
    _say("full population P50M2-0")
    prob, n, M, packed = ref_instance("P50M2-0")
    np.random.seed(42)
    pop = np.random.uniform(0.0, 1.0, size=(500, 100)).astype(np.float32)
    ref_ms, ref_log, _ = oracle.decode(packed, pop, n_threads=oracle_threads(), with_log=True)
    dec = _decoder(packed)
    ms, log = dec.evaluate(pop, with_log=True)
    np.testing.assert_array_equal(ms, ref_ms)
    assert feasible_logs_equal(ref_ms, log, ref_log) is None
    dec.close()


def test_determinism_and_batch_invariance():
    """Same batch twice; permuted batch; single rows (CROSS_ROW_LEAKAGE_FINDINGS.md:12-47)."""
    _say("determinism")
    prob, n, M, packed = syn_instance("C2", seed=3)
    keys = syn_keys("C2", 40, n, seed=5)
    dec = _decoder(packed)
    a, la = dec.evaluate(keys, with_log=True)
    b, lb = dec.evaluate(keys, with_log=True)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(la, lb)
    perm = np.random.default_rng(0).permutation(len(keys))
    c, _ = dec.evaluate(keys[perm])
    np.testing.assert_array_equal(c, a[perm])
    for i in (0, 17, 39):
        d, _ = dec.evaluate(keys[i:i + 1])
        assert d[0] == a[i]
    dec.close()


def test_chromosome_462_batch_sweep_tiling_and_poison_rows():
    """verify_native_vs_pp.py:120-201 on the seed-42 P50M2-0 population: chromosome
    462 decodes to the same makespan (the golden 448166.8336) at the end of batches
    of 1-500 rows (A), at its own row or swapped to row 0, and in a batch of 500
    copies of itself (B), twice in a row (C), and in that tiled batch with one row
    replaced by another chromosome (D).  Batch sizes span fewer, as many and more
    contexts than the persistent workgroups, so both the first-round assignment and
    the queue decode it."""
    _say("chromosome 462 sweep / tiling / poison rows")
    prob, n, M, packed = ref_instance("P50M2-0")
    np.random.seed(42)
    pop = np.random.uniform(0.0, 1.0, size=(500, 100)).astype(np.float32)
    chosen = pop[462].copy()
    want = oracle.decode(packed, pop[462:463])[0][0]
    assert abs(want - 448166.8336) < 1e-3
    dec = _decoder(packed)
    for N in (1, 100, 200, 250, 256, 257, 300, 400, 463, 500):   # A
        batch = pop[:N].copy()
        batch[N - 1] = chosen
        assert dec.evaluate(batch)[0][N - 1] == want, N
    full = dec.evaluate(pop)[0]                                  # B
    assert full[462] == want
    swapped = pop.copy()
    swapped[[0, 462]] = swapped[[462, 0]]
    assert dec.evaluate(swapped)[0][0] == want
    tiled = np.tile(chosen, (500, 1))
    t = dec.evaluate(tiled)[0]
    assert np.all(t == want)
    np.testing.assert_array_equal(dec.evaluate(pop)[0], full)    # C
    for k in (0, 1, 50, 100, 200, 300, 400, 461, 463):           # D
        batch = tiled.copy()
        batch[k] = pop[k]
        got = dec.evaluate(batch)[0]
        assert got[462] == want and got[k] == full[k], k
    dec.close()


def test_bin_capacity_growth():
    """A workspace of one bin per context overflows and is grown transparently."""
    _say("capacity growth")
    prob, n, M, packed = ref_instance("P25M2-4")
    fp = load_fingerprints()["P25M2-4"]
    pop = fingerprint_population(fp, n)[:64]
    ref_ms, _, _ = oracle.decode(packed, pop)
    dec = _decoder(packed, max_bins=1)
    ms, _ = dec.evaluate(pop)
    np.testing.assert_array_equal(ms, ref_ms)
    dec.close()


def test_float64_keys_and_device_entry_point():
    _say("f64 + device entry")
    import torch
    prob, n, M, packed = syn_instance("C1", seed=2)
    keys = syn_keys("C1", 32, n, seed=9).astype(np.float64)
    ref_ms, _, _ = oracle.decode(packed, keys)
    dec = _decoder(packed)
    ms, _ = dec.evaluate(keys)
    np.testing.assert_array_equal(ms, ref_ms)
    dk = torch.from_numpy(keys).cuda()
    out = torch.empty(len(keys), dtype=torch.float64, device="cuda")
    dec.evaluate_device(dk, out)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), ref_ms)
    dec.close()


def test_empty_population_and_shape_checks():
    """An empty population decodes to nothing on both entry points; a mismatched
    output buffer is refused before any launch; the decoder still agrees with the
    oracle afterwards."""
    _say("empty population + shape checks")
    import torch
    from brkga_fft2_amd.native import DecoderError
    prob, n, M, packed = syn_instance("C1", seed=2)
    dec = _decoder(packed)
    ms, log = dec.evaluate(np.zeros((0, 2 * n), np.float32), with_log=True)
    assert ms.shape == (0,) and log.shape == (0, n, 6)
    dk = torch.zeros((0, 2 * n), dtype=torch.float32, device="cuda")
    out = torch.empty(0, dtype=torch.float64, device="cuda")
    assert dec.evaluate_device(dk, out) is out
    keys = syn_keys("C1", 16, n, seed=4)
    dk = torch.from_numpy(keys).cuda()
    with pytest.raises(ValueError):
        dec.evaluate_device(dk, torch.empty(8, dtype=torch.float64, device="cuda"))
    with pytest.raises(DecoderError):
        dec.evaluate(np.zeros((4, 2 * n + 1), np.float32))
    ref_ms, _, _ = oracle.decode(packed, keys)
    out = torch.empty(16, dtype=torch.float64, device="cuda")
    dec.evaluate_device(dk, out)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), ref_ms)
    dec.close()
