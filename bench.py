"""Benchmark: decoded chromosomes/s of the MI355X population decoder.

Workload (default): 50 parts, 2 identical printer machines with 300x300-cell
plates, seeded polygon masks of 20-100 cells per side, 4 rotations, population
256 (one BRKGA generation's decode = one step).  Keys are float32 and resident
in HBM before the timed region; every step decodes a different population.

Prints one JSON line with the throughput, the roofline of the decode kernel
at full shape (every placement row, no early exit: the algorithmic int8 ops
of its collision tests -- the dense valid-region correlation the reference's
FFT evaluates -- over its measured duration; the timed kernel's figures are
reported beside it), the CPU oracle timed on
this host's cores on a bounded sample of the same workload, the plain
single-chromosome GPU path, and the wall time to a target makespan of
surrogate-screened vs plain BRKGA on the same instance (population 1280, so
1024 non-elite chromosomes are decoded per generation; pool of 4096 screened;
100 generations; 10 seeds).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

INT8_PEAK_TOPS = 5000.0   # MI355X dense int8 MFMA peak, 5.0 POPS


def _traffic_from_profile(config, pop, kind):
    """HBM bytes per launch of the decode kernel (``kind`` "dense" or "pruned"),
    from the committed profiler summary (FETCH_SIZE + WRITE_SIZE per dispatch,
    scaled by what the same counters read for a copy of known size), or None
    when there is none."""
    path = os.path.join(ROOT, "profiles", "decode_kernel_%s_pop%d_traffic.md" % (config, pop))
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch_%s:" % kind):
                    return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _batched_vs_single(config, seed, dev, stream, pop=1024, launches=3, singles=8):
    """Chromosomes/s of the batched persistent decode (population ``pop`` per
    launch) and of the same decoder fed one chromosome per launch."""
    import numpy as np
    import torch
    from brkga_fft2_amd.native import Decoder
    from brkga_fft2_amd.problem import pack_problem, synthetic_population, synthetic_problem, thresholds_for
    prob, n, M = synthetic_problem(config, seed=seed)
    packed = pack_problem(prob, n, M, thresholds_for(M), prob.instance_parts)
    keys = [torch.from_numpy(synthetic_population(pop, n, seed=2000 + s, dtype=np.float32)).to(dev)
            for s in range(launches + 1)]
    out = torch.empty(pop, dtype=torch.float64, device=dev)
    d = Decoder(packed, device=0)
    d.evaluate_device(keys[0], out, stream=stream, check_overflow=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in keys[1:]:
        d.evaluate_device(k, out, stream=stream, check_overflow=False)
    torch.cuda.synchronize()
    batched = pop * launches / (time.perf_counter() - t0)
    one = [keys[1][i:i + 1].contiguous() for i in range(singles + 1)]
    o1 = torch.empty(1, dtype=torch.float64, device=dev)
    d.evaluate_device(one[0], o1, stream=stream, check_overflow=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in one[1:]:
        d.evaluate_device(k, o1, stream=stream, check_overflow=False)
    torch.cuda.synchronize()
    single = singles / (time.perf_counter() - t0)
    d.close()
    return {"workload": "C3: %d parts, %d machines, plates 250/300/350/400, population %d per launch"
                        % (n, M, pop),
            "batched_per_s": round(batched, 2), "single_per_s": round(single, 2),
            "batched_over_single": round(batched / single, 2)}


# The reference's published decoder timings (NVIDIA RTX A4000, its native
# C++/CUDA decoder with the cuFFT collision test): one evaluate_batch call on a
# population of 10 n chromosomes drawn by np.random.default_rng(127), mean of 5
# calls, the faster of its 1- and 2-worker runs.  Source:
# benchmarks/20260422_163051_NVIDIA_RTX_A4000_w1-2_stage3-a4000-baseline.json
# (mean_s at :29,58 P25M2-2; :87,116 P25M2-4; :145,174 P50M2-1; :203,232
# P50M2-3; :261,290 P75M2-0; :319,348 P75M2-4; :377,406 P100M4-0; :435,464
# P100M4-3).
REFERENCE_S_PER_CALL = {"P25M2-2": 0.21187, "P25M2-4": 0.19037, "P50M2-1": 0.75892, "P50M2-3": 0.68457,
                        "P75M2-0": 1.60600, "P75M2-4": 1.78770, "P100M4-0": 5.47184, "P100M4-3": 5.40797}


def _reference_instances(reps=5):
    """The same workloads on this GPU: each instance's fingerprint population
    (host float32 keys in, host makespans out, as evaluate_batch takes and
    returns them), mean of ``reps`` calls after one warm-up call; the first five
    feasible makespans and the feasible count are checked against the values
    the reference recorded for the same population (reported; the GPU tests
    are the parity gate)."""
    import numpy as np
    from brkga_fft2_amd.native import Decoder
    from brkga_fft2_amd.problem import load_reference_instance, pack_problem, thresholds_for
    golden = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(golden, "fingerprints.json")) as f:
        fps = json.load(f)
    out = []
    for name, ref_s in REFERENCE_S_PER_CALL.items():
        fp = fps[name]
        prob, n, M = load_reference_instance(golden, name)
        packed = pack_problem(prob, n, M, thresholds_for(M), prob.instance_parts)
        rng = np.random.default_rng(fp["seed"])
        keys = rng.random((fp["num_individuals"], 2 * n), dtype=np.float64).astype(np.float32)
        d = Decoder(packed, device=0)
        ms, _ = d.evaluate(keys)
        t0 = time.perf_counter()
        for _ in range(reps):
            ms, _ = d.evaluate(keys)
        sec = (time.perf_counter() - t0) / reps
        d.close()
        feas = [float(m) for m in ms if m < 1e15]
        ok = ["%.6f" % m for m in feas[:5]] == fp["fingerprint"] and len(feas) == fp["n_feasible"]
        pop = len(keys)
        out.append({"instance": name, "population": pop, "s_per_call": round(sec, 5),
                    "chromosomes_per_s": round(pop / sec, 1),
                    "reference_a4000_s_per_call": ref_s,
                    "reference_chromosomes_per_s": round(pop / ref_s, 1),
                    "speedup": round(ref_s / sec, 1), "fingerprint_ok": bool(ok)})
        print("bench: %s population %d: %.5f s/call (%.0f chrom/s) vs reference A4000 %.5f s/call, "
              "%.0fx, fingerprint %s" % (name, pop, sec, pop / sec, ref_s, ref_s / sec, "ok" if ok else "MISMATCH"),
              file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="C2", choices=["C1", "C2", "C3", "C5"])
    ap.add_argument("--pop", type=int, default=0, help="population per step (default: config's)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--roofline-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the plain single-chromosome path (population 1 per launch)")
    ap.add_argument("--no-ttt", action="store_true", help="skip the time-to-target comparison")
    ap.add_argument("--ttt-generations", type=int, default=100)
    ap.add_argument("--ttt-seeds", type=int, default=10)
    ap.add_argument("--ttt-out", default="", help="write the full time-to-target record (JSON) here")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile table instead of tuning it before the timed steps")
    ap.add_argument("--no-init", action="store_true",
                    help="skip timing the greedy initial solution")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the C3 batched vs single-chromosome comparison")
    ap.add_argument("--no-ref", action="store_true",
                    help="skip timing the reference's own instances and populations")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB ten times (calibrates the "
                         "profiler's HBM byte counters against a known size) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the decoder shards by chromosome; see the design notes)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if args.copy_probe:
        nb = args.copy_probe << 20
        a = torch.ones(nb, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nb, "launches": 10}), flush=True)
        return
    from brkga_fft2_amd.native import Decoder
    from brkga_fft2_amd.problem import (SYNTHETIC_CONFIGS, pack_problem, synthetic_population,
                                        synthetic_problem, thresholds_for)

    cfg = SYNTHETIC_CONFIGS[args.config]
    pop = args.pop or cfg["pop"]
    key_dt = np.float64 if cfg["key_dtype"] == "float64" else np.float32
    prob, n, M = synthetic_problem(args.config, seed=args.seed)
    packed = pack_problem(prob, n, M, thresholds_for(M), prob.instance_parts)
    total = args.warmup + args.steps
    print("bench: config %s n=%d M=%d pop=%d steps=%d warmup=%d" %
          (args.config, n, M, pop, args.steps, args.warmup), file=sys.stderr, flush=True)
    host_pops = [synthetic_population(pop, n, seed=1000 + s, dtype=key_dt) for s in range(total)]
    dev = torch.device("cuda:0")
    keys = [torch.from_numpy(p).to(dev) for p in host_pops]
    outs = [torch.empty(pop, dtype=torch.float64, device=dev) for _ in range(total)]
    stream = torch.cuda.current_stream(dev)
    dec = Decoder(packed, device=0)
    torch.cuda.synchronize()
    tuned = None
    if not args.no_autotune:
        from brkga_fft2_amd import autotune
        t_def, w_def = dec.get_tiles()
        rec = autotune.autotune(dec, packed, keys[0], reps=3)
        tuned = {k: rec[k] for k in ("default_ms", "tuned_ms", "trials")}
        tuned["changed_entries"] = int(np.sum(np.asarray(rec["tiles_per_pass"]) != t_def) +
                                       np.sum(np.asarray(rec["window_tiles"]) != w_def))
        print("bench: autotune %d trials, default %.3f ms -> tuned %.3f ms, %d table entries changed" %
              (rec["trials"], rec["default_ms"], rec["tuned_ms"], tuned["changed_entries"]),
              file=sys.stderr, flush=True)

    for s in range(args.warmup):
        dec.evaluate_device(keys[s], outs[s], stream=stream, check_overflow=False)
    torch.cuda.synchronize()
    dec.reset_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.warmup, total):
        dec.evaluate_device(keys[s], outs[s], stream=stream, check_overflow=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = dec.stats()
    if st["overflows"]:
        raise RuntimeError("bin workspace overflowed during the timed steps")
    elapsed = t1 - t0
    ms_per_step = elapsed * 1e3 / args.steps
    value = pop * args.steps / elapsed
    print("bench: %.3f ms/step, %.1f chrom/s, decode kernel %.3f ms/launch, tests/chrom %.1f" %
          (ms_per_step, value, st["kernel_ms"] / args.steps, st["tests"] / (pop * args.steps)),
          file=sys.stderr, flush=True)
    # spot-check one step against a fresh decode of the same keys (full check is in tests)
    chk = torch.empty(pop, dtype=torch.float64, device=dev)
    dec.evaluate_device(keys[-1], chk, stream=stream)
    torch.cuda.synchronize()
    if not torch.equal(chk, outs[-1]):
        raise RuntimeError("decode is not deterministic")
    # host-buffer entry (keys over PCIe in, makespans out) on the same populations
    nh = min(10, args.steps)
    dec.evaluate(host_pops[args.warmup])
    t0 = time.perf_counter()
    for s in range(args.warmup, args.warmup + nh):
        dec.evaluate(host_pops[s])
    host_rate = pop * nh / (time.perf_counter() - t0)
    pruned = dict(tests=st["tests"], algo_ops=st["algo_ops"], mfma_ops=st["mfma_ops"],
                  kernel_ms=st["kernel_ms"])
    tot = sum(st["phase_us"].values()) or 1.0
    print("bench: workgroup time by phase: " +
          ", ".join("%s %.1f%%" % (k, 100 * v / tot) for k, v in st["phase_us"].items()) +
          "; longest context (max over launches) %.0f us, mean context %.0f us, row groups/test %.2f, "
          "MFMA %.1f TOPS issued" %
          (st["context_us_max"], st["context_us_sum"] / max(1, st["contexts"]),
           st["row_groups"] / max(1, st["tests"]), st["mfma_ops"] / (st["kernel_ms"] * 1e-3) / 1e12),
          file=sys.stderr, flush=True)
    dec.close()

    # ---- roofline: the decode kernel at full shape (every placement row, no
    # early exit), so the algorithmic int8 ops of its tests --
    # 2*h*w*(H-h+1)*(W-w+1) per test, the dense valid-region correlation the
    # reference's FFT evaluates -- are all issued; HIP events on the launch
    # stream give its duration.  The same tests' ops over the timed (pruned)
    # kernel's duration are reported beside it: pruning skips most of them, so
    # that rate is an equivalent rate and may exceed the peak. ----
    kern_ms = pruned["kernel_ms"] / args.steps
    algo_launch = pruned["algo_ops"] / args.steps
    timed_equiv = algo_launch / (kern_ms * 1e-3) / 1e12
    dd = Decoder(packed, device=0, dense=True)
    dd.evaluate_device(keys[0], outs[0], stream=stream, check_overflow=False)
    torch.cuda.synchronize()
    dd.reset_stats()
    rs = max(1, args.roofline_steps)
    for s in range(rs):
        dd.evaluate_device(keys[args.warmup + s % args.steps], outs[0], stream=stream,
                           check_overflow=False)
    torch.cuda.synchronize()
    ds = dd.stats()
    dd.close()
    tot = sum(ds["phase_us"].values()) or 1.0
    print("bench: dense workgroup time by phase: " +
          ", ".join("%s %.1f%%" % (k, 100 * v / tot) for k, v in ds["phase_us"].items()),
          file=sys.stderr, flush=True)
    dense_ms = ds["kernel_ms"] / rs
    achieved = ds["algo_ops"] / rs / (dense_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": INT8_PEAK_TOPS,
                "unit": "TFLOP/s", "frac": round(achieved / INT8_PEAK_TOPS, 4),
                "traffic": _traffic_from_profile(args.config, pop, "dense"),
                "kernel": "decode_kernel<true> (full shape)", "kernel_ms": round(dense_ms, 4),
                "algo_ops_per_launch": ds["algo_ops"] / rs,
                "issued_mfma_tops": round(ds["mfma_ops"] / rs / (dense_ms * 1e-3) / 1e12, 2),
                "timed_kernel": {
                    "kernel": "decode_kernel<false> (row pruning, early exit)",
                    "kernel_ms": round(kern_ms, 4), "algo_ops_per_launch": algo_launch,
                    "algo_equivalent_tops": round(timed_equiv, 2),
                    "issued_mfma_tops": round(pruned["mfma_ops"] / args.steps / (kern_ms * 1e-3) / 1e12, 2),
                    "traffic": _traffic_from_profile(args.config, pop, "pruned")}}
    print("bench: full-shape decode %.3f ms/launch, %.1f TOPS algorithmic (%.1f%% of int8 peak), issued "
          "MFMA %.1f TOPS; timed decode %.3f ms/launch, %.1f TOPS algorithmic-equivalent, issued MFMA "
          "%.1f TOPS" % (dense_ms, achieved, 100 * achieved / INT8_PEAK_TOPS,
                         roofline["issued_mfma_tops"], kern_ms, timed_equiv,
                         roofline["timed_kernel"]["issued_mfma_tops"]), file=sys.stderr, flush=True)

    single = None
    if not args.no_single:
        ds1 = Decoder(packed, device=0)
        one = [k[:1].contiguous() for k in keys[:8]]
        o1 = torch.empty(1, dtype=torch.float64, device=dev)
        ds1.evaluate_device(one[0], o1, stream=stream, check_overflow=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in one:
            ds1.evaluate_device(k, o1, stream=stream, check_overflow=False)
        torch.cuda.synchronize()
        single = len(one) / (time.perf_counter() - t0)
        ds1.close()

    # ---- greedy initial solution (the seed chromosome) on the same instance ----
    init = None
    if not args.no_init:
        from brkga_fft2_amd.problem import initial_machine_keys, type_heights
        th = type_heights(prob, packed)
        mk = initial_machine_keys(thresholds_for(M), M, np.random.default_rng(args.seed))
        dg = Decoder(packed, device=0)
        dg.initial_solution(th, mk, stream=stream)
        t0 = time.perf_counter()
        g_keys, g_ms, _ = dg.initial_solution(th, mk, stream=stream)
        g_s = time.perf_counter() - t0
        dg.close()
        init = {"workload": "%s instance, %d parts, %d machines, one run" % (args.config, n, M),
                "gpu_s": round(g_s, 4), "makespan": g_ms}
        print("bench: greedy initial solution %.4f s, makespan %.6g" % (g_s, g_ms),
              file=sys.stderr, flush=True)

    # ---- CPU baseline: the C oracle on this host's cores, bounded sample ----
    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        cores = min(16, len(os.sched_getaffinity(0)))
        sample = np.concatenate(host_pops[args.warmup:])
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < args.cpu_seconds and done < len(sample):
            oracle.decode(packed, sample[done:done + pop], n_threads=cores)
            done += min(pop, len(sample) - done)
        cpu_t = time.perf_counter() - t0
        want, _, _ = oracle.decode(packed, sample[:8], n_threads=cores)
        if not np.array_equal(want, outs[args.warmup][:8].cpu().numpy()):
            raise RuntimeError("GPU makespans differ from the CPU oracle")
        t0 = time.perf_counter()
        done1 = 0
        while time.perf_counter() - t0 < args.cpu_seconds / 2 and done1 < len(sample):
            oracle.decode(packed, sample[done1:done1 + 16], n_threads=1)
            done1 += min(16, len(sample) - done1)
        cpu1 = done1 / (time.perf_counter() - t0)
        cpu = {"value": round(done / cpu_t, 2), "unit": "chromosomes/s", "cores": cores,
               "kind": "port",
               "sample": "%d chromosomes (the first %d timed steps' populations), same instance, "
                         "C oracle with %d OpenMP threads, %.1f s" % (done, done // pop, cores, cpu_t),
               "value_1_thread": round(cpu1, 2), "cpu_model": _cpu_model()}
        if init is not None:   # the greedy's C restatement, one thread
            from brkga_fft2_amd.problem import initial_solution_order
            t0 = time.perf_counter()
            w_keys, w_ms, _ = oracle.initial_solution(packed, initial_solution_order(packed, th), mk)
            init["cpu_oracle_s"] = round(time.perf_counter() - t0, 4)
            if w_ms != g_ms or not np.array_equal(w_keys, g_keys):
                raise RuntimeError("GPU greedy initial solution differs from the CPU oracle")

    # ---- C3: batched persistent decode vs the single-chromosome path ----
    c3 = None
    if not args.no_c3:
        c3 = _batched_vs_single("C3", args.seed, dev, stream)
        print("bench: C3 batched %.1f chrom/s, single %.1f chrom/s, ratio %.1f" %
              (c3["batched_per_s"], c3["single_per_s"], c3["batched_over_single"]),
              file=sys.stderr, flush=True)

    ref_inst = None if args.no_ref else _reference_instances()

    # ---- wall time to a target makespan: surrogate-screened vs plain BRKGA ----
    ttt = None
    if not args.no_ttt:
        from brkga_fft2_amd.surrogate import time_to_target
        dt = Decoder(packed, device=0)
        rec = time_to_target(dt, prob, prob.instance_parts, generations=args.ttt_generations,
                             seeds=args.ttt_seeds)
        dt.close()
        if args.ttt_out:
            with open(args.ttt_out, "w") as f:
                json.dump(rec, f)
        rec.pop("runs")
        ttt = dict(workload="%s instance, population 1280 (1024 non-elite decoded per generation), "
                            "pool 4096 screened, %d generations, %d seeds; target = median over seeds "
                            "0-%d of plain's best after %d generations" %
                            (args.config, rec["generations"], rec["seeds"], rec["target_seeds"] - 1,
                             rec["generations"] // 2), **rec)
        print("bench: time to target %.5f: plain %s s (%d/%d reached), surrogate %s s (%d/%d), "
              "speedup %s" % (rec["target"], rec["plain"]["median_s"], rec["plain"]["reached"], rec["seeds"],
                              rec["surrogate"]["median_s"], rec["surrogate"]["reached"], rec["seeds"],
                              rec["speedup"]), file=sys.stderr, flush=True)

    line = {
        "metric": "decoded chromosomes/s (exact placements); wall time to target makespan, 1 GPU",
        "value": round(value, 2),
        "unit": "chromosomes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": "%s: %d parts, %d machines, plates %s cells, masks %d-%d cells, "
                               "%d rotations, population %d per step, %s keys" %
                               (args.config, n, M, "/".join("%dx%d" % (p, p) for p in cfg["plates"]),
                                cfg["bbox"][0], cfg["bbox"][1], cfg["nrot"], pop, cfg["key_dtype"]),
                   "population": pop, "parts": n, "machines": M},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "decode_kernel_ms": round(pruned["kernel_ms"] / args.steps, 4),
        "tests_per_chromosome": round(pruned["tests"] / (pop * args.steps), 2),
    }
    if single is not None:
        line["single_chromosome_per_s"] = round(single, 2)
        line["batched_over_single"] = round(value / single, 2)
    line["host_buffers_per_s"] = round(host_rate, 2)
    line["autotune"] = tuned
    line["initial_solution"] = init
    line["c3_batched_vs_single"] = c3
    line["reference_instances"] = ref_inst
    line["time_to_target"] = ttt
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
