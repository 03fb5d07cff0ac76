"""Benchmark: decoded chromosomes/s of the MI355X population decoder.

Headline workload (default, ``--config C5``): 200 parts, 4 printer machines
with 500x500-cell plates (0.5 mm cells), seeded polygon masks of 40-200 cells
per side, 4 rotations, population 2048 (one BRKGA generation's decode = one
step), float32 keys resident in HBM before the timed region; every step
decodes a different population.  ``--config C2`` (50 parts, 2 machines,
300x300, population 256) and the others are available; C2 is also timed as
an extra record.

Prints one JSON line: the throughput; the roofline of the decode kernel that
the timed steps launch (int8 MFMA ops it issues per launch over its HIP-event
duration on the launch stream, against the 5.0 POP/s dense int8 peak), with
the full-shape decode beside it (every placement row, no early exit: the
algorithmic int8 ops of the collision tests -- the dense valid-region
correlation the reference's FFT evaluates -- over its duration); the CPU
oracle timed on this host's cores on a bounded sample of the same workload;
the plain single-chromosome GPU path; the device generation loop (graph
replay against eager launches); and the wall time to a target makespan of
surrogate-screened vs plain BRKGA (population 1280, 1024 non-elite
chromosomes decoded per generation, pool of 4096 screened, 10 seeds) on the
C2 and C3 instances.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X dense MFMA peaks (TOP/s) of the collision correlation's operand formats
PEAK_TOPS = {"fp4": 10100.0, "int8": 5000.0}


def _say(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


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


def _cpu_info():
    """CPU model and topology of this host (lscpu, else /proc/cpuinfo)."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            k, v = k.strip(), v.strip()
            if k in ("Model name", "CPU family", "Model", "Stepping", "Thread(s) per core",
                     "Core(s) per socket", "Socket(s)", "CPU max MHz", "L3 cache"):
                info[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    if "Model name" not in info:
        try:
            with open("/proc/cpuinfo") as f:
                for ln in f:
                    if ln.startswith("model name"):
                        info["Model name"] = ln.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return info


def _workload(config, pop, n, M, cfg):
    return ("%s: %d parts, %d machines, plates %s cells, masks %d-%d cells, %d rotations, population %d "
            "per step, %s keys" % (config, n, M, "/".join("%dx%d" % (p, p) for p in cfg["plates"]),
                                   cfg["bbox"][0], cfg["bbox"][1], cfg["nrot"], pop, cfg["key_dtype"]))


def _timed_decode(config, pop, steps, warmup, seed, dev, autotune_reps=0, only_timed=False, dense=False,
                  mfma="fp4"):
    """Decode ``warmup`` then ``steps`` device-resident populations of ``config``
    back to back on one stream; returns the record and the state later legs use."""
    import numpy as np
    import torch
    from brkga_fft2_amd.native import Decoder
    from brkga_fft2_amd.problem import (SYNTHETIC_CONFIGS, pack_problem, synthetic_population,
                                        synthetic_problem, thresholds_for)
    cfg = SYNTHETIC_CONFIGS[config]
    key_dt = np.float64 if cfg["key_dtype"] == "float64" else np.float32
    prob, n, M = synthetic_problem(config, seed=seed)
    packed = pack_problem(prob, n, M, thresholds_for(M), prob.instance_parts)
    total = warmup + steps
    host_pops = [synthetic_population(pop, n, seed=1000 + s, dtype=key_dt) for s in range(total)]
    keys = [torch.from_numpy(p).to(dev) for p in host_pops]
    outs = [torch.empty(pop, dtype=torch.float64, device=dev) for _ in range(total)]
    stream = torch.cuda.current_stream(dev)
    dec = Decoder(packed, device=0, dense=dense, mfma=mfma)
    torch.cuda.synchronize()
    tuned = None
    if autotune_reps:
        from brkga_fft2_amd import autotune
        t_def, w_def = dec.get_tiles()
        s_def = dec.get_screen()
        rec = autotune.autotune(dec, packed, keys[0], reps=autotune_reps)
        tuned = {k: rec[k] for k in ("default_ms", "tuned_ms", "trials")}
        tuned["changed_entries"] = int(np.sum(np.asarray(rec["tiles_per_pass"]) != t_def) +
                                       np.sum(np.asarray(rec["window_tiles"]) != w_def) +
                                       np.sum(np.asarray(rec["screen_rows"]) != s_def))
        tuned["changed"] = autotune.changed_entries(packed, (t_def, w_def, s_def), rec)
        _say("%s autotune %d trials, default %.3f ms -> tuned %.3f ms, %d table entries changed" %
             (config, rec["trials"], rec["default_ms"], rec["tuned_ms"], tuned["changed_entries"]))
    for s in range(warmup):
        dec.evaluate_device(keys[s], outs[s], stream=stream, check_overflow=False)
    torch.cuda.synchronize()
    dec.reset_stats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(warmup, total):
        dec.evaluate_device(keys[s], outs[s], stream=stream, check_overflow=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = dec.stats()
    if st["overflows"]:
        raise RuntimeError("bin workspace overflowed during the timed steps")
    elapsed = t1 - t0
    rec = {"config": config, "population": pop, "parts": n, "machines": M, "mfma": mfma,
           "workload": _workload(config, pop, n, M, cfg),
           "value": pop * steps / elapsed, "ms_per_step": elapsed * 1e3 / steps, "stats": st,
           "autotune": tuned}
    kern_ms = st["kernel_ms"] / steps
    _say("%s %.3f ms/step, %.1f chrom/s, decode kernel %.3f ms/launch, issued MFMA %.1f TOP/s, tests/chrom %.1f"
         % (config, rec["ms_per_step"], rec["value"], kern_ms, st["mfma_ops"] / steps / (kern_ms * 1e-3) / 1e12,
            st["tests"] / (pop * steps)))
    state = dict(prob=prob, packed=packed, n=n, M=M, host_pops=host_pops, keys=keys, outs=outs,
                 stream=stream, dec=dec)
    if only_timed:
        return rec, state
    # spot-check: a fresh decode of the last step's keys gives the same makespans
    chk = torch.empty(pop, dtype=torch.float64, device=dev)
    dec.evaluate_device(keys[-1], chk, stream=stream)
    torch.cuda.synchronize()
    if not torch.equal(chk, outs[-1]):
        raise RuntimeError("decode is not deterministic")
    # host-buffer entry (keys over PCIe in, makespans out) on the same populations
    nh = min(5 if pop >= 2048 else 10, steps)
    dec.evaluate(host_pops[warmup])
    t0 = time.perf_counter()
    for s in range(warmup, warmup + nh):
        dec.evaluate(host_pops[s])
    rec["host_buffers_per_s"] = pop * nh / (time.perf_counter() - t0)
    tot = sum(st["phase_us"].values()) or 1.0
    _say("%s workgroup time by phase: " % config +
         ", ".join("%s %.1f%%" % (k, 100 * v / tot) for k, v in st["phase_us"].items()) +
         "; longest context (max over launches) %.0f us, mean context %.0f us, row groups/test %.2f" %
         (st["context_us_max"], st["context_us_sum"] / max(1, st["contexts"]),
          st["row_groups"] / max(1, st["tests"])))
    return rec, state


def _roofline(rec, state, steps, warmup, roofline_steps, dev):
    """Roofline of the timed decode kernel, with the full-shape decode beside it."""
    import torch
    from brkga_fft2_amd.native import Decoder
    st = rec["stats"]
    config, pop = rec["config"], rec["population"]
    kern_ms = st["kernel_ms"] / steps
    mfma_launch = st["mfma_ops"] / steps
    algo_launch = st["algo_ops"] / steps
    issued = mfma_launch / (kern_ms * 1e-3) / 1e12
    keys, outs, stream = state["keys"], state["outs"], state["stream"]
    peak = PEAK_TOPS[rec["mfma"]]
    full = None
    if roofline_steps > 0:
        dd = Decoder(state["packed"], device=0, dense=True, mfma=rec["mfma"])
        dd.evaluate_device(keys[0], outs[0], stream=stream, check_overflow=False)
        torch.cuda.synchronize()
        dd.reset_stats()
        for s in range(roofline_steps):
            dd.evaluate_device(keys[warmup + s % steps], outs[0], stream=stream, check_overflow=False)
        torch.cuda.synchronize()
        ds = dd.stats()
        dd.close()
        dense_ms = ds["kernel_ms"] / roofline_steps
        achieved = ds["algo_ops"] / roofline_steps / (dense_ms * 1e-3) / 1e12
        full = {"kernel": "decode_kernel<true, %s> (full shape: every placement row, no early exit)"
                          % ("true" if rec["mfma"] == "fp4" else "false"),
                "basis": "algorithmic ops of the tests, 2*h*w*(H-h+1)*(W-w+1) per test (the dense valid-region "
                         "correlation), over the launch's HIP-event duration, against the operand format's peak",
                "achieved": round(achieved, 2), "frac": round(achieved / peak, 4),
                "kernel_ms": round(dense_ms, 4), "algo_ops_per_launch": ds["algo_ops"] / roofline_steps,
                "issued_mfma_tops": round(ds["mfma_ops"] / roofline_steps / (dense_ms * 1e-3) / 1e12, 2),
                "traffic": _traffic_from_profile(config, pop, "dense")}
        tot = sum(ds["phase_us"].values()) or 1.0
        _say("%s full-shape decode %.3f ms/launch, %.1f TOP/s algorithmic (%.1f%% of the %s peak), issued MFMA "
             "%.1f TOP/s; phases: %s" % (config, dense_ms, achieved, 100 * achieved / peak, rec["mfma"],
                                         full["issued_mfma_tops"],
                                         ", ".join("%s %.1f%%" % (k, 100 * v / tot)
                                                   for k, v in ds["phase_us"].items())))
    roof = {"bound": "mfma", "achieved": round(issued, 2), "peak": peak, "unit": "TOP/s",
            "frac": round(issued / peak, 4), "format": rec["mfma"],
            "traffic": _traffic_from_profile(config, pop, "pruned"),
            "kernel": "decode_kernel<false, %s> (the timed steps' kernel: row pruning, early exit)"
                      % ("true" if rec["mfma"] == "fp4" else "false"),
            "basis": "MFMA ops the kernel issues per launch (fp4: 32x32x64 x 2 per instruction, int8: "
                     "32x32x32 x 2) over its HIP-event duration on the launch stream, against the dense peak "
                     "of the operand format",
            "kernel_ms": round(kern_ms, 4), "mfma_ops_per_launch": mfma_launch,
            "algo_ops_per_launch": algo_launch,
            "algo_equivalent_tops": round(algo_launch / (kern_ms * 1e-3) / 1e12, 2),
            "full_shape": full}
    _say("%s timed decode %.3f ms/launch, issued MFMA %.1f TOP/s (%.1f%% of the %s peak), algorithmic-equivalent "
         "%.1f TOP/s" % (config, kern_ms, issued, 100 * issued / peak, rec["mfma"], roof["algo_equivalent_tops"]))
    return roof


def _single_chromosome(state, dev, n_single=8):
    """The same decoder fed one chromosome per launch (the plain GPU path)."""
    import torch
    from brkga_fft2_amd.native import Decoder
    ds1 = Decoder(state["packed"], device=0)
    one = [state["keys"][i % len(state["keys"])][i:i + 1].contiguous() for i in range(n_single + 1)]
    o1 = torch.empty(1, dtype=torch.float64, device=dev)
    ds1.evaluate_device(one[0], o1, stream=state["stream"], check_overflow=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in one[1:]:
        ds1.evaluate_device(k, o1, stream=state["stream"], check_overflow=False)
    torch.cuda.synchronize()
    single = n_single / (time.perf_counter() - t0)
    ds1.close()
    return single


def _cpu_baseline(state, pop, warmup, outs, seconds):
    """The C oracle on every core this process may use, over a bounded sample of
    the timed populations; a one-thread figure beside it."""
    import numpy as np
    from oracle import oracle
    packed = state["packed"]
    cores = len(os.sched_getaffinity(0))
    sample = np.concatenate(state["host_pops"][warmup:])
    chunk = max(2 * cores, 16)
    t0 = time.perf_counter()
    done = 0
    while time.perf_counter() - t0 < seconds and done < len(sample):
        oracle.decode(packed, sample[done:done + chunk], n_threads=cores)
        done += min(chunk, len(sample) - done)
    cpu_t = time.perf_counter() - t0
    want, _, _ = oracle.decode(packed, sample[:8], n_threads=cores)
    if not np.array_equal(want, outs[warmup][:8].cpu().numpy()):
        raise RuntimeError("GPU makespans differ from the CPU oracle")
    t0 = time.perf_counter()
    done1 = 0
    step1 = 2 if pop >= 2048 else 16
    while time.perf_counter() - t0 < seconds / 3 and done1 < len(sample):
        oracle.decode(packed, sample[done1:done1 + step1], n_threads=1)
        done1 += min(step1, len(sample) - done1)
    cpu1 = done1 / (time.perf_counter() - t0)
    info = _cpu_info()
    return {"value": round(done / cpu_t, 2), "unit": "chromosomes/s", "cores": cores, "kind": "port",
            "sample": "the first %d chromosomes of the timed steps' populations (same instance), decoded by "
                      "the C oracle with %d OpenMP threads over (chromosome, machine) contexts, %.1f s"
                      % (done, cores, cpu_t),
            "value_1_thread": round(cpu1, 3), "cpu_model": info.get("Model name"), "cpu_topology": info}


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


def _device_ga(config, seed, N, gens):
    """The device generation loop (sort, breed, decode, elite fitness, commit,
    best) per generation: one hipGraph replay against the same kernels launched
    eagerly, same seed, same trajectory."""
    import math
    import numpy as np
    from brkga_fft2_amd.native import Decoder, DeviceGA
    from brkga_fft2_amd.problem import pack_problem, synthetic_problem, thresholds_for
    prob, n, M = synthetic_problem(config, seed=seed)
    packed = pack_problem(prob, n, M, thresholds_for(M), prob.instance_parts)
    ne, nm = int(math.ceil(0.2 * N)), int(math.ceil(0.15 * N))
    out = {"workload": "%s instance, population %d (%d elites, %d mutants), %d generations"
                       % (config, N, ne, nm, gens)}
    best = {}
    for mode, graph in (("graph", True), ("eager", False)):
        dec = Decoder(packed, device=0, max_bins=n)
        ga = DeviceGA(dec, N, ne, nm, elite_prob=0.7, seed=77, history_capacity=gens + 2)
        ga.run(1, use_graph=graph)   # capture (graph) and first-launch set-up
        t0 = time.perf_counter()
        ga.run(gens, use_graph=graph)
        out["%s_ms_per_generation" % mode] = round(1e3 * (time.perf_counter() - t0) / gens, 4)
        best[mode] = ga.best()[2]
        ga.close()
        dec.close()
    out["same_trajectory"] = bool(np.array_equal(best["graph"], best["eager"]))
    out["graph_over_eager"] = round(out["eager_ms_per_generation"] / out["graph_ms_per_generation"], 3)
    _say("device GA %s: graph %.3f ms/generation, eager %.3f ms/generation, same trajectory %s" %
         (config, out["graph_ms_per_generation"], out["eager_ms_per_generation"], out["same_trajectory"]))
    return out


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
        _say("%s population %d: %.5f s/call (%.0f chrom/s) vs reference A4000 %.5f s/call, %.0fx, fingerprint %s"
             % (name, pop, sec, pop / sec, ref_s, ref_s / sec, "ok" if ok else "MISMATCH"))
    return out


def _time_to_target(config, seed, generations, seeds, out_path=""):
    """Surrogate-screened vs plain BRKGA on one synthetic instance."""
    from brkga_fft2_amd.native import Decoder
    from brkga_fft2_amd.problem import pack_problem, synthetic_problem, thresholds_for
    from brkga_fft2_amd.surrogate import time_to_target
    prob, n, M = synthetic_problem(config, seed=seed)
    packed = pack_problem(prob, n, M, thresholds_for(M), prob.instance_parts)
    dt = Decoder(packed, device=0, max_bins=n)   # device loops cannot grow the workspace
    rec = time_to_target(dt, prob, prob.instance_parts, generations=generations, seeds=seeds)
    st = dt.stats()
    dt.close()
    if out_path:
        with open(out_path, "w") as f:
            json.dump(rec, f)
    rec.pop("runs")
    rec = dict(workload="%s instance, population 1280 (1024 non-elite decoded per generation), pool 4096 "
                        "screened, %d generations, %d seeds; target = median over seeds 0-%d of plain's best "
                        "after %d generations" % (config, rec["generations"], rec["seeds"],
                                                  rec["target_seeds"] - 1, rec["generations"] // 2), **rec)
    rec["decode_issued_mfma_tops"] = round(st["mfma_ops"] / max(st["kernel_ms"], 1e-9) / 1e9, 2)
    _say("%s time to target %.5f: plain %s s (%d/%d reached), surrogate %s s (%d/%d), speedup %s" %
         (config, rec["target"], rec["plain"]["median_s"], rec["plain"]["reached"], rec["seeds"],
          rec["surrogate"]["median_s"], rec["surrogate"]["reached"], rec["seeds"], rec["speedup"]))
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="C5", choices=["C1", "C2", "C3", "C5"])
    ap.add_argument("--pop", type=int, default=0, help="population per step (default: config's)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--roofline-steps", type=int, default=2,
                    help="full-shape launches timed beside the timed kernel (0 skips them)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--only-timed", action="store_true",
                    help="stop after the timed steps (for profiler runs: every decode launch is then "
                         "a warm-up or timed one)")
    ap.add_argument("--mfma", default="fp4", choices=["fp4", "int8"],
                    help="operand format of the collision correlation's MFMAs (same results)")
    ap.add_argument("--dense", action="store_true",
                    help="with --only-timed: time the full-shape decode instead (every placement row, no "
                         "early exit; profiler runs of the roofline kernel)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the plain single-chromosome path (population 1 per launch)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the extra C2 record when the headline is another configuration")
    ap.add_argument("--no-ttt", action="store_true", help="skip the time-to-target comparisons")
    ap.add_argument("--ttt-configs", default="C2,C3,C5", help="instances of the time-to-target comparisons")
    ap.add_argument("--ttt-generations", type=int, default=100)
    ap.add_argument("--ttt-seeds", type=int, default=10)
    ap.add_argument("--ttt-c5-generations", type=int, default=100)
    ap.add_argument("--ttt-c5-seeds", type=int, default=3)
    ap.add_argument("--ttt-out", default="", help="write the full time-to-target records (JSON) under this prefix")
    ap.add_argument("--autotune", type=int, default=-1, metavar="REPS",
                    help="tune the tile table before the timed steps with REPS timings per trial "
                         "(default: 3 on C1-C3, off on C5, whose trials take a third of a second each)")
    ap.add_argument("--no-init", action="store_true", help="skip timing the greedy initial solution")
    ap.add_argument("--no-c3", action="store_true", help="skip the C3 batched vs single-chromosome comparison")
    ap.add_argument("--no-ref", action="store_true",
                    help="skip timing the reference's own instances and populations")
    ap.add_argument("--no-ga", action="store_true", help="skip timing the device generation loop")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB ten times (calibrates the "
                         "profiler's HBM byte counters against a known size) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the decoder shards by chromosome; see the design notes)" % args.gpus, file=sys.stderr)
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
    from brkga_fft2_amd.problem import SYNTHETIC_CONFIGS, initial_machine_keys, thresholds_for, type_heights

    cfg = SYNTHETIC_CONFIGS[args.config]
    pop = args.pop or cfg["pop"]
    dev = torch.device("cuda:0")
    reps = args.autotune if args.autotune >= 0 else (2 if args.config == "C5" else 3)
    _say("config %s pop=%d steps=%d warmup=%d" % (args.config, pop, args.steps, args.warmup))
    if args.dense and not args.only_timed:
        ap.error("--dense needs --only-timed (the full-shape decode is not the headline)")
    rec, state = _timed_decode(args.config, pop, args.steps, args.warmup, args.seed, dev,
                               autotune_reps=reps, only_timed=args.only_timed, dense=args.dense, mfma=args.mfma)
    if args.only_timed:
        state["dec"].close()
        print(json.dumps({"config": rec["workload"], "value": round(rec["value"], 2),
                          "ms_per_step": round(rec["ms_per_step"], 4),
                          "decode_kernel_ms": round(rec["stats"]["kernel_ms"] / args.steps, 4)}), flush=True)
        return
    state["dec"].close()
    roofline = _roofline(rec, state, args.steps, args.warmup, args.roofline_steps, dev)
    st = rec["stats"]

    single = None if args.no_single else _single_chromosome(state, dev)

    # ---- greedy initial solution (the seed chromosome) on the same instance ----
    init = None
    if not args.no_init:
        from brkga_fft2_amd.native import Decoder
        prob, packed, n, M = state["prob"], state["packed"], state["n"], state["M"]
        th = type_heights(prob, packed)
        mk = initial_machine_keys(thresholds_for(M), M, np.random.default_rng(args.seed))
        dg = Decoder(packed, device=0)
        dg.initial_solution(th, mk, stream=state["stream"])
        t0 = time.perf_counter()
        g_keys, g_ms, _ = dg.initial_solution(th, mk, stream=state["stream"])
        g_s = time.perf_counter() - t0
        dg.close()
        init = {"workload": "%s instance, %d parts, %d machines, one run" % (args.config, n, M),
                "gpu_s": round(g_s, 4), "makespan": g_ms}
        _say("greedy initial solution %.4f s, makespan %.6g" % (g_s, g_ms))

    # ---- CPU baseline: the C oracle on this host's cores, bounded sample ----
    cpu = None
    if not args.no_cpu:
        cpu = _cpu_baseline(state, pop, args.warmup, state["outs"], args.cpu_seconds)
        _say("CPU oracle %.2f chrom/s on %d threads (%s), %.3f chrom/s on one" %
             (cpu["value"], cpu["cores"], cpu["cpu_model"], cpu["value_1_thread"]))
        if init is not None:   # the greedy's C restatement, one thread
            from oracle import oracle
            from brkga_fft2_amd.problem import initial_solution_order
            t0 = time.perf_counter()
            w_keys, w_ms, _ = oracle.initial_solution(state["packed"], initial_solution_order(state["packed"], th), mk)
            init["cpu_oracle_s"] = round(time.perf_counter() - t0, 4)
            if w_ms != g_ms or not np.array_equal(w_keys, g_keys):
                raise RuntimeError("GPU greedy initial solution differs from the CPU oracle")
    state.clear()
    torch.cuda.empty_cache()

    extra = None
    if not args.no_extra and args.config != "C2":
        r2, s2 = _timed_decode("C2", SYNTHETIC_CONFIGS["C2"]["pop"], 50, 5, args.seed, dev,
                               autotune_reps=3 if reps else 0, mfma=args.mfma)
        s2["dec"].close()
        roof2 = _roofline(r2, s2, 50, 5, 1, dev)
        extra = {"workload": r2["workload"], "value": round(r2["value"], 2),
                 "ms_per_step": round(r2["ms_per_step"], 4), "roofline": roof2,
                 "autotune": r2["autotune"], "host_buffers_per_s": round(r2["host_buffers_per_s"], 2)}
        s2.clear()

    c3 = None
    if not args.no_c3:
        c3 = _batched_vs_single("C3", args.seed, dev, torch.cuda.current_stream(dev))
        _say("C3 batched %.1f chrom/s, single %.1f chrom/s, ratio %.1f" %
             (c3["batched_per_s"], c3["single_per_s"], c3["batched_over_single"]))

    ref_inst = None if args.no_ref else _reference_instances()

    ga = None
    if not args.no_ga:
        ga = {"C2": _device_ga("C2", args.seed, 1280, 20)}
        if args.config != "C2":
            ga[args.config] = _device_ga(args.config, args.seed, SYNTHETIC_CONFIGS[args.config]["pop"],
                                         3 if args.config == "C5" else 10)

    ttt = None
    if not args.no_ttt:
        ttt = {}
        for c in [c for c in args.ttt_configs.split(",") if c]:
            big = c == "C5"   # a C5 generation decodes 1024 chromosomes of 200 parts: fewer, shorter runs
            ttt[c] = _time_to_target(c, args.seed, args.ttt_c5_generations if big else args.ttt_generations,
                                     args.ttt_c5_seeds if big else args.ttt_seeds,
                                     (args.ttt_out + "_%s.json" % c) if args.ttt_out else "")

    line = {
        "metric": "decoded chromosomes/s (exact placements); wall time to target makespan, 1 GPU",
        "value": round(rec["value"], 2),
        "unit": "chromosomes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(rec["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.mfma,
        "data": "synthetic",
        "config": {"workload": rec["workload"], "population": pop, "parts": rec["parts"],
                   "machines": rec["machines"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "decode_kernel_ms": round(st["kernel_ms"] / args.steps, 4),
        "tests_per_chromosome": round(st["tests"] / (pop * args.steps), 2),
        "context_us_mean": round(st["context_us_sum"] / max(1, st["contexts"]), 1),
        "context_us_max": round(st["context_us_max"], 1),
    }
    if single is not None:
        line["single_chromosome_per_s"] = round(single, 2)
        line["batched_over_single"] = round(rec["value"] / single, 2)
    line["host_buffers_per_s"] = round(rec["host_buffers_per_s"], 2)
    line["autotune"] = rec["autotune"]
    line["initial_solution"] = init
    line["c2"] = extra
    line["c3_batched_vs_single"] = c3
    line["reference_instances"] = ref_inst
    line["device_ga"] = ga
    line["time_to_target"] = ttt
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
