"""Surrogate-screened BRKGA: a BF16 MLP ranks oversized offspring pools.

The network (``brkga_surrogate.hip``: 4n + 4M -> 512 -> 512 -> 1, BF16
weights and activations, FP32 accumulation, Adam) is trained online on the
decoder's exact makespans.  Its inputs are the keys as the decoder reads them
and the parts' areas and heights in sequence order (4n, scaled by 1/8), and
per machine the processing time, plate-relative area, count and tallest part
the chromosome sends to it (4M): a machine's makespan is mostly its load plus
setup per batch, so on several machines these aggregates carry most of the
ranking signal.  Each generation after a warm-up:

1. the population is sorted by exact fitness (``brkga_sort_fitness``);
2. a pool of ``pool`` candidates is bred -- mutants and biased-crossover
   offspring in the plain method's proportion (``brkga_breed_pool``);
3. the network predicts every candidate; the best-predicted ``num_mutants``
   mutants and ``num_offspring`` offspring are decoded exactly;
4. the next population is the elites (exact fitness, never predicted) plus
   the decoded candidates; the decoded pairs feed the training buffer.

During the warm-up, or when the network's rank correlation on the last
decoded batch falls below ``fallback_rho``, step 3 takes the first
candidates of each kind instead -- the plain BRKGA generation (the pool is
i.i.d., so its first rows are exactly what plain BRKGA would breed).

The reference has no surrogate; the exact decoder is the only source of
fitness values, so every reported makespan is exact.
"""
import ctypes
import math
import time

import numpy as np

from .native import Decoder, DecoderError, load_library

_P = ctypes.c_void_p


def _lib():
    lib = load_library()
    if not hasattr(lib, "_mlp_bound"):
        lib.brkga_mlp_create.argtypes = [ctypes.c_int32, ctypes.c_int32, _P, ctypes.c_int32, _P, _P, _P,
                                         ctypes.c_uint64, ctypes.c_float, ctypes.c_int32, ctypes.POINTER(_P)]
        lib.brkga_mlp_predict.argtypes = [_P, _P, ctypes.c_int64, _P, _P]
        lib.brkga_mlp_gradients.argtypes = [_P, _P, _P, ctypes.c_int64, _P, _P]
        lib.brkga_mlp_train_step.argtypes = [_P, _P, _P, ctypes.c_int64, _P, _P]
        lib.brkga_mlp_num_params.argtypes = [_P, _P, _P]
        lib.brkga_mlp_get_params.argtypes = [_P, _P]
        lib.brkga_mlp_set_params.argtypes = [_P, _P]
        lib.brkga_mlp_features.argtypes = [_P, _P, ctypes.c_int64, _P, _P]
        lib.brkga_mlp_last_error.argtypes = []
        lib.brkga_mlp_last_error.restype = ctypes.c_char_p
        lib.brkga_mlp_destroy.argtypes = [_P]
        lib.brkga_mlp_destroy.restype = None
        lib.brkga_sort_fitness.argtypes = [_P, ctypes.c_int32, _P, _P]
        lib.brkga_breed_pool.argtypes = [_P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                         ctypes.c_uint64, ctypes.c_uint32, _P, _P]
        for f in ("brkga_mlp_create", "brkga_mlp_predict", "brkga_mlp_gradients",
                  "brkga_mlp_train_step", "brkga_mlp_num_params", "brkga_mlp_get_params",
                  "brkga_mlp_set_params", "brkga_mlp_features", "brkga_sort_fitness",
                  "brkga_breed_pool"):
            getattr(lib, f).restype = ctypes.c_int
        lib._mlp_bound = True
    return lib


def slot_features(problem, instance_parts, machine=0):
    """Per-slot (area, height) scaled to [0, 1]: part area, and the height-driven
    processing time of machine 0 (proportional to the part height)."""
    area = np.array([problem.parts[int(p)].area for p in instance_parts], np.float64)
    height = np.array([problem.machines[machine].parts[int(p)].proc_time_height
                       for p in instance_parts], np.float64)
    return ((area / area.max()).astype(np.float32),
            (height / max(height.max(), 1e-300)).astype(np.float32))


def machine_terms(problem, instance_parts, n_machines):
    """Per-machine terms of the aggregate features, float32 ``[M][3][n]``: the
    processing time of slot j's part on machine m, its area over m's plate
    area, and its height-driven processing time on m.  The first two are
    scaled so that an even split of the parts over the machines sums to about
    1 per machine, the third by its largest value.  A machine's makespan is the
    sum over its batches of processing time, setup and the batch's largest
    height-driven time (``placement.py:86-89``), so its load and its area
    relative to the plate (about the number of batches it needs) are what a
    chromosome's makespan mostly depends on."""
    parts = [int(p) for p in instance_parts]
    M = int(n_machines)
    pt = np.array([[problem.machines[m].parts[p].proc_time for p in parts] for m in range(M)], np.float64)
    pth = np.array([[problem.machines[m].parts[p].proc_time_height for p in parts] for m in range(M)],
                   np.float64)
    area = np.array([problem.parts[p].area for p in parts], np.float64)
    plate = np.array([problem.machines[m].bin_area for m in range(M)], np.float64)
    t = np.zeros((M, 3, len(parts)), np.float64)
    t[:, 0] = pt / max(pt.mean(axis=0).sum() / M, 1e-300)
    t[:, 1] = (area[None, :] / plate[:, None]) / max((area.sum() / plate.mean()) / M, 1e-300)
    t[:, 2] = pth / max(pth.max(), 1e-300)
    return t.astype(np.float32)


def _stream_ptr(stream):
    return ctypes.c_void_p(stream.cuda_stream)


class MLPSurrogate:
    """ctypes handle of one surrogate network on one GPU."""

    def __init__(self, n_parts, slot_area, slot_height, thresholds=(), hidden=512, seed=0, lr=1e-3,
                 device=0, machine_terms=None):
        self._lib = _lib()
        sa = np.ascontiguousarray(slot_area, np.float32)
        sh = np.ascontiguousarray(slot_height, np.float32)
        thr = np.ascontiguousarray(thresholds, np.float64).reshape(-1)
        self.thresholds = thr.copy()
        self.n_machines = thr.size + 1
        mt = None
        if machine_terms is not None:
            mt = np.ascontiguousarray(machine_terms, np.float32)
            if mt.shape != (self.n_machines, 3, int(n_parts)):
                raise ValueError("machine_terms must be [n_machines][3][n_parts]")
        self.machine_terms = mt
        h = _P()
        rc = self._lib.brkga_mlp_create(int(n_parts), int(self.n_machines),
                                        thr.ctypes.data if thr.size else None, int(hidden),
                                        sa.ctypes.data, sh.ctypes.data, mt.ctypes.data if mt is not None else None,
                                        int(seed), float(lr), int(device), ctypes.byref(h))
        if rc != 0:
            raise DecoderError(rc, "brkga_mlp_create: " + self._lib.brkga_mlp_last_error().decode())
        self._h = h
        self.n_parts, self.hidden, self.device = int(n_parts), int(hidden), int(device)
        np_, fp = ctypes.c_int64(), ctypes.c_int32()
        self._lib.brkga_mlp_num_params(self._h, ctypes.byref(np_), ctypes.byref(fp))
        self.n_params, self.fp = np_.value, fp.value

    def _check(self, rc, what):
        if rc != 0:
            raise DecoderError(rc, "%s: %s" % (what, self._lib.brkga_mlp_last_error().decode()))

    def _stream(self, stream):
        import torch
        return stream or torch.cuda.current_stream(torch.device("cuda", self.device))

    def predict(self, keys, out=None, stream=None):
        import torch
        keys = keys.contiguous()
        if out is None:
            out = torch.empty(keys.shape[0], dtype=torch.float32, device=keys.device)
        s = self._stream(stream)
        self._check(self._lib.brkga_mlp_predict(self._h, keys.data_ptr(), keys.shape[0], out.data_ptr(),
                                                _stream_ptr(s)), "brkga_mlp_predict")
        return out

    def gradients(self, keys, targets, stream=None):
        import torch
        g = torch.empty(self.n_params, dtype=torch.float32, device=keys.device)
        s = self._stream(stream)
        self._check(self._lib.brkga_mlp_gradients(self._h, keys.contiguous().data_ptr(),
                                                  targets.contiguous().data_ptr(), keys.shape[0],
                                                  g.data_ptr(), _stream_ptr(s)), "brkga_mlp_gradients")
        return g

    def train_step(self, keys, targets, stream=None, want_loss=True):
        """One Adam step; returns the batch loss (a host sync) or None."""
        loss = ctypes.c_float()
        s = self._stream(stream)
        self._check(self._lib.brkga_mlp_train_step(self._h, keys.contiguous().data_ptr(),
                                                   targets.contiguous().data_ptr(), keys.shape[0],
                                                   ctypes.byref(loss) if want_loss else None,
                                                   _stream_ptr(s)),
                    "brkga_mlp_train_step")
        return loss.value if want_loss else None

    def features(self, keys, stream=None):
        import torch
        x = torch.empty((keys.shape[0], self.fp), dtype=torch.int16, device=keys.device)
        s = self._stream(stream)
        self._check(self._lib.brkga_mlp_features(self._h, keys.contiguous().data_ptr(), keys.shape[0],
                                                 x.data_ptr(), _stream_ptr(s)), "brkga_mlp_features")
        return x

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        self._check(self._lib.brkga_mlp_get_params(self._h, p.ctypes.data), "brkga_mlp_get_params")
        return p

    def set_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        self._check(self._lib.brkga_mlp_set_params(self._h, p.ctypes.data), "brkga_mlp_set_params")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.brkga_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _average_ranks(v):
    """0-based ranks, ties given their average rank (float64)."""
    srt = np.sort(v)
    return (np.searchsorted(srt, v, "left") + np.searchsorted(srt, v, "right") - 1) * 0.5


def spearman(a, b):
    """Spearman rank correlation (average ranks on ties); NaN for fewer than 3
    values or a constant input."""
    a, b = np.asarray(a), np.asarray(b)
    if a.size < 3 or np.all(a == a[0]) or np.all(b == b[0]):
        return float("nan")
    ra, rb = _average_ranks(a), _average_ranks(b)
    ra -= ra.mean()
    rb -= rb.mean()
    return float(ra @ rb / np.sqrt((ra @ ra) * (rb @ rb)))


class SurrogateBRKGA:
    """Device-resident BRKGA with optional surrogate screening (see module doc)."""

    def __init__(self, decoder, problem, instance_parts, num_individuals=1280, elite_frac=0.2,
                 mutant_frac=0.15, elite_prob=0.7, pool=4096, warmup=5, seed=0,
                 use_surrogate=True, hidden=512, lr=1e-3, train_steps=16, batch=1024,
                 buffer=16384, fallback_rho=0.2, audit_every=0, target="rank", aggregates=True):
        import torch
        self.torch = torch
        self.dec = decoder
        self.n = decoder.n_parts
        self.G2 = 2 * self.n
        self.N = int(num_individuals)
        self.ne = int(math.ceil(elite_frac * self.N))
        self.nm = int(math.ceil(mutant_frac * self.N))
        self.no = self.N - self.ne - self.nm
        self.rho = float(elite_prob)
        self.pool = max(int(pool), self.nm + self.no)
        self.pm = int(round(self.pool * self.nm / (self.nm + self.no)))
        self.po = self.pool - self.pm
        self.warmup = int(warmup)
        self.seed = int(seed)
        self.use_surrogate = bool(use_surrogate)
        self.train_steps, self.batch, self.buffer = int(train_steps), int(batch), int(buffer)
        self.fallback_rho = float(fallback_rho)
        self.audit_every = int(audit_every)
        if target not in ("rank", "zscore"):
            raise ValueError("target must be 'rank' or 'zscore'")
        self.target = target
        self.dev = torch.device("cuda", decoder.device)
        self.stream = torch.cuda.current_stream(self.dev)
        self._lib = _lib()
        sa, sh = slot_features(problem, instance_parts)
        mt = machine_terms(problem, instance_parts, decoder.n_machines) if aggregates else None
        self.net = MLPSurrogate(self.n, sa, sh, decoder.thresholds, hidden=hidden, seed=seed, lr=lr,
                                device=decoder.device, machine_terms=mt) if self.use_surrogate else None
        g = torch.Generator(device="cpu").manual_seed(self.seed)
        self.pop = torch.rand((self.N, self.G2), generator=g, dtype=torch.float32).to(self.dev)
        self.fit = torch.empty(self.N, dtype=torch.float64, device=self.dev)
        self.order = torch.empty(self.N, dtype=torch.int32, device=self.dev)
        self.poolbuf = torch.empty((self.pool, self.G2), dtype=torch.float32, device=self.dev)
        self.buf_x = torch.empty((self.buffer, self.G2), dtype=torch.float32, device=self.dev)
        self.buf_y = torch.empty(self.buffer, dtype=torch.float64, device=self.dev)
        self.buf_n = 0
        self.buf_pos = 0
        self.gen = 0
        self.rng = np.random.default_rng(self.seed + 1)
        self.log = {"min": [], "best": [], "time": [], "spearman": [], "screened": [],
                    "loss": [], "regret": [], "audit_spearman": [], "recall": []}
        self.decoder_calls = 0
        # device decodes cannot grow the bin workspace: a context that runs out
        # of bins decodes to NaN, so every run checks the overflow counter
        self._overflows = decoder.stats()["overflows"]
        self._decode(self.pop, self.fit)
        self._check_overflow()
        self._add_training(self.pop, self.fit)
        self.best = float(self.fit.min())

    def _decode(self, keys, out):
        self.dec.evaluate_device(keys, out, stream=self.stream, check_overflow=False)
        self.decoder_calls += keys.shape[0]

    def _check_overflow(self):
        ov = self.dec.stats()["overflows"]
        if ov != self._overflows:
            raise DecoderError(-4, "a context needed more bins than the decoder's max_bins (%d); create "
                                   "the decoder with max_bins=n_parts" % self.dec.max_bins)

    def _add_training(self, keys, fit, feasible=None):
        """Append the feasible (keys, fitness) pairs to the ring buffer.  With
        ``feasible`` (a host mask already at hand) no device-to-host read is needed."""
        torch = self.torch
        if feasible is None:
            ok = fit < 1e15
            k, f = keys[ok], fit[ok]
        elif feasible.all():
            k, f = keys, fit
        else:
            sel = torch.from_numpy(np.nonzero(feasible)[0]).to(self.dev)
            k, f = keys.index_select(0, sel), fit.index_select(0, sel)
        m = k.shape[0]
        if m == 0:
            return
        idx = (torch.arange(m, device=self.dev) + self.buf_pos) % self.buffer
        self.buf_x[idx] = k
        self.buf_y[idx] = f
        self.buf_pos = (self.buf_pos + m) % self.buffer
        self.buf_n = min(self.buffer, self.buf_n + m)

    def _train(self):
        torch = self.torch
        n = self.buf_n
        if n < 32:
            return float("nan")
        y = self.buf_y[:n]
        if self.target == "rank":
            # normal scores of the ranks: equal resolution for every quantile,
            # so the near-best region is not squeezed by bad outliers
            r = torch.empty(n, dtype=torch.float64, device=self.dev)
            r[torch.argsort(y, stable=True)] = torch.arange(n, dtype=torch.float64, device=self.dev)
            tt = (math.sqrt(2.0) * torch.erfinv(2.0 * (r + 0.5) / n - 1.0)).float()
        else:
            tt = ((y - y.mean()) / y.std().clamp_min(1e-12)).float()
        b = min(self.batch, n)
        idx_all = torch.from_numpy(self.rng.integers(0, n, size=(self.train_steps, b))).to(self.dev)
        # every step's batch gathered at once; step k trains on rows k*b .. k*b+b-1
        xs = self.buf_x[idx_all.reshape(-1)]
        ts = tt[idx_all.reshape(-1)]
        loss = float("nan")
        for k in range(self.train_steps):
            loss = self.net.train_step(xs[k * b:(k + 1) * b], ts[k * b:(k + 1) * b], stream=self.stream,
                                       want_loss=(k == self.train_steps - 1))
        return loss

    def step(self):
        torch = self.torch
        lib = self._lib
        s = _stream_ptr(self.stream)
        lib.brkga_sort_fitness(self.fit.data_ptr(), self.N, self.order.data_ptr(), s)
        screen = (self.use_surrogate and self.gen >= self.warmup and
                  not (self.log["spearman"] and self.log["spearman"][-1] < self.fallback_rho))
        pm, po = (self.pm, self.po) if screen else (self.nm, self.no)
        rc = lib.brkga_breed_pool(self.pop.data_ptr(), self.order.data_ptr(), self.N, self.G2, self.ne,
                                  pm, po, self.rho, self.seed, self.gen, self.poolbuf.data_ptr(), s)
        if rc != 0:
            raise DecoderError(rc, "brkga_breed_pool")
        if screen:
            pred = self.net.predict(self.poolbuf[:pm + po], stream=self.stream)
            sel_m = torch.topk(pred[:pm], self.nm, largest=False).indices
            sel_o = torch.topk(pred[pm:pm + po], self.no, largest=False).indices + pm
            chosen = torch.cat([self.poolbuf[sel_m], self.poolbuf[sel_o]])
            chosen_pred = torch.cat([pred[sel_m], pred[sel_o]])
        else:
            chosen = self.poolbuf[:self.nm + self.no].clone()
            chosen_pred = None
            if self.use_surrogate and self.gen >= 1:
                chosen_pred = self.net.predict(chosen, stream=self.stream)
        fit_c = torch.empty(chosen.shape[0], dtype=torch.float64, device=self.dev)
        self._decode(chosen, fit_c)
        if self.audit_every and screen and self.gen % self.audit_every == 0:
            self._audit(pred, pm, po, sel_m, sel_o)
        elite = self.order[:self.ne].long()
        self.pop = torch.cat([self.pop[elite], chosen])
        self.fit = torch.cat([self.fit[elite], fit_c])
        rho_s = float("nan")
        ok = None
        if chosen_pred is not None:
            pf = torch.stack([chosen_pred.double(), fit_c]).cpu().numpy()   # one copy to the host
            ok = pf[1] < 1e15
            rho_s = spearman(pf[0][ok], pf[1][ok])
        if self.use_surrogate:
            self.log["spearman"].append(rho_s)
            self._add_training(chosen, fit_c, ok)
            self.log["loss"].append(self._train())
        self.log["screened"].append(bool(screen))
        self.gen += 1
        m = float(self.fit.min())
        self.best = min(self.best, m)
        self.log["min"].append(m)
        self.log["best"].append(self.best)

    def _audit(self, pred, pm, po, sel_m, sel_o):
        """Exact decode of the whole pool: regret of the screened pick and pool-wide rho."""
        torch = self.torch
        allf = torch.empty(pm + po, dtype=torch.float64, device=self.dev)
        self.dec.evaluate_device(self.poolbuf[:pm + po].contiguous(), allf, stream=self.stream,
                                 check_overflow=False)
        f = allf.cpu().numpy()
        chosen = np.concatenate([sel_m.cpu().numpy(), sel_o.cpu().numpy()])
        self.log["regret"].append(float(f[chosen].min() - f.min()))
        # share of the pool's true best len(chosen) that the screen picked
        best = np.argsort(f, kind="stable")[:chosen.size]
        self.log["recall"].append(float(np.intersect1d(best, chosen).size) / chosen.size)
        ok = f < 1e15
        self.log["audit_spearman"].append(spearman(pred.cpu().numpy()[ok], f[ok]))

    def run(self, generations, target=None):
        """Run up to ``generations``; stop early once best <= target.  Returns seconds."""
        t0 = time.perf_counter()
        self.time_to_target = None
        if target is not None and self.best <= target:
            self.time_to_target = 0.0
            return 0.0
        for _ in range(generations):
            self.step()
            self.log["time"].append(time.perf_counter() - t0)
            if target is not None and self.best <= target:
                self.time_to_target = time.perf_counter() - t0
                break
        elapsed = time.perf_counter() - t0
        self._check_overflow()
        return elapsed

    def close(self):
        if self.net is not None:
            self.net.close()


def time_to_target(decoder, problem, instance_parts, generations=100, seeds=10, target_seeds=5,
                   audit_every=10, **kw):
    """Wall time to a target makespan, plain vs surrogate-screened BRKGA.

    Both methods share everything but the screening (``SurrogateBRKGA`` with
    ``use_surrogate`` False / True), run ``generations`` generations for each
    seed and log the best exact makespan after every generation.  The target
    is the median, over seeds ``0 .. target_seeds-1``, of the best makespan
    plain BRKGA holds after ``generations // 2`` generations (``final_target``
    repeats the comparison for the median of plain's best after the last
    generation, a target plain reaches only late).  A run's time to
    target is the wall time (from its first generation; set-up and the initial
    population's decode excluded) of the first generation whose best reaches
    the target, infinite when none does.  One extra surrogate run (seed 0,
    not timed) decodes the whole pool every ``audit_every`` screened
    generations for the regret and pool-wide rank correlation.
    """
    import torch
    w = SurrogateBRKGA(decoder, problem, instance_parts, seed=10 ** 6, warmup=1, **kw)
    w.run(3)   # first-call set-up of every kernel, outside the timed runs
    w.close()
    torch.cuda.synchronize()
    runs = {"plain": [], "surrogate": []}
    dec_stats = {}
    for kind in runs:
        s0 = decoder.stats()
        for seed in range(seeds):
            ga = SurrogateBRKGA(decoder, problem, instance_parts, seed=seed,
                                use_surrogate=(kind == "surrogate"), **kw)
            ga.run(generations)
            runs[kind].append({"best": list(ga.log["best"]), "time": list(ga.log["time"]),
                               "spearman": list(ga.log["spearman"]),
                               "screened": list(ga.log["screened"])})
            ga.close()
        s1 = decoder.stats()
        dec_stats[kind] = {k: s1[k] - s0[k] for k in ("kernel_ms", "mfma_ops", "contexts")}
    half = max(1, generations // 2)
    target = float(np.median([r["best"][half - 1] for r in runs["plain"][:target_seeds]]))

    # a second, harder target: the median of plain's final best over the same seeds
    target_final = float(np.median([r["best"][-1] for r in runs["plain"][:target_seeds]]))

    def ttt(r, tg=None):
        b = np.asarray(r["best"])
        hit = np.nonzero(b <= (target if tg is None else tg))[0]
        return float(r["time"][hit[0]]) if hit.size else float("inf")

    out = {"target": target, "generations": generations, "seeds": seeds, "target_seeds": min(target_seeds, seeds)}
    for kind, rs in runs.items():
        t = [ttt(r) for r in rs]
        med = float(np.median(t))
        out[kind] = {"median_s": med if math.isfinite(med) else None,
                     "per_seed_s": [x if math.isfinite(x) else None for x in t],
                     "reached": int(np.sum(np.isfinite(t))),
                     "ms_per_generation": 1e3 * float(np.mean([r["time"][-1] / generations for r in rs])),
                     "final_best_mean": float(np.mean([r["best"][-1] for r in rs]))}
        ds = dec_stats[kind]
        # the decode kernel's share of a generation and its issued MFMA rate
        out[kind]["decode_kernel_ms_per_generation"] = ds["kernel_ms"] / max(1, seeds * generations)
        out[kind]["decode_issued_mfma_tops"] = ds["mfma_ops"] / max(ds["kernel_ms"], 1e-9) / 1e9
    sp = [x for r in runs["surrogate"] for x, sc in zip(r["spearman"], r["screened"])
          if sc and math.isfinite(x)]
    out["surrogate"]["spearman_decoded_mean"] = float(np.mean(sp)) if sp else None
    a, b = out["plain"]["median_s"], out["surrogate"]["median_s"]
    out["speedup"] = (a / b) if (a is not None and b) else None
    fin = {"target": target_final}
    for kind, rs in runs.items():
        t = [ttt(r, target_final) for r in rs]
        med = float(np.median(t))
        fin[kind] = {"median_s": med if math.isfinite(med) else None, "reached": int(np.sum(np.isfinite(t)))}
    a, b = fin["plain"]["median_s"], fin["surrogate"]["median_s"]
    fin["speedup"] = (a / b) if (a is not None and b) else None
    out["final_target"] = fin
    if audit_every:
        ga = SurrogateBRKGA(decoder, problem, instance_parts, seed=0, audit_every=audit_every, **kw)
        ga.run(generations)
        out["surrogate"]["regret_mean"] = float(np.mean(ga.log["regret"])) if ga.log["regret"] else None
        out["surrogate"]["regret_max"] = float(np.max(ga.log["regret"])) if ga.log["regret"] else None
        ps = [x for x in ga.log["audit_spearman"] if math.isfinite(x)]
        out["surrogate"]["spearman_pool_mean"] = float(np.mean(ps)) if ps else None
        out["surrogate"]["recall_mean"] = float(np.mean(ga.log["recall"])) if ga.log["recall"] else None
        out["surrogate"]["audits"] = len(ga.log["regret"])
        ga.close()
    out["runs"] = runs
    return out
