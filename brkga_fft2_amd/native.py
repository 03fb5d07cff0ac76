"""ctypes binding of the C ABI in ``include/brkga_mi355x.h``.

The library is loaded from ``brkga_fft2_amd/_lib/libbrkga_mi355x.so``.  There
is no fallback: if the library is missing or the device is unavailable the
calls raise, naming what is missing.
"""
import ctypes
import os

import numpy as np

from .build import LIB_PATH

BRKGA_KEYS_F32 = 0
BRKGA_KEYS_F64 = 1
INFEASIBLE = 1e16
_ERRORS = {-1: "EINVAL", -2: "EDEVICE", -3: "ENOMEM", -4: "EOVERFLOW", -5: "ESTATE"}

_P = ctypes.c_void_p


class BrkgaProblem(ctypes.Structure):
    _fields_ = [
        ("n_parts", ctypes.c_int32), ("n_machines", ctypes.c_int32),
        ("n_types", ctypes.c_int32), ("n_rots", ctypes.c_int32),
        ("thresholds", _P), ("slot_type", _P), ("type_id", _P), ("area", _P),
        ("nrot", _P), ("best_rot", _P), ("rot_base", _P), ("rot_h", _P), ("rot_w", _P),
        ("mask_off", _P), ("masks", _P), ("plate_h", _P), ("plate_w", _P), ("setup", _P),
        ("proc_time", _P), ("proc_time_height", _P),
    ]


class BrkgaOptions(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("wgs_per_cu", ctypes.c_int32),
                ("max_bins", ctypes.c_int32), ("dense", ctypes.c_int32),
                ("tiles_per_pass", ctypes.c_int32), ("mfma_format", ctypes.c_int32),
                ("screen_rows", ctypes.c_int32), ("reserved", ctypes.c_int32)]


MFMA_FORMATS = {"fp4": 0, "int8": 1}   # BRKGA_MFMA_FP4, BRKGA_MFMA_INT8


class BrkgaStats(ctypes.Structure):
    _fields_ = [("calls", ctypes.c_int64), ("chromosomes", ctypes.c_int64),
                ("contexts", ctypes.c_int64), ("placements", ctypes.c_int64),
                ("bins_opened", ctypes.c_int64), ("tests", ctypes.c_int64),
                ("row_groups", ctypes.c_int64), ("mfma_ops", ctypes.c_int64),
                ("algo_ops", ctypes.c_int64), ("overflows", ctypes.c_int64),
                ("kernel_ms", ctypes.c_double), ("last_kernel_ms", ctypes.c_double),
                ("phase_us", ctypes.c_double * 8), ("context_us_max", ctypes.c_double),
                ("context_us_sum", ctypes.c_double), ("growths", ctypes.c_int64),
                ("max_bins", ctypes.c_int32), ("n_workgroups", ctypes.c_int32),
                ("machine_contexts", ctypes.c_int64 * 64), ("machine_bins", ctypes.c_int64 * 64),
                ("machine_peak_bins", ctypes.c_int32 * 64)]

MAX_MACHINES = 64   # BRKGA_MAX_MACHINES

PHASES = ["sequence", "vacancy_filter", "band_staging", "correlation", "screen",
          "stamp_update", "new_bin", "makespan"]


class BrkgaGaParams(ctypes.Structure):
    _fields_ = [("num_individuals", ctypes.c_int32), ("num_elites", ctypes.c_int32),
                ("num_mutants", ctypes.c_int32), ("history_capacity", ctypes.c_int32),
                ("elite_prob", ctypes.c_double), ("seed", ctypes.c_uint64)]


TILE_CLASSES = 16   # BRKGA_TILE_CLASSES: mask width classes nc = ceil((w + 31) / 32)

EXPORTED = ["brkga_create", "brkga_evaluate", "brkga_evaluate_device", "brkga_get_stats",
            "brkga_reset_stats", "brkga_set_tiles", "brkga_get_tiles", "brkga_set_screen",
            "brkga_get_screen", "brkga_initial_solution",
            "brkga_evaluate_device_noalloc",
            "brkga_last_error",
            "brkga_destroy",
            "brkga_ga_create", "brkga_ga_run", "brkga_ga_best", "brkga_ga_population",
            "brkga_ga_last_error", "brkga_ga_destroy", "brkga_sort_fitness", "brkga_breed_pool",
            "brkga_mlp_create", "brkga_mlp_predict", "brkga_mlp_gradients", "brkga_mlp_train_step",
            "brkga_mlp_num_params", "brkga_mlp_get_params", "brkga_mlp_set_params",
            "brkga_mlp_features", "brkga_mlp_last_error", "brkga_mlp_destroy"]

_lib = None


def load_library(path=LIB_PATH):
    """Load the decoder library (cached).  Raises OSError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError("decoder library %s is not built; run __graft_entry__.build() "
                      "or python -m brkga_fft2_amd.build" % path)
    # torch ships its own HIP runtime.  Loading torch first makes the library
    # bind to that same runtime (one runtime per process), so device pointers
    # and streams of torch tensors are valid inside the library.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    PP = ctypes.POINTER(ctypes.c_void_p)
    lib.brkga_create.argtypes = [ctypes.POINTER(BrkgaProblem), ctypes.POINTER(BrkgaOptions), PP]
    lib.brkga_create.restype = ctypes.c_int
    lib.brkga_evaluate.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, _P, _P, _P]
    lib.brkga_evaluate.restype = ctypes.c_int
    lib.brkga_evaluate_device.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                          _P, _P, _P, ctypes.c_int]
    lib.brkga_evaluate_device.restype = ctypes.c_int
    lib.brkga_evaluate_device_noalloc.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int64, ctypes.c_int64,
                                                  _P, _P, _P, ctypes.c_int]
    lib.brkga_evaluate_device_noalloc.restype = ctypes.c_int
    lib.brkga_get_stats.argtypes = [_P, ctypes.POINTER(BrkgaStats)]
    lib.brkga_get_stats.restype = ctypes.c_int
    lib.brkga_reset_stats.argtypes = [_P]
    lib.brkga_reset_stats.restype = ctypes.c_int
    lib.brkga_set_tiles.argtypes = [_P, _P, _P]
    lib.brkga_set_tiles.restype = ctypes.c_int
    lib.brkga_get_tiles.argtypes = [_P, _P, _P]
    lib.brkga_get_tiles.restype = ctypes.c_int
    lib.brkga_set_screen.argtypes = [_P, _P]
    lib.brkga_set_screen.restype = ctypes.c_int
    lib.brkga_get_screen.argtypes = [_P, _P]
    lib.brkga_get_screen.restype = ctypes.c_int
    lib.brkga_initial_solution.argtypes = [_P, _P, _P, _P, _P, _P, _P]
    lib.brkga_initial_solution.restype = ctypes.c_int
    lib.brkga_last_error.argtypes = [_P]
    lib.brkga_last_error.restype = ctypes.c_char_p
    lib.brkga_destroy.argtypes = [_P]
    lib.brkga_destroy.restype = None
    lib.brkga_ga_create.argtypes = [_P, ctypes.POINTER(BrkgaGaParams), _P, PP]
    lib.brkga_ga_create.restype = ctypes.c_int
    lib.brkga_ga_run.argtypes = [_P, ctypes.c_int32, ctypes.c_int, _P]
    lib.brkga_ga_run.restype = ctypes.c_int
    lib.brkga_ga_best.argtypes = [_P, _P, _P, _P, _P]
    lib.brkga_ga_best.restype = ctypes.c_int
    lib.brkga_ga_population.argtypes = [_P, _P, _P]
    lib.brkga_ga_population.restype = ctypes.c_int
    lib.brkga_ga_last_error.argtypes = []
    lib.brkga_ga_last_error.restype = ctypes.c_char_p
    lib.brkga_ga_destroy.argtypes = [_P]
    lib.brkga_ga_destroy.restype = None
    _lib = lib
    return lib


_ARRAYS = [("thresholds", np.float64), ("slot_type", np.int32), ("type_id", np.int32),
           ("area", np.float64), ("nrot", np.int32), ("best_rot", np.int32),
           ("rot_base", np.int32), ("rot_h", np.int32), ("rot_w", np.int32),
           ("mask_off", np.int64), ("masks", np.uint8), ("plate_h", np.int32),
           ("plate_w", np.int32), ("setup", np.float64), ("proc_time", np.float64),
           ("proc_time_height", np.float64)]


class DecoderError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%s)" % (msg, _ERRORS.get(code, code)))
        self.code = code


class Decoder:
    """One decoder handle: a problem uploaded to one GPU."""

    def __init__(self, packed, device=0, dense=False, tiles_per_pass=0, wgs_per_cu=0,
                 max_bins=0, mfma="fp4", screen_rows=0):
        self._lib = load_library()
        self.n_parts = int(packed["n_parts"])
        self.n_machines = int(packed["n_machines"])
        self.n_types = int(packed["n_types"])
        self.thresholds = np.asarray(packed["thresholds"], np.float64).reshape(-1)[:max(self.n_machines - 1, 0)]
        keep = []
        prob = BrkgaProblem(self.n_parts, self.n_machines, int(packed["n_types"]),
                            int(packed["n_rots"]))
        for name, dt in _ARRAYS:
            a = np.ascontiguousarray(packed[name], dtype=dt)
            keep.append(a)
            setattr(prob, name, a.ctypes.data if a.size else None)
        if mfma not in MFMA_FORMATS:
            raise ValueError("mfma must be one of %s" % sorted(MFMA_FORMATS))
        self.mfma = mfma
        opt = BrkgaOptions(int(device), int(wgs_per_cu), int(max_bins), int(bool(dense)),
                           int(tiles_per_pass), MFMA_FORMATS[mfma], int(screen_rows))
        h = ctypes.c_void_p()
        rc = self._lib.brkga_create(ctypes.byref(prob), ctypes.byref(opt), ctypes.byref(h))
        if rc != 0:
            raise DecoderError(rc, "brkga_create: " + self._lib.brkga_last_error(None).decode())
        self._h = h
        self.device = int(device)
        # bins per context in the workspace (brkga_evaluate grows it on demand;
        # device-side loops cannot, and need max_bins = n_parts to be safe)
        self.max_bins = int(max_bins) if max_bins > 0 else min(self.n_parts, 32)

    def _check(self, rc, what):
        if rc != 0:
            raise DecoderError(rc, "%s: %s" % (what, self._lib.brkga_last_error(self._h).decode()))

    def evaluate(self, keys, with_log=False):
        """Decode host keys ``[N, 2n]`` (float32 or float64).

        Returns ``(makespans float64[N], log int32[N, n, 6] or None)``.
        """
        keys = np.asarray(keys)
        if keys.dtype not in (np.float32, np.float64):
            keys = keys.astype(np.float32)
        keys = np.ascontiguousarray(keys)
        if keys.ndim != 2:
            raise ValueError("keys must be a 2-D array [N, 2*n_parts]")
        n = keys.shape[0]
        out = np.empty(n, np.float64)
        log = np.empty((n, self.n_parts, 6), np.int32) if with_log else None
        if n == 0:
            return out, log
        dt = BRKGA_KEYS_F64 if keys.dtype == np.float64 else BRKGA_KEYS_F32
        rc = self._lib.brkga_evaluate(self._h, keys.ctypes.data, dt, n, keys.shape[1],
                                      out.ctypes.data, log.ctypes.data if with_log else None,
                                      None)
        self._check(rc, "brkga_evaluate")
        return out, log

    def evaluate_device(self, keys, out, log=None, stream=None, check_overflow=True):
        """Decode device-resident keys (torch tensors on this GPU) into ``out`` (float64)."""
        import torch
        if keys.device.type != "cuda" or out.device.type != "cuda":
            raise ValueError("evaluate_device needs GPU tensors")
        if not keys.is_contiguous() or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError("keys must be contiguous; out must be contiguous float64")
        if keys.dtype not in (torch.float32, torch.float64):
            raise ValueError("keys must be float32 or float64")
        if keys.dim() != 2 or out.dim() != 1 or out.shape[0] != keys.shape[0]:
            raise ValueError("keys must be [N, 2*n_parts] and out [N]")
        if keys.shape[0] == 0:
            return out
        dt = BRKGA_KEYS_F64 if keys.dtype == torch.float64 else BRKGA_KEYS_F32
        if stream is None:
            stream = torch.cuda.current_stream(keys.device)
        rc = self._lib.brkga_evaluate_device(
            self._h, keys.data_ptr(), dt, keys.shape[0], keys.shape[1], out.data_ptr(),
            log.data_ptr() if log is not None else None, ctypes.c_void_p(stream.cuda_stream),
            int(bool(check_overflow)))
        self._check(rc, "brkga_evaluate_device")
        return out

    def evaluate_device_noalloc(self, keys, out, log=None, stream=None, check_overflow=False):
        """``evaluate_device`` through the allocation-free entry point: nothing is
        allocated (the per-call buffers are the library's static scratch on this
        GPU, for up to 65536 (chromosome, machine) contexts), so it may run inside
        a captured stream; calls on one GPU must not overlap."""
        import torch
        if keys.device.type != "cuda" or out.device.type != "cuda" or keys.dtype not in (torch.float32, torch.float64):
            raise ValueError("keys (float32/float64) and out must be GPU tensors")
        if (not keys.is_contiguous() or keys.dim() != 2 or out.dtype != torch.float64 or not out.is_contiguous()
                or out.dim() != 1 or out.shape[0] != keys.shape[0]):
            raise ValueError("keys must be contiguous [N, 2*n_parts] and out contiguous float64 [N]")
        if keys.shape[0] == 0:
            return out
        dt = BRKGA_KEYS_F64 if keys.dtype == torch.float64 else BRKGA_KEYS_F32
        if stream is None:
            stream = torch.cuda.current_stream(keys.device)
        rc = self._lib.brkga_evaluate_device_noalloc(
            self._h, keys.data_ptr(), dt, keys.shape[0], keys.shape[1], out.data_ptr(),
            log.data_ptr() if log is not None else None, ctypes.c_void_p(stream.cuda_stream),
            int(bool(check_overflow)))
        self._check(rc, "brkga_evaluate_device_noalloc")
        return out

    def stats(self):
        st = BrkgaStats()
        self._check(self._lib.brkga_get_stats(self._h, ctypes.byref(st)), "brkga_get_stats")
        d = {k: getattr(st, k) for k, _ in BrkgaStats._fields_}
        d["phase_us"] = dict(zip(PHASES, list(st.phase_us)))
        M = self.n_machines
        for k in ("machine_contexts", "machine_bins", "machine_peak_bins"):
            d[k] = list(getattr(st, k))[:M]
        return d

    def bin_stats(self):
        """Per machine ``(peak_bins_per_sol, total_bins, total_sols, num_batches,
        peak_slots, num_growth_events)``, the tuple of the reference's
        ``get_bin_stats`` (``full_native_decoder.py:2286-2301``): the most bins one
        context of the machine opened, the bins its contexts opened, its
        contexts, the decode calls, the bin slots one context may use (each
        persistent workgroup reuses its own ``max_bins`` slots), and the
        workspace growths."""
        d = self.stats()
        return [(d["machine_peak_bins"][m], d["machine_bins"][m], d["machine_contexts"][m], d["calls"],
                 d["max_bins"], d["growths"]) for m in range(self.n_machines)]

    def reset_stats(self):
        self._check(self._lib.brkga_reset_stats(self._h), "brkga_reset_stats")

    def initial_solution(self, type_height, machine_keys, with_log=False, stream=None):
        """Greedy seed chromosome on the GPU: ``(keys float32[2n], makespan,
        log int32[n, 6] or None)`` (see ``brkga_initial_solution``)."""
        th = np.ascontiguousarray(type_height, np.float64)
        mk = np.ascontiguousarray(machine_keys, np.float64)
        if th.size != self.n_types or mk.size != self.n_machines:
            raise ValueError("type_height needs n_types values and machine_keys n_machines values")
        keys = np.zeros(2 * self.n_parts, np.float32)
        ms = ctypes.c_double()
        log = np.full((self.n_parts, 6), -1, np.int32) if with_log else None
        self._check(self._lib.brkga_initial_solution(self._h, th.ctypes.data, mk.ctypes.data, keys.ctypes.data,
                                                     ctypes.byref(ms), log.ctypes.data if with_log else None,
                                                     ctypes.c_void_p(stream.cuda_stream) if stream is not None
                                                     else None), "brkga_initial_solution")
        return keys, ms.value, log

    def get_tiles(self):
        """Tile table ``(tiles_per_pass, window_tiles)``, int32 ``[n_machines, TILE_CLASSES]``."""
        t = np.zeros((self.n_machines, TILE_CLASSES), np.int32)
        w = np.zeros((self.n_machines, TILE_CLASSES), np.int32)
        self._check(self._lib.brkga_get_tiles(self._h, t.ctypes.data, w.ctypes.data), "brkga_get_tiles")
        return t, w

    def set_tiles(self, tiles_per_pass=None, window_tiles=None):
        """Set the tile table (either may be None = unchanged); results never change."""
        def arr(a):
            if a is None:
                return None, None
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.int32),
                                                     (self.n_machines, TILE_CLASSES)))
            return a, a.ctypes.data
        t, tp = arr(tiles_per_pass)
        w, wp = arr(window_tiles)
        self._check(self._lib.brkga_set_tiles(self._h, tp, wp), "brkga_set_tiles")

    def get_screen(self):
        """Screening rows per tile class, int32 ``[n_machines, TILE_CLASSES]`` (0: none)."""
        s = np.zeros((self.n_machines, TILE_CLASSES), np.int32)
        self._check(self._lib.brkga_get_screen(self._h, s.ctypes.data), "brkga_get_screen")
        return s

    def set_screen(self, screen_rows):
        """Set the screening rows (a scalar or ``[n_machines, TILE_CLASSES]``, 0..16); results
        never change."""
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(screen_rows, np.int32),
                                                 (self.n_machines, TILE_CLASSES)))
        self._check(self._lib.brkga_set_screen(self._h, s.ctypes.data), "brkga_set_screen")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.brkga_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGA:
    """Device-resident BRKGA loop over a :class:`Decoder` (one hipGraph per generation)."""

    def __init__(self, decoder, num_individuals, num_elites, num_mutants, elite_prob=0.7, seed=0,
                 history_capacity=1024, init_population=None):
        self._lib = load_library()
        self.decoder = decoder
        self.n_genes = 2 * decoder.n_parts
        self.num_individuals = int(num_individuals)
        p = BrkgaGaParams(int(num_individuals), int(num_elites), int(num_mutants),
                          int(history_capacity), float(elite_prob), int(seed))
        init = None
        if init_population is not None:
            init = np.ascontiguousarray(init_population, dtype=np.float32)
            if init.shape != (self.num_individuals, self.n_genes):
                raise ValueError("init_population must be [num_individuals, 2*n_parts]")
        h = ctypes.c_void_p()
        rc = self._lib.brkga_ga_create(decoder._h, ctypes.byref(p),
                                       init.ctypes.data if init is not None else None,
                                       ctypes.byref(h))
        if rc != 0:
            raise DecoderError(rc, "brkga_ga_create: " + self._lib.brkga_ga_last_error().decode())
        self._h = h

    def run(self, generations, use_graph=True, stream=None):
        """Run generations on ``stream`` (a torch stream).  Graph replay needs a
        non-null stream; by default the loop uses a stream of its own."""
        import torch
        if stream is None:
            if getattr(self, "_stream", None) is None:
                self._stream = torch.cuda.Stream(device=torch.device("cuda", self.decoder.device))
            stream = self._stream
        rc = self._lib.brkga_ga_run(self._h, int(generations), int(bool(use_graph)),
                                    ctypes.c_void_p(stream.cuda_stream))
        if rc != 0:
            raise DecoderError(rc, "brkga_ga_run: " + self._lib.brkga_ga_last_error().decode())

    def best(self):
        """``(best_fitness, best_keys float32[2n], history float64[done+1], generations_done)``."""
        bf = ctypes.c_double()
        keys = np.empty(self.n_genes, np.float32)
        done = ctypes.c_int32()
        rc = self._lib.brkga_ga_best(self._h, None, None, None, ctypes.byref(done))
        if rc != 0:
            raise DecoderError(rc, "brkga_ga_best: " + self._lib.brkga_ga_last_error().decode())
        hist = np.empty(done.value + 1, np.float64)
        rc = self._lib.brkga_ga_best(self._h, ctypes.byref(bf), keys.ctypes.data, hist.ctypes.data,
                                     ctypes.byref(done))
        if rc != 0:
            raise DecoderError(rc, "brkga_ga_best: " + self._lib.brkga_ga_last_error().decode())
        return bf.value, keys, hist, done.value

    def population(self):
        keys = np.empty((self.num_individuals, self.n_genes), np.float32)
        fit = np.empty(self.num_individuals, np.float64)
        rc = self._lib.brkga_ga_population(self._h, keys.ctypes.data, fit.ctypes.data)
        if rc != 0:
            raise DecoderError(rc, "brkga_ga_population: " + self._lib.brkga_ga_last_error().decode())
        return keys, fit

    def close(self):
        if getattr(self, "_h", None):
            self._lib.brkga_ga_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
