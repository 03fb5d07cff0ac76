"""ctypes front end of the CPU oracle (``pp_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, as the checker or the timed CPU
baseline.  The product decoder never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpp_oracle.so")
_lib = None


class _Problem(ctypes.Structure):
    _fields_ = [
        ("n_parts", ctypes.c_int32), ("n_machines", ctypes.c_int32),
        ("n_types", ctypes.c_int32), ("n_rots", ctypes.c_int32),
        ("thresholds", ctypes.c_void_p), ("slot_type", ctypes.c_void_p),
        ("type_id", ctypes.c_void_p), ("area", ctypes.c_void_p),
        ("nrot", ctypes.c_void_p), ("best_rot", ctypes.c_void_p),
        ("rot_base", ctypes.c_void_p), ("rot_h", ctypes.c_void_p),
        ("rot_w", ctypes.c_void_p), ("mask_off", ctypes.c_void_p),
        ("masks", ctypes.c_void_p), ("plate_h", ctypes.c_void_p),
        ("plate_w", ctypes.c_void_p), ("setup", ctypes.c_void_p),
        ("proc_time", ctypes.c_void_p), ("proc_time_height", ctypes.c_void_p),
    ]


class _Stats(ctypes.Structure):
    _fields_ = [("tests", ctypes.c_int64), ("positions", ctypes.c_int64),
                ("bins_opened", ctypes.c_int64), ("vacancy_rejects", ctypes.c_int64)]


def build(force=False):
    """Compile ``libpp_oracle.so`` with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "libpp_oracle.so"],
                       check=True, stdout=subprocess.DEVNULL)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        lib.oracle_decode.argtypes = [ctypes.POINTER(_Problem), ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_int, ctypes.POINTER(_Stats)]
        lib.oracle_decode.restype = ctypes.c_int
        lib.oracle_overlap_count.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int]
        lib.oracle_overlap_count.restype = ctypes.c_int64
        lib.oracle_bottom_left_bruteforce.argtypes = [
            ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        lib.oracle_bottom_left_bruteforce.restype = ctypes.c_int
        lib.oracle_initial_solution.argtypes = [ctypes.POINTER(_Problem), ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p]
        lib.oracle_initial_solution.restype = ctypes.c_int
        _lib = lib
    return _lib


_ARRAYS = [("thresholds", np.float64), ("slot_type", np.int32), ("type_id", np.int32),
           ("area", np.float64), ("nrot", np.int32), ("best_rot", np.int32),
           ("rot_base", np.int32), ("rot_h", np.int32), ("rot_w", np.int32),
           ("mask_off", np.int64), ("masks", np.uint8), ("plate_h", np.int32),
           ("plate_w", np.int32), ("setup", np.float64), ("proc_time", np.float64),
           ("proc_time_height", np.float64)]


def _problem(packed):
    keep = {}
    s = _Problem(packed["n_parts"], packed["n_machines"], packed["n_types"], packed["n_rots"])
    for name, dt in _ARRAYS:
        a = np.ascontiguousarray(packed[name], dtype=dt)
        keep[name] = a
        setattr(s, name, a.ctypes.data if a.size else None)
    return s, keep


def decode(packed, keys, n_threads=0, with_log=False):
    """Decode a population with the oracle.

    Returns ``(makespans float64[N], log int32[N, n_parts, 6] or None, stats dict)``;
    log rows are ``(machine, part_id, bin, col, row, rot)`` in placement order.
    """
    lib = _load()
    keys = np.ascontiguousarray(keys)
    if keys.dtype not in (np.float32, np.float64):
        keys = keys.astype(np.float32)
    if keys.ndim != 2 or keys.shape[1] != 2 * packed["n_parts"]:
        raise ValueError("keys must be [N, 2*n_parts]")
    n = keys.shape[0]
    out = np.zeros(n, np.float64)
    log = np.zeros((n, packed["n_parts"], 6), np.int32) if with_log else None
    prob, keep = _problem(packed)
    st = _Stats()
    rc = lib.oracle_decode(ctypes.byref(prob), keys.ctypes.data, int(keys.dtype == np.float64),
                           n, out.ctypes.data, log.ctypes.data if with_log else None,
                           int(n_threads), ctypes.byref(st))
    del keep
    if rc != 0:
        raise RuntimeError("oracle_decode failed with status %d" % rc)
    stats = {k: int(getattr(st, k)) for k, _ in _Stats._fields_}
    return out, log, stats


def overlap_count(grid, mask, r, c):
    lib = _load()
    g = np.ascontiguousarray(grid, np.uint8)
    m = np.ascontiguousarray(mask, np.uint8)
    return int(lib.oracle_overlap_count(g.ctypes.data, g.shape[0], g.shape[1], m.ctypes.data,
                                        m.shape[0], m.shape[1], r, c))


def bottom_left_bruteforce(grid, mask):
    """``(has, col, row)`` as collision_backend.py:109-128 reports it."""
    lib = _load()
    g = np.ascontiguousarray(grid, np.uint8)
    m = np.ascontiguousarray(mask, np.uint8)
    col, row = ctypes.c_int(), ctypes.c_int()
    has = lib.oracle_bottom_left_bruteforce(g.ctypes.data, g.shape[0], g.shape[1], m.ctypes.data,
                                            m.shape[0], m.shape[1], ctypes.byref(col),
                                            ctypes.byref(row))
    return (True, col.value, row.value) if has else (False, None, None)


def initial_solution(packed, order, machine_keys, with_log=False):
    """Greedy initial solution (BRKGA_alg3.py:445-603): ``(keys float32[2n],
    makespan, log [n, 6] or None)`` for the slot order ``order`` and the
    per-machine machine keys ``machine_keys``."""
    lib = _load()
    s, keep = _problem(packed)
    n = int(packed["n_parts"])
    order = np.ascontiguousarray(order, np.int32)
    mk = np.ascontiguousarray(machine_keys, np.float64)
    keys = np.zeros(2 * n, np.float32)
    ms = ctypes.c_double()
    log = np.full((n, 6), -1, np.int32) if with_log else None
    rc = lib.oracle_initial_solution(ctypes.byref(s), order.ctypes.data, mk.ctypes.data,
                                     keys.ctypes.data, ctypes.byref(ms),
                                     log.ctypes.data if with_log else None)
    del keep
    if rc != 0:
        raise RuntimeError("oracle_initial_solution: a part fits no machine")
    return keys, ms.value, log
