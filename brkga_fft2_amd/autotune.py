"""Tile autotuner for the decoder's collision passes.

The exact correlation runs in passes of up to 8 MFMA x-tiles (32 placements
each) over LDS windows of up to 64 x-tiles (``csrc/brkga_decoder.hip``).  The
best pass and window sizes depend on the mask width (how many 32-column
Toeplitz chunks, ceil((w + 31) / 32), feed each tile, hence register pressure) and on the plate width (how
many tiles a row has).  ``autotune`` times the decoder on a sample population
for each (plate size, mask width class) entry of the tile table and keeps the
fastest setting, coordinate by coordinate; ``apply_cached`` reloads a result
from a JSON cache keyed by the instance's shape signature and the GPU name.

Every table entry gives bit-identical results (the tiling only changes which
MFMAs compute the same exact counts); ``autotune`` re-checks that on the
sample before returning.
"""
import hashlib
import json
import os

import numpy as np

from .native import TILE_CLASSES

TILE_CHOICES = (1, 2, 4, 8)
WINDOW_CHOICES = (4, 8, 16, 32, 64)
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "brkga_mi355x", "tiles.json")


def chunk_class(w):
    """Tile-table class of a mask w columns wide: the Toeplitz chunk count
    ceil((w + 31) / 32) the kernel uses, clamped to the last class."""
    return int(min((int(w) + 62) // 32, TILE_CLASSES - 1))


def _default_tiles(nc):
    nc = max(nc, 1)
    return 8 if nc <= 5 else (4 if nc <= 8 else 2)


def tile_classes(packed):
    """``{(plate_h, plate_w): sorted list of width classes}`` present in the instance."""
    ncs = {chunk_class(w) for w in np.asarray(packed["rot_w"], np.int64)}
    out = {}
    for m in range(int(packed["n_machines"])):
        key = (int(packed["plate_h"][m]), int(packed["plate_w"][m]))
        out.setdefault(key, set()).update(ncs)
    return {k: sorted(v) for k, v in out.items()}


def signature(packed, population, device_name=""):
    """Shape signature of an instance: plates, rotation sizes, population, GPU."""
    h = hashlib.sha1()
    for name in ("plate_h", "plate_w", "rot_h", "rot_w"):
        h.update(np.ascontiguousarray(packed[name], np.int64).tobytes())
    h.update(str((int(packed["n_machines"]), int(population), device_name)).encode())
    return h.hexdigest()[:20]


def _machines_of(packed, plate):
    return [m for m in range(int(packed["n_machines"]))
            if (int(packed["plate_h"][m]), int(packed["plate_w"][m])) == plate]


def _time(dec, keys, out, stream, reps):
    """Median kernel time (ms, HIP events on the launch stream) of ``reps`` decodes."""
    import torch
    ms = []
    for _ in range(reps):
        dec.evaluate_device(keys, out, stream=stream, check_overflow=False)
        torch.cuda.synchronize()
        ms.append(dec.stats()["last_kernel_ms"])
    return float(np.median(ms))


def autotune(dec, packed, keys, reps=3, min_gain=0.02, cache=None, tiles=TILE_CHOICES,
             windows=WINDOW_CHOICES, log=None):
    """Tune ``dec``'s tile table on the device population ``keys``; returns a record.

    Coordinate descent over (plate size, width class, {tiles per pass, window
    tiles}): a candidate replaces the current value when its median kernel
    time is at least ``min_gain`` lower than the best so far and, measured
    again, than the incumbent re-measured beside it (kernel times vary by a
    few percent between launches).  The tuned table stays set on
    ``dec``; with ``cache`` (a JSON path) the record is stored under the
    instance signature.
    """
    import torch
    stream = torch.cuda.current_stream(keys.device)
    n = keys.shape[0]
    out = torch.empty(n, dtype=torch.float64, device=keys.device)
    ref = torch.empty(n, dtype=torch.float64, device=keys.device)
    t_tab, w_tab = dec.get_tiles()
    dec.evaluate_device(keys, ref, stream=stream)   # also warms the kernel up
    torch.cuda.synchronize()
    base_ms = best_ms = _time(dec, keys, out, stream, reps)
    trials = 0
    for plate, ncs in sorted(tile_classes(packed).items()):
        ms_rows = _machines_of(packed, plate)
        for nc in ncs:
            for which, choices in ((0, tiles), (1, windows)):
                tab = t_tab if which == 0 else w_tab
                for v in choices:
                    if which == 0 and v > _default_tiles(nc):
                        continue   # the width's register budget caps the pass size
                    cur = int(tab[ms_rows[0], nc])
                    if v == cur:
                        continue
                    tab[ms_rows, nc] = v
                    dec.set_tiles(t_tab, w_tab)
                    ms = _time(dec, keys, out, stream, reps)
                    trials += 1
                    if log:
                        log("autotune: plate %dx%d nc %d %s %d -> %.3f ms (best %.3f)" %
                            (plate[0], plate[1], nc, "tiles" if which == 0 else "window", v, ms, best_ms))
                    if ms < best_ms * (1.0 - min_gain):
                        # confirm against the incumbent, measured again right now
                        tab[ms_rows, nc] = cur
                        dec.set_tiles(t_tab, w_tab)
                        inc = _time(dec, keys, out, stream, reps)
                        tab[ms_rows, nc] = v
                        dec.set_tiles(t_tab, w_tab)
                        ms = _time(dec, keys, out, stream, reps)
                        trials += 2
                        if ms < inc * (1.0 - min_gain):
                            best_ms = ms
                            continue
                    tab[ms_rows, nc] = cur
                dec.set_tiles(t_tab, w_tab)
    dec.evaluate_device(keys, out, stream=stream)
    torch.cuda.synchronize()
    if not torch.equal(out, ref):
        raise RuntimeError("tile table changed decoded makespans")
    final_ms = _time(dec, keys, out, stream, reps)
    name = torch.cuda.get_device_name(keys.device)
    rec = {"signature": signature(packed, n, name), "device": name, "population": int(n),
           "default_ms": base_ms, "tuned_ms": final_ms, "trials": trials,
           "tiles_per_pass": t_tab.tolist(), "window_tiles": w_tab.tolist()}
    if cache:
        save(cache, rec)
    return rec


def load(cache):
    try:
        with open(cache) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save(cache, rec):
    d = load(cache)
    d[rec["signature"]] = rec
    os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
    tmp = cache + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1)
    os.replace(tmp, cache)


def apply_cached(dec, packed, population, cache=DEFAULT_CACHE):
    """Set ``dec``'s tile table from the cache; returns the record or None."""
    import torch
    name = torch.cuda.get_device_name(dec.device)
    rec = load(cache).get(signature(packed, population, name))
    if rec is None:
        return None
    dec.set_tiles(rec["tiles_per_pass"], rec["window_tiles"])
    return rec
