"""Tile autotuner for the decoder's collision passes.

The exact correlation runs in passes of up to 8 MFMA x-tiles (32 placements
each) over LDS windows of up to 64 x-tiles (``csrc/brkga_decoder.hip``),
after a screening pass that correlates a few evenly spaced mask rows over
every tile and lets the full passes skip the tiles with no zero count.  The
best pass and window sizes depend on the mask width (how many 32-column
Toeplitz chunks, ceil((w + 31) / 32), feed each tile, hence register
pressure) and on the plate width (how many tiles a row has); the best number
of screening rows on how cluttered the plates get (more rows reject more
tiles, and cost more).  ``autotune`` times the decoder on a sample
population for each (plate size, mask width class) entry of the tile table
and keeps the fastest setting, coordinate by coordinate; ``apply_cached``
reloads a result from a JSON cache keyed by the instance's shape signature
and the GPU name.

Every table entry gives bit-identical results (the tiling only changes which
MFMAs compute the same exact counts, the screen only skips tiles that hold
no placement); ``autotune`` re-checks that on the sample before returning.
"""
import hashlib
import json
import os

import numpy as np

from .native import TILE_CLASSES

TILE_CHOICES = (1, 2, 4, 8)
WINDOW_CHOICES = (4, 8, 16, 32, 64)
SCREEN_CHOICES = (0, 3, 4, 5, 6, 8, 10)
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
             windows=WINDOW_CHOICES, screens=SCREEN_CHOICES, log=None):
    """Tune ``dec``'s tile table on the device population ``keys``; returns a record.

    Coordinate descent over the knobs {screening rows, tiles per pass, window
    tiles}: first each knob for every width class of a plate size at once,
    then per (plate size, width class).  A candidate replaces the current
    value when its median kernel time is at least ``min_gain`` lower than the
    best so far and, measured again, than the incumbent re-measured beside it
    (kernel times vary by a few percent between launches).  The tuned table
    stays set on ``dec``; with ``cache`` (a JSON path) the record is stored
    under the instance signature.
    """
    import torch
    stream = torch.cuda.current_stream(keys.device)
    n = keys.shape[0]
    out = torch.empty(n, dtype=torch.float64, device=keys.device)
    ref = torch.empty(n, dtype=torch.float64, device=keys.device)
    t_tab, w_tab = dec.get_tiles()
    s_tab = dec.get_screen()
    tabs = (t_tab, w_tab, s_tab)
    names = ("tiles", "window", "screen")

    def apply():
        dec.set_tiles(t_tab, w_tab)
        dec.set_screen(s_tab)

    def value(which, nc, v):
        # a pass size above the width's register budget means that budget
        return min(v, _default_tiles(nc)) if which == 0 else v

    dec.evaluate_device(keys, ref, stream=stream)   # also warms the kernel up
    torch.cuda.synchronize()
    base_ms = _time(dec, keys, out, stream, reps)
    state = {"best": base_ms, "trials": 0}

    def attempt(which, rows, ncs, v, label):
        """Set knob `which` to v for classes ncs of machines rows; keep it if faster."""
        tab = tabs[which]
        old = {nc: int(tab[rows[0], nc]) for nc in ncs}
        new = {nc: value(which, nc, v) for nc in ncs}
        if new == old:
            return
        for nc in ncs:
            tab[rows, nc] = new[nc]
        apply()
        ms = _time(dec, keys, out, stream, reps)
        state["trials"] += 1
        if log:
            log("autotune: %s %s %d -> %.3f ms (best %.3f)" % (label, names[which], v, ms, state["best"]))
        if ms < state["best"] * (1.0 - min_gain):
            # confirm against the incumbent, measured again right now
            for nc in ncs:
                tab[rows, nc] = old[nc]
            apply()
            inc = _time(dec, keys, out, stream, reps)
            for nc in ncs:
                tab[rows, nc] = new[nc]
            apply()
            ms = _time(dec, keys, out, stream, reps)
            state["trials"] += 2
            if ms < inc * (1.0 - min_gain):
                state["best"] = ms
                return
        for nc in ncs:
            tab[rows, nc] = old[nc]
        apply()

    knobs = ((2, screens), (0, tiles), (1, windows))
    for plate, ncs in sorted(tile_classes(packed).items()):
        rows = _machines_of(packed, plate)
        label = "plate %dx%d" % plate
        for which, choices in knobs:          # every class of the plate at once
            for v in choices:
                attempt(which, rows, ncs, v, label + " all classes")
        for nc in ncs:                        # then class by class
            for which, choices in knobs:
                for v in choices:
                    if which == 0 and v > _default_tiles(nc):
                        continue
                    attempt(which, rows, [nc], v, label + " nc %d" % nc)
    dec.evaluate_device(keys, out, stream=stream)
    torch.cuda.synchronize()
    if not torch.equal(out, ref):
        raise RuntimeError("tile table changed decoded makespans")
    final_ms = _time(dec, keys, out, stream, reps)
    name = torch.cuda.get_device_name(keys.device)
    rec = {"signature": signature(packed, n, name), "device": name, "population": int(n),
           "default_ms": base_ms, "tuned_ms": final_ms, "trials": state["trials"],
           "tiles_per_pass": t_tab.tolist(), "window_tiles": w_tab.tolist(), "screen_rows": s_tab.tolist()}
    if cache:
        save(cache, rec)
    return rec


def changed_entries(packed, defaults, rec):
    """The entries ``rec`` moved from ``defaults`` (tiles, windows, screens), as
    ``[plate, width class, knob, default, tuned]`` for the classes the instance uses."""
    out = []
    tuned = (rec["tiles_per_pass"], rec["window_tiles"], rec["screen_rows"])
    for plate, ncs in sorted(tile_classes(packed).items()):
        m = _machines_of(packed, plate)[0]
        for nc in ncs:
            for name, d, t in zip(("tiles", "window", "screen"), defaults, tuned):
                if int(d[m][nc]) != int(t[m][nc]):
                    out.append(["%dx%d" % plate, int(nc), name, int(d[m][nc]), int(t[m][nc])])
    return out


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
    if "screen_rows" in rec:
        dec.set_screen(rec["screen_rows"])
    return rec
