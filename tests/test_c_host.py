"""A plain-C host (tests/c_host/decode_c.c) drives the decoder through the C ABI
alone; its makespans must equal the oracle's on the same problem and keys."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c_host", "decode_c")


def _problem():
    """The problem decode_c.c builds, as the oracle's packed arrays."""
    def rect(h, w):
        return np.ones((h, w), np.uint8)

    def ell(s):
        y, x = np.mgrid[0:s, 0:s]
        return ((x < 3) | (y >= s - 3)).astype(np.uint8)

    def turn(m):   # quarter turn: out[x][h-1-y] = in[y][x]
        return np.rot90(m, -1).copy()

    rots = [rect(8, 12), rect(12, 8), ell(10)]
    for _ in range(3):
        rots.append(turn(rots[-1]))
    rots.append(rect(6, 6))
    off = np.cumsum([0] + [m.size for m in rots])[:-1]
    return {
        "n_parts": 8, "n_machines": 2, "n_types": 3, "n_rots": len(rots),
        "thresholds": np.array([0.5]), "slot_type": np.array([0, 1, 2, 0, 1, 2, 0, 2]),
        "type_id": np.array([10, 11, 12]), "area": np.array([96.0, 51.0, 36.0]),
        "nrot": np.array([2, 4, 1]), "best_rot": np.array([0, 0, 0]), "rot_base": np.array([0, 2, 6]),
        "rot_h": np.array([m.shape[0] for m in rots]), "rot_w": np.array([m.shape[1] for m in rots]),
        "mask_off": off, "masks": np.concatenate([m.ravel() for m in rots]),
        "plate_h": np.array([30, 25]), "plate_w": np.array([30, 40]), "setup": np.array([1.0, 2.0]),
        "proc_time": np.array([1.5, 2.0, 0.5, 1.2, 2.5, 0.7]),
        "proc_time_height": np.array([0.1, 0.2, 0.05, 0.15, 0.1, 0.02]),
    }


def test_ell_masks_and_areas_agree():
    p = _problem()
    m = p["masks"]
    for r in range(p["n_rots"]):
        blk = m[p["mask_off"][r]:p["mask_off"][r] + p["rot_h"][r] * p["rot_w"][r]]
        ty = int(np.searchsorted(p["rot_base"], r, side="right") - 1)
        assert blk.sum() == p["area"][ty]


@pytest.mark.gpu
def test_c_host_matches_oracle():
    print("C host decode", flush=True)
    sys.stdout.flush()
    if not os.path.exists(EXE):   # a test program, not the product: skip rather than stop the run
        pytest.skip("tests/c_host/decode_c is not built: run __graft_entry__.build()")
    res = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.split("\n")
    head = lines[0].split()
    assert head[0] == "problem"
    n, N = int(head[1]), int(head[2])
    keys = np.zeros((N, 2 * n), np.float32)
    ms = np.zeros(N)
    for ln in lines[1:]:
        f = ln.split()
        if not f:
            continue
        if f[0] == "keys":
            keys[int(f[1])] = np.array([float(v) for v in f[2:]], np.float32)
        elif f[0] == "makespan":
            ms[int(f[1])] = float(f[2])
    want, _, _ = oracle.decode(_problem(), keys)
    np.testing.assert_array_equal(ms, want)
    assert np.all(want < 1e15)   # every chromosome is feasible on these plates
