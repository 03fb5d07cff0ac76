"""Per-context bin state for introspection (``_last_contexts``).

The reference's evaluators keep, when ``_collect_contexts`` is set, one
``BatchPlacementContext`` per (machine, solution) with its ``open_bins``
(``wave_batch_evaluator.py:74-106,251-252``).  Its own tools read them:
``create_golden.py:170-172`` (bins with ``area > 0``, machine makespan as the
sum of ``proc_time + proc_time_height + setup_time``), and
``inspect_chromosome.py:251-261,389,473,538-566`` (``parts_assigned``,
``area``, ``proc_time``, ``proc_time_height``, ``grid``, ``bin_idx``,
``parts_sequence``).

Here the decoder runs on the GPU and hands back only the placement log
(machine, part id, bin, col, bottom row, rotation per part).  The bin state
is rebuilt on the host from that log and the problem data, applying each
placement as ``wave_batch_evaluator.py:761-775`` (existing bin) and ``:855-867``
(new bin) do: stamp the rotation's mask, ``area += part area``, occupied-row
bounds, ``enclosure_box_length = bin_length - min_occupied_row``,
``proc_time += pt``, ``proc_time_height = max(.., pth)``; the free-run
("vacancy") vector is the longest run of empty cells per row
(``numba_utils.py:161-188``).  The float64 sums are taken in placement order,
so the per-machine makespan of the rebuilt bins equals the decoder's.

A context that holds a part fitting the machine in neither orientation is
infeasible and has no bins, as ``placement.py:46-53`` returns it; a feasible
context the log does not cover (rows of -1) is left undecoded (no bins,
``is_done`` False).
"""
from dataclasses import dataclass, field
from typing import Dict, List

import numpy as np


@dataclass
class BinState:
    """One bin (build plate) of a context (``wave_batch_evaluator.py:74-90``)."""
    bin_idx: int
    grid: np.ndarray
    vacancy_vector: np.ndarray
    grid_state_idx: int
    area: float
    enclosure_box_length: int
    min_occupied_row: int
    max_occupied_row: int
    proc_time: float
    proc_time_height: float
    grid_fft_valid: bool
    parts_assigned: List[int]
    bin_length: int
    bin_width: int


@dataclass
class BatchPlacementContext:
    """One (solution, machine) decode (``wave_batch_evaluator.py:93-106``)."""
    solution_idx: int
    machine_idx: int
    parts_sequence: np.ndarray
    current_part_idx: int
    bin_length: int
    bin_width: int
    bin_area: float
    open_bins: List[BinState] = field(default_factory=list)
    next_grid_idx: int = 0
    is_done: bool = False
    is_feasible: bool = True


def free_runs(grid):
    """Longest run of empty cells in each row of a {0,1} grid."""
    g = np.asarray(grid)
    out = np.zeros(g.shape[0], np.int32)
    run = np.zeros(g.shape[0], np.int32)
    for c in range(g.shape[1]):
        run = np.where(g[:, c] == 0, run + 1, 0)
        out = np.maximum(out, run)
    return out


def fits_plate(shape0, H, W):
    """A part whose rotation-0 shape fits an H x W plate in some orientation
    (``placement.py:46-50``)."""
    return not ((shape0[0] > H or shape0[1] > W) and (shape0[1] > H or shape0[0] > W))


def machine_sequences(keys, n_parts, thresholds, n_machines, instance_parts):
    """Per chromosome, per machine: the instance slots of the machine in
    ascending sequence-key order (ties by slot) and their part ids
    (``wave_batch_evaluator.py:256-275``)."""
    keys = np.asarray(keys)
    thr = np.asarray(thresholds, np.float64)
    inst = np.asarray(instance_parts)
    out = []
    for c in range(keys.shape[0]):
        sv, mv = keys[c, :n_parts], keys[c, n_parts:].astype(np.float64)
        if n_machines <= 1:
            mach = np.zeros(n_parts, np.int64)
        else:
            mach = np.searchsorted(thr[:n_machines - 1], mv, side="left")
        per = []
        for m in range(n_machines):
            slots = np.flatnonzero(mach == m)
            slots = slots[np.argsort(sv[slots], kind="stable")]
            per.append((slots, inst[slots]))
        out.append(per)
    return out


def build_contexts(problem, n_parts, n_machines, thresholds, instance_parts, keys, log):
    """``{machine: [BatchPlacementContext per solution]}`` from a decode's keys
    ``[N, 2n]`` and placement log ``[N, n_parts, 6]`` (rows in machine order,
    then placement order)."""
    keys = np.asarray(keys)
    log = np.asarray(log)
    seqs = machine_sequences(keys, n_parts, thresholds, n_machines, instance_parts)
    out: Dict[int, List[BatchPlacementContext]] = {m: [] for m in range(n_machines)}
    for c in range(keys.shape[0]):
        off = 0
        for m in range(n_machines):
            md = problem.machines[m]
            H, W = int(md.bin_length), int(md.bin_width)
            slots, seq = seqs[c][m]
            rows = log[c, off:off + len(slots)]
            off += len(slots)
            ctx = BatchPlacementContext(solution_idx=c, machine_idx=m, parts_sequence=seq,
                                        current_part_idx=0, bin_length=H, bin_width=W,
                                        bin_area=float(md.bin_area),
                                        next_grid_idx=c * n_parts)
            unfit = [i for i, pid in enumerate(seq) if not fits_plate(problem.parts[int(pid)].shapes[0], H, W)]
            if unfit:
                # placement.py:46-53: the first part that fits neither way ends the context
                ctx.current_part_idx = unfit[0]
                ctx.is_feasible = False
                ctx.is_done = True
                out[m].append(ctx)
                continue
            if any(int(r[0]) != m for r in rows):
                # not in the log (a decoder may stop a chromosome at its first
                # infeasible machine): not decoded, no bins
                out[m].append(ctx)
                continue
            bins = ctx.open_bins
            for r in rows:
                _, pid, b, col, row, rot = (int(v) for v in r)
                part = problem.parts[pid]
                mp = md.parts[pid]
                mask = np.asarray(part.rotations[rot], np.uint8)
                h, w = mask.shape
                y0 = row - h + 1
                if b == len(bins):   # new bin (wave_batch_evaluator.py:784-799,855-867)
                    bins.append(BinState(bin_idx=b, grid=np.zeros((H, W), np.uint8),
                                         vacancy_vector=np.full(H, W, np.int32),
                                         grid_state_idx=ctx.next_grid_idx, area=0.0, enclosure_box_length=0,
                                         min_occupied_row=H, max_occupied_row=-1, proc_time=0.0,
                                         proc_time_height=0.0, grid_fft_valid=False, parts_assigned=[],
                                         bin_length=H, bin_width=W))
                    ctx.next_grid_idx += 1
                bs = bins[b]
                bs.grid[y0:row + 1, col:col + w] += mask
                bs.parts_assigned.append(pid)
                bs.area += part.area
                bs.min_occupied_row = min(bs.min_occupied_row, y0)
                bs.max_occupied_row = max(bs.max_occupied_row, row)
                bs.enclosure_box_length = H - bs.min_occupied_row
                bs.proc_time += mp.proc_time
                bs.proc_time_height = max(bs.proc_time_height, mp.proc_time_height)
            for bs in bins:
                bs.vacancy_vector = free_runs(bs.grid)
            ctx.current_part_idx = len(seq)
            ctx.is_done = True
            out[m].append(ctx)
    return out


def machine_makespan(ctx, setup_time):
    """``create_golden.py:171-175``: sum over bins with area > 0 of
    ``proc_time + proc_time_height + setup_time``, in bin order."""
    total = 0
    for b in ctx.open_bins:
        if b.area > 0:
            total = total + (b.proc_time + b.proc_time_height + setup_time)
    return total
