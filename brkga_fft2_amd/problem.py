"""Problem data: parts, printer machines, and the flat layout the decoder consumes.

The dataclasses mirror the reference's input contract
(``data_structures.py:14-105``) so that code written against
``ProblemData`` keeps working; the per-machine FFT spectra of the reference
(``MachinePartData.ffts``) have no counterpart here because the collision test
is an exact integer correlation that works on the masks directly.

Part preprocessing follows ``BRKGA_alg3.py:331-386``:

* rotations are ``np.rot90(mask, k)`` (counter-clockwise) for ``k < nrot``;
* ``nrot`` is 2 when the mask equals its 180-degree rotation, else 4 (real
  instances); synthetic instances fix the rotation set instead;
* the density of a rotation is, per mask row, the longest run of 1-cells;
* ``best_rotation`` is the first rotation of minimum height.

Machine-specific times follow ``BRKGA_alg3.py:411-417``:
``proc_time = volume * VT + support * SPT`` and
``proc_time_height = height * HT`` (float64, in that order).
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

INFEASIBLE_PENALTY = 1e16  # placement.py:9 (10000000000000000)


@dataclass
class PartData:
    """One part type (``data_structures.py:14-59``)."""
    id: int
    area: float
    nrot: int
    rotations: List[np.ndarray]
    shapes: List[Tuple[int, int]]
    densities: List[np.ndarray]
    best_rotation: int
    height: float = 0.0   # part height in mm (orders the greedy initial solution)

    @property
    def lengths(self) -> List[int]:
        return [s[0] for s in self.shapes]


@dataclass
class MachinePartData:
    """Machine-specific data of a part (``data_structures.py:62-70``)."""
    proc_time: float
    proc_time_height: float
    ffts: Optional[list] = None  # kept for API shape; unused (no FFT here)


@dataclass
class MachineData:
    """One printer machine (``data_structures.py:73-82``)."""
    bin_length: int   # plate rows (L)
    bin_width: int    # plate columns (W)
    bin_area: int
    setup_time: float
    parts: Dict[int, MachinePartData] = field(default_factory=dict)


@dataclass
class ProblemData:
    """All problem data (``data_structures.py:85-105``)."""
    parts: Dict[int, PartData]
    machines: List[MachineData]
    instance_parts: np.ndarray
    instance_parts_unique: np.ndarray

    def get_part(self, part_id: int) -> PartData:
        return self.parts[part_id]

    def get_machine(self, machine_idx: int) -> MachineData:
        return self.machines[machine_idx]

    def get_machine_part(self, machine_idx: int, part_id: int) -> MachinePartData:
        return self.machines[machine_idx].parts[part_id]


def row_max_run(mat: np.ndarray) -> np.ndarray:
    """Longest run of 1-cells in each row (``BRKGA_alg3.py:351-360``)."""
    m = np.asarray(mat).astype(np.int8)
    out = np.zeros(m.shape[0], dtype=np.int32)
    run = np.zeros(m.shape[0], dtype=np.int32)
    for c in range(m.shape[1]):
        run = np.where(m[:, c] != 0, run + 1, 0)
        out = np.maximum(out, run)
    return out


def prepare_part(part_id: int, mask: np.ndarray, area: float,
                 nrot: Optional[int] = None) -> PartData:
    """Rotations, shapes, densities and best rotation of one part type.

    ``nrot=None`` applies the reference's symmetry rule (``BRKGA_alg3.py:337-340``).
    """
    m = np.ascontiguousarray(np.asarray(mask).astype(np.int32))
    if m.ndim != 2 or m.size == 0:
        raise ValueError("part %d: mask must be a non-empty 2-D array" % part_id)
    if not np.isin(m, (0, 1)).all():
        raise ValueError("part %d: mask must be binary" % part_id)
    if nrot is None:
        nrot = 2 if np.array_equal(m, np.rot90(m, 2)) else 4
    if nrot not in (1, 2, 3, 4):
        raise ValueError("nrot must be 1..4")
    rotations, shapes, densities = [], [], []
    for k in range(nrot):
        r = np.ascontiguousarray(np.rot90(m, k))
        rotations.append(r)
        shapes.append((int(r.shape[0]), int(r.shape[1])))
        densities.append(row_max_run(r))
    best = int(np.argmin([s[0] for s in shapes]))
    return PartData(id=int(part_id), area=float(area), nrot=nrot, rotations=rotations,
                    shapes=shapes, densities=densities, best_rotation=best)


def build_problem(instance_parts: Sequence[int], masks: Dict[int, np.ndarray],
                  areas: Dict[int, float], volume: Dict[int, float],
                  support: Dict[int, float], height: Dict[int, float],
                  machines: Sequence[dict], nrot: Optional[Dict[int, int]] = None
                  ) -> ProblemData:
    """Assemble a :class:`ProblemData` (``BRKGA_alg3.py:331-439``).

    ``machines`` holds dicts with keys ``L`` (rows), ``W`` (cols), ``ST``,
    ``VT``, ``SPT``, ``HT``.
    """
    inst = np.asarray(instance_parts, dtype=np.int64)
    uniq = np.unique(inst)
    parts = {int(p): prepare_part(int(p), masks[int(p)], areas[int(p)],
                                  None if nrot is None else nrot[int(p)])
             for p in uniq}
    for p in uniq:
        parts[int(p)].height = float(height[int(p)])
    mlist = []
    for m in machines:
        L, W = int(m["L"]), int(m["W"])
        mp = {}
        for p in uniq:
            p = int(p)
            pt = float(volume[p]) * float(m["VT"]) + float(support[p]) * float(m["SPT"])
            pth = float(height[p]) * float(m["HT"])
            mp[p] = MachinePartData(proc_time=pt, proc_time_height=pth)
        mlist.append(MachineData(bin_length=L, bin_width=W, bin_area=L * W,
                                 setup_time=m["ST"], parts=mp))
    return ProblemData(parts=parts, machines=mlist, instance_parts=inst,
                       instance_parts_unique=uniq)


def thresholds_for(nb_machines: int) -> List[float]:
    """Machine-assignment thresholds ``t / M`` (``BRKGA_alg3.py:442``)."""
    return [t / nb_machines for t in range(1, nb_machines)]


def type_heights(problem, packed) -> np.ndarray:
    """Part height of each packed part type: ``PartData.height`` when set,
    else the height time on machine 0 (height x HT, the same order)."""
    out = np.empty(int(packed["n_types"]), np.float64)
    for t, pid in enumerate(np.asarray(packed["type_id"])):
        part = problem.parts[int(pid)]
        h = float(getattr(part, "height", 0.0) or 0.0)
        out[t] = h if h > 0 else float(problem.machines[0].parts[int(pid)].proc_time_height)
    return out


def initial_solution_order(packed, type_height) -> np.ndarray:
    """Slots in the greedy initial solution's order: decreasing height, then
    decreasing polygon area, ties in instance order (a stable sort, as
    ``sort_values(['height(mm)', 'Area'], ascending=False)`` of
    ``BRKGA_alg3.py:450-452``)."""
    st = np.asarray(packed["slot_type"])
    h = np.asarray(type_height, np.float64)[st]
    a = np.asarray(packed["area"], np.float64)[st]
    return np.lexsort((np.arange(st.size), -a, -h)).astype(np.int32)


def initial_machine_keys(thresholds, nb_machines, rng) -> np.ndarray:
    """One machine key per machine inside its threshold bucket, drawn as
    ``BRKGA_alg3.py:585-593`` draws them (``random.uniform`` there, the
    generator ``rng`` here): machine 0 in (0, t0), the last in
    (t_{M-2} + 1e-4, 1 - 1e-4), the others in (t_{m-1} + 1e-4, t_m - 1e-4).
    Any key inside the bucket decodes to the same machine."""
    t = list(thresholds)
    out = np.empty(nb_machines, np.float64)
    for m in range(nb_machines):
        if nb_machines == 1:
            lo, hi = 0.0, 1.0
        elif m == 0:
            lo, hi = 0.0, t[0]
        elif m == nb_machines - 1:
            lo, hi = t[m - 1] + 0.0001, 1 - 0.0001
        else:
            lo, hi = t[m - 1] + 0.0001, t[m] - 0.0001
        out[m] = rng.uniform(lo, hi)
    return out


# ---------------------------------------------------------------------------
# Reference-format instances (fixture directory written from the reference's
# data/ tree: bit-packed masks, spec JSON, instance lists).
# ---------------------------------------------------------------------------

def load_reference_instance(fixture_dir: str, name: str) -> Tuple[ProblemData, int, int]:
    """Load instance ``P{n}M{m}-{i}`` from a fixture directory.

    Returns ``(problem_data, nb_parts, nb_machines)``.  Machines are the first
    ``m`` rows of the machine sheet, as ``BRKGA_alg3.py:388-393`` indexes them.
    """
    with open(os.path.join(fixture_dir, "ref_instances.json")) as f:
        inst = json.load(f)[name]
    with open(os.path.join(fixture_dir, "ref_specs.json")) as f:
        specs = json.load(f)
    z = np.load(os.path.join(fixture_dir, "ref_parts.npz"), allow_pickle=False)
    shapes, offs, bits = z["shapes"], z["offsets"], z["bits"]
    masks = {}
    for p in set(inst):
        h, w = (int(v) for v in shapes[p])
        masks[p] = np.unpackbits(bits[offs[p]:offs[p + 1]])[: h * w].reshape(h, w)
    nb_parts = len(inst)
    nb_machines = int(name.split("M")[1].split("-")[0])
    ids = range(len(specs["area"]))
    problem = build_problem(
        inst, masks,
        areas={i: specs["area"][i] for i in ids},
        volume={i: specs["volume"][i] for i in ids},
        support={i: specs["support"][i] for i in ids},
        height={i: specs["height"][i] for i in ids},
        machines=specs["machines"][:nb_machines])
    return problem, nb_parts, nb_machines


# ---------------------------------------------------------------------------
# Synthetic instances (seeded random simple polygons rasterised to cells).
# ---------------------------------------------------------------------------

SYNTHETIC_CONFIGS = {
    # name: parts, plates (rows=cols), cell mm, bbox range, rotations, height range, pop
    "C1": dict(n=25, plates=[250], cell_mm=1.0, bbox=(20, 80), nrot=2,
               height=(5.0, 50.0), pop=64, key_dtype="float64", gens=20),
    "C2": dict(n=50, plates=[300, 300], cell_mm=1.0, bbox=(20, 100), nrot=4,
               height=(5.0, 50.0), pop=256, key_dtype="float32", gens=50),
    "C3": dict(n=100, plates=[250, 300, 350, 400], cell_mm=1.0, bbox=(15, 120), nrot=4,
               height=(5.0, 80.0), pop=1024, key_dtype="float32", gens=100),
    "C5": dict(n=200, plates=[500, 500, 500, 500], cell_mm=0.5, bbox=(40, 200), nrot=4,
               height=(5.0, 80.0), pop=2048, key_dtype="float32", gens=1),
}
# Machine time constants of the synthetic configs: scan 0.03 h/cm^3,
# recoat 0.007 h/mm of the batch's max part height, setup 2 h per batch.
SYN_VT = 0.03 / 1000.0   # h per mm^3
SYN_HT = 0.007           # h per mm
SYN_ST = 2.0             # h


def random_polygon_mask(rng: np.random.Generator, lo: int, hi: int) -> np.ndarray:
    """A star-shaped (hence simple) polygon of 5-12 vertices rasterised at cell centres.

    The bounding box is drawn uniformly in ``[lo, hi]`` cells per side, the
    mask is cropped to its tight box, so every row and column holds a 1-cell.
    """
    k = int(rng.integers(5, 13))
    ang = np.sort(rng.uniform(0.0, 2.0 * math.pi, size=k))
    rad = rng.uniform(0.45, 1.0, size=k)
    bh = int(rng.integers(lo, hi + 1))
    bw = int(rng.integers(lo, hi + 1))
    px = np.cos(ang) * rad
    py = np.sin(ang) * rad
    # scale vertices to fill the drawn box
    px = (px - px.min()) / max(px.max() - px.min(), 1e-9) * bw
    py = (py - py.min()) / max(py.max() - py.min(), 1e-9) * bh
    yy, xx = np.mgrid[0:bh, 0:bw]
    cx = xx + 0.5
    cy = yy + 0.5
    inside = np.zeros((bh, bw), dtype=bool)
    j = k - 1
    for i in range(k):   # even-odd ray casting
        xi, yi, xj, yj = px[i], py[i], px[j], py[j]
        cond = (yi > cy) != (yj > cy)
        xint = (xj - xi) * (cy - yi) / np.where(yj - yi == 0, 1e-12, yj - yi) + xi
        inside ^= cond & (cx < xint)
        j = i
    if not inside.any():
        inside[bh // 2, bw // 2] = True
    rows = np.flatnonzero(inside.any(axis=1))
    cols = np.flatnonzero(inside.any(axis=0))
    return inside[rows[0]:rows[-1] + 1, cols[0]:cols[-1] + 1].astype(np.uint8)


def synthetic_problem(config: str, seed: int = 0) -> Tuple[ProblemData, int, int]:
    """Seeded synthetic instance of one of the benchmark configurations.

    Part ``i`` draws everything from ``np.random.default_rng(seed * 100003 + i)``.
    Part area is the pixel count times the cell area (mm^2); volume is area
    times height; support is 0.  Times are in hours.
    """
    cfg = SYNTHETIC_CONFIGS[config]
    n = cfg["n"]
    cell_area = cfg["cell_mm"] ** 2
    masks, areas, vol, sup, hgt = {}, {}, {}, {}, {}
    for i in range(n):
        rng = np.random.default_rng(seed * 100003 + i)
        m = random_polygon_mask(rng, *cfg["bbox"])
        masks[i] = m
        cells = float(m.sum())
        areas[i] = cells
        h_mm = float(rng.uniform(*cfg["height"]))
        hgt[i] = h_mm
        vol[i] = cells * cell_area * h_mm
        sup[i] = 0.0
    machines = [dict(L=p, W=p, ST=SYN_ST, VT=SYN_VT, SPT=0.0, HT=SYN_HT)
                for p in cfg["plates"]]
    problem = build_problem(np.arange(n), masks, areas, vol, sup, hgt, machines,
                            nrot={i: cfg["nrot"] for i in range(n)})
    return problem, n, len(machines)


def synthetic_population(n_chrom: int, n_parts: int, seed: int,
                         dtype=np.float32) -> np.ndarray:
    """Uniform [0,1) keys with no ties inside a chromosome's sequence half."""
    rng = np.random.default_rng(seed)
    pop = rng.random((n_chrom, 2 * n_parts)).astype(dtype)
    for r in range(n_chrom):
        sv = pop[r, :n_parts]
        while np.unique(sv).size != n_parts:
            sv[:] = rng.random(n_parts).astype(dtype)
    return pop


# ---------------------------------------------------------------------------
# Flat layout handed across the C ABI.
# ---------------------------------------------------------------------------

def pack_problem(problem: ProblemData, nb_parts: int, nb_machines: int,
                 thresholds: Sequence[float], instance_parts: Sequence[int]) -> dict:
    """Flatten a problem into the structure-of-arrays the decoder library takes.

    Part types are renumbered ``0..P-1`` in ``instance_parts_unique`` order;
    ``slot_type[j]`` maps instance slot ``j`` to its type.  Rotation ``r`` of
    type ``t`` is row ``rot_base[t] + r`` of the rotation table; its mask is
    ``masks[mask_off : mask_off + h*w]`` row-major uint8.
    """
    uniq = [int(p) for p in problem.instance_parts_unique]
    tindex = {p: i for i, p in enumerate(uniq)}
    inst = np.asarray(instance_parts, dtype=np.int64)
    if inst.size != nb_parts:
        raise ValueError("instance_parts has %d entries, nb_parts=%d" % (inst.size, nb_parts))
    if len(problem.machines) < nb_machines:
        raise ValueError("problem has fewer machines than nb_machines")
    if nb_machines > 1 and len(thresholds) != nb_machines - 1:
        raise ValueError("need nb_machines-1 thresholds")
    P = len(uniq)
    area = np.zeros(P, np.float64)
    nrot = np.zeros(P, np.int32)
    best = np.zeros(P, np.int32)
    rot_base = np.zeros(P, np.int32)
    rh, rw, moff, chunks = [], [], [], []
    off = 0
    for i, p in enumerate(uniq):
        pd_ = problem.parts[p]
        area[i] = pd_.area
        nrot[i] = pd_.nrot
        best[i] = pd_.best_rotation
        rot_base[i] = len(rh)
        for r in range(pd_.nrot):
            m = np.ascontiguousarray(pd_.rotations[r], dtype=np.uint8)
            rh.append(m.shape[0])
            rw.append(m.shape[1])
            moff.append(off)
            chunks.append(m.ravel())
            off += m.size
    M = nb_machines
    H = np.array([problem.machines[m].bin_length for m in range(M)], np.int32)
    W = np.array([problem.machines[m].bin_width for m in range(M)], np.int32)
    setup = np.array([float(problem.machines[m].setup_time) for m in range(M)], np.float64)
    pt = np.zeros((M, P), np.float64)
    pth = np.zeros((M, P), np.float64)
    for m in range(M):
        for i, p in enumerate(uniq):
            mp = problem.machines[m].parts[p]
            pt[m, i] = mp.proc_time
            pth[m, i] = mp.proc_time_height
    return dict(
        n_parts=int(nb_parts), n_machines=int(M), n_types=int(P), n_rots=len(rh),
        thresholds=np.asarray(thresholds if M > 1 else [], np.float64),
        slot_type=np.array([tindex[int(p)] for p in inst], np.int32),
        type_id=np.array(uniq, np.int32),
        area=area, nrot=nrot, best_rot=best, rot_base=rot_base,
        rot_h=np.array(rh, np.int32), rot_w=np.array(rw, np.int32),
        mask_off=np.array(moff, np.int64),
        masks=np.concatenate(chunks).astype(np.uint8) if chunks else np.zeros(0, np.uint8),
        plate_h=H, plate_w=W, setup=setup,
        proc_time=np.ascontiguousarray(pt), proc_time_height=np.ascontiguousarray(pth),
    )
