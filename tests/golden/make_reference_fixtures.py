"""Write the reference-held data and golden vectors as small fixtures.

Run once in a container where the reference tree is mounted read-only:

    python tests/golden/make_reference_fixtures.py /root/reference

Nothing here imports or runs reference code.  It reads data files only
(with ``brkga_fft2_amd.instances``):

* part masks ``data/partsMatrices/matrix_<id>.npy`` (int64 {0,1}) with
  ``numpy.load(allow_pickle=False)``; stored here bit-packed;
* the spec workbooks ``data/PartsMachines/*.xlsx`` parsed as zipped XML with
  the standard library (the pandas pickle ``cached_specs.pkl`` is never
  loaded);
* instance part lists ``data/Instances/P*M*-*.txt``;
* the golden placement log ``results/golden/P75M2-0`` (JSON);
* the placement-procedure bin grids ``results/inspection_output/pp_bin_*.txt``
  and ``solution_summary.txt`` for P50M2-0 chromosome 462;
* the makespan fingerprints of the benchmark JSON
  ``benchmarks/20260422_163051_NVIDIA_RTX_A4000_w1-2_stage3-a4000-baseline.json``
  and of ``PARALLEL_MACHINES_PLAN.md:468-473`` (P50M2-0).

Outputs (all under ``tests/golden/``):

* ``ref_parts.npz``      masks of parts 0..99, bit-packed, plus shapes;
* ``ref_specs.json``     part volume/support/height/area and machine L/W/ST/VT/SPT/HT;
* ``ref_instances.json`` part-id lists of all 30 instances;
* ``golden_P75M2-0.json`` chromosome + per-machine placement log + makespans;
* ``golden_P50M2-0_c462.npz`` the 8 PP bin grids (bit-packed) + part orders;
* ``fingerprints.json``  first-5-feasible makespans and feasible counts.
"""
import json
import os
import re
import sys

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(OUT)))
from brkga_fft2_amd.instances import instance_names, read_instance, read_mask, read_specs  # noqa: E402


def main(ref):
    data = os.path.join(ref, "data")
    # ---- masks -----------------------------------------------------------
    specs = read_specs(data)
    assert specs["part_ids"] == list(range(100))
    shapes, bits, offs = [], [], [0]
    for pid in specs["part_ids"]:
        m = read_mask(data, pid)
        shapes.append(m.shape)
        packed = np.packbits(m.ravel())
        bits.append(packed)
        offs.append(offs[-1] + packed.size)
    np.savez_compressed(os.path.join(OUT, "ref_parts.npz"),
                        shapes=np.array(shapes, dtype=np.int32),
                        offsets=np.array(offs, dtype=np.int64),
                        bits=np.concatenate(bits))

    # ---- specs -----------------------------------------------------------
    specs["source"] = "data/PartsMachines/{part-machine-information,polygon_areas}.xlsx"
    with open(os.path.join(OUT, "ref_specs.json"), "w") as f:
        json.dump(specs, f, indent=1)

    # ---- instances (all of data/Instances) -------------------------------
    inst = {name: read_instance(data, name) for name in instance_names(data)}
    with open(os.path.join(OUT, "ref_instances.json"), "w") as f:
        json.dump(inst, f)

    # ---- P75M2-0 placement golden ----------------------------------------
    with open(os.path.join(ref, "results/golden/P75M2-0")) as f:
        g = json.load(f)
    g["source"] = "results/golden/P75M2-0"
    with open(os.path.join(OUT, "golden_P75M2-0.json"), "w") as f:
        json.dump(g, f, indent=1)

    # ---- P50M2-0 chromosome 462 PP grids -----------------------------------
    io = os.path.join(ref, "results/inspection_output")
    summary = open(os.path.join(io, "solution_summary.txt")).read()
    pp = summary.split("PLACEMENTPROCEDURE DECODER", 1)[1]
    orders = {}
    mach = None
    for line in pp.splitlines():
        mm = re.match(r"Machine (\d+)\s+\(bin size", line)
        if mm:
            mach = int(mm.group(1))
        mb = re.match(r"\s+Bin (\d+):", line)
        if mb:
            b = int(mb.group(1))
        mo = re.match(r"\s+Parts \(placement order\): \[(.*)\]", line)
        if mo and mach is not None:
            orders["m%d_b%d" % (mach, b)] = [int(x) for x in mo.group(1).split(",")]
    grids = {}
    for key in sorted(orders):
        arr = np.loadtxt(os.path.join(io, "pp_bin_%s.txt" % key), dtype=np.int64)
        assert set(np.unique(arr)) <= {0, 1}
        grids["grid_" + key] = np.packbits(arr.astype(np.uint8).ravel())
        grids["shape_" + key] = np.array(arr.shape, dtype=np.int32)
        grids["order_" + key] = np.array(orders[key], dtype=np.int32)
    mk = re.search(r"Makespan \(placementProcedure\):\s+([0-9.]+)s", summary)
    grids["makespan"] = np.array(float(mk.group(1)))
    grids["chromosome_idx"] = np.array(462)
    np.savez_compressed(os.path.join(OUT, "golden_P50M2-0_c462.npz"), **grids)

    # ---- fingerprints ------------------------------------------------------
    bj = os.path.join(ref, "benchmarks/20260422_163051_NVIDIA_RTX_A4000_w1-2_stage3-a4000-baseline.json")
    with open(bj) as f:
        bench = json.load(f)
    fps = {}
    for r in bench["results"]:
        if r["num_workers"] != 1:
            continue
        fps[r["instance"]] = {"num_individuals": r["num_individuals"],
                              "seed": bench["seed"] + bench["n_reps"] - 1,
                              "fingerprint": r["fingerprint"],
                              "n_feasible": r["n_feasible"],
                              "source": os.path.basename(bj)}
    plan = open(os.path.join(ref, "PARALLEL_MACHINES_PLAN.md")).read()
    blk = plan.split("**P50M2-0** (seed 123, 5 reps, pop=500):", 1)[1].split("**P75M2-0**", 1)[0]
    feas = re.search(r"Feasible: (\d+) / (\d+)", blk)
    fp = re.findall(r"'([0-9.]+)'", blk)
    fps["P50M2-0"] = {"num_individuals": int(feas.group(2)), "seed": 127,
                      "fingerprint": fp, "n_feasible": int(feas.group(1)),
                      "source": "PARALLEL_MACHINES_PLAN.md:468-473"}
    with open(os.path.join(OUT, "fingerprints.json"), "w") as f:
        json.dump(fps, f, indent=1)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
