"""Reference instance files -> ``ProblemData``.

Reads the reference's ``data/`` tree directly, with loaders that execute
nothing from the files:

* instance part lists ``data/Instances/P<n>M<m>-<i>.txt`` (whitespace
  separated part ids; ``BRKGA_alg3.py:299-305``);
* part masks ``data/partsMatrices/matrix_<id>.npy`` (``numpy.load`` with
  ``allow_pickle=False``; ``BRKGA_alg3.py:331-336``);
* the spec workbooks ``data/PartsMachines/part-machine-information.xlsx``
  (sheets ``part`` and ``machine``) and ``polygon_areas.xlsx``, parsed as
  zipped XML with the standard library (``BRKGA_alg3.py:307-321``; the
  reference's pickled spec cache is never read).

``problem_from_data(data_dir, name)`` builds the same ``ProblemData`` as
``problem.load_reference_instance`` builds from the repository's fixtures
(which ``tests/golden/make_reference_fixtures.py`` writes with these
functions), so every one of the reference's instances can be decoded.
"""
import os
import re
import zipfile
import xml.etree.ElementTree as ET
from typing import Dict, List, Tuple

import numpy as np

from .problem import ProblemData, build_problem

_NS = {"m": "http://schemas.openxmlformats.org/spreadsheetml/2006/main"}
_RELNS = "http://schemas.openxmlformats.org/officeDocument/2006/relationships"
_NAME = re.compile(r"P(\d+)M(\d+)-(\d+)$")


def xlsx_sheets(path: str) -> Dict[str, List[list]]:
    """``{sheet name: rows}``; each row is a list of cell texts (None for empty)."""
    z = zipfile.ZipFile(path)
    shared = []
    if "xl/sharedStrings.xml" in z.namelist():
        root = ET.fromstring(z.read("xl/sharedStrings.xml"))
        for si in root.findall("m:si", _NS):
            shared.append("".join(t.text or "" for t in si.iter("{%s}t" % _NS["m"])))
    wb = ET.fromstring(z.read("xl/workbook.xml"))
    rels = ET.fromstring(z.read("xl/_rels/workbook.xml.rels"))
    target = {r.get("Id"): r.get("Target") for r in rels}
    out = {}
    for s in wb.find("m:sheets", _NS):
        sheet = ET.fromstring(z.read("xl/" + target[s.get("{%s}id" % _RELNS)]))
        rows = []
        for row in sheet.iter("{%s}row" % _NS["m"]):
            cells = {}
            for c in row.findall("m:c", _NS):
                col = re.match(r"[A-Z]+", c.get("r")).group(0)
                v = c.find("m:v", _NS)
                val = None if v is None else v.text
                if c.get("t") == "s" and val is not None:
                    val = shared[int(val)]
                cells[col] = val
            ncol = max((ord(k) - 64 for k in cells), default=0)
            rows.append([cells.get(chr(65 + i)) for i in range(ncol)])
        out[s.get("name")] = rows
    return out


def _num(text):
    """Spreadsheet number text -> int when the text is integral, else float."""
    if re.fullmatch(r"-?\d+", text):
        return int(text)
    return float(text)


def read_specs(data_dir: str) -> dict:
    """Part volume/support/height/area and machine L/W/ST/VT/SPT/HT.

    Same layout as the repository's ``tests/golden/ref_specs.json``.
    """
    pm = xlsx_sheets(os.path.join(data_dir, "PartsMachines", "part-machine-information.xlsx"))
    part_rows = pm["part"]
    hdr = part_rows[0]
    parts = {}
    for r in part_rows[1:]:
        rec = {hdr[i]: _num(r[i]) for i in range(len(hdr)) if i < len(r) and r[i] is not None}
        parts[int(rec["part ID"])] = rec
    mach_rows = pm["machine"]
    mh = mach_rows[0]
    machines = [{mh[i]: _num(r[i]) for i in range(len(mh))} for r in mach_rows[1:]]
    arows = list(xlsx_sheets(os.path.join(data_dir, "PartsMachines", "polygon_areas.xlsx")).values())[0]
    if arows[0] != ["part ID", "Area"]:
        raise ValueError("polygon_areas.xlsx: unexpected header %r" % (arows[0],))
    ids = sorted(parts)
    area = {int(r[0]): float(r[1]) for r in arows[1:]}
    if sorted(area) != ids:
        raise ValueError("polygon_areas.xlsx and part-machine-information.xlsx list different parts")
    return {
        "part_ids": ids,
        "volume": [float(parts[i]["volume(mm3)"]) for i in ids],
        "support": [float(parts[i]["support(mm3)"]) for i in ids],
        "height": [float(parts[i]["height(mm)"]) for i in ids],
        "area": [area[i] for i in ids],
        "machines": [{"L": int(m["L(mm)"]), "W": int(m["W(mm)"]), "ST": m["ST(s)"],
                      "VT": float(m["VT(s/mm3)"]), "SPT": float(m["SPT(s/mm3)"]),
                      "HT": float(m["HT(s/mm3)"])} for m in machines],
    }


def read_mask(data_dir: str, part_id: int) -> np.ndarray:
    """Part mask as uint8 {0, 1} (``matrix_<id>.npy``, no pickles)."""
    m = np.load(os.path.join(data_dir, "partsMatrices", "matrix_%d.npy" % int(part_id)),
                allow_pickle=False)
    if m.ndim != 2 or not set(np.unique(m)) <= {0, 1}:
        raise ValueError("matrix_%d.npy is not a 2-D {0,1} mask" % part_id)
    return m.astype(np.uint8)


def instance_names(data_dir: str) -> List[str]:
    """Instance names in ``data/Instances``, ordered by (parts, machines, index)."""
    names = [f[:-4] for f in os.listdir(os.path.join(data_dir, "Instances"))
             if f.endswith(".txt") and _NAME.match(f[:-4])]
    return sorted(names, key=lambda s: tuple(int(g) for g in _NAME.match(s).groups()))


def read_instance(data_dir: str, name: str) -> List[int]:
    """Part ids of instance ``name`` (one entry per slot, repeats allowed)."""
    with open(os.path.join(data_dir, "Instances", name + ".txt")) as f:
        return [int(x) for x in f.read().split()]


def machines_of(name: str) -> int:
    m = _NAME.match(name)
    if not m:
        raise ValueError("instance names look like P<parts>M<machines>-<index>: %r" % name)
    return int(m.group(2))


def problem_from_data(data_dir: str, name: str) -> Tuple[ProblemData, int, int]:
    """``(problem_data, nb_parts, nb_machines)`` of instance ``name`` read from
    the reference's ``data/`` directory.  Machines are the first ``m`` rows of
    the machine sheet, as ``BRKGA_alg3.py:388-393`` indexes them."""
    inst = read_instance(data_dir, name)
    specs = read_specs(data_dir)
    nb_machines = machines_of(name)
    ids = specs["part_ids"]
    masks = {p: read_mask(data_dir, p) for p in set(inst)}
    problem = build_problem(
        inst, masks,
        areas=dict(zip(ids, specs["area"])),
        volume=dict(zip(ids, specs["volume"])),
        support=dict(zip(ids, specs["support"])),
        height=dict(zip(ids, specs["height"])),
        machines=specs["machines"][:nb_machines])
    return problem, len(inst), nb_machines
