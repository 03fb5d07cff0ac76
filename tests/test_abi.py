"""C ABI: the library loads, exports what include/*.h declares, and rejects bad input
(host-side validation runs before any device call, so this needs no GPU)."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from brkga_fft2_amd import native
from tests.helpers import syn_instance

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for hdr in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(hdr).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(brkga_[a-z_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = native.load_library()
    declared = _declared_functions()
    assert declared == set(native.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name


def _create(packed, opt=None):
    lib = native.load_library()
    keep = []
    prob = native.BrkgaProblem(packed["n_parts"], packed["n_machines"], packed["n_types"],
                               packed["n_rots"])
    for name, dt in native._ARRAYS:
        a = np.ascontiguousarray(packed[name], dtype=dt)
        keep.append(a)
        setattr(prob, name, a.ctypes.data if a.size else None)
    h = ctypes.c_void_p()
    rc = lib.brkga_create(ctypes.byref(prob), ctypes.byref(opt) if opt is not None else None,
                          ctypes.byref(h))
    return rc, lib.brkga_last_error(None).decode()


@pytest.mark.parametrize("mutate,msg", [
    (lambda p: p.__setitem__("masks", np.where(p["masks"] > 0, 2, 0).astype(np.uint8)), "binary"),
    (lambda p: p.__setitem__("area", -p["area"]), "area"),
    (lambda p: p.__setitem__("slot_type", p["slot_type"] + 1000), "slot_type"),
    (lambda p: p.__setitem__("nrot", p["nrot"] * 0 + 5), "nrot"),
    (lambda p: p.__setitem__("plate_h", p["plate_h"] * 0), "plate"),
])
def test_create_rejects_bad_problems(mutate, msg):
    _, n, m, packed = syn_instance("C1")
    packed = dict(packed)
    mutate(packed)
    rc, err = _create(packed)
    assert rc == -1
    assert msg in err


@pytest.mark.parametrize("field,value,msg", [
    ("screen_rows", 17, "screen_rows"), ("screen_rows", -2, "screen_rows"),
    ("tiles_per_pass", 3, "tiles_per_pass"), ("mfma_format", 2, "mfma_format"),
])
def test_create_rejects_bad_options(field, value, msg):
    """Options are checked before any device call (no GPU needed)."""
    _, n, m, packed = syn_instance("C1")
    opt = native.BrkgaOptions()
    setattr(opt, field, value)
    rc, err = _create(packed, opt)
    assert rc == -1
    assert msg in err


def test_create_null_arguments():
    lib = native.load_library()
    h = ctypes.c_void_p()
    assert lib.brkga_create(None, None, ctypes.byref(h)) == -1
    assert "NULL" in lib.brkga_last_error(None).decode()


@pytest.mark.parametrize("compiler,flags", [("gcc", ["-std=c99", "-x", "c"]), ("g++", ["-std=c++11", "-x", "c++"])])
def test_header_compiles_as_plain_c_and_cpp(tmp_path, compiler, flags):
    """The boundary header is plain C (what a cgo / JNI / ctypes binding includes)
    and also valid C++; every declared entry point is callable by name."""
    import shutil
    import subprocess
    if shutil.which(compiler) is None:
        pytest.skip("%s not installed" % compiler)
    names = sorted(_declared_functions())
    src = tmp_path / "use_header.c"
    src.write_text('#include "brkga_mi355x.h"\n'
                   "typedef void (*entry_fn)(void);\n"
                   "entry_fn entry_points[] = {\n" + "".join("    (entry_fn)&%s,\n" % n for n in names) + "};\n")
    cmd = [compiler] + flags + ["-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
                                "-c", str(src), "-o", str(tmp_path / "use_header.o")]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert res.returncode == 0, res.stdout
