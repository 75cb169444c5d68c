"""CPU-side checks of the C ABI: the HIP library loads and exports every
entry point include/tissue_nr.h declares, and the stats struct layout of the
Python binding matches the header.  No device computation is attempted."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tissue_nr.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tf_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ["tf_nr_create", "tf_nr_set_material", "tf_nr_set_growth", "tf_nr_set_visc_surface",
              "tf_nr_step", "tf_csr_pattern", "tf_nr_destroy"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from tissueorigami_amd import nr
    lib = nr.load()
    out = subprocess.check_output(["nm", "-D", "--defined-only", nr.LIB_PATH]).decode()
    exported = set(re.findall(r" T (tf_[a-z_0-9]+)", out))
    for name in declared():
        assert name in exported, name
        assert hasattr(lib, name)
    assert set(nr.EXPORTS) <= exported


def _c_layout(tmp_path, struct, fields):
    """sizeof and offsetof of a header struct, from gcc."""
    src = tmp_path / "sz.c"
    args = ", ".join(["sizeof(%s)" % struct] + ["offsetof(%s, %s)" % (struct, f) for f in fields])
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "tissue_nr.h"\n'
                   'int main(void){printf("' + " ".join(["%zu"] * (len(fields) + 1)) +
                   '\\n", ' + args + ');return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    return list(map(int, subprocess.check_output([str(exe)]).split()))


@pytest.mark.parametrize("struct,pyname", [("tf_step_stats", "StepStats"),
                                           ("tf_kernel_times", "KernelTimes")])
def test_struct_layouts_match_header(tmp_path, struct, pyname):
    from tissueorigami_amd import nr
    cls = getattr(nr, pyname)
    fields = [f[0] for f in cls._fields_]
    size, *offs = _c_layout(tmp_path, struct, fields)
    assert ctypes.sizeof(cls) == size
    for f, off in zip(fields, offs):
        assert getattr(cls, f).offset == off, f


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from tissueorigami_amd import nr
    monkeypatch.setattr(nr, "_lib", None)
    monkeypatch.setattr(nr, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError):
        nr.load()
