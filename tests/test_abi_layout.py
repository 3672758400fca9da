"""The Python bindings' ctypes structures against the C header: a small C
program compiled against include/illum_hiprt.h prints the size of every
structure and the offset of every field the bindings mirror, and each must
equal what ctypes computes (a field added, dropped or reordered on one side
only would pass arguments to the library silently wrong).  CPU only."""
import ctypes
import os
import shutil
import subprocess

import pytest

import hiprt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (ctypes class in hiprt.py, C typedef, prefix of the C field names)
STRUCTS = [
    ("_BvhParams", "hiprt_bvh_params", ""),
    ("_Rays", "hiprt_rays", "d_"),
    ("_Hits", "hiprt_hits", "d_"),
    ("_TracerOpts", "hiprt_tracer_opts", ""),
    ("_BatchEntry", "hiprt_batch_entry", ""),
    ("_TuneResult", "hiprt_tune_result", ""),
    ("_PtTuneResult", "hiprt_pt_tune_result", ""),
    ("_PtParams", "hiprt_pt_params", ""),
]
C_NAME = {("_BatchEntry", "occluded"): "d_occluded"}


def test_ctypes_structures_match_the_header(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    body = []
    for py, c, prefix in STRUCTS:
        cls = getattr(hiprt, py)
        body.append(f'printf("{py} %zu\\n", sizeof({c}));')
        for name, _ in cls._fields_:
            cf = C_NAME.get((py, name), prefix + name)
            body.append(f'printf("{py}.{name} %zu\\n", offsetof({c}, {cf}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"illum_hiprt.h\"\nint main(void) {\n" +
                   "\n".join(body) + "\nreturn 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                               text=True).stdout.splitlines())
    for py, c, _ in STRUCTS:
        cls = getattr(hiprt, py)
        assert int(got[py]) == ctypes.sizeof(cls), f"sizeof({c})"
        for name, _ in cls._fields_:
            assert int(got[f"{py}.{name}"]) == getattr(cls, name).offset, f"{c}.{name}"
