"""The ctypes structures of the Python binding have the C layout of
include/snowmpm.h: a small C program compiled with gcc prints sizeof and
offsetof of every field, and each must equal what ctypes computes.  A
mismatch would make the library read arguments from the wrong bytes."""
import ctypes as C
import os
import subprocess

import pytest

import snowmpm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "mpm_params": snowmpm.Params,
    "mpm_collider": snowmpm.Collider,
    "mpm_timings": snowmpm.StepTimings,
    "mpm_config": snowmpm.Config,
    "mpm_request": snowmpm.Request,
    "mpm_serve_stats": snowmpm.ServeStats,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("layout")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "snowmpm.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'    printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["    return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)], check=True,
                   capture_output=True, timeout=120)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout
    res = {}
    for line in out.splitlines():
        s, f, v = line.split()
        res[(s, f)] = int(v)
    return res


@pytest.mark.parametrize("cname", sorted(STRUCTS))
def test_ctypes_layout_matches_header(c_layout, cname):
    cls = STRUCTS[cname]
    assert C.sizeof(cls) == c_layout[(cname, "sizeof")], cname
    for f in cls._fields_:
        assert getattr(cls, f[0]).offset == c_layout[(cname, f[0])], (cname, f[0])
