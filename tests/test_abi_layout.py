"""Struct layouts of the C ABI: the C compiler's sizeof/offsetof for every struct that
include/tamc_gpu.h declares, against the ctypes mirrors in tamc_amd/_lib.py and against the Rust
repr(C) binding shown in INTEGRATION.md.  A field added to one side and not the other makes the
library read past (or short of) a caller's struct, so every mirror is checked field by field."""
import ctypes as C
import os
import re
import subprocess

import pytest

from tamc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = {
    "tamc_instance_info": _lib.InstanceInfo,
    "tamc_instance_desc": _lib.InstanceDesc,
    "tamc_pt_params": _lib.PtParams,
    "tamc_pt_result": _lib.PtResult,
    "tamc_pt_state": _lib.PtState,
    "tamc_batch_options": _lib.BatchOptions,
    "tamc_batch_stats": _lib.BatchStats,
    "tamc_server_options": _lib.ServerOptions,
    "tamc_request_times": _lib.RequestTimes,
    "tamc_server_stats": _lib.ServerStats,
    "tamc_sa_params": _lib.SaParams,
    "tamc_sa_result": _lib.SaResult,
    "tamc_partition": _lib.Partition,
    "tamc_gla_result": _lib.GlaResult,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    d = tmp_path_factory.mktemp("layout")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "tamc_gpu.h"', "int main(void) {"]
    for cname, cls in MIRRORS.items():
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in cls._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines += ["  return 0;", "}"]
    src = d / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "probe"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for ln in out.splitlines():
        name, field, val = ln.split()
        layout.setdefault(name, {})[field] = int(val)
    return layout


def _c_fields(header, cname):
    """Field names of a typedef'd struct in the header, in declaration order."""
    m = re.search(r"typedef struct \{([^{}]*)\}\s*%s;" % cname, header)
    assert m, cname
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    names = []
    for decl in body.split(";"):
        for part in decl.split(","):
            if part.strip():
                names.append(re.search(r"(\w+)\s*(?:\[[^\]]*\])?\s*$", part).group(1))
    return names


@pytest.mark.parametrize("cname", sorted(MIRRORS))
def test_ctypes_mirror_matches_c(c_layout, cname):
    cls = MIRRORS[cname]
    header = open(os.path.join(ROOT, "include", "tamc_gpu.h")).read()
    assert [f[0] for f in cls._fields_] == _c_fields(header, cname), cname
    assert C.sizeof(cls) == c_layout[cname]["size"], cname
    for f in cls._fields_:
        assert getattr(cls, f[0]).offset == c_layout[cname][f[0]], (cname, f[0])


_RUST = {"u8": (1, 1), "i8": (1, 1), "u32": (4, 4), "i32": (4, 4), "f32": (4, 4), "u64": (8, 8),
         "i64": (8, 8), "f64": (8, 8)}


def _rust_layout(body):
    """repr(C) offsets of `pub name: type` fields."""
    off, align_max, out = 0, 1, {}
    for name, ty in re.findall(r"pub (\w+): ([^,\n]+?)\s*(?:,|\n|$)", body):
        ty = ty.strip()
        size, align = (8, 8) if ty.startswith("*") else _RUST[ty]
        off = (off + align - 1) // align * align
        out[name] = off
        off += size
        align_max = max(align_max, align)
    return out, (off + align_max - 1) // align_max * align_max


def test_rust_binding_in_integration_notes_matches_c(c_layout):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    structs = dict(re.findall(r"pub struct (\w+) \{[ \t]*(?://[^\n]*)?\n(.*?)\n\}", text, re.S))
    checked = 0
    for rname, cname in [("TamcPtParams", "tamc_pt_params"), ("TamcSaParams", "tamc_sa_params"),
                         ("TamcInstanceDesc", "tamc_instance_desc"), ("TamcPtState", "tamc_pt_state"),
                         ("TamcBatchOptions", "tamc_batch_options")]:
        if rname not in structs:
            continue
        body = re.sub(r"//[^\n]*", "", structs[rname])
        offs, size = _rust_layout(body)
        assert list(offs) == [f[0] for f in MIRRORS[cname]._fields_], rname
        assert size == c_layout[cname]["size"], rname
        for k, v in offs.items():
            assert v == c_layout[cname][k], (rname, k)
        checked += 1
    assert checked >= 1
