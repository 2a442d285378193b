"""The ctypes mirrors in farm_amd/capi.py against the structs that
include/farm_capi.h declares: same field names in the same order, and the
same size as the C compiler lays the struct out (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from farm_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "farm_capi.h")
MIRRORS = {
    "farm_raster_hdr": capi.RasterHdr,
    "farm_layer": capi.Layer,
    "farm_polyset": capi.PolySet,
    "farm_opts": capi.Opts,
    "farm_stats": capi.Stats,
    "farm_tuning": capi.Tuning,
}


def header_fields(name: str) -> list[str]:
    text = open(HEADER).read()
    m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), text, re.S)
    assert m, name
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        for part in decl.split(","):  # "int32_t dimx, dimy" declares two fields
            ident = re.findall(r"([A-Za-z_]\w*)\s*(?:\[[^\]]*\])?\s*$", part.strip())
            assert ident, decl
            fields.append(ident[0])
    return fields


@pytest.mark.parametrize("name", sorted(MIRRORS))
def test_field_names_and_order(name):
    assert [f for f, _ in MIRRORS[name]._fields_] == header_fields(name)


def test_sizes_match_the_c_layout(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include "farm_capi.h"\n#include <stdio.h>\nint main(void) {\n' + "".join(
        f'    printf("{n} %zu\\n", sizeof({n}));\n' for n in sorted(MIRRORS)) + "    return 0;\n}\n")
    exe = tmp_path / "sz"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True)
    sizes = {line.split()[0]: int(line.split()[1]) for line in out.splitlines()}
    for n, cls in MIRRORS.items():
        assert C.sizeof(cls) == sizes[n], n
