"""CPU checks of the C ABI: the library loads, exports every entry point
include/mdgpu.h declares, and the struct mirrors have the reference's
layout (global.h) and match the Python ctypes mirrors.  No compute call."""
import ctypes as C
import os
import re
import subprocess

import pytest

import mdbas_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "mdgpu.h")
REF_INC = "/root/reference/include"

STRUCTS = [("MDG_PBC", "PBC"), ("MDG_CTRL", "CTRL"), ("MDG_PARAM", "PARAM"), ("MDG_BATH", "BATH"),
           ("MDG_BOND", "BOND"), ("MDG_ANGLE", "ANGLE"), ("MDG_DIHE", "DIHE"),
           ("MDG_NEIGH", "NEIGH"), ("MDG_ENERGY", "ENERGY"), ("MDG_CONSTRAINT", "CONSTRAINT"),
           ("MDG_EWALD", "EWALD"), ("MDG_ATOM", "ATOM"), ("MDG_IO", "IO")]


def declared():
    src = open(HDR).read()
    return sorted(set(re.findall(r"\b(mdg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_entry_point():
    assert os.path.exists(M.LIB_PATH), "build the library first"
    L = C.CDLL(M.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(M.EXPORTS)


def _c_sizes(prog: str, tmp_path, extra_inc=()):
    src = tmp_path / "l.c"
    exe = tmp_path / "l"
    src.write_text(prog)
    cmd = ["gcc", "-std=gnu99", "-I" + os.path.join(ROOT, "include"), *("-I" + p for p in extra_inc),
           str(src), "-o", str(exe)]
    subprocess.run(cmd, check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return dict(line.split() for line in out.strip().splitlines())


def test_ctypes_mirrors_match_c_header(tmp_path):
    prog = "#include <stdio.h>\n#include \"mdgpu.h\"\nint main(void){\n"
    for s, _ in STRUCTS + [("MDG_STEP", None)]:
        prog += f'printf("{s} %zu\\n", sizeof({s}));\n'
    prog += "return 0;}\n"
    sizes = _c_sizes(prog, tmp_path)
    for s, _ in STRUCTS + [("MDG_STEP", None)]:
        py = getattr(M, s, None)
        if py is not None:
            assert C.sizeof(py) == int(sizes[s]), s


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference headers not present")
def test_struct_mirrors_have_reference_layout(tmp_path):
    """Every member of every mirror sits at the reference struct's offset."""
    ref = open(os.path.join(REF_INC, "global.h")).read()
    prog = ("#include <stdio.h>\n#include <stddef.h>\n#include \"global.h\"\n#include \"mdgpu.h\"\n"
            "int main(void){\n")
    for mine, theirs in STRUCTS:
        end = re.search(r"\}\s*" + theirs + r"\s*;", ref).start()
        start = ref.rfind("typedef struct", 0, end)
        body = ref[ref.index("{", start) + 1:end]
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            parts = decl.replace("\n", " ").split()
            names = " ".join(parts[1:]) if parts[0] not in ("enum",) else " ".join(parts[2:])
            for nm in names.split(","):
                nm = nm.strip().lstrip("*").split("[")[0]
                if nm:
                    fields.append(nm)
        prog += f'printf("{mine} %zu %zu\\n", sizeof({mine}), sizeof({theirs}));\n'
        for fld in fields:
            prog += (f'printf("{mine}.{fld} %zu %zu\\n", offsetof({mine},{fld}), '
                     f"offsetof({theirs},{fld}));\n")
    prog += "return 0;}\n"
    src = tmp_path / "o.c"
    exe = tmp_path / "o"
    src.write_text(prog)
    subprocess.run(["gcc", "-std=gnu99", "-I" + REF_INC, "-I" + os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    assert len(rows) > 100
    for name, a, b in rows:
        assert a == b, name
