"""The C ABI without a GPU: libkodama_hip.so loads, exports every function that
include/kodama_hip.h declares, and the ctypes binding lays out the option and
result structs exactly as the header does (checked against a C program built
from the header). No compute entry point is called here."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "kodama_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return sorted(set(re.findall(r"\b(kodama_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("kodama_hip_knncv", "kodama_hip_core_knn", "kodama_hip_core_knn_graph",
                 "kodama_hip_plsldacv", "kodama_hip_core_plslda", "kodama_hip_knn_graph",
                 "kodama_hip_dissimilarity_inplace", "kodama_hip_core_knn_batch", "kodama_hip_last_error"):
        assert must in names, must


def test_library_exports_every_declared_function():
    from kodama_amd import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libkodama_hip.so is not built (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(N.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing
    lib.kodama_hip_version.restype = C.c_char_p
    assert lib.kodama_hip_version().decode().startswith("kodama_hip")


@pytest.mark.parametrize("struct,cls", [("kodama_hip_core_options", "CoreOptions"),
                                        ("kodama_hip_core_result", "CoreResult"),
                                        ("kodama_hip_matrix_options", "MatrixOptions")])
def test_ctypes_structs_match_the_header(tmp_path, struct, cls):
    from kodama_amd import _native as N
    ct = getattr(N, cls)
    lines = [f'  printf("size %zu\\n", sizeof({struct}));']
    for name, _ in ct._fields_:
        lines.append(f'  printf("{name} %zu\\n", offsetof({struct}, {name}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"kodama_hip.h\"\nint main(void) {\n"
                   + "\n".join(lines) + "\n  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                       text=True).stdout.splitlines())
    assert int(out["size"]) == C.sizeof(ct)
    for name, _ in ct._fields_:
        assert int(out[name]) == getattr(ct, name).offset, name
