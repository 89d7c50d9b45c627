"""The C-ABI library loads without a GPU and exports exactly what include/edge_hip.h
declares, with the argument types the ctypes binding uses.  No compute calls."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "edge_hip.h")


def parse_header():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    decls = {}
    for m in re.finditer(r"(?:const\s+)?\w+\s*\*?\s*(edge_\w+)\s*\(([^)]*)\)\s*;", txt):
        name, args = m.group(1), m.group(2).strip()
        codes = ""
        if args and args != "void":
            for a in args.split(","):
                a = a.strip()
                if "*" in a:
                    codes += "p"
                elif re.match(r"(const\s+)?int\b", a):
                    codes += "i"
                elif re.match(r"(const\s+)?long\b", a):
                    codes += "l"
                elif re.match(r"(const\s+)?float\b", a):
                    codes += "f"
                elif re.match(r"(const\s+)?double\b", a):
                    codes += "d"
                else:
                    raise AssertionError(f"unparsed argument {a!r} of {name}")
        decls[name] = codes
    return decls


def test_header_matches_binding():
    from edge_hip._lib import SIGNATURES
    decls = parse_header()
    assert set(decls) == set(SIGNATURES), (set(decls) ^ set(SIGNATURES))
    for name, codes in decls.items():
        assert SIGNATURES[name] == codes, (name, SIGNATURES[name], codes)


def test_library_exports_every_declared_symbol():
    from edge_hip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libedge_hip.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}
    missing = set(parse_header()) - exported
    assert not missing, missing


def test_library_loads_and_reports_version():
    from edge_hip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libedge_hip.so not built")
    L = _lib.lib()
    assert L.edge_abi_version() == 1
    assert _lib.last_error() == ""


def test_bad_arguments_are_rejected_without_gpu_work():
    # argument validation happens on the host before any launch
    from edge_hip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libedge_hip.so not built")
    with pytest.raises(RuntimeError, match="bad dims"):
        _lib.call("edge_march_edge", *([None] * 12), 2, 2, 2, 0.1, 0.1, 0.01, None)
    with pytest.raises(RuntimeError, match="null"):
        _lib.call("edge_map_u_psi", *([None] * 7), 8, 8, 8, 0.1, None)
    # the gathers' 32-bit buffer ranges: fewer than 2^27 cells (1024 x 512 x 256 is 2^27)
    with pytest.raises(RuntimeError, match="bad dims"):
        _lib.call("edge_split_psi_hermite", *([None] * 7), 1024, 512, 256, 0.1, None)
    with pytest.raises(RuntimeError, match="bad dims"):
        _lib.call("edge_backtrack_faces", None, None, None, None, 0, *([None] * 6), 2, 8, 8, 0.1, None)
