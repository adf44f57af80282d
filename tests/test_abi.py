"""The C-ABI library loads and exports every function include/tcme.h declares.
No compute calls: this runs without a GPU."""
import ctypes
import os
import re
import subprocess

from conftest import ROOT

LIB = os.path.join(ROOT, "tcodec_amd", "lib", "libtcme.so")
HEADER = os.path.join(ROOT, "include", "tcme.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(",
                         text, flags=re.M):
        if m.group(1) not in ("if", "while", "defined"):
            names.add(m.group(1))
    return names


def test_header_parses_to_expected_set():
    names = declared_functions()
    assert "tc_motion_est" in names and "tcme_search" in names
    from tcodec_amd import me
    assert names == set(me.EXPORTS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libtcme.so not built"
    ctypes.CDLL(LIB)  # loads without a GPU
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = declared_functions() - exported
    assert not missing, missing


def test_library_is_gfx950_code_object():
    """The embedded offload bundle targets gfx950."""
    blob = open(LIB, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_product_does_not_import_oracle():
    """The shipped package never reaches for the checker."""
    pkg = os.path.join(ROOT, "tcodec_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".c")) or f == "Makefile":
                src = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "import oracle" not in src and "tcme_oracle" not in src, f
