"""The C-ABI library loads and exports every symbol include/viennals_gpu.h
declares (no GPU calls)."""
import os
import re
import subprocess

import viennals_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "viennals_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lsg_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared() == sorted(L.EXPORTED)


def test_library_exports_all_symbols():
    assert os.path.exists(L.LIB_PATH), "build with __graft_entry__.build() first"
    out = subprocess.check_output(["nm", "-D", "--defined-only", L.LIB_PATH], text=True)
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared() if s not in syms]
    assert not missing, missing


def test_library_loads_and_binds():
    lib = L.lib()
    for name in L.EXPORTED:
        assert getattr(lib, name) is not None
