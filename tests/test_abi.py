"""The C-ABI library loads and exports every entry point of include/hybird_amd.h.

No compute call is made (no GPU needed)."""
import os
import re
import subprocess

import pytest

from helpers import ROOT
from hybird_amd import runtime

HEADER = os.path.join(ROOT, "include", "hybird_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int|void)\s+(hyb_\w+)\s*\(", text, flags=re.M)))


def test_header_and_runtime_list_agree():
    assert declared() == sorted(runtime.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(runtime.LIB_PATH):
        pytest.fail("libhybird_amd.so is not built")
    out = subprocess.run(["nm", "-D", "--defined-only", runtime.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared() if s not in syms]
    assert not missing, missing
    lib = runtime.load()
    for s in declared():
        assert hasattr(lib, s)
