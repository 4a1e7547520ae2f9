"""The ctypes mirror of `struct hyb_config` (hybird_amd/_abi.py) against the C
header: every field's offset and the struct size as gcc lays them out, so a
field added on one side only fails here, without a GPU."""
import os
import shutil
import subprocess
import tempfile

import pytest

from hybird_amd._abi import HybConfig, config_from_cfg
from hybird_amd.cfg import base_config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_hyb_config_layout_matches_header():
    names = [f[0] for f in HybConfig._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hybird_amd.h"', 'int main(void) {',
             '  printf("size %zu\\n", sizeof(struct hyb_config));']
    lines += [f'  printf("{n} %zu\\n", offsetof(struct hyb_config, {n}));' for n in names]
    lines += ['  return 0;', '}']
    with tempfile.TemporaryDirectory() as tmp:
        src, exe = os.path.join(tmp, "layout.c"), os.path.join(tmp, "layout")
        with open(src, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.split() for l in out if l.strip())
    assert int(got["size"]) == __import__("ctypes").sizeof(HybConfig)
    for n in names:
        assert int(got[n]) == getattr(HybConfig, n).offset, n


def test_drum_keys_reach_the_config():
    c = config_from_cfg(base_config(problemName="DRUM", drumSpeed=0.3, fluidMass=800.0))
    assert (c.problem, c.drumSpeed, c.fluidMass) == (4, 0.3, 800.0)
