"""The oracle (FP64 CPU restatement) against the golden fixtures.

The fixtures were written by tests/golden/make_golden.py from the reference
compiled from its own sources (oracle/_ref). Every compared field must be
bit-identical: types, flags, neighbour tables, pair lists (integer) and
populations, velocities, forces, element states (FP64).
"""
import numpy as np
import pytest

from helpers import fixture_config, load_case, objects_of
from oracle.oracle import OracleSim

CASES = ["lb16", "chan", "bingham", "turb", "turb_bingham", "couple2", "dense", "dem60", "fs", "walls", "walls_fs", "io", "io_fs", "objects", "clumps"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_bitwise(name):
    cfg, rec, z = load_case(name)
    sim = OracleSim(fixture_config(cfg), rec, objects_of(z))
    prev = 0
    checked = 0
    for c in (int(c) for c in z["cycles"]):
        sim.cycle(c - prev)
        prev = c
        for key in z.files:
            if not key.startswith(f"c{c}/"):
                continue
            field = key.split("/", 1)[1]
            ref = z[key]
            mine = sim.get(field)
            assert ref.size == mine.size, (name, c, field)
            if ref.dtype.kind == "f":
                assert np.array_equal(ref.ravel(), mine.ravel(), equal_nan=True), (name, c, field)
            else:
                assert np.array_equal(ref.ravel(), mine.ravel().astype(ref.dtype)), (name, c, field)
            checked += 1
    assert checked > 0
