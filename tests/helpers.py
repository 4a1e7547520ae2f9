"""Shared helpers for the test-suite: fixture loading and comparisons."""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from hybird_amd._abi import config_from_cfg, element_records  # noqa: E402


def load_case(name: str):
    """Returns (cfg dict, element records, fixture arrays) of a golden case."""
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    cfg = {str(k): float(v) for k, v in zip(z["cfg_keys"], z["cfg_vals"])}
    cfg["problemName"] = str(z["cfg_problem"])
    cfg["contactModel"] = "HERTZIAN"
    opt = {k: z[f"bed_{n}"] for k, n in (("size", "sizes"), ("w", "omegas"), ("q", "quats")) if f"bed_{n}" in z.files}
    rec = element_records(z["bed_x"], z["bed_r"], z["bed_v"], **opt)
    return cfg, rec, z


def objects_of(z):
    """Object records (HYB_OBJECT_RECORD rows) of a fixture, or None."""
    from hybird_amd.particles import object_records
    if "objects" not in z.files or z["objects"].size == 0:
        return None
    return object_records(z["objects"])


def fixture_config(cfg: dict, **kw):
    return config_from_cfg(cfg, **kw)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d
