"""Checkpoint and restart through hyb_state_save / hyb_state_load.

A run is saved after c1 cycles and continued for c2 more; a second run made
from the same configuration restores the file and steps c2 cycles. Every
lattice and element field must be bit-identical (NaN patterns of inactive
nodes included): the file holds the complete state, so the restart is exact.
Cases cover particles across periodic planes, wall contacts, the free surface
with and without slip walls, and the FP32 form.
"""
import numpy as np
import pytest

from helpers import fixture_config, load_case

pytestmark = pytest.mark.gpu

LB_KEYS = ("f", "types", "pflag", "solidIndex", "mass", "fluidMass")
DEM_KEYS = ("e_x", "e_q", "e_w", "e_forces", "e_misc", "neighborTable", "nearWallTable", "ghostIndex", "ghostVec",
            "dem_scalars", "energy")


def _runtime():
    from hybird_amd import runtime
    return runtime


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)
    return np.array_equal(a, b)


@pytest.mark.parametrize("name,c1,c2,prec", [("couple2", 5, 7, "FP64"), ("dense", 15, 10, "FP64"),
                                             ("fs", 10, 12, "FP64"), ("walls_fs", 12, 9, "FP64"),
                                             ("walls", 6, 8, "FP32")])
def test_restart_continues_bit_identically(tmp_path, name, c1, c2, prec):
    rt = _runtime()
    cfg, rec, _ = load_case(name)
    conf = fixture_config(cfg, precision=prec)
    print(f"[gpu state] case={name} {prec} save after {c1}, compare after {c2} more", flush=True)
    path = tmp_path / f"{name}.state"
    a = rt.Simulation(conf, rec)
    a.cycle(c1)
    a.save_state(path)
    a.cycle(c2)
    b = rt.Simulation(conf, rec)
    b.cycle(1)  # the restored state replaces whatever the handles hold
    b.load_state(path)
    b.cycle(c2)
    for key in LB_KEYS:
        assert _same(a.get(key), b.get(key)), f"{name}: {key} differs after restart"
    if rec.shape[0]:
        for key in DEM_KEYS:
            assert _same(a.get(key), b.get(key)), f"{name}: {key} differs after restart"
    a.close()
    b.close()


def test_restart_rejects_other_configurations(tmp_path):
    rt = _runtime()
    cfg, rec, _ = load_case("couple2")
    a = rt.Simulation(fixture_config(cfg), rec)
    a.cycle(2)
    path = tmp_path / "couple2.state"
    a.save_state(path)
    other = dict(cfg, lbSizeX=cfg["lbSizeX"] + 2)
    b = rt.Simulation(fixture_config(other), rec)
    with pytest.raises(rt.HybError, match="differs"):
        b.load_state(path)
    c = rt.Simulation(fixture_config(cfg, precision="FP32"), rec)
    with pytest.raises(rt.HybError, match="differs"):
        c.load_state(path)
    with pytest.raises(rt.HybError, match="cannot open"):
        a.load_state(tmp_path / "missing.state")
    bad = tmp_path / "bad.state"
    bad.write_bytes(b"not a state file at all")
    with pytest.raises(rt.HybError, match="not a state file"):
        a.load_state(bad)
    for s in (a, b, c):
        s.close()


def test_restart_with_other_particles_leaves_both_handles_unusable(tmp_path):
    # same lattice, one more particle: the lattice section loads, the DEM section is rejected, and the pair
    # (lattice with the file's state, DEM without) refuses to step instead of running inconsistently
    rt = _runtime()
    cfg, rec, _ = load_case("couple2")
    a = rt.Simulation(fixture_config(cfg), rec)
    a.cycle(2)
    path = tmp_path / "couple2.state"
    a.save_state(path)
    extra = rec[:1].copy()
    extra[0, 0] = rec.shape[0]
    extra[0, 3:6] = [12.0, 5.0, 18.0]  # clear of the other two spheres and of the walls
    b = rt.Simulation(fixture_config(cfg), np.concatenate([rec, extra]))
    with pytest.raises(rt.HybError, match="particle set"):
        b.load_state(path)
    with pytest.raises(rt.HybError):
        b.cycle(1)
    a.close()
    b.close()
