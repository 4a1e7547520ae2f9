"""The reference's own main (hybird.cpp, unchanged) driving the GPU through the
C++ host layer (hybird_amd/host): the binary hybird_gpu is built from the
reference sources by `make -C hybird_amd/host dropin`. The run writes its final
state (HYBIRD_AMD_DUMP_DIR) and is compared with the oracle after the same
number of cycles."""
import os
import subprocess

import numpy as np
import pytest

from helpers import ROOT, fixture_config, load_case, rel_l2
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "hybird_amd", "host", "hybird_gpu")


def _write_case(tmp, cfg, bed_x, bed_r, cycles):
    from hybird_amd.cfg import write_cfg
    from hybird_amd.particles import Bed, write_empty_object_file, write_particle_file
    pf, of = os.path.join(tmp, "p.dat"), os.path.join(tmp, "o.dat")
    write_particle_file(pf, Bed(np.asarray(bed_x), np.asarray(bed_r)))
    write_empty_object_file(of)
    cfg = dict(cfg, particleFile=pf, objectFile=of, maximumTimeSteps=cycles, screenExpTime=0)
    path = os.path.join(tmp, "case.cfg")
    write_cfg(path, cfg)
    return path


def _load(d, name, dtype=np.float64):
    return np.fromfile(os.path.join(d, name + ".bin"), dtype=dtype)


@pytest.mark.skipif(not os.path.exists(BIN), reason="hybird_gpu not built (needs the reference's hybird.cpp)")
def test_dropin_main_matches_oracle(tmp_path):
    cfg, rec, z = load_case("couple2")
    cycles = 12
    tmp = str(tmp_path)
    path = _write_case(tmp, cfg, z["bed_x"], z["bed_r"], cycles)
    dump = os.path.join(tmp, "dump")
    os.makedirs(dump)
    env = dict(os.environ, HYBIRD_AMD_DUMP_DIR=dump)
    print("[gpu dropin] hybird_gpu couple2 12 cycles", flush=True)
    r = subprocess.run([BIN, "-c", path, "-d", os.path.join(tmp, "out")], env=env, capture_output=True, text=True,
                       timeout=300, cwd=tmp)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    orc = OracleSim(fixture_config(cfg), rec)
    orc.cycle(cycles)
    t = orc.get("types")
    assert np.array_equal(_load(dump, "lb_types", np.uint8), t)
    act = (t == 0) | (t == 3)
    f = _load(dump, "lb_f").reshape(-1, 19)
    assert rel_l2(f[act], orc.get("f")[act]) <= 1e-12
    ex = _load(dump, "dem_e_x").reshape(-1, 6, 3)
    assert rel_l2(ex, orc.get("e_x")) <= 1e-11
    assert np.array_equal(_load(dump, "dem_neighborTable", np.uint32), orc.get("neighborTable"))


@pytest.mark.skipif(not os.path.exists(BIN), reason="hybird_gpu not built (needs the reference's hybird.cpp)")
def test_dropin_restart_matches_straight_run(tmp_path):
    """saveStateFile / restartFile (host-layer keys): 5 cycles, save, 7 cycles
    from the file equals 12 cycles in one run, bit for bit."""
    cfg, rec, z = load_case("couple2")
    tmp = str(tmp_path)

    def run(tag, cycles, **keys):
        d = os.path.join(tmp, tag)
        os.makedirs(d)
        path = _write_case(d, dict(cfg, **keys), z["bed_x"], z["bed_r"], cycles)
        env = dict(os.environ, HYBIRD_AMD_DUMP_DIR=d)
        print(f"[gpu dropin] hybird_gpu couple2 {tag} {cycles} cycles", flush=True)
        r = subprocess.run([BIN, "-c", path, "-d", os.path.join(d, "out")], env=env, capture_output=True, text=True,
                           timeout=300, cwd=d)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        return d

    state = os.path.join(tmp, "couple2.state")
    straight = run("straight", 12)
    run("first", 5, saveStateFile=state)
    assert os.path.getsize(state) > 0
    resumed = run("resumed", 7, restartFile=state)
    for name, dt in (("lb_types", np.uint8), ("lb_f", np.float64), ("dem_e_x", np.float64),
                     ("dem_e_forces", np.float64), ("dem_neighborTable", np.uint32)):
        a, b = _load(straight, name, dt), _load(resumed, name, dt)
        assert np.array_equal(a, b, equal_nan=(dt == np.float64)), name
