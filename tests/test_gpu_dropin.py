"""The reference's own main (hybird.cpp, unchanged) driving the GPU through the
C++ host layer (hybird_amd/host): the binary hybird_gpu is built from the
reference sources by `make -C hybird_amd/host dropin`. The run writes its final
state (HYBIRD_AMD_DUMP_DIR) and is compared with the oracle after the same
number of cycles."""
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import ROOT, fixture_config, load_case, rel_l2
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "hybird_amd", "host", "hybird_gpu")


def _write_case(tmp, cfg, bed_x, bed_r, cycles, objects=None, bed_v=None, z=None):
    """The cfg, particle and object files of a run; z: a fixture whose clump fields (sizes, spins,
    orientations) go into the particle file."""
    from hybird_amd.cfg import write_cfg
    from hybird_amd.particles import Bed, write_empty_object_file, write_object_file, write_particle_file
    pf, of = os.path.join(tmp, "p.dat"), os.path.join(tmp, "o.dat")
    opt = {k: z[f"bed_{k}"] for k in ("sizes", "omegas", "quats") if z is not None and f"bed_{k}" in z.files}
    write_particle_file(pf, Bed(np.asarray(bed_x), np.asarray(bed_r)), bed_v, **opt)
    if objects is not None and len(objects):
        write_object_file(of, objects)
    else:
        write_empty_object_file(of)
    cfg = dict(cfg, particleFile=pf, objectFile=of, maximumTimeSteps=cycles, screenExpTime=cfg.get("screenExpTime", 0))
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


def _arrays(text):
    """Numbers of every named DataArray of a VTK text file (ASCII), by name."""
    import re
    out, name = {}, None
    for part in re.split(r"(<[^>]*>)", text):
        if part.startswith("<"):
            m = re.search(r'Name="([^"]+)"', part)
            name = m.group(1) if m else None
        elif name is not None and part.strip():
            out.setdefault(name, []).extend(float(v) for v in part.split())
    return {k: np.array(v) for k, v in out.items()}


def _close(a, b, what):
    assert a.shape == b.shape, (what, a.shape, b.shape)
    scale = max(float(np.max(np.abs(a))) if a.size else 0.0, 1e-300)
    # six significant digits as printed, plus a floor for components that are
    # zero up to rounding in the reference's arithmetic
    bad = np.abs(a - b) > 1e-5 * np.maximum(np.abs(a), np.abs(b)) + 1e-9 * scale
    assert not bad.any(), (what, np.flatnonzero(bad)[:5], a[bad][:5], b[bad][:5])


@pytest.mark.skipif(not os.path.exists(BIN), reason="hybird_gpu not built (needs the reference's hybird.cpp)")
@pytest.mark.parametrize("name", ["io", "io_fs"])
def test_dropin_output_files_match_reference(tmp_path, name):
    """Fluid .vti, particle .vtu and recycle .dat files, and the status line
    (export.dat) with its side files (energy.dat, maxVel.dat, force.dat,
    plasticity.dat), written by the host layer from the GPU state against the files
    the reference wrote for the same run (the fixture holds them as bytes): same
    file names, same structure, numbers equal to the printed six digits.
    solidIndex is compared on particle and wall nodes (elsewhere the reference
    keeps the index of the last particle there)."""
    cfg, rec, z = load_case(name)
    cycles = int(max(z["cycles"]))
    tmp = str(tmp_path)
    path = _write_case(tmp, cfg, z["bed_x"], z["bed_r"], cycles)
    print(f"[gpu dropin] hybird_gpu {name} output files", flush=True)
    os.makedirs(os.path.join(tmp, "out"))  # the driver creates only the last level (hybird.cpp:113-115)
    r = subprocess.run([BIN, "-c", path, "-d", os.path.join(tmp, "out"), "-n", "run"], capture_output=True, text=True,
                       timeout=300, cwd=tmp)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    mine = {}
    for dirpath, _, names in os.walk(os.path.join(tmp, "out")):
        top = os.path.basename(dirpath) not in ("fluidData", "particleData")
        for fn in names:
            if not top or fn.endswith(".dat"):
                mine[fn] = open(os.path.join(dirpath, fn), "rb").read().decode()
    ref = {k.split("/", 1)[1]: bytes(z[k]).decode() for k in z.files if k.startswith("file/")}
    assert {"export.dat", "energy.dat", "maxVel.dat", "force.dat", "plasticity.dat"} <= set(ref)
    assert sorted(mine) == sorted(ref), (sorted(set(mine) - set(ref)), sorted(set(ref) - set(mine)))
    for fn, text in ref.items():
        if fn in STATUS_FILES:
            _status_close(text, mine[fn], fn)
            continue
        if fn.endswith(".dat"):
            a = np.array([float(v) for v in text.split()])
            b = np.array([float(v) for v in mine[fn].split()])
            _close(a, b, fn)
            continue
        tags_ref = [t for t in text.split() if t.startswith("<")]
        tags_mine = [t for t in mine[fn].split() if t.startswith("<")]
        assert tags_ref == tags_mine, fn
        ar, am = _arrays(text), _arrays(mine[fn])
        assert sorted(ar) == sorted(am), fn
        for key in ar:
            if key == "solidIndex":
                sel = np.isin(ar["type"], [1, 5, 6, 7, 8])
                assert np.array_equal(ar[key][sel], am[key][sel]), (fn, key)
            elif key == "type":
                assert np.array_equal(ar[key], am[key]), (fn, key)
            else:
                _close(ar[key], am[key], f"{fn}:{key}")


STATUS_FILES = ("export.dat", "energy.dat", "maxVel.dat", "force.dat", "plasticity.dat")


def _status_close(text, mine, fn):
    """Words equal, numbers to the six printed digits ("MaxFSpeed= 0.02(2%)" splits at the brackets)."""
    split = lambda t: [w for line in t.splitlines() for w in re.split(r"[\s()%;]+", line) if w] + ["|"]
    wr, wm = split(text), split(mine)
    assert len(text.splitlines()) == len(mine.splitlines()) and len(wr) == len(wm), fn
    for x, y in zip(wr, wm):
        try:
            fx, fy = float(x), float(y)
        except ValueError:
            assert x == y, (fn, x, y)
            continue
        assert abs(fx - fy) <= 1e-5 * max(abs(fx), abs(fy)) + 1e-12, (fn, x, y)


def _tokens_close(ref, mine, what, rtol=1e-9):
    """Line by line and word by word: words equal, or numbers equal to rtol."""
    lr, lm = ref.splitlines(), mine.splitlines()
    assert len(lr) == len(lm), f"{what}: {len(lm)} lines, reference {len(lr)}"
    for n, (a, b) in enumerate(zip(lr, lm)):
        wa, wb = a.split(), b.split()
        assert len(wa) == len(wb), f"{what} line {n + 1}: {b!r} vs reference {a!r}"
        for x, y in zip(wa, wb):
            try:
                fx, fy = float(x), float(y)
            except ValueError:
                assert x == y, f"{what} line {n + 1}: {y!r} vs reference {x!r}"
                continue
            assert abs(fx - fy) <= rtol * max(abs(fx), abs(fy)) + 1e-14, f"{what} line {n + 1}: {y} vs reference {x}"


@pytest.mark.skipif(not os.path.exists(BIN), reason="hybird_gpu not built (needs the reference's hybird.cpp)")
def test_dropin_mercury_files_match_reference(tmp_path):
    """MercuryDPM .data, .restart and .fstat files (DEM.cpp:2280-2565) written by the
    host layer from the sub-step captures of the GPU DEM step, against the files the
    reference wrote for the same run: saving every 3 sub-steps (inside LB steps of
    10), with particle, wall and object contacts and a pair through a ghost (which
    the reference does not write). Same lines and words; numbers to 1e-9. The same
    run writes object .vtu files (IO.cpp:422-480) with the objects' contact force,
    printed to six digits."""
    cfg, rec, z = load_case("mercury")
    cycles = int(max(z["cycles"]))
    tmp = str(tmp_path)
    path = _write_case(tmp, cfg, z["bed_x"], z["bed_r"], cycles, objects=z["objects"], bed_v=z["bed_v"])
    print("[gpu dropin] hybird_gpu mercury files", flush=True)
    os.makedirs(os.path.join(tmp, "out"))
    r = subprocess.run([BIN, "-c", path, "-d", os.path.join(tmp, "out"), "-n", "run"], capture_output=True, text=True,
                       timeout=300, cwd=tmp)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    mine = {}
    for dirpath, _, names in os.walk(os.path.join(tmp, "out")):  # -n run: files go to out/run
        for fn in names:
            if fn.endswith((".data", ".restart", ".fstat")) or fn.startswith("object"):
                mine[fn] = open(os.path.join(dirpath, fn)).read()
    ref = {k.split("/", 1)[1]: bytes(z[k]).decode() for k in z.files if k.startswith("file/")}
    ref = {k: v for k, v in ref.items() if k.endswith((".data", ".restart", ".fstat")) or k.startswith("object")}
    assert {"NONE.data", "NONE.fstat", "NONE.restart"} <= set(ref)
    assert sum(fn.startswith("object") for fn in ref) == 3
    assert sorted(mine) == sorted(ref)
    for fn, text in ref.items():
        _tokens_close(text, mine[fn], fn, rtol=2e-5 if fn.endswith(".vtu") else 1e-9)


@pytest.mark.skipif(not os.path.exists(BIN), reason="hybird_gpu not built (needs the reference's hybird.cpp)")
def test_dropin_clumps_match_oracle_and_reference_files(tmp_path):
    """Multi-sphere elements through the reference's main: the state after 15
    cycles against the oracle, and the output files of an 8-cycle run
    (MercuryDPM .data, .restart, .fstat per sphere every 3 sub-steps, particle
    .vtu per sphere, recycle .dat per element, the status line and its side
    files) against the files the reference wrote for it."""
    cfg, rec, z = load_case("clumps")
    tmp = str(tmp_path)
    d = os.path.join(tmp, "state")
    os.makedirs(d)
    path = _write_case(d, cfg, z["bed_x"], z["bed_r"], 15, bed_v=z["bed_v"], z=z)
    env = dict(os.environ, HYBIRD_AMD_DUMP_DIR=d)
    print("[gpu dropin] hybird_gpu clumps 15 cycles", flush=True)
    r = subprocess.run([BIN, "-c", path, "-d", os.path.join(d, "out")], env=env, capture_output=True, text=True,
                       timeout=300, cwd=d)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    orc = OracleSim(fixture_config(cfg), rec)
    orc.cycle(15)
    t = orc.get("types")
    assert np.array_equal(_load(d, "lb_types", np.uint8), t)
    act = (t == 0) | (t == 3)
    assert rel_l2(_load(d, "lb_f").reshape(-1, 19)[act], orc.get("f")[act]) <= 1e-12
    assert rel_l2(_load(d, "dem_e_x").reshape(-1, 6, 3), orc.get("e_x")) <= 1e-11
    assert np.array_equal(_load(d, "dem_neighborTable", np.uint32), orc.get("neighborTable"))

    cfg, rec, z = load_case("clumps_io")
    d = os.path.join(tmp, "files")
    os.makedirs(os.path.join(d, "out"))
    path = _write_case(d, cfg, z["bed_x"], z["bed_r"], int(max(z["cycles"])), bed_v=z["bed_v"], z=z)
    print("[gpu dropin] hybird_gpu clumps output files", flush=True)
    r = subprocess.run([BIN, "-c", path, "-d", os.path.join(d, "out"), "-n", "run"], capture_output=True, text=True,
                       timeout=300, cwd=d)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    mine = {}
    for dirpath, _, names in os.walk(os.path.join(d, "out")):
        for fn in names:
            mine[fn] = open(os.path.join(dirpath, fn)).read()
    ref = {k.split("/", 1)[1]: bytes(z[k]).decode() for k in z.files if k.startswith("file/")}
    assert {"NONE.data", "NONE.fstat", "NONE.restart", "part0000000004.vtu", "recycle_4.dat"} <= set(ref)
    for fn, text in ref.items():
        assert fn in mine, fn
        if fn in STATUS_FILES:
            _status_close(text, mine[fn], fn)
        elif fn.endswith((".data", ".restart", ".fstat")):
            _tokens_close(text, mine[fn], fn, rtol=1e-9)
        elif fn.endswith(".vtu"):
            ar, am = _arrays(text), _arrays(mine[fn])
            assert sorted(ar) == sorted(am), fn
            for key in ar:
                _close(ar[key], am[key], f"{fn}:{key}")
        else:  # recycle files: six printed digits
            a = np.array([float(v) for v in text.split()])
            b = np.array([float(v) for v in mine[fn].split()])
            _close(a, b, fn)


@pytest.mark.skipif(not os.path.exists(BIN), reason="hybird_gpu not built (needs the reference's hybird.cpp)")
def test_dropin_object_file_matches_oracle(tmp_path):
    """An object file with static spheres (one near a periodic plane, so it gets
    periodic copies): the reference's main over the GPU against the oracle."""
    from helpers import objects_of
    cfg, rec, z = load_case("objects")
    cycles = int(max(z["cycles"]))
    tmp = str(tmp_path)
    path = _write_case(tmp, cfg, z["bed_x"], z["bed_r"], cycles, objects=z["objects"], bed_v=z["bed_v"])
    dump = os.path.join(tmp, "dump")
    os.makedirs(dump)
    env = dict(os.environ, HYBIRD_AMD_DUMP_DIR=dump)
    print(f"[gpu dropin] hybird_gpu objects {cycles} cycles", flush=True)
    r = subprocess.run([BIN, "-c", path, "-d", os.path.join(tmp, "out")], env=env, capture_output=True, text=True,
                       timeout=300, cwd=tmp)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    orc = OracleSim(fixture_config(cfg), rec, objects_of(z))
    orc.cycle(cycles)
    t = orc.get("types")
    assert np.array_equal(_load(dump, "lb_types", np.uint8), t)
    act = (t == 0) | (t == 3)
    f = _load(dump, "lb_f").reshape(-1, 19)
    assert rel_l2(f[act], orc.get("f")[act]) <= 1e-12
    ex = _load(dump, "dem_e_x").reshape(-1, 6, 3)
    assert rel_l2(ex, orc.get("e_x")) <= 1e-11


def test_dropin_drum_matches_reference(tmp_path):
    """problemName DRUM through the reference's own main (drumSpeed and fluidMass
    parsed by its driver, hybird.cpp:199-205) against the reference's dump of the
    same run (tests/golden/drum.npz) at the last fixture cycle."""
    cfg, rec, z = load_case("drum")
    cycles = int(max(z["cycles"]))
    tmp = str(tmp_path)
    path = _write_case(tmp, cfg, z["bed_x"], z["bed_r"], cycles, bed_v=z["bed_v"])
    dump = os.path.join(tmp, "dump")
    os.makedirs(dump)
    env = dict(os.environ, HYBIRD_AMD_DUMP_DIR=dump)
    print(f"[gpu dropin] hybird_gpu drum {cycles} cycles", flush=True)
    r = subprocess.run([BIN, "-c", path, "-d", os.path.join(tmp, "out")], env=env, capture_output=True, text=True,
                       timeout=300, cwd=tmp)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    t = z[f"c{cycles}/types"].ravel()
    assert np.array_equal(_load(dump, "lb_types", np.uint8).ravel(), t)
    act = (t == 0) | (t == 3)
    f = _load(dump, "lb_f").reshape(-1, 19)
    assert rel_l2(f[act], z[f"c{cycles}/f"].reshape(-1, 19)[act]) <= 1e-12
    # positions and velocities (the higher Gear slots carry 1/dt^k times rounding)
    ex = _load(dump, "dem_e_x").reshape(-1, 6, 3)
    assert rel_l2(ex[:, :2], z[f"c{cycles}/e_x"][:, :2]) <= 1e-11
