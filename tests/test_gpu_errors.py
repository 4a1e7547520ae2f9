"""Device-side failures reach the caller through the C ABI.

A DEM list that overflows its capacity, or a non-finite density or element
velocity, is found by a kernel while the stream runs ahead of the host; the
kernel records it in the handle's error word and every later DEM kernel does
nothing. These tests force each condition from the host with valid inputs (a
tiny pair capacity, a non-finite initial population, an element record whose
velocity overflows) and check that the failure is returned, with its code, by
the next calls and by a synchronising call, and that the DEM state stays frozen
at the failing sub-step. No kernel faults: the kernels' own checks turn each
condition into an error.
"""
import numpy as np
import pytest

from helpers import fixture_config, load_case

pytestmark = pytest.mark.gpu

HYB_E_CAPACITY = -4
HYB_E_NUMERIC = -6


def _runtime():
    from hybird_amd import runtime
    return runtime


def _expect(code, fn):
    rt = _runtime()
    with pytest.raises(rt.HybError) as ei:
        fn()
    assert ei.value.code == code, str(ei.value)


def test_pair_capacity_overflow_is_returned_and_freezes_the_dem():
    rt = _runtime()
    cfg, rec, _ = load_case("dem60")
    conf = fixture_config(cfg)
    conf.pairCapacity = 8  # the initial rebuild lists far more pairs
    print("[gpu errors] pair capacity 8 on dem60", flush=True)
    sim = rt.Simulation(conf, rec)
    _expect(HYB_E_CAPACITY, sim.sync)
    _expect(HYB_E_CAPACITY, lambda: sim.cycle(1))
    _expect(HYB_E_CAPACITY, lambda: sim.get("e_x"))
    sim.close()
    # with the automatic capacity the same run steps normally
    ok = rt.Simulation(fixture_config(cfg), rec)
    ok.cycle(2)
    ok.sync()
    ok.close()


def test_non_finite_element_velocity_stops_the_dem():
    rt = _runtime()
    cfg, rec, _ = load_case("dem60")
    rec = rec.copy()
    rec[3, 6] = 1.0e300  # |v|^2 overflows to inf
    print("[gpu errors] infinite element speed on dem60", flush=True)
    sim = rt.Simulation(fixture_config(cfg), rec)
    codes = []
    for _ in range(3):  # returned by the first call after the device ran the check
        try:
            sim.cycle(1)
        except rt.HybError as e:
            codes.append(e.code)
            break
    try:
        sim.sync()
    except rt.HybError as e:
        codes.append(e.code)
    assert codes and all(c == HYB_E_NUMERIC for c in codes), codes
    sim.close()


def test_non_finite_density_is_returned():
    rt = _runtime()
    cfg, rec, _ = load_case("lb16")
    print("[gpu errors] NaN population on lb16", flush=True)
    sim = rt.Simulation(fixture_config(cfg), rec)
    f = sim.get("f")
    t = sim.get("types")
    node = int(np.flatnonzero((t == 0))[100])
    f[node, 5] = np.nan
    sim.set_populations(f)
    sim.cycle(1)
    _expect(HYB_E_NUMERIC, sim.sync)
    _expect(HYB_E_NUMERIC, lambda: sim.cycle(1))
    sim.close()
