"""The no-allocation entry point hyb_lb_wall_fhydro_into against the wallFHydro view.

hyb_lb_wall_fhydro_into writes walls[].FHydro (LB.cpp:1321-1341) into device
memory the caller owns, using a fixed device array for its exact sums instead
of allocating. Both paths sum exactly (fixed point), so the values must be
bit-identical to the view's. Two calls in a row on different handles, the
second on a smaller lattice with fewer moving walls, check that the array is
cleared between calls (it keeps the last call's contents otherwise).
"""
import ctypes as C

import numpy as np
import pytest

from helpers import fixture_config, load_case

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime (for the caller-side device buffer of the test)."""
    import os

    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return C.CDLL(name if os.path.isabs(name) else name)
        except OSError:
            continue
    raise RuntimeError("libamdhip64.so not found")


def _into(sim, walls):
    """hyb_lb_wall_fhydro_into into a device buffer the test allocates, copied back after hyb_sync."""
    from hybird_amd import runtime

    hip = _hip()
    n = 3 * walls
    dptr = C.c_void_p()
    assert hip.hipMalloc(C.byref(dptr), C.c_size_t(8 * n)) == 0
    try:
        nan = np.full(n, np.nan)
        assert hip.hipMemcpy(dptr, nan.ctypes.data_as(C.c_void_p), C.c_size_t(8 * n), 1) == 0  # host to device
        rc = runtime.load().hyb_lb_wall_fhydro_into(sim._lb, dptr, n)
        assert rc == 0, runtime.load().hyb_last_error().decode()
        sim.sync()
        out = np.empty(n)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), dptr, C.c_size_t(8 * n), 2) == 0  # device to host
    finally:
        hip.hipFree(dptr)
    return out.reshape(walls, 3)


def test_no_alloc_wall_force_matches_view_on_two_calls():
    from hybird_amd import runtime

    # first call: the walls case (moving no-slip floor among six box walls), after 10 cycles
    cfg, rec, _ = load_case("walls")
    a = runtime.Simulation(fixture_config(cfg), rec)
    a.cycle(10)
    va = a.get("wallFHydro")
    print(f"[gpu no-alloc] walls case, {va.shape[0]} walls", flush=True)
    ga = _into(a, va.shape[0])
    assert np.abs(va).max() > 0.0
    assert np.array_equal(ga, va)
    a.close()
    # second call: a smaller lattice with two moving no-slip walls (z faces), sheared flow
    cfg2 = dict(cfg, lbSizeX=10, lbSizeY=8, lbSizeZ=9, boundary0=4, boundary1=4, boundary2=4, boundary3=4,
                boundary4=8, boundary5=8)
    b = runtime.Simulation(fixture_config(cfg2), rec[:1] * np.array([1, 1, 0.5] + [1.0] * 17))
    b.cycle(6)
    vb = b.get("wallFHydro")
    gb = _into(b, vb.shape[0])
    assert vb.shape[0] < va.shape[0] and np.abs(vb).max() > 0.0
    assert np.array_equal(gb, vb)
    b.close()
