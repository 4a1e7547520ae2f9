"""Run-to-run determinism of the coupled step on the HIP path.

Every order-dependent sum of the step is either sequential in a fixed order or
an exact fixed-point sum, so two runs of the same input must agree bit for bit.
The case has more than 1024 particles (the dense suspension's radii and solid
fraction on a 160^3 lattice), so that the per-element hydrodynamic reduction
runs over many workgroups; a race there shows as element forces that differ
between the runs after the first LB step.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(cfg, rec, cycles, precision):
    from hybird_amd import runtime
    from hybird_amd._abi import config_from_cfg

    sim = runtime.Simulation(config_from_cfg(cfg, collision="TRT", precision=precision), rec)
    sim.cycle(cycles)
    out = {k: sim.get(k).copy() for k in ("e_forces", "e_x", "e_w", "solidIndex")}
    sim.close()
    return out


@pytest.mark.parametrize("precision", ["FP64", "FP32"])
def test_two_runs_are_bit_identical(precision):
    import bench
    from hybird_amd._abi import element_records

    cfg, bed = bench.workload(160)
    assert bed.count > 1024
    rec = element_records(bed.x, bed.r)
    print(f"[gpu determinism] {precision} 160^3, {bed.count} spheres, 3 cycles, twice", flush=True)
    a = _run(cfg, rec, 3, precision)
    b = _run(cfg, rec, 3, precision)
    for k in a:
        assert np.array_equal(a[k], b[k]), (k, np.nonzero(np.any((a[k] != b[k]).reshape(a[k].shape[0], -1), axis=1))[0][:10])
