"""Python access to libhybird_amd.so through its C ABI (ctypes).

``Simulation`` owns one LB handle and one DEM handle on the GPU and runs the
coupled cycle of the reference (goCycle, hybird.cpp:35-66). The library is
required: nothing here falls back to a CPU path, and loading fails loudly when
the shared library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._abi import HybConfig

HERE = os.path.dirname(os.path.abspath(__file__))
# HYBIRD_AMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("HYBIRD_AMD_LIB", os.path.join(HERE, "libhybird_amd.so"))

_lib = None

LB_FIELDS = {"f": (np.float64, (-1, 19)), "n": (np.float64, None), "u": (np.float64, (-1, 3)),
             "hydroForce": (np.float64, (-1, 3)), "mass": (np.float64, None), "visc": (np.float64, None),
             "types": (np.uint8, None), "pflag": (np.uint8, None), "solidIndex": (np.uint32, None),
             "totalMass": (np.float64, None), "fluidMass": (np.float64, None), "wallFHydro": (np.float64, (-1, 3))}
DEM_FIELDS = {"walls": (np.float64, (-1, 12)), "objectForces": (np.float64, (-1, 3)),
              "e_x": (np.float64, (-1, 6, 3)), "e_xp": (np.float64, (-1, 6, 3)), "e_q": (np.float64, (-1, 6, 4)),
              "e_qp": (np.float64, (-1, 6, 4)), "e_w": (np.float64, (-1, 6, 3)), "e_wp": (np.float64, (-1, 6, 3)),
              "e_forces": (np.float64, (-1, 7, 3)), "e_misc": (np.float64, (-1, 19)),
              "p_data": (np.float64, (-1, 10)), "p_idx": (np.uint32, (-1, 4)), "neighborTable": (np.uint32, None),
              "nearWallTable": (np.uint32, None), "ghostIndex": (np.uint32, None), "ghostVec": (np.float64, (-1, 3)),
              "dem_scalars": (np.float64, None), "rebuilds": (np.float64, None), "energy": (np.float64, None),
              "nearObjectTable": (np.uint32, None), "objects": (np.float64, (-1, 8))}

# every extern "C" entry point declared in include/hybird_amd.h
EXPORTS = ["hyb_config_defaults", "hyb_last_error", "hyb_device_count", "hyb_sync", "hyb_dem_create", "hyb_dem_init",
           "hyb_dem_step", "hyb_dem_destroy", "hyb_lb_create", "hyb_lb_init", "hyb_lb_fs_step", "hyb_lb_couple",
           "hyb_lb_step", "hyb_lb_destroy", "hyb_cycle", "hyb_lb_get", "hyb_dem_get", "hyb_lb_set_profiling",
           "hyb_lb_kernel_time", "hyb_lb_active_nodes", "hyb_lb_set_capture", "hyb_dem_init_lattice",
           "hyb_lb_set_force_field", "hyb_lb_set_populations", "hyb_adj_forward", "hyb_adj_backward",
           "hyb_state_save", "hyb_state_load", "hyb_dem_set_objects", "hyb_dem_set_save_count", "hyb_dem_saved",
           "hyb_dem_saved_get", "hyb_copy_bandwidth", "hyb_lb_wall_fhydro_into", "hyb_adj_device_time"]


class HybError(RuntimeError):
    """A failed C-ABI call; ``code`` is its HYB_E_* return value."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


def load():
    """Load libhybird_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HybError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.hyb_last_error.restype = C.c_char_p
    L.hyb_config_defaults.argtypes = [vp]
    L.hyb_config_defaults.restype = None
    for name, args in [("hyb_device_count", [C.POINTER(C.c_int)]), ("hyb_sync", [vp]),
                       ("hyb_dem_create", [vp, i32, vp, C.POINTER(vp)]), ("hyb_dem_init", [vp, vp]),
                       ("hyb_dem_init_lattice", [vp, vp, vp, C.c_double, C.c_double, C.c_double, vp]),
                       ("hyb_dem_step", [vp]), ("hyb_dem_destroy", [vp]), ("hyb_lb_create", [vp, C.POINTER(vp)]),
                       ("hyb_lb_init", [vp, vp]), ("hyb_lb_fs_step", [vp]),
                       ("hyb_lb_couple", [vp, vp, C.POINTER(i32)]), ("hyb_lb_step", [vp, vp]),
                       ("hyb_lb_destroy", [vp]), ("hyb_cycle", [vp, vp, i32]),
                       ("hyb_lb_get", [vp, C.c_char_p, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
                       ("hyb_dem_get", [vp, C.c_char_p, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
                       ("hyb_lb_set_profiling", [vp, i32]), ("hyb_lb_set_capture", [vp, i32]),
                       ("hyb_lb_set_force_field", [vp, vp]), ("hyb_lb_set_populations", [vp, vp]),
                       ("hyb_adj_forward", [vp, vp, i32, i32]), ("hyb_adj_backward", [vp, vp, vp, vp, vp, vp]),
                       ("hyb_lb_kernel_time", [vp, C.POINTER(C.c_double), C.POINTER(i64)]),
                       ("hyb_lb_active_nodes", [vp, C.POINTER(i64)]),
                       ("hyb_state_save", [vp, vp, C.c_char_p]), ("hyb_state_load", [vp, vp, C.c_char_p]),
                       ("hyb_dem_set_objects", [vp, i32, vp]), ("hyb_dem_set_save_count", [vp, i32]),
                       ("hyb_dem_saved", [vp, C.POINTER(i32)]),
                       ("hyb_dem_saved_get", [vp, i32, C.c_char_p, vp, C.c_size_t, C.POINTER(C.c_size_t)]),
                       ("hyb_copy_bandwidth", [C.c_size_t, i32, i32, C.POINTER(C.c_double)]),
                       ("hyb_adj_device_time", [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
                       ("hyb_lb_wall_fhydro_into", [vp, vp, i32])]:
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise HybError(f"{what} failed ({rc}): {load().hyb_last_error().decode()}", rc)


def copy_bandwidth(nbytes: int, width: int, reps: int = 20) -> float:
    """GB/s (read + written) of a plain non-temporal device copy with `width`-byte lanes."""
    g = C.c_double(0.0)
    _check(load().hyb_copy_bandwidth(int(nbytes), int(width), int(reps), C.byref(g)), "hyb_copy_bandwidth")
    return g.value


def default_config() -> HybConfig:
    c = HybConfig()
    load().hyb_config_defaults(C.byref(c))
    return c


class Simulation:
    """Coupled LB + DEM state on the GPU, created as the reference's main() does."""

    def __init__(self, config: HybConfig, records: np.ndarray, objects=None):
        """objects: optional (n, 9) object records (index, size, r, x0, x1)."""
        L = load()
        rec = np.ascontiguousarray(records, dtype=np.float64).reshape(-1, 20)
        self.config = config
        self._dem = C.c_void_p()
        self._lb = C.c_void_p()
        _check(L.hyb_dem_create(C.byref(config), int(rec.shape[0]), rec.ctypes.data if rec.size else None,
                                C.byref(self._dem)), "hyb_dem_create")
        if objects is not None and len(objects):
            obj = np.ascontiguousarray(objects, dtype=np.float64).reshape(-1, 9)
            _check(L.hyb_dem_set_objects(self._dem, int(obj.shape[0]), obj.ctypes.data), "hyb_dem_set_objects")
        _check(L.hyb_lb_create(C.byref(config), C.byref(self._lb)), "hyb_lb_create")
        _check(L.hyb_dem_init(self._dem, self._lb), "hyb_dem_init")
        _check(L.hyb_lb_init(self._lb, self._dem), "hyb_lb_init")

    def cycle(self, n: int = 1) -> None:
        _check(load().hyb_cycle(self._lb, self._dem, int(n)), "hyb_cycle")

    def sync(self) -> None:
        _check(load().hyb_sync(self._lb), "hyb_sync")

    def get(self, name: str) -> np.ndarray:
        L = load()
        if name in LB_FIELDS:
            fn, h, (dt, shp) = L.hyb_lb_get, self._lb, LB_FIELDS[name]
        elif name in DEM_FIELDS:
            fn, h, (dt, shp) = L.hyb_dem_get, self._dem, DEM_FIELDS[name]
        else:
            raise KeyError(name)
        need = C.c_size_t(0)
        _check(fn(h, name.encode(), None, 0, C.byref(need)), f"get {name}")
        out = np.empty(need.value // np.dtype(dt).itemsize, dtype=dt)
        if need.value:
            _check(fn(h, name.encode(), out.ctypes.data, need.value, C.byref(need)), f"get {name}")
        return out.reshape(shp) if shp else out

    def set_capture(self, on: bool) -> None:
        """Keep n, u and hydroForce of every collision (needed for those views with free surface)."""
        _check(load().hyb_lb_set_capture(self._lb, 1 if on else 0), "hyb_lb_set_capture")

    def set_profiling(self, on: bool) -> None:
        _check(load().hyb_lb_set_profiling(self._lb, 1 if on else 0), "hyb_lb_set_profiling")

    def kernel_time(self):
        ms = C.c_double(0.0)
        n = C.c_int64(0)
        _check(load().hyb_lb_kernel_time(self._lb, C.byref(ms), C.byref(n)), "hyb_lb_kernel_time")
        return ms.value, n.value

    # ---- adjoint (FP64, particles frozen)
    def set_force_field(self, F) -> None:
        """Per-node body force [N, 3] (float32, lattice units); None removes it."""
        if F is None:
            _check(load().hyb_lb_set_force_field(self._lb, None), "hyb_lb_set_force_field")
            return
        a = np.ascontiguousarray(F, dtype=np.float32).reshape(-1, 3)
        _check(load().hyb_lb_set_force_field(self._lb, a.ctypes.data), "hyb_lb_set_force_field")

    def set_populations(self, f) -> None:
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1, 19)
        _check(load().hyb_lb_set_populations(self._lb, a.ctypes.data), "hyb_lb_set_populations")

    def adj_forward(self, steps: int, every: int) -> None:
        _check(load().hyb_adj_forward(self._lb, self._dem, int(steps), int(every)), "hyb_adj_forward")

    def adj_backward(self, seed):
        """Returns (dL/dF [N,3], dL/dtau, dL/df0 [N,19]) for seed = dL/df_T [N,19]."""
        sd = np.ascontiguousarray(seed, dtype=np.float64).reshape(-1, 19)
        n = sd.shape[0]
        gF = np.zeros((n, 3))
        gT = np.zeros(1)
        g0 = np.zeros((n, 19))
        _check(load().hyb_adj_backward(self._lb, self._dem, sd.ctypes.data, gF.ctypes.data, gT.ctypes.data,
                                       g0.ctypes.data), "hyb_adj_backward")
        return gF, float(gT[0]), g0

    def adj_device_time(self):
        """Device milliseconds of the last adjoint forward loop and reverse loop (HIP events)."""
        f, b = C.c_double(0.0), C.c_double(0.0)
        _check(load().hyb_adj_device_time(self._lb, C.byref(f), C.byref(b)), "hyb_adj_device_time")
        return f.value, b.value

    def step_lb(self, n: int = 1) -> None:
        """LB steps only (no DEM sub-steps, no coupling): particles stay frozen."""
        L = load()
        for _ in range(int(n)):
            _check(L.hyb_lb_step(self._lb, self._dem), "hyb_lb_step")

    def save_state(self, path) -> None:
        """Checkpoint of the lattice and DEM state (full precision)."""
        _check(load().hyb_state_save(self._lb, self._dem, str(path).encode()), "hyb_state_save")

    def load_state(self, path) -> None:
        """Restores a checkpoint written from the same configuration and elements."""
        _check(load().hyb_state_load(self._lb, self._dem, str(path).encode()), "hyb_state_load")

    def set_save_count(self, save_count: int) -> None:
        """MercuryDPM-format saving every save_count DEM sub-steps (0: off)."""
        _check(load().hyb_dem_set_save_count(self._dem, int(save_count)), "hyb_dem_set_save_count")

    def saved(self):
        """Captures of the saving sub-steps of the last DEM step: a list of dicts with
        info (demTime, demTimeStep, pairs, ...), elements [E][9] and the pairs / walls /
        objects statistics rows [.][12]."""
        L = load()
        n = C.c_int32(0)
        _check(L.hyb_dem_saved(self._dem, C.byref(n)), "hyb_dem_saved")
        out = []
        for ev in range(n.value):
            d = {}
            for part, cols in (("info", 10), ("elements", 9), ("pairs", 12), ("walls", 12), ("objects", 12)):
                need = C.c_size_t(0)
                _check(L.hyb_dem_saved_get(self._dem, ev, part.encode(), None, 0, C.byref(need)), "hyb_dem_saved_get")
                a = np.empty(need.value // 8, dtype=np.float64)
                if need.value:
                    _check(L.hyb_dem_saved_get(self._dem, ev, part.encode(), a.ctypes.data, need.value, C.byref(need)),
                           "hyb_dem_saved_get")
                d[part] = a if part == "info" else a.reshape(-1, cols)
            out.append(d)
        return out

    def active_nodes(self) -> int:
        n = C.c_int64(0)
        _check(load().hyb_lb_active_nodes(self._lb, C.byref(n)), "hyb_lb_active_nodes")
        return n.value

    def close(self) -> None:
        if self._lb:
            load().hyb_lb_destroy(self._lb)
            self._lb = C.c_void_p()
        if self._dem:
            load().hyb_dem_destroy(self._dem)
            self._dem = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
