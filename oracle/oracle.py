"""ORACLE — test infrastructure, not the product.

ctypes wrapper of oracle/liboracle.so, the FP64 CPU restatement of the
reference's coupled LBM-DEM cycle. Only tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg use it, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "hybird_ref_dump")

_lib = None

_DTYPES = {
    "types": np.uint8, "pflag": np.uint8, "solidIndex": np.uint32, "activeNodes": np.uint32, "fluidNodes": np.uint32,
    "interfaceNodes": np.uint32, "particleNodes": np.uint32, "neighbors": np.uint32, "lbSize": np.uint32,
    "p_idx": np.uint32, "neighborTable": np.uint32, "nearWallTable": np.uint32, "cellTable": np.int32,
    "ghostIndex": np.uint32, "e_comp": np.uint32, "e_comp_off": np.uint32, "nearObjectTable": np.uint32,
}
_SHAPES = {"f": (-1, 19), "u": (-1, 3), "hydroForce": (-1, 3), "neighbors": (-1, 19), "e_x": (-1, 6, 3),
           "e_xp": (-1, 6, 3), "e_q": (-1, 6, 4), "e_qp": (-1, 6, 4), "e_w": (-1, 6, 3), "e_wp": (-1, 6, 3),
           "e_misc": (-1, 19), "e_forces": (-1, 7, 3), "p_data": (-1, 10), "p_idx": (-1, 4), "ghostVec": (-1, 3),
           "walls": (-1, 12), "objects": (-1, 11)}


def build() -> None:
    subprocess.run(["make", "-C", HERE, "oracle"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle oracle`")
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_create.restype = C.c_void_p
        _lib.orc_create.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_char_p, C.c_int]
        _lib.orc_destroy.argtypes = [C.c_void_p]
        _lib.orc_cycle.argtypes = [C.c_void_p, C.c_int]
        _lib.orc_cycle.restype = C.c_int
        _lib.orc_lb_step.argtypes = [C.c_void_p, C.c_int]
        _lib.orc_lb_step.restype = C.c_int
        _lib.orc_error.argtypes = [C.c_void_p]
        _lib.orc_error.restype = C.c_char_p
        _lib.orc_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_long]
        _lib.orc_get.restype = C.c_long
    return _lib


class OracleSim:
    """One coupled simulation in the CPU restatement."""

    def __init__(self, config, records: np.ndarray, objects=None):
        L = lib()
        rec = np.ascontiguousarray(records, dtype=np.float64).reshape(-1, 20)
        obj = np.ascontiguousarray(np.zeros((0, 9)) if objects is None else objects, dtype=np.float64).reshape(-1, 9)
        err = C.create_string_buffer(512)
        self._h = L.orc_create(C.byref(config), int(rec.shape[0]), rec.ctypes.data if rec.size else None,
                               int(obj.shape[0]), obj.ctypes.data if obj.size else None, err, 512)
        if not self._h:
            raise RuntimeError("oracle create failed: " + err.value.decode())

    def cycle(self, n: int = 1) -> None:
        rc = lib().orc_cycle(self._h, int(n))
        if rc != 0:
            raise RuntimeError(f"oracle cycle failed ({rc}): {lib().orc_error(self._h).decode()}")

    def lb_step(self, n: int = 1) -> None:
        """LB steps only: particles frozen (no DEM sub-steps, no coupling)."""
        rc = lib().orc_lb_step(self._h, int(n))
        if rc != 0:
            raise RuntimeError(f"oracle LB step failed ({rc}): {lib().orc_error(self._h).decode()}")

    def get(self, name: str) -> np.ndarray:
        L = lib()
        need = L.orc_get(self._h, name.encode(), None, 0)
        if need < 0:
            raise KeyError(name)
        dt = np.dtype(_DTYPES.get(name, np.float64))
        out = np.empty(need // dt.itemsize, dtype=dt)
        if need:
            L.orc_get(self._h, name.encode(), out.ctypes.data, need)
        shp = _SHAPES.get(name)
        return out.reshape(shp) if shp else out

    def close(self) -> None:
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
