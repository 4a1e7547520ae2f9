"""ctypes declarations of the C ABI in include/hybird_amd.h.

``HybConfig`` mirrors ``struct hyb_config`` field for field; ``config_from_cfg``
maps reference ``.cfg`` keys (hybird.cpp:136-190, LB.cpp:93-185,
DEM.cpp:15-181) onto it.
"""
from __future__ import annotations

import ctypes as C

HYB_PROBLEM_NONE = 0
HYB_PROBLEM_SPLASH = 2
HYB_PROBLEM_DRUM = 4
HYB_PROBLEM_NET = 6
HYB_PROBLEM_LIQUID_BOX = 100
HYB_CONTACT_LINEAR = 0
HYB_CONTACT_HERTZIAN = 1
HYB_COLLISION_BGK = 0
HYB_COLLISION_TRT = 1
HYB_FP64 = 0
HYB_FP32_SHIFTED = 1
HYB_RHEOLOGY_BINGHAM = 0
HYB_RHEOLOGY_FRICTIONAL = 1
HYB_TANGENTIAL_VISCOUS = 0
HYB_TANGENTIAL_SPRING = 1
HYB_WALL_CONTACTS_FIRST = 0
HYB_WALL_CONTACTS_ALL = 1

ELEMENT_RECORD = 20

_PROBLEMS = {"NONE": HYB_PROBLEM_NONE, "SPLASH": HYB_PROBLEM_SPLASH, "DRUM": HYB_PROBLEM_DRUM, "NET": HYB_PROBLEM_NET,
             "LIQUID_BOX": HYB_PROBLEM_LIQUID_BOX}


class HybConfig(C.Structure):
    _fields_ = [
        ("problem", C.c_int32),
        ("demSolve", C.c_int32),
        ("lbSolve", C.c_int32),
        ("freeSurface", C.c_int32),
        ("forceField", C.c_int32),
        ("nonNewtonian", C.c_int32),
        ("turbulence", C.c_int32),
        ("demInitialRepeat", C.c_double),
        ("unitLength", C.c_double),
        ("unitTime", C.c_double),
        ("unitDensity", C.c_double),
        ("lbSize", C.c_double * 3),
        ("plasticVisc", C.c_double),
        ("yieldStress", C.c_double),
        ("initVisc", C.c_double),
        ("initVelocity", C.c_double * 3),
        ("lbF", C.c_double * 3),
        ("boundary", C.c_int32 * 6),
        ("slipCoefficient", C.c_double),
        ("turbConst", C.c_double),
        ("density", C.c_double),
        ("contactModel", C.c_int32),
        ("youngMod", C.c_double),
        ("poisson", C.c_double),
        ("linearStiff", C.c_double),
        ("restitution", C.c_double),
        ("viscTang", C.c_double),
        ("frictionCoefPart", C.c_double),
        ("frictionCoefWall", C.c_double),
        ("numVisc", C.c_double),
        ("multiStep", C.c_int32),
        ("criticalRatio", C.c_double),
        ("translate", C.c_double * 3),
        ("scale", C.c_double),
        ("collision", C.c_int32),
        ("trtMagic", C.c_double),
        ("precision", C.c_int32),
        ("liquidBox", C.c_double * 6),
        ("avalanchePosit", C.c_double),
        ("drumSpeed", C.c_double),
        ("fluidMass", C.c_double),
        ("pairCapacity", C.c_int32),
        ("rheology", C.c_int32),
        ("frictionCoefFluid", C.c_double),
        ("tangentialModel", C.c_int32),
        ("wallContacts", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


def config_from_cfg(cfg: dict, collision: str = "BGK", precision: str = "FP64", trt_magic: float = 0.25,
                    liquid_box=None) -> HybConfig:
    """Build a HybConfig from reference cfg keys (see hybird_amd.cfg.base_config)."""
    c = HybConfig()
    g = cfg.get
    c.problem = _PROBLEMS[str(g("problemName", "NONE")).upper()] if not isinstance(g("problemName"), int) else g(
        "problemName")
    c.demSolve = int(g("demSolve", 0))
    c.lbSolve = int(g("lbSolve", 0))
    c.freeSurface = int(g("freeSurfaceSolve", 0))
    c.forceField = int(g("forceFieldSolve", 0))
    c.nonNewtonian = int(g("nonNewtonianSolve", 0))
    c.turbulence = int(g("turbulenceSolve", 0))
    c.demInitialRepeat = float(g("demInitialRepeat", 0.0))
    c.unitLength = float(g("unitLength", 1.0))
    c.unitTime = float(g("unitTime", 1.0))
    c.unitDensity = float(g("unitDensity", 1.0))
    c.lbSize[:] = [float(g("lbSizeX")), float(g("lbSizeY")), float(g("lbSizeZ"))]
    c.plasticVisc = float(g("plasticVisc", 1.0))
    c.yieldStress = float(g("yieldStress", 1.0))
    c.initVisc = float(g("initVisc", 1.0))
    c.initVelocity[:] = [float(g("initVelocityX", 0.0)), float(g("initVelocityY", 0.0)), float(g("initVelocityZ", 0.0))]
    c.lbF[:] = [float(g("lbFX", 0.0)), float(g("lbFY", 0.0)), float(g("lbFZ", 0.0))]
    c.boundary[:] = [int(g(f"boundary{k}", 1)) for k in range(6)]
    c.slipCoefficient = float(g("slipCoefficient", 0.0))
    c.turbConst = float(g("turbConst", 0.0))
    c.density = float(g("density", 1.0))
    cm = str(g("contactModel", "HERTZIAN")).upper()
    c.contactModel = HYB_CONTACT_HERTZIAN if cm == "HERTZIAN" else HYB_CONTACT_LINEAR
    c.youngMod = float(g("youngMod", 1.0))
    c.poisson = float(g("poisson", 1.0))
    c.linearStiff = float(g("linearStiff", 1.0))
    c.restitution = float(g("restitution", 1.0))
    c.viscTang = float(g("viscTang", 1.0))
    c.frictionCoefPart = float(g("frictionCoefPart", 1.0))
    c.frictionCoefWall = float(g("frictionCoefWall", 1.0))
    c.numVisc = float(g("numVisc", 0.0))
    c.multiStep = int(g("multiStep", 1))
    c.criticalRatio = float(g("criticalRatio", 0.1))
    c.translate[:] = [float(g("translateX", 0.0)), float(g("translateY", 0.0)), float(g("translateZ", 0.0))]
    c.scale = float(g("scale", 1.0))
    c.collision = HYB_COLLISION_TRT if collision.upper() == "TRT" else HYB_COLLISION_BGK
    c.trtMagic = float(trt_magic)
    c.precision = HYB_FP32_SHIFTED if precision.upper() in ("FP32", "FP32_SHIFTED") else HYB_FP64
    if liquid_box is not None:
        c.liquidBox[:] = [float(v) for v in liquid_box]
    c.avalanchePosit = float(g("avalanchePosit", 0.0))
    c.drumSpeed = float(g("drumSpeed", 0.0))
    c.fluidMass = float(g("fluidMass", 0.0))
    c.rheology = HYB_RHEOLOGY_FRICTIONAL if str(g("rheology", "BINGHAM")).upper() == "FRICTIONAL" else HYB_RHEOLOGY_BINGHAM
    c.frictionCoefFluid = float(g("frictionCoefFluid", 0.0))
    c.tangentialModel = (HYB_TANGENTIAL_SPRING if str(g("tangentialModel", "VISCOUS")).upper() == "SPRING"
                         else HYB_TANGENTIAL_VISCOUS)
    c.wallContacts = (HYB_WALL_CONTACTS_ALL if str(g("wallContacts", "FIRST")).upper() == "ALL"
                      else HYB_WALL_CONTACTS_FIRST)
    return c


def element_records(x, r, v=None, size=None, w=None, q=None):
    """Particle-file records (DEM.cpp:85-120): centres x, radii r and optional
    velocities v, sphere counts size (1-4), angular velocities w and
    orientation quaternions q (defaults: at rest, single spheres, identity)."""
    import numpy as np

    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    n = x.shape[0]
    rec = np.zeros((n, ELEMENT_RECORD), dtype=np.float64)
    rec[:, 0] = np.arange(n)
    rec[:, 1] = 1 if size is None else np.asarray(size, dtype=np.float64).reshape(-1)
    rec[:, 2] = r
    rec[:, 3:6] = x
    if v is not None:
        rec[:, 6:9] = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    if w is not None:
        rec[:, 9:12] = np.asarray(w, dtype=np.float64).reshape(-1, 3)
    rec[:, 12] = 1.0
    if q is not None:
        rec[:, 12:16] = np.asarray(q, dtype=np.float64).reshape(-1, 4)
    return np.ascontiguousarray(rec)
