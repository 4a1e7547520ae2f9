"""Reference-compatible ``.cfg`` files (GetPot ``key = value`` syntax).

The keys and their meaning follow the reference's parse sites:
hybird.cpp:136-190 (run control), LB.cpp:93-185 (lattice), DEM.cpp:15-181
(material, particles).  ``base_config`` returns the lattice-unit defaults used
by every workload here (unitLength = unitTime = unitDensity = 1, so every unit
conversion in the reference is an exact multiplication by one).
"""
from __future__ import annotations


def base_config(**over) -> dict:
    cfg = {
        "problemName": "NONE",
        "demSolve": 1,
        "lbSolve": 1,
        "freeSurfaceSolve": 0,
        "forceFieldSolve": 1,
        "nonNewtonianSolve": 0,
        "turbulenceSolve": 0,
        "maximumTimeSteps": 10,
        "maxTime": 1.0e9,
        "saveCount": 0,
        "screenExpTime": 0,
        "fluidExpTime": 0,
        "partExpTime": 0,
        "recycleExpTime": 0,
        "objectExpTime": 0,
        "demInitialRepeat": 0,
        "lbmInitialRepeat": 0,
        "unitLength": 1,
        "unitTime": 1,
        "unitDensity": 1,
        "lbSizeX": 16,
        "lbSizeY": 16,
        "lbSizeZ": 16,
        "initVisc": 0.1,
        "initVelocityX": 0,
        "initVelocityY": 0,
        "initVelocityZ": 0,
        "lbFX": 0,
        "lbFY": 0,
        "lbFZ": 0,
        "boundary0": 4,
        "boundary1": 4,
        "boundary2": 4,
        "boundary3": 4,
        "boundary4": 7,
        "boundary5": 7,
        "slipCoefficient": 0,
        "turbConst": 0,
        "plasticVisc": 0.1,
        "yieldStress": 0,
        "density": 2.5,
        "contactModel": "HERTZIAN",
        "youngMod": 1.0e5,
        "poisson": 0.3,
        "linearStiff": 0,
        "restitution": 0.88,
        "viscTang": 0.2,
        "frictionCoefPart": 0.5,
        "frictionCoefWall": 0.5,
        "particleFile": "particles.dat",
        "objectFile": "objects.dat",
        "translateX": 0,
        "translateY": 0,
        "translateZ": 0,
        "scale": 1,
        "numVisc": 0,
        "multiStep": 10,
        "criticalRatio": 0.1,
    }
    cfg.update(over)
    return cfg


def _fmt(v) -> str:
    if isinstance(v, bool):
        return "1" if v else "0"
    if isinstance(v, float):
        return repr(float(v))
    return str(v)


def write_cfg(path: str, cfg: dict) -> None:
    with open(path, "w") as fh:
        for k, v in cfg.items():
            fh.write(f"{k} = {_fmt(v)}\n")
