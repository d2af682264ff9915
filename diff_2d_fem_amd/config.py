"""Scene parameters the solver reads (reference core/config.py:15-29). Contact parameters
left at None are scaled to the scene at Model.finalize()."""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Optional


@dataclass
class SimConfig:
    young: float = 1.0e6                 # Young's modulus of a body added without one [Pa]
    nu: float = 0.40                     # Poisson ratio
    rho_mass: float = 1000.0             # density [kg/m^3]
    damping: float = 0.0                 # velocity drag [1/s]: v scaled by exp(-dt*damping)
    ipc_d_hat: Optional[float] = None            # barrier reach [m]; None: 1/8 pixel
    obstacle_stiffness: Optional[float] = None   # collider barrier kappa [Pa]; None: 0.004 E px/dhat
    contact_stiffness: Optional[float] = None    # "node" contact penalty [Pa]; None: 0.2 E
    obstacle_friction: float = 0.5               # Coulomb mu of a collider added without one
    obstacle_friction_epsv: float = 1.0e-3       # friction smoothing speed [m/s]


def default_precision():
    """Process default arithmetic: "double" (FP64 everywhere) or "mixed" (FP32 element
    arithmetic, FP64 state, dot products and energies). Read from DIFF2DFEM_PRECISION; the
    reference's "single" is accepted as a name for "mixed"."""
    p = os.environ.get("DIFF2DFEM_PRECISION", "double").strip().lower()
    if p in ("double", "float64", "64", "d"):
        return "double"
    if p in ("mixed", "single", "float32", "32", "f"):
        return "mixed"
    raise ValueError(f"DIFF2DFEM_PRECISION must be 'double' or 'mixed', not {p!r}")
