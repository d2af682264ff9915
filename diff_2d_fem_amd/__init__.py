"""diff_2d_fem_amd — the implicit PNCG-IPC step of Diff_2D_FEM on AMD MI355X (HIP, gfx950).

The public names match the reference package (reference __init__.py:70-83):

    from diff_2d_fem_amd import Model, ImplicitIntegrator
"""
from .config import SimConfig, default_precision
from .model import Model
from .integrator import ImplicitIntegrator
from .mesh import build_scene, check_mesh, mask_to_mesh

PRECISION = default_precision()

__all__ = ["Model", "ImplicitIntegrator", "SimConfig", "build_scene", "check_mesh", "mask_to_mesh",
           "PRECISION"]
