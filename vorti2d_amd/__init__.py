"""vorti2d_amd: MI355X (gfx950) implementation of vorti2d's metric, assembly and
Newton/Krylov hot path, behind the reference's API (vorti2d/__init__.py:13-35).

``core`` is the drop-in for the reference's compiled f2py module
(``vorti2d.core``); ``Config`` / ``Solver`` / ``run`` mirror the reference
driver.  All compute runs in libvorti2d.so (HIP); importing fails if the library
has not been built.
"""
from __future__ import annotations

from . import _core  # noqa: F401  (ctypes binding of libvorti2d.so)

core = _core.vorti2d_core

from . import mesh  # noqa: E402
from .config import Config  # noqa: E402
from .device import DeviceContext  # noqa: E402
from .forces import ForceCoeffs, compute_force_coeffs  # noqa: E402
from .mesh import generate_annulus, generate_cylinder, load_cgns_ogrid, load_mesh, save_mesh  # noqa: E402
from .solver import Solver, run  # noqa: E402
from .airfoil import AirfoilGrid  # noqa: E402
from .adjoint import SteadyForces, force_gradients, metrics_vjp, steady_solve  # noqa: E402
from .adjoint_unsteady import BDF2Run, unsteady_forces, unsteady_forces_gradient  # noqa: E402

__all__ = [
    "Config", "Solver", "run", "generate_cylinder", "generate_annulus", "load_mesh",
    "load_cgns_ogrid", "save_mesh", "compute_force_coeffs", "ForceCoeffs", "mesh", "core",
    "DeviceContext",
]

__version__ = "0.1.0"
