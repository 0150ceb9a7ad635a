"""Mesh generation and CSV interchange (mirror of vorti2d/mesh.py).

Meshes are (imax, jmax) arrays: i circumferential with the seam stored twice
(i = 0 and i = imax-1 are one line), j radial (j = 0 wall, j = jmax-1 far
field).
"""
from __future__ import annotations

import numpy as np


def generate_cylinder(imax: int = 181, jmax: int = 181, inner_rad: float = 0.5,
                      outer_rad: float = 50.0):
    """Cylinder O-grid with the reference clustering (vorti2d/mesh.py:14-50).

    Radial arc length: 0, 1, 2, 3 then increments growing by one per cell
    (j - 2 for 0-based j >= 4), normalised.  Circumferential arc length: running
    sum of min(i, imax - i)^0.6, normalised; theta_i = 2 pi (1 - s_i)."""
    a = np.zeros(jmax)
    if jmax > 1:
        a[1:min(jmax, 4)] = np.arange(1, min(jmax, 4), dtype=float)
    for j in range(4, jmax):
        a[j] = a[j - 1] + (j - 2)
    a /= a[jmax - 1]
    s = np.zeros(imax)
    for i in range(1, imax):
        s[i] = s[i - 1] + min(i, imax - i) ** 0.6
    s /= s[imax - 1]
    theta = 2.0 * np.pi * (1.0 - s)
    rad = inner_rad + (outer_rad - inner_rad) * a
    return np.outer(np.cos(theta), rad), np.outer(np.sin(theta), rad)


def generate_annulus(imax: int, jmax: int, inner_rad: float = 0.5, outer_rad: float = 50.0):
    """O-grid with geometric radial stretching and uniform theta.

    r_j = r_in q^j with q = (r_out / r_in)^(1/(jmax-1)) (so every cell is q times
    the previous one), theta_i = 2 pi (1 - i/(imax-1)), same orientation as
    generate_cylinder.  The metrics have closed forms (a log-polar map)."""
    q = (outer_rad / inner_rad) ** (1.0 / (jmax - 1))
    rad = inner_rad * q ** np.arange(jmax)
    rad[-1] = outer_rad
    theta = 2.0 * np.pi * (1.0 - np.arange(imax) / (imax - 1))
    return np.outer(np.cos(theta), rad), np.outer(np.sin(theta), rad)


def stretch_ratio(jmax: int, inner_rad: float = 0.5, outer_rad: float = 50.0) -> float:
    """Cell-to-cell radial growth factor of generate_annulus."""
    return (outer_rad / inner_rad) ** (1.0 / (jmax - 1))


def save_mesh(xg: np.ndarray, yg: np.ndarray, xg_path: str, yg_path: str):
    """Write a mesh as two CSV files (vorti2d/mesh.py:53-57)."""
    np.savetxt(xg_path, xg, delimiter=",")
    np.savetxt(yg_path, yg, delimiter=",")


def load_mesh(xg_path: str, yg_path: str):
    """Read a mesh from two CSV files (vorti2d/mesh.py:60-68)."""
    xg = np.loadtxt(xg_path, delimiter=",")
    yg = np.loadtxt(yg_path, delimiter=",")
    if xg.shape != yg.shape:
        raise ValueError(f"xg shape {xg.shape} != yg shape {yg.shape}")
    if xg.ndim != 2:
        raise ValueError("mesh CSVs must be 2-D (imax x jmax)")
    return (np.ascontiguousarray(xg, dtype=np.float64),
            np.ascontiguousarray(yg, dtype=np.float64))


def load_cgns_ogrid(cgns_path: str, plane: int = 0, block: int = 0, verbose: bool = False):
    """CGNS import (vorti2d/mesh.py:93-140) needs cgnsutilities, which this
    build does not ship; convert the grid to the CSV pair instead."""
    raise NotImplementedError(
        "CGNS O-grid import requires cgnsutilities; export the grid to xg/yg CSV "
        "and use load_mesh")
