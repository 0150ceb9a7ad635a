"""NACA 4-digit airfoil O-grids with Hicks-Henne surface bumps.

The reference has no airfoil generator; this module builds the grids of the
adjoint workload: an analytic NACA 4-digit section (blunt trailing edge, the
classical thickness polynomial with coefficient -0.1015) on an algebraically
generated O-grid, layout as vorti2d/mesh.py:14-50 produces for the cylinder:
arrays of shape (imax, jmax), i around the body with the seam stored twice
(i = 0 and i = imax-1 at the trailing-edge base midpoint), j = 0 on the wall,
j = jmax-1 on a far-field circle.  Nodes run clockwise (trailing edge ->
lower surface -> leading edge -> upper surface -> trailing edge), which gives
a positive metric Jacobian with j outward.  The chord runs from x = -0.5 to
0.5: the wall-vorticity row normalises the wall node position by its radius
(vorti2d_core.f90:343-349), so no wall node may sit at the origin.

Each grid line i is the straight segment from the surface point S_i to the
far-field point F_i = R (cos phi_i, sin phi_i), phi_i = -2 pi i / (imax-1),
about the mid-chord, with a geometric radial distribution of first spacing
`first_spacing` (in units of |F - S|).  Hicks-Henne bumps
    b_k(x) = sin(pi x^(ln 0.5 / ln h_k))^width,   h_k in (0, 1)
displace the upper (k < K) or lower (k >= K) surface in y by amplitude a_k; the
displacement decays linearly to zero at the far field, so the grid is affine in
the amplitudes and dY/da_k is a fixed field (`AirfoilGrid.basis`).
"""
from __future__ import annotations

import math

import numpy as np


def naca4(x, m=0.02, p=0.4, t=0.12):
    """Camber line y_c, its slope and half thickness y_t at chord stations x."""
    x = np.asarray(x, dtype=np.float64)
    yt = 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x**2 + 0.2843 * x**3 - 0.1015 * x**4)
    if m == 0.0:
        return np.zeros_like(x), np.zeros_like(x), yt
    fwd = x < p
    yc = np.where(fwd, m / p**2 * (2 * p * x - x**2), m / (1 - p)**2 * ((1 - 2 * p) + 2 * p * x - x**2))
    dyc = np.where(fwd, 2 * m / p**2 * (p - x), 2 * m / (1 - p)**2 * (p - x))
    return yc, dyc, yt


def surface(imax, m=0.02, p=0.4, t=0.12):
    """Clockwise closed surface of imax points: i = 0 and i = imax-1 are the
    midpoint of the blunt trailing-edge base, i = 1..nh the lower surface from
    the trailing edge to the leading edge (i = nh), i = nh+1..imax-2 the upper
    surface back to the trailing edge, with nh = (imax-1)/2 and cosine-clustered
    chord stations x_q = (1 + cos(pi q / nh)) / 2.  Returns xs, ys, the chord
    station of every point and its side (+1 upper, -1 lower, 0 base / nose)."""
    if (imax - 1) % 2 or imax < 9:
        raise ValueError("imax must be odd and >= 9")
    nh = (imax - 1) // 2
    q = np.arange(nh + 1)
    xc = 0.5 * (1.0 + np.cos(np.pi * q / nh))
    yc, dyc, yt = naca4(xc, m, p, t)
    th = np.arctan(dyc)
    xl, yl = xc + yt * np.sin(th), yc - yt * np.cos(th)     # lower surface
    xu, yu = xc - yt * np.sin(th), yc + yt * np.cos(th)     # upper surface
    xs = np.empty(imax)
    ys = np.empty(imax)
    xcs = np.empty(imax)
    side = np.zeros(imax, dtype=int)
    xs[0] = xs[-1] = 0.5 * (xl[0] + xu[0])
    ys[0] = ys[-1] = 0.5 * (yl[0] + yu[0])
    xcs[0] = xcs[-1] = 1.0
    lo = np.arange(1, nh + 1)                 # lower: q = i
    xs[lo], ys[lo], xcs[lo], side[lo] = xl[lo], yl[lo], xc[lo], -1
    up = np.arange(nh + 1, imax - 1)          # upper: q = 2 nh - i
    qu = 2 * nh - up
    xs[up], ys[up], xcs[up], side[up] = xu[qu], yu[qu], xc[qu], 1
    side[nh] = 0                              # the leading edge
    return xs, ys, xcs, side


def radial_distribution(jmax, first_spacing=1.5e-4):
    """r_j in [0, 1], geometric with r_1 = first_spacing."""
    n = jmax - 1
    if first_spacing * n >= 1.0:
        return np.linspace(0.0, 1.0, jmax)
    lo, hi = 1.0 + 1e-12, 2.0
    for _ in range(200):                 # (q^n - 1)/(q - 1) * first_spacing = 1
        q = 0.5 * (lo + hi)
        if first_spacing * (q**n - 1.0) / (q - 1.0) > 1.0:
            hi = q
        else:
            lo = q
    q = 0.5 * (lo + hi)
    r = first_spacing * (q ** np.arange(jmax) - 1.0) / (q - 1.0)
    r[-1] = 1.0
    return r


def hicks_henne(x, h, width=3.0):
    """b(x) = sin(pi x^(ln 0.5 / ln h))^width, zero at x = 0 and x = 1."""
    x = np.clip(np.asarray(x, dtype=np.float64), 0.0, 1.0)
    e = math.log(0.5) / math.log(h)
    return np.sin(np.pi * x**e) ** width


def bump_peaks(n_per_side=16):
    """Peak stations h_k of the bumps on one side, evenly in (0.05, 0.95)."""
    return np.linspace(0.05, 0.95, n_per_side)


class AirfoilGrid:
    """NACA 4-digit O-grid, affine in the Hicks-Henne amplitudes."""

    def __init__(self, imax=257, jmax=129, m=0.02, p=0.4, t=0.12, radius=30.0,
                 first_spacing=1.5e-4, n_bumps=32, width=3.0):
        if n_bumps % 2:
            raise ValueError("n_bumps must be even (half upper, half lower)")
        self.imax, self.jmax = imax, jmax
        xs, ys, xcs, side = surface(imax, m, p, t)
        xs = xs - 0.5                      # mid-chord at the origin
        phi = -2.0 * np.pi * np.arange(imax) / (imax - 1)
        xf = radius * np.cos(phi)
        yf = radius * np.sin(phi)
        r = radial_distribution(jmax, first_spacing)
        self.r = r
        self.x0 = xs[:, None] + (xf - xs)[:, None] * r[None, :]
        self.y0 = ys[:, None] + (yf - ys)[:, None] * r[None, :]
        K = n_bumps // 2
        peaks = bump_peaks(K)
        fade = (1.0 - r)[None, :]
        basis = np.zeros((n_bumps, imax, jmax))
        for k in range(n_bumps):
            on = side == (1 if k < K else -1)
            b = np.where(on, hicks_henne(xcs, peaks[k % K], width), 0.0)
            basis[k] = b[:, None] * fade
        self.basis = basis                 # dY/da_k; dX/da_k = 0
        self.n_bumps = n_bumps

    def grid(self, amplitudes=None):
        """(xg, yg) for the given bump amplitudes (None: the clean section)."""
        y = self.y0.copy()
        if amplitudes is not None:
            a = np.asarray(amplitudes, dtype=np.float64)
            if a.shape != (self.n_bumps,):
                raise ValueError(f"expected {self.n_bumps} amplitudes")
            y += np.tensordot(a, self.basis, axes=1)
        return self.x0.copy(), y

    def bump_gradient(self, gx, gy):
        """dJ/da_k = sum over the grid of dJ/dy * dy/da_k (the grid x does not
        depend on the amplitudes)."""
        return np.tensordot(self.basis, np.asarray(gy), axes=([1, 2], [0, 1]))
