"""Force / moment coefficients (mirror of vorti2d/forces.py).

The wall-traction integral of compute_force_coeffs (vorti2d/forces.py:82-194)
runs as one FP64 workgroup reduction on the GPU (k_forces in libvorti2d).
"""
from __future__ import annotations

from dataclasses import asdict, dataclass

import numpy as np

from ._lib import check, f64c, lib, ptr


@dataclass
class ForceCoeffs:
    """Force / moment coefficients and their pressure / friction split
    (vorti2d/forces.py:61-75)."""
    cd: float
    cl: float
    cm: float
    cdp: float
    cdf: float
    clp: float
    clf: float
    cmp: float
    cmf: float
    ref_length: float

    def to_dict(self) -> dict:
        return asdict(self)

    @classmethod
    def from_array(cls, out) -> "ForceCoeffs":
        return cls(*[float(v) for v in out[:10]])


FORCE_COLUMNS = ("t", "cd", "cl", "cm", "cdp", "cdf", "clp", "clf", "cmp", "cmf")


def compute_force_coeffs(imax: int, jmax: int, re: float, dksi: float, deta: float,
                         beta, gama, detadx, detady, xphys, yphys, ome,
                         ref_length: float | None = None,
                         moment_ref: tuple[float, float] = (0.0, 0.0)) -> ForceCoeffs:
    """Lift / drag / moment of a flattened vorticity field (forces.py:82-194).

    Pressure force by parts, Fxp = (1/Re) sum y dw/dn ds, Fyp = -(1/Re) sum x
    dw/dn ds; friction -(1/Re) w t; dw/dn = (gama w_xi + beta w_eta)/sqrt(beta)
    with w_xi periodic-central and w_eta one-sided; periodic trapezoid arc
    weights over the imax-1 distinct wall nodes; ref_length None -> twice the
    arc-weighted mean wall radius."""
    arrs = [f64c(a) for a in (beta, gama, detadx, detady, xphys, yphys, ome)]
    out = np.zeros(10)
    d = -1.0 if ref_length is None else float(ref_length)
    check(lib.v2d_force_coeffs(int(imax), int(jmax), float(re), float(dksi), float(deta),
                               *[ptr(a) for a in arrs], d, float(moment_ref[0]),
                               float(moment_ref[1]), ptr(out)))
    return ForceCoeffs.from_array(out)
