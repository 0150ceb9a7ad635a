"""The multifrontal analysis of a grid, on the host (v2d_mf_analyse and
v2d_mf_export of the C ABI): the nested-dissection fronts the device
factorisation sweeps, for inspection.  No device call.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, ptr

FRONT_FIELDS = ("fo", "vo", "zo", "uo", "mo", "ns", "nsp", "nb", "nf", "ld", "nch", "co", "big",
                "height")


def analyse(imax: int, jmax: int, ff_flags=None) -> dict:
    """Fronts of the grid's nested dissection.  ff_flags[i]: far-field outflow
    bits of node (i, jmax-1) (1: the omega row couples j-1, j-2; 2: the psi
    row does), None for none.  Returns the sizes, the front table (one row of
    FRONT_FIELDS per front, fronts ordered by height) and the maps."""
    h = C.c_void_p()
    fl = None if ff_flags is None else np.ascontiguousarray(ff_flags, dtype=np.uint8)
    check(lib.v2d_mf_analyse(imax, jmax, ptr(fl), C.byref(h)))
    try:
        sz = np.zeros(16, np.int64)
        flops = C.c_double()
        check(lib.v2d_mf_sizes(h, ptr(sz), C.byref(flops)))
        nfront, _, _, _, _, _, nfu, nmap, nch, npos, nvals = (int(v) for v in sz[:11])
        fronts = np.zeros((nfront, len(FRONT_FIELDS)), np.int64)
        fu = np.zeros(nfu, np.int32)
        ch = np.zeros(max(nch, 1), np.int32)
        mp = np.zeros(max(nmap, 1), np.int32)
        aoff = np.zeros(npos, np.int64)
        aseg = np.zeros(npos + 1, np.int32)
        avid = np.zeros(nvals, np.int32)
        check(lib.v2d_mf_export(h, ptr(fronts), ptr(fu), ptr(ch), ptr(mp), ptr(aoff), ptr(aseg),
                                ptr(avid)))
    finally:
        lib.v2d_mf_destroy(h)
    keys = ("fronts", "heights", "big_fronts", "front_elems", "vector_elems", "pivot_elems")
    out = dict(zip(keys, (int(v) for v in sz[:6])))
    out.update(flops=flops.value, table=fronts, fu=fu, ch=ch[:nch], map=mp[:nmap], aoff=aoff,
               aseg=aseg, avid=avid)
    return out
