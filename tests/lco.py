"""Shedding frequency of a saturated Cl(t) limit cycle, as measured by the
reference's post-processing (examples/lco_frequency.py:19-79): find the trailing
window where successive half-swing amplitudes stay within `tol` of the late
amplitude, then average the frequencies from peak, trough and rising
zero-crossing times.  Test helper only."""
from __future__ import annotations

import numpy as np
from scipy.signal import find_peaks


def saturated_start(y, tol=0.12, min_swings=4):
    """Index where the saturated cycle begins (examples/lco_frequency.py:19-40)."""
    yc = y - y.mean()
    ext = np.sort(np.concatenate([find_peaks(yc)[0], find_peaks(-yc)[0]]))
    if len(ext) < min_swings:
        return 0
    swings = np.abs(np.diff(yc[ext]))
    ref = np.median(swings[-max(min_swings, len(swings) // 4):])
    j = len(swings) - 1
    while j >= 0 and abs(swings[j] - ref) <= tol * ref:
        j -= 1
    return int(ext[j + 1])


def _rate(times):
    return (len(times) - 1) / (times[-1] - times[0]) if len(times) >= 2 else np.nan


def strouhal(t, y, tol=0.12):
    """Mean of the peak / trough / rising-zero-crossing frequencies over the
    saturated window (examples/lco_frequency.py:49-79); with D = U = 1 this is
    the Strouhal number."""
    i0 = saturated_start(y, tol)
    ts, ys = t[i0:], y[i0:] - np.mean(y[i0:])
    if ys.size < 4 or ys.max() <= 0:
        return dict(st=np.nan, i0=i0, amp=0.0)
    h = 0.3 * ys.max()
    pk = find_peaks(ys, height=h)[0]
    tr = find_peaks(-ys, height=h)[0]
    rise = np.where(np.diff(np.signbit(ys).astype(int)) == -1)[0]
    f = [x for x in (_rate(ts[pk]), _rate(ts[tr]), _rate(ts[rise])) if not np.isnan(x)]
    return dict(st=float(np.mean(f)) if f else np.nan, i0=i0, n_periods=max(len(pk) - 1, 0),
                amp=float(0.5 * (ys.max() - ys.min())))
