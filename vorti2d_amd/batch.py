"""Batched sweeps of independent steady solves on one GPU.

Every case owns a device context and therefore its own HIP stream; a host
thread pool drives up to `workers` contexts at once (the ctypes calls release
the GIL), so the small solves of a sweep overlap on the GPU instead of each
leaving most of it idle.  A case runs exactly the kernels of a single solve,
and its results equal the single-solve results bit for bit in both precisions:
the mixed solver's refactorisation decision is, by default, a function of the
iterates alone (DeviceContext.set_refactor_policy("fixed")), so concurrency
does not change a trajectory.  A case that misses its stopping rule within
max_iter Newton steps raises.

The reference has no batched driver; a case is the steady loop of
vorti2d/solver.py:136-145 (stopping rule res = ||delta||_2 / (2N) <= tol) on an
O-grid around an elliptic body, with the force coefficients of
vorti2d/forces.py:82-194 and, optionally, the adjoint dCd/dalpha.
"""
from __future__ import annotations

import math
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from .device import DeviceContext


def elliptic_grid(imax, jmax, aspect, inner=0.5, outer=50.0):
    """O-grid around the ellipse with semi-axes (inner, aspect * inner) out to
    the circle of radius `outer`: x = a_j cos(theta), y = b_j sin(theta) with
    a_j = r_j geometric from inner to outer (as generate_annulus) and
    b_j = r_j (aspect + (1 - aspect) s_j), s_j = log(r_j / inner) / log(outer / inner);
    theta_i = 2 pi (1 - i / (imax - 1)) (clockwise, seam stored twice)."""
    if not 0.0 < aspect <= 1.0:
        raise ValueError("aspect must be in (0, 1]")
    q = (outer / inner) ** (1.0 / (jmax - 1))
    r = inner * q ** np.arange(jmax)
    r[-1] = outer
    s = np.log(r / inner) / math.log(outer / inner)
    a = r
    b = r * (aspect + (1.0 - aspect) * s)
    theta = 2.0 * np.pi * (1.0 - np.arange(imax) / (imax - 1))
    return np.outer(np.cos(theta), a), np.outer(np.sin(theta), b)


@dataclass
class Case:
    imax: int
    jmax: int
    aspect: float
    re: float
    alpha_deg: float


def sweep_cases(n=64, seed=0, imax_range=(129, 257), jmax_range=(129, 385), aspect_range=(0.5, 1.0),
                re_range=(20.0, 45.0), alpha_range=(-10.0, 10.0)):
    """Seeded sweep: angular size, radial size, aspect ratio, Re and angle of
    attack drawn uniformly (sizes inclusive), numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        out.append(Case(int(rng.integers(imax_range[0], imax_range[1] + 1)),
                        int(rng.integers(jmax_range[0], jmax_range[1] + 1)),
                        float(rng.uniform(*aspect_range)), float(rng.uniform(*re_range)),
                        float(rng.uniform(*alpha_range))))
    return out


@dataclass
class CaseResult:
    case: Case
    cd: float
    cl: float
    newton_iters: int
    res: float
    setup_s: float
    solve_s: float
    adjoint_s: float = 0.0
    dcd_dalpha: float = float("nan")
    history: list = field(default_factory=list)
    rnorm: float = float("nan")      # FP64 ||R||_2 at the converged state
    cm: float = float("nan")
    psi: np.ndarray | None = None
    ome: np.ndarray | None = None


def solve_case(case: Case, precision="mixed", tol=1e-10, max_iter=50, forcing=0.1, ref_length=1.0,
               adjoint=False, rms_tol=1e-10, keep_state=False, graph="apply"):
    """One steady solve (and, with adjoint=True, dCd/dalpha per degree from an
    FP64 context at the converged state).  Stopping rules: res = ||delta||_2 /
    (2N) <= tol (vorti2d/solver.py:141-144) and, unless rms_tol is None, the
    final FP64 residual ||R||_2 / sqrt(2N) < rms_tol; the mixed GMRES forcing
    aims at that residual (eta_k >= 0.5 rms_tol sqrt(2N) / ||R_k||).
    keep_state: the result also holds the converged psi, omega.
    graph: the mixed GMRES launch mode (DeviceContext.set_graph)."""
    xg, yg = elliptic_grid(case.imax, case.jmax, case.aspect)
    a = math.radians(case.alpha_deg)
    t0 = time.perf_counter()
    ctx = DeviceContext(xg, yg, 0, math.cos(a), math.sin(a), precision)
    if precision == "mixed":
        ctx.set_solver(refactor_every=0, gmres_rtol=1e-6, gmres_max_iter=30)
        n2 = 2 * case.imax * case.jmax
        ctx.set_forcing(forcing, (rms_tol if rms_tol is not None else tol * math.sqrt(n2))
                        * math.sqrt(n2))
        ctx.set_graph(graph)
    ctx.init_state()
    t1 = time.perf_counter()
    r = ctx.steady_newton(case.re, tol=tol, rms_tol=rms_tol, max_iter=max_iter, record=True)
    t2 = time.perf_counter()
    if not r.converged:
        ctx.close()
        raise RuntimeError(f"case {case}: no convergence in {r.iters} Newton steps "
                           f"(res {r.res:.3e}, ||R||_2 {r.rnorm:.3e})")
    f = ctx.forces(case.re, ref_length)
    out = CaseResult(case, float(f[0]), float(f[1]), r.iters, r.res, t1 - t0, t2 - t1,
                     history=r.res_hist, rnorm=r.rnorm, cm=float(f[2]))
    if keep_state:
        out.psi, out.ome = ctx.get_state(("psi", "ome"))
    if adjoint:
        t3 = time.perf_counter()
        if precision == "fp64":
            adj = ctx.adjoint_forces(case.re, ref_length, 1.0, 0.0)
        else:
            c64 = ctx.like("fp64")   # mesh, state and pattern copied on the device
            adj = c64.adjoint_forces(case.re, ref_length, 1.0, 0.0)
            c64.close()
        out.adjoint_s = time.perf_counter() - t3
        out.dcd_dalpha = adj["dalpha"]
    ctx.close()
    return out


def solve_batch(cases, workers=8, **kw):
    """Solve every case; up to `workers` contexts (streams) in flight.  Returns
    (results in case order, stats dict: wall seconds, solves per second, p50 and
    p99 of the per-case setup + solve (+ adjoint) time)."""
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=max(1, int(workers))) as ex:
        results = list(ex.map(lambda c: solve_case(c, **kw), cases))
    wall = time.perf_counter() - t0
    per = np.array([r.setup_s + r.solve_s + r.adjoint_s for r in results])
    stats = dict(wall_s=wall, solves_per_s=len(cases) / wall, p50_s=float(np.percentile(per, 50)),
                 p99_s=float(np.percentile(per, 99)), workers=workers, cases=len(cases))
    return results, stats


# ------------------------------------------------------------ lockstep --
class DeviceBatch:
    """Handle on a lockstep batch of device contexts (v2d_batch_* in
    include/vorti2d.h): up to 16 slots of one precision stepped together with
    shared assembly, SpMV, multi-dot / multi-axpy and update launches, the
    factorisations and preconditioner applies on the cases' own streams."""

    MAX = 16

    def __init__(self, precision="mixed", slots=8):
        import ctypes as C
        from ._lib import check, lib
        from .device import PRECISIONS
        self._C, self._lib, self._check = C, lib, check
        self.slots = int(slots)
        h = C.c_void_p()
        check(lib.v2d_batch_create(C.byref(h), PRECISIONS[precision], self.slots))
        self._h = h
        self._ctx = [None] * self.slots

    def set_case(self, slot, ctx):
        self._check(self._lib.v2d_batch_set_case(self._h, int(slot), None if ctx is None else ctx._h))
        self._ctx[slot] = ctx

    def newton_step(self, re, active=None):
        """One Newton step of the active filled slots; re: per-slot Reynolds
        numbers.  Returns per-slot arrays (res, ||R||_2 before the step, GMRES
        iterations)."""
        from ._lib import ptr
        n = self.slots
        re = np.ascontiguousarray(re, dtype=np.float64)
        act = None if active is None else np.ascontiguousarray(active, dtype=np.int32)
        res = np.full(n, np.nan)
        rn = np.full(n, np.nan)
        lin = np.zeros(n, np.int32)
        self._check(self._lib.v2d_batch_newton_step(self._h, ptr(act), ptr(re), None, None, None,
                                                    ptr(res), ptr(rn), ptr(lin)))
        return res, rn, lin

    def blocks(self):
        from ._lib import ptr
        out = np.zeros(3, np.int32)
        self._check(self._lib.v2d_batch_get_blocks(self._h, ptr(out)))
        return dict(zip(("assembly", "spmv", "vector"), (int(x) for x in out)))

    def set_blocks(self, assembly=256, spmv=256, vector=256):
        from ._lib import ptr
        b = np.array([assembly, spmv, vector], np.int32)
        self._check(self._lib.v2d_batch_set_blocks(self._h, ptr(b)))

    def autotune(self, re):
        """Time 128 .. 1024-thread workgroups for the batched pointwise launches
        on the filled slots and keep the fastest (results do not depend on
        them)."""
        from ._lib import ptr
        out = np.zeros(3, np.int32)
        self._check(self._lib.v2d_batch_autotune(self._h, float(re), ptr(out)))
        return dict(zip(("assembly", "spmv", "vector"), (int(x) for x in out)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.v2d_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _case_context(case: Case, precision, tol, rms_tol, forcing):
    xg, yg = elliptic_grid(case.imax, case.jmax, case.aspect)
    a = math.radians(case.alpha_deg)
    ctx = DeviceContext(xg, yg, 0, math.cos(a), math.sin(a), precision)
    if precision == "mixed":
        n2 = 2 * case.imax * case.jmax
        ctx.set_solver(refactor_every=0, gmres_rtol=1e-6, gmres_max_iter=30)
        ctx.set_forcing(forcing, (rms_tol if rms_tol is not None else tol * math.sqrt(n2))
                        * math.sqrt(n2))
    ctx.init_state()
    return ctx


def solve_batch_lockstep(cases, inflight=8, precision="mixed", tol=1e-10, rms_tol=1e-10, max_iter=50,
                         forcing=0.1, ref_length=1.0, adjoint=False, autotune=True, keep_state=False):
    """Solve every case with up to `inflight` of them stepped in lockstep on
    one DeviceBatch; a finished case leaves its slot and the next case takes
    it.  Host threads build the next cases' contexts ahead of time and finish
    the solved ones (forces, optional adjoint dCd/dalpha from an FP64 copy of
    the context) while the batch steps on.  Stopping rules and knobs are
    solve_case's, and every case's result equals its solve_case result bit for
    bit.  Returns (results in case order, stats as solve_batch's plus the tuned
    workgroup sizes)."""
    t0 = time.perf_counter()
    n = len(cases)
    if n == 0:
        return [], dict(wall_s=0.0, solves_per_s=0.0, p50_s=0.0, p99_s=0.0, cases=0)
    slots = max(1, min(int(inflight), DeviceBatch.MAX, n))
    results = [None] * n

    def build(q):
        ts = time.perf_counter()
        return q, _case_context(cases[q], precision, tol, rms_tol, forcing), time.perf_counter() - ts

    def finish(q, ctx, it, res, rfin, hist, setup, solve):
        c = cases[q]
        try:
            f = ctx.forces(c.re, ref_length)
            r = CaseResult(c, float(f[0]), float(f[1]), it, res, setup, solve, history=hist,
                           rnorm=rfin, cm=float(f[2]))
            if keep_state:
                r.psi, r.ome = ctx.get_state(("psi", "ome"))
            if adjoint:
                ta = time.perf_counter()
                if precision == "fp64":
                    adj = ctx.adjoint_forces(c.re, ref_length, 1.0, 0.0)
                else:
                    c64 = ctx.like("fp64")
                    adj = c64.adjoint_forces(c.re, ref_length, 1.0, 0.0)
                    c64.close()
                r.adjoint_s = time.perf_counter() - ta
                r.dcd_dalpha = adj["dalpha"]
        finally:
            ctx.close()
        results[q] = r

    batch = DeviceBatch(precision, slots)
    st = [None] * slots           # per slot: dict(q, ctx, setup, t_in, it, hist)
    pending = []                  # futures of contexts being built, in case order
    done_f = []
    nxt = 0
    tuned = None
    with ThreadPoolExecutor(max_workers=slots) as ex:
        try:
            while nxt < n and len(pending) < slots:
                pending.append(ex.submit(build, nxt))
                nxt += 1
            while True:
                for s in range(slots):
                    if st[s] is None and pending:
                        q, ctx, setup = pending.pop(0).result()
                        batch.set_case(s, ctx)
                        st[s] = dict(q=q, ctx=ctx, setup=setup, t_in=time.perf_counter(), it=0, hist=[])
                        if nxt < n:
                            pending.append(ex.submit(build, nxt))
                            nxt += 1
                if all(x is None for x in st):
                    break
                if autotune and tuned is None:
                    tuned = batch.autotune(cases[next(x for x in st if x)["q"]].re)
                re = np.array([cases[x["q"]].re if x else 0.0 for x in st])
                res, _, _ = batch.newton_step(re, [1 if x else 0 for x in st])
                for s in range(slots):
                    x = st[s]
                    if x is None:
                        continue
                    c = cases[x["q"]]
                    x["it"] += 1
                    x["hist"].append(float(res[s]))
                    if not np.isfinite(res[s]):
                        raise RuntimeError(f"case {c}: non-finite Newton update")
                    done, rfin = False, float("nan")
                    if res[s] <= tol:
                        x["ctx"].assemble(c.re)
                        rfin = x["ctx"].rhs_norm()
                        done = rms_tol is None or rfin / math.sqrt(2 * c.imax * c.jmax) < rms_tol
                    if not done and x["it"] >= max_iter:
                        raise RuntimeError(f"case {c}: no convergence in {x['it']} Newton steps "
                                           f"(res {res[s]:.3e})")
                    if not done:
                        continue
                    batch.set_case(s, None)
                    st[s] = None
                    done_f.append(ex.submit(finish, x["q"], x["ctx"], x["it"], float(res[s]), rfin,
                                            x["hist"], x["setup"], time.perf_counter() - x["t_in"]))
            for f in done_f:
                f.result()
        finally:
            for x in st:
                if x is not None:
                    x["ctx"].close()
            for f in pending:
                try:
                    f.result()[1].close()
                except Exception:
                    pass
            batch.close()
    wall = time.perf_counter() - t0
    per = np.array([r.setup_s + r.solve_s + r.adjoint_s for r in results])
    stats = dict(wall_s=wall, solves_per_s=n / wall, p50_s=float(np.percentile(per, 50)),
                 p99_s=float(np.percentile(per, 99)), workers=slots, cases=n, blocks=tuned)
    return results, stats
