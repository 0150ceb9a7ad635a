"""Python handle on a libvorti2d device context (v2d_ctx_* in include/vorti2d.h).

The context owns the mesh, metrics, state, Jacobian slot planes and the block-LU
factorisation in HBM; methods move only scalars unless they are explicit
uploads / downloads.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, f64c, lib, ptr

PRECISIONS = {"fp64": 0, "mixed": 1}
REFACTOR_POLICIES = {"fixed": 0, "measured": 1}
LINEAR_SOLVERS = {"block_lu": 0, "multifrontal": 1}
UPDATE_KERNELS = {"valu64": 0, "valu128": 1, "mfma64": 2}
SWEEP_KERNELS = {"levels": 0, "persistent": 1, "fused": 2}


@dataclass
class NewtonSolve:
    """Outcome of DeviceContext.steady_newton."""
    converged: bool
    iters: int
    res: float          # ||delta||_2 / (2N) of the last step (vorti2d/solver.py:134)
    rnorm: float        # FP64 ||R||_2 at the final state
    res_hist: list = field(default_factory=list)
    rnorm_hist: list = field(default_factory=list)   # ||R||_2 before each step
    lin_iters: list = field(default_factory=list)


class DeviceContext:
    def __init__(self, xg, yg, ff_bc: int = 0, ca: float = 1.0, sa: float = 0.0,
                 precision: str | int = "fp64"):
        xg = np.asarray(xg, dtype=np.float64)
        yg = np.asarray(yg, dtype=np.float64)
        if xg.ndim != 2 or xg.shape != yg.shape:
            raise ValueError("xg, yg must be 2-D arrays of equal shape (imax, jmax)")
        self.imax, self.jmax = xg.shape
        self.N = self.imax * self.jmax
        self.precision = precision if isinstance(precision, int) else PRECISIONS[precision]
        h = C.c_void_p()
        # (imax, jmax) arrays in i-fastest memory order
        xf = np.ascontiguousarray(xg.T)
        yf = np.ascontiguousarray(yg.T)
        check(lib.v2d_ctx_create(C.byref(h), self.imax, self.jmax, ptr(xf), ptr(yf), int(ff_bc),
                                 float(ca), float(sa), self.precision))
        self._h = h

    def like(self, precision: str | int = "fp64") -> "DeviceContext":
        """A new context in `precision` on this mesh and far-field setting,
        with this context's state (after its queued work); the mesh, state and
        Jacobian pattern are copied on the device, not rebuilt."""
        prec = precision if isinstance(precision, int) else PRECISIONS[precision]
        h = C.c_void_p()
        check(lib.v2d_ctx_create_like(C.byref(h), self._h, prec))
        other = DeviceContext.__new__(DeviceContext)
        other.imax, other.jmax, other.N = self.imax, self.jmax, self.N
        other.precision = prec
        other._h = h
        return other

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.v2d_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ state ----
    def device_bytes(self) -> int:
        b = C.c_int64()
        check(lib.v2d_ctx_device_bytes(self._h, C.byref(b)))
        return int(b.value)

    def metrics(self):
        out = np.empty((10, self.N))
        check(lib.v2d_ctx_metrics(self._h, ptr(out)))
        return tuple(np.ascontiguousarray(out[q]) for q in range(10))

    def set_state(self, psi=None, ome=None, omeold=None, omeoldold=None):
        arrs = [None if a is None else f64c(a) for a in (psi, ome, omeold, omeoldold)]
        for a in arrs:
            if a is not None and a.size != self.N:
                raise ValueError("state arrays must have length imax*jmax")
        check(lib.v2d_ctx_set_state(self._h, *[ptr(a) for a in arrs]))

    def get_state(self, which=("psi", "ome", "omeold", "omeoldold")):
        names = ("psi", "ome", "omeold", "omeoldold")
        outs = {n: (np.empty(self.N) if n in which else None) for n in names}
        check(lib.v2d_ctx_get_state(self._h, *[ptr(outs[n]) for n in names]))
        return tuple(outs[n] for n in which)

    def init_state(self):
        check(lib.v2d_ctx_init_state(self._h))

    def set_solver(self, refactor_every: int = 1, gmres_rtol: float = 1e-6,
                   gmres_max_iter: int = 20):
        check(lib.v2d_ctx_set_solver(self._h, int(refactor_every), float(gmres_rtol),
                                     int(gmres_max_iter)))

    def set_forcing(self, eta_max: float, r_target: float = 0.0):
        """Eisenstat-Walker inexact-Newton tolerance for the mixed GMRES
        (eta_max <= 0: fixed gmres_rtol); r_target > 0: an absolute ||R||_2
        the solve is content with (no step solves more accurately than it
        needs)."""
        check(lib.v2d_ctx_set_forcing(self._h, float(eta_max), float(r_target)))

    def set_update_kernel(self, kind: str | int):
        """Gauss-Jordan update kernel of the block LU: "valu64", "valu128" or
        "mfma64" (the default)."""
        check(lib.v2d_ctx_set_update_kernel(self._h, UPDATE_KERNELS.get(kind, kind)))

    def set_sweep_kernel(self, kind: str | int):
        """Block-LU solve sweeps: "levels" (two launches per level step),
        "persistent" (one launch per solve) or "fused" (one launch per FP32
        level step whose right-hand side couples one neighbouring level,
        launch pairs for the others and in FP64).  Default: "persistent" for
        FP32 solves on meshes whose persistent launch fills the GPU (2 imax >=
        16 x compute units, e.g. imax 1025 on 256 units), "levels" otherwise."""
        check(lib.v2d_ctx_set_sweep_kernel(self._h, SWEEP_KERNELS.get(kind, kind)))

    def set_sweep_spin_limit(self, polls: int):
        """Bound of each persistent-sweep wait, in polls (default 2**21).  An
        expired wait makes the solve be redone with launch pairs, which the
        context then keeps; 0 expires every wait at once (forces that path)."""
        check(lib.v2d_ctx_set_sweep_spin_limit(self._h, int(polls)))

    def sweep_kernel(self) -> tuple[str, int]:
        """(sweep kernel the solves use now, persistent solves redone with
        launch pairs so far)."""
        k, n = C.c_int(), C.c_int64()
        check(lib.v2d_ctx_get_sweep_kernel(self._h, C.byref(k), C.byref(n)))
        names = {v: k for k, v in SWEEP_KERNELS.items()}
        return names[int(k.value)], int(n.value)

    def set_refactor_policy(self, policy: str | int):
        """Mixed, refactor_every = 0: "fixed" (the default) gives up on a stored
        FP32 factorisation once GMRES projects more than gmres_max_iter
        iterations, a function of the iterates only; "measured" weighs the
        factorisation and iteration times measured on the context instead
        (load-dependent trajectories)."""
        check(lib.v2d_ctx_set_refactor_policy(self._h, REFACTOR_POLICIES.get(policy, policy)))

    def set_linear_solver(self, kind: str | int):
        """Linear solver of the Newton systems: "multifrontal" (nested-dissection
        multifrontal LU, the structure of the reference's MUMPS solve) or
        "block_lu" (block-tridiagonal LU over radial levels).  Drops a stored
        factorisation."""
        check(lib.v2d_ctx_set_linear_solver(self._h, LINEAR_SOLVERS.get(kind, kind)))

    def set_step_pairing(self, paired: bool):
        """Multifrontal factorisation: Gauss-Jordan steps in light / heavy pairs
        (True, the default; the bulk of every front is read and written once
        per two steps) or every step over the whole front (False).  Same
        factorisation up to rounding; applies from the next factorisation."""
        check(lib.v2d_ctx_set_step_pairing(self._h, 1 if paired else 0))

    def set_front_fusion(self, fused: bool):
        """Multifrontal factorisation: fronts of a single 64-pivot step
        assembled and swept by one workgroup each in one launch per tree
        height (True, the default) or through the level assembly, entry
        scatter and step launches (False).  Same factorisation up to rounding;
        applies from the next factorisation."""
        check(lib.v2d_ctx_set_front_fusion(self._h, 1 if fused else 0))

    def linear_solver(self) -> str:
        k = C.c_int()
        check(lib.v2d_ctx_get_linear_solver(self._h, C.byref(k)))
        return {v: n for n, v in LINEAR_SOLVERS.items()}[k.value]

    def mf_info(self) -> dict:
        """Sizes of the multifrontal analysis of this grid (see v2d_ctx_mf_info)."""
        out = np.zeros(8, np.int64)
        fl = C.c_double()
        check(lib.v2d_ctx_mf_info(self._h, ptr(out), C.byref(fl)))
        keys = ("fronts", "heights", "big_fronts", "front_elems", "vector_elems", "panel_elems",
                "entries", "positions")
        return dict(zip(keys, (int(v) for v in out)), flops=fl.value)

    MF_GROUPS = ("assemble", "scatter", "small", "gj_begin", "gj_full", "gj_light", "gj_heavy", "front1",
                 "gather", "fwd", "bwd")
    # kernel of each group, as rocprofv3 names it ({T}: float or double);
    # "small" and "assemble" are several instances of one template, "front1"
    # the instances of two (head, then row tiles)
    MF_KERNELS = {"assemble": "k_mf_assemble<{T}, *>", "scatter": "k_mf_scatter<{T}>",
                  "small": "k_mf_small_fused<{T}, *>",
                  "gj_begin": "k_gj_step_mfma_b<{T}, TableBatch<{T}>, 3>",
                  "gj_full": "k_gj_step_mfma_b<{T}, TableBatch<{T}>, 0>",
                  "gj_light": "k_gj_step_mfma_b<{T}, TableBatch<{T}>, 1>",
                  "gj_heavy": "k_gj_step_mfma_b<{T}, TableBatch<{T}>, 2>",
                  "front1": "k_mf_front1_*<{T}, *>",
                  "gather": "k_mf_fgather<{T}>", "fwd": "k_mf_fgemv<{T}, *>", "bwd": "k_mf_bwd<{T}, *>"}

    def profile_mf(self) -> dict:
        """Factorise the last assembled Jacobian and solve once with the
        multifrontal LU, every launch timed with HIP events on the context's
        stream: per kernel group ms, launches, algorithmic flops and bytes, and
        the factorisation / solve spans."""
        ng = len(self.MF_GROUPS)
        out = np.zeros(4 * ng + 2)
        check(lib.v2d_ctx_profile_mf(self._h, ptr(out)))
        t = "double" if self.precision == PRECISIONS["fp64"] else "float"
        g = {k: dict(ms=float(out[4 * i]), launches=int(out[4 * i + 1]), flops=float(out[4 * i + 2]),
                     bytes=float(out[4 * i + 3]), kernel=self.MF_KERNELS[k].format(T=t))
             for i, k in enumerate(self.MF_GROUPS)}
        return dict(groups=g, factor_ms=float(out[4 * ng]), solve_ms=float(out[4 * ng + 1]))

    def mf_fronts(self) -> np.ndarray:
        """The factored multifrontal front storage (FP64 copy), for inspection."""
        out = np.empty(self.mf_info()["front_elems"])
        check(lib.v2d_ctx_mf_fronts(self._h, ptr(out)))
        return out

    GRAPH_MODES = {"off": 0, "apply": 1, "iteration": 2}

    def set_graph(self, mode):
        """Mixed precision: how GMRES launches its device work. "apply" (or
        True, the default): the preconditioner apply of each iteration is one
        captured HIP graph replay; "iteration": the whole iteration (apply,
        SpMV, Gram-Schmidt, norm, read-backs) is one replay of a graph captured
        per basis index; "off" (or False): every kernel launched eagerly.
        Results are the same bit for bit in every mode."""
        if mode is True:
            mode = "apply"
        elif mode is False:
            mode = "off"
        if mode not in self.GRAPH_MODES:
            raise ValueError(f"graph mode must be one of {sorted(self.GRAPH_MODES)}, True or False")
        check(lib.v2d_ctx_set_graph(self._h, self.GRAPH_MODES[mode]))

    def autotune_update_kernel(self, re: float, kinds=("mfma64", "valu64")) -> dict:
        """Time one block-LU factorisation of the current state's Jacobian with
        each update kernel and keep the fastest; returns {kind: seconds}."""
        import time
        self.assemble(re)
        times = {}
        for k in kinds:
            self.set_update_kernel(k)
            self.factor()   # first use of a kernel variant pays its load
            t0 = time.perf_counter()
            self.factor()
            times[k] = time.perf_counter() - t0
        self.set_update_kernel(min(times, key=times.get))
        return times

    def profile_factor(self) -> dict:
        """One factorisation of the last assembly with the update launches
        timed by HIP events (see v2d_ctx_profile_factor)."""
        out = np.zeros(5)
        check(lib.v2d_ctx_profile_factor(self._h, ptr(out)))
        return dict(launches=int(out[0]), update_ms=float(out[1]), flops=float(out[2]),
                    bytes=float(out[3]), factor_ms=float(out[4]))

    # ----------------------------------------------------------- system ----
    def assemble(self, re, invdtau=0.0, inv2dt=0.0, urot=0.0):
        check(lib.v2d_ctx_assemble(self._h, float(re), float(invdtau), float(inv2dt), float(urot)))

    def rhs_norm(self) -> float:
        v = C.c_double()
        check(lib.v2d_ctx_rhs_norm(self._h, C.byref(v)))
        return float(v.value)

    def nnz(self) -> int:
        v = C.c_int64()
        check(lib.v2d_ctx_nnz(self._h, C.byref(v)))
        return int(v.value)

    def export_coo(self):
        nnz = self.nnz()
        ci = np.empty(nnz, np.int32)
        cj = np.empty(nnz, np.int32)
        cv = np.empty(nnz, np.float64)
        b = np.empty(2 * self.N, np.float64)
        check(lib.v2d_ctx_export_coo(self._h, ptr(ci), ptr(cj), ptr(cv), ptr(b)))
        return ci, cj, cv, b

    def export_csr(self):
        nnz = self.nnz()
        ip = np.empty(2 * self.N + 1, np.int64)
        ix = np.empty(nnz, np.int32)
        dv = np.empty(nnz, np.float64)
        check(lib.v2d_ctx_export_csr(self._h, ptr(ip), ptr(ix), ptr(dv)))
        return ip, ix, dv

    def spmv(self, x):
        x = f64c(x)
        if x.size != 2 * self.N:
            raise ValueError("x must have length 2*imax*jmax")
        y = np.empty(2 * self.N)
        check(lib.v2d_ctx_spmv(self._h, ptr(x), ptr(y)))
        return y

    def factor(self):
        check(lib.v2d_ctx_factor(self._h))

    def solve(self, b):
        b = f64c(b)
        if b.size != 2 * self.N:
            raise ValueError("b must have length 2*imax*jmax")
        x = np.empty(2 * self.N)
        check(lib.v2d_ctx_solve(self._h, ptr(b), ptr(x)))
        return x

    # ----------------------------------------------------------- Newton ----
    def newton_step(self, re, invdtau=0.0, inv2dt=0.0, urot=0.0):
        """One assemble + solve + update; returns (res, ||b||, lin_iters)."""
        res, rn, its = C.c_double(), C.c_double(), C.c_int()
        check(lib.v2d_ctx_newton_step(self._h, float(re), float(invdtau), float(inv2dt),
                                      float(urot), C.byref(res), C.byref(rn), C.byref(its)))
        return float(res.value), float(rn.value), int(its.value)

    def solve_transpose(self, b):
        """x = A^-T b with the stored factorisation."""
        b = f64c(b)
        if b.size != 2 * self.N:
            raise ValueError("b must have length 2*imax*jmax")
        x = np.empty(2 * self.N)
        check(lib.v2d_ctx_solve_transpose(self._h, ptr(b), ptr(x)))
        return x

    def residual_vjp(self, lam, re, inv2dt=0.0, urot=0.0):
        """lambda^T db/d(metric planes) (10 x N), lambda . db/dRe, lambda . db/dalpha [rad]."""
        lam = f64c(lam)
        if lam.size != 2 * self.N:
            raise ValueError("lam must have length 2*imax*jmax")
        gm = np.empty((10, self.N))
        gs = np.zeros(2)
        check(lib.v2d_ctx_residual_vjp(self._h, float(re), float(inv2dt), float(urot), ptr(lam),
                                       ptr(gm), ptr(gs)))
        return gm, float(gs[0]), float(gs[1])

    def forces_vjp(self, re, ref_length, wd=1.0, wl=0.0):
        """dJ/d(omega) (N), dJ/d(metric planes) (10 x N) and (Cd, Cl, explicit dJ/dRe)
        for J = wd Cd + wl Cl at the current state."""
        go = np.empty(self.N)
        gm = np.empty((10, self.N))
        out = np.zeros(3)
        check(lib.v2d_ctx_forces_vjp(self._h, float(re), float(ref_length), float(wd), float(wl),
                                     ptr(go), ptr(gm), ptr(out)))
        return go, gm, out

    def adjoint_forces(self, re, ref_length, wd=1.0, wl=0.0, urot=0.0):
        """Adjoint gradient of J = wd Cd + wl Cl at the stored steady state:
        returns dict(cd, cl, dalpha (per degree), dre, gx, gy (imax x jmax))."""
        gx = np.empty(self.N)
        gy = np.empty(self.N)
        gs = np.zeros(2)
        cf = np.zeros(2)
        check(lib.v2d_ctx_adjoint_forces(self._h, float(re), float(urot), float(ref_length), float(wd),
                                         float(wl), ptr(gx), ptr(gy), ptr(gs), ptr(cf)))
        return dict(cd=float(cf[0]), cl=float(cf[1]), dalpha=float(gs[0]), dre=float(gs[1]),
                    gx=gx.reshape(self.jmax, self.imax).T.copy(),
                    gy=gy.reshape(self.jmax, self.imax).T.copy())

    def steady_newton(self, re, tol=1e-10, rms_tol=None, max_iter=50, urot=0.0,
                      record=False) -> "NewtonSolve":
        """Pure Newton iterations from the current state (the steady inner loop
        of vorti2d/solver.py:136-145): step while res = ||delta||_2 / (2N) > tol
        and fewer than max_iter steps.  rms_tol: also require ||R||_2 / sqrt(2N)
        < rms_tol at the final state (an FP64 assembly after the last step),
        stepping on until it holds.  record: keep per-step histories.  The
        result's `converged` says whether the rules were met."""
        n2 = 2 * self.N
        res, it = 1.0e5, 0
        rs, rns, lin = [], [], []
        rfinal = float("nan")
        while it < max_iter:
            res, rn, li = self.newton_step(re, 0.0, 0.0, urot)
            it += 1
            if record:
                rs.append(res)
                rns.append(rn)
                lin.append(li)
            if not np.isfinite(res):
                break
            if res <= tol:
                if rms_tol is None:
                    break
                self.assemble(re, 0.0, 0.0, urot)
                rfinal = self.rhs_norm()
                if rfinal / np.sqrt(n2) < rms_tol:
                    break
        if np.isnan(rfinal) and np.isfinite(res):
            self.assemble(re, 0.0, 0.0, urot)
            rfinal = self.rhs_norm()
        ok = bool(np.isfinite(res) and res <= tol
                  and (rms_tol is None or rfinal / np.sqrt(n2) < rms_tol))
        return NewtonSolve(ok, it, float(res), float(rfinal), rs, rns, lin)

    def adjoint_bdf2_begin(self):
        """Start the reverse sweep of an unsteady adjoint (FP64 context)."""
        check(lib.v2d_ctx_adjoint_bdf2_begin(self._h))

    def adjoint_bdf2_step(self, re, inv2dt, urot, ref_length, wd=0.0, wl=0.0):
        """One level of the reverse sweep at the current state (the level's
        converged psi, omega and its history); returns (Cd_n, Cl_n)."""
        cf = np.zeros(2)
        check(lib.v2d_ctx_adjoint_bdf2_step(self._h, float(re), float(inv2dt), float(urot),
                                            float(ref_length), float(wd), float(wl), ptr(cf)))
        return float(cf[0]), float(cf[1])

    def adjoint_bdf2_end(self):
        """Finish the sweep: dict(dalpha (per degree), dre, gx, gy (imax x jmax))."""
        gx = np.empty(self.N)
        gy = np.empty(self.N)
        gs = np.zeros(2)
        check(lib.v2d_ctx_adjoint_bdf2_end(self._h, ptr(gx), ptr(gy), ptr(gs)))
        return dict(dalpha=float(gs[0]), dre=float(gs[1]),
                    gx=gx.reshape(self.jmax, self.imax).T.copy(),
                    gy=gy.reshape(self.jmax, self.imax).T.copy())

    def bdf2_shift(self):
        check(lib.v2d_ctx_bdf2_shift(self._h))

    def forces(self, re, ref_length=None, moment_ref=(0.0, 0.0)):
        out = np.zeros(10)
        d = -1.0 if ref_length is None else float(ref_length)
        check(lib.v2d_ctx_forces(self._h, float(re), d, float(moment_ref[0]),
                                 float(moment_ref[1]), ptr(out)))
        return out

    def timers(self, reset: bool = False) -> dict:
        ms = np.zeros(4)
        cnt = np.zeros(3, np.int64)
        check(lib.v2d_ctx_timers(self._h, ptr(ms), ptr(cnt), int(reset)))
        return dict(assemble_ms=ms[0], factor_ms=ms[1], solve_ms=ms[2], step_ms=ms[3],
                    steps=int(cnt[0]), factorizations=int(cnt[1]), gmres_iters=int(cnt[2]))

    def time_assembly(self, re, invdtau=0.0, inv2dt=0.0, urot=0.0, launches=50) -> float:
        v = C.c_double()
        check(lib.v2d_ctx_time_assembly(self._h, float(re), float(invdtau), float(inv2dt),
                                        float(urot), int(launches), C.byref(v)))
        return float(v.value)

    def time_spmv(self, launches=50) -> float:
        v = C.c_double()
        check(lib.v2d_ctx_time_spmv(self._h, int(launches), C.byref(v)))
        return float(v.value)
