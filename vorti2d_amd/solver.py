"""Unified steady / unsteady dual-time Newton driver (mirror of vorti2d/solver.py).

Outer loop: physical time (BDF2; a single level when steady).  Inner loop:
Newton / pseudo-time iterations until ||d||_2/(2N) <= pseudo_tol
(vorti2d/solver.py:136-145).  Each inner iteration is one device call
(v2d_ctx_newton_step): assembly, block-LU solve and state update all run on the
GPU; the host sees two scalars per iteration.
"""
from __future__ import annotations

import math
import os
import time
import warnings

import numpy as np

from . import fields_io
from . import mesh as meshmod
from . import restart as restartmod
from .config import Config
from .device import DeviceContext
from .forces import FORCE_COLUMNS, ForceCoeffs

_PROFILE = bool(os.environ.get("VORTI2D_TIMING"))
_XDMF_WARNED = False
_METRIC_NAMES = ("jac", "alfa", "beta", "gama", "pmet", "qmet", "detadx", "detady",
                 "xphys", "yphys")


class Solver:
    """Replicated-state solver of vorti2d/solver.py:32-283, state resident on
    the GPU.  ``mesh=(xg, yg)`` supplies the grid in memory instead of the CSV
    pair named in the config."""

    def __init__(self, config: Config, mesh=None):
        self.cfg = config
        self.rank, self.size = 0, 1
        self.restarted = False
        self.res_history: list[list[float]] = []
        self.rnorm_history: list[list[float]] = []
        self._setup(mesh)

    def _log(self, msg: str):
        if self.cfg.verbose:
            print(msg, flush=True)

    # -------------------------------------------------------------- setup --
    def _setup(self, mesh):
        cfg = self.cfg
        cfg.check_linsolve()
        for msg in cfg.ignored_reference_knobs():
            warnings.warn(msg, stacklevel=3)
        if mesh is not None:
            xg, yg = mesh
        elif cfg.mesh_cgns:
            xg, yg = meshmod.load_cgns_ogrid(cfg.mesh_cgns, verbose=cfg.verbose)
        else:
            xg, yg = meshmod.load_mesh(cfg.mesh_xg, cfg.mesh_yg)
        self.xg = np.ascontiguousarray(xg, dtype=np.float64)
        self.yg = np.ascontiguousarray(yg, dtype=np.float64)
        self.imax, self.jmax = self.xg.shape
        self.ndof = self.imax * self.jmax
        self.dksi = 1.0 / (self.imax - 1)
        self.deta = 1.0 / (self.jmax - 1)
        self._ff_bc = cfg.ff_bc_code()
        self._ca = cfg.cos_alpha()
        self._sa = cfg.sin_alpha()
        self.ctx = DeviceContext(self.xg, self.yg, self._ff_bc, self._ca, self._sa,
                                 cfg.precision_code())
        if cfg.precision_code() == 1:
            self.ctx.set_solver(cfg.refactor_every, cfg.gmres_rtol, cfg.gmres_max_iter)
            self.ctx.set_forcing(cfg.gmres_forcing, cfg.gmres_forcing_target)
        for name, arr in zip(_METRIC_NAMES, self.ctx.metrics()):
            setattr(self, name, arr)
        self._init_state()
        self._log(f"vorti2d: mesh {self.imax}x{self.jmax}  ndof={self.ndof}  ranks=1  "
                  f"mode={'steady' if cfg.steady else 'unsteady'}  precision={cfg.precision}")

    def _init_state(self):
        cfg = self.cfg
        if cfg.restart_in:
            st = restartmod.load_restart(cfg.restart_in, self.imax, self.jmax)
            self.ctx.set_state(st.psi, st.ome, st.omeold, st.omeoldold)
            self.t = float(st.t)
            self.step = int(st.step)
            self.restarted = True
            self._log(f"  restarted from {cfg.restart_in}: t={self.t}, step={self.step}")
        else:
            self.ctx.init_state()
            self.t = cfg.t_start
            self.step = 0

    # -------------------------------------------------------------- state --
    @property
    def psi(self):
        return self.ctx.get_state(("psi",))[0]

    @property
    def ome(self):
        return self.ctx.get_state(("ome",))[0]

    @property
    def omeold(self):
        return self.ctx.get_state(("omeold",))[0]

    @property
    def omeoldold(self):
        return self.ctx.get_state(("omeoldold",))[0]

    # ------------------------------------------------------------ solving --
    def _assemble_and_solve_once(self, invdtau, inv2dt, urot):
        res, rnorm, _ = self.ctx.newton_step(self.cfg.re, invdtau, inv2dt, urot)
        self._last_rnorm = rnorm
        return res

    def _pseudo_solve(self, t, invdtau, inv2dt, urot):
        cfg = self.cfg
        res, it = 1.0e5, 0
        res_log, rn_log = [], []
        while res > cfg.pseudo_tol and it < cfg.max_pseudo_iter:
            res = self._assemble_and_solve_once(invdtau, inv2dt, urot)
            it += 1
            res_log.append(math.log10(res) if res > 0 else -300.0)
            rn_log.append(self._last_rnorm)
        self.res_history.append(res_log)
        self.rnorm_history.append(rn_log)
        return res, it, res_log

    # ---------------------------------------------------------------- run --
    def run(self):
        cfg = self.cfg
        fields_io.ensure_dirs(cfg.out_dir)
        fields_io.write_grid(cfg.out_dir, self.xg, self.yg)
        forces_path = os.path.join(cfg.out_dir, "forces.csv")
        if cfg.compute_forces and (not self.restarted or not os.path.exists(forces_path)):
            fields_io.init_force_history(cfg.out_dir, FORCE_COLUMNS)
        global _XDMF_WARNED
        if cfg.write_xdmf and not _XDMF_WARNED:
            _XDMF_WARNED = True   # the reference default is True: say it once per process
            warnings.warn("write_xdmf: vorti2d_amd has no XDMF/HDF5 output; "
                          "fields are written as CSV only", stacklevel=2)

        invdtau = cfg.inv_dtau()
        inv2dt = cfg.inv_2dt()
        if cfg.steady:
            times = [cfg.t_start]
        else:
            start = self.t + cfg.dt_phys if self.restarted else cfg.t_start
            nsteps = int(round((cfg.t_end - start) / cfg.dt_phys)) + 1
            times = [start + m * cfg.dt_phys for m in range(max(nsteps, 0))]

        t_wall0 = time.time()
        local_step = 0
        for t in times:
            urot = cfg.u_rot(t)
            res, it, res_log = self._pseudo_solve(t, invdtau, inv2dt, urot)
            if not cfg.steady:
                self.ctx.bdf2_shift()
            self.t = t
            self.step += 1
            local_step += 1
            time_idx = 0 if cfg.steady else int(round(t / cfg.dt_phys))
            self._log(f"  t={t:8.3f}  inner_iters={it:3d}  res={res:.3e}  step={self.step}")
            save_now = cfg.steady or local_step % cfg.save_fields_every == 0
            if cfg.compute_forces:
                fc = self.force_coeffs()
                fields_io.append_force_history(cfg.out_dir, FORCE_COLUMNS,
                                               {"t": t, **fc.to_dict()})
                self._log(f"           Cd={fc.cd:+.5f}  Cl={fc.cl:+.5f}  Cm={fc.cm:+.5f}")
            if save_now and cfg.write_csv:
                psi, ome = self.ctx.get_state(("psi", "ome"))
                fields_io.write_fields(cfg.out_dir, time_idx, psi, ome)
                fields_io.write_residual_history(cfg.out_dir, time_idx, res_log)
            if cfg.restart_every and not cfg.steady and local_step % cfg.restart_every == 0:
                self._write_restart()
        self._write_restart()
        total = time.time() - t_wall0
        self._log(f"vorti2d: done in {total:.2f}s, final t={self.t}")
        if _PROFILE:
            self._report_timings(total)
        return self

    def _report_timings(self, total: float):
        tm = self.ctx.timers()
        n = max(tm["steps"], 1)
        print(f"\n  --- vorti2d device phase breakdown ({n} Newton steps, "
              f"{self.imax}x{self.jmax}, {self.cfg.precision}) ---", flush=True)
        for name, key in (("assembly", "assemble_ms"), ("block-LU factor", "factor_ms"),
                          ("linear solve", "solve_ms"), ("whole step", "step_ms")):
            print(f"  {name:<18}{tm[key] / 1e3:>10.3f} s{tm[key] / n:>11.3f} ms/step")
        print(f"  factorisations {tm['factorizations']}, GMRES iterations {tm['gmres_iters']}, "
              f"wall {total:.2f} s\n", flush=True)

    def _write_restart(self):
        path = self.cfg.restart_out
        if not os.path.isabs(path):
            path = os.path.join(self.cfg.out_dir, path)
        psi, ome, omeold, omeoldold = self.ctx.get_state()
        st = restartmod.RestartState(psi, ome, omeold, omeoldold, self.t, self.step,
                                     self.imax, self.jmax, self.cfg.re)
        restartmod.save_restart(path, st)

    # ---------------------------------------------------------- accessors --
    def field_2d(self, name: str) -> np.ndarray:
        """psi or ome reshaped to (imax, jmax) (vorti2d/solver.py:267-270)."""
        arr = {"psi": self.psi, "ome": self.ome}[name]
        return arr.reshape(self.jmax, self.imax).T

    def force_coeffs(self) -> ForceCoeffs:
        """Force coefficients of the current state on the device
        (vorti2d/solver.py:272-283)."""
        cfg = self.cfg
        return ForceCoeffs.from_array(self.ctx.forces(cfg.re, cfg.ref_length, cfg.moment_center))


def run(config: Config, mesh=None):
    """Build a solver and run it (vorti2d/solver.py:286-296).  The distributed
    DMDA path of the reference has no separate form on one GPU: the device
    solver holds the whole mesh."""
    return Solver(config, mesh=mesh).run()
