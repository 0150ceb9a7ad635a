"""Case configuration (mirror of vorti2d/config.py:13-169).

The reference fields keep their names, defaults and meaning; the fields after
``restart_every`` are additive knobs of the device solver.
"""
from __future__ import annotations

import math
from dataclasses import asdict, dataclass

_FF_CODES = {"dirichlet": 0, "outflow": 1, "outflow_psi": 2}
_LINSOLVE = ("mumps", "gmres_asm", "gmres_jacobi", "gpu", "gmres_fs", "fieldsplit",
             "gmres_amgx", "gpu_amgx")


@dataclass
class Config:
    # physics (config.py:52)
    re: float = 60.0
    # time integration (config.py:60-64)
    steady: bool = False
    dt_phys: float = 0.2
    t_start: float = 0.0
    t_end: float = 50.0
    dtau: float = math.inf
    # inner controls (config.py:67-68)
    pseudo_tol: float = 1.0e-10
    max_pseudo_iter: int = 200
    # rotation schedule (config.py:71-77)
    rot_speed: float = 0.0
    rot_until: float = 0.0
    # far field / angle of attack (config.py:80-100)
    farfield_bc: str = "dirichlet"
    alpha_deg: float = 0.0
    # parallel / linear solver (config.py:103-144).  On one GPU every choice
    # runs the device block-LU Newton solve selected by `precision`; the names
    # are accepted so reference configurations load unchanged.
    distributed: bool = False
    linsolve: str = "mumps"
    ksp_rtol: float = 1.0e-10
    ksp_restart: int = 200
    ilu_fill: int = 1
    asm_overlap: int = 1
    # mesh (config.py:147-155)
    mesh_xg: str = "xg.csv"
    mesh_yg: str = "yg.csv"
    mesh_cgns: str | None = None
    inner_rad: float = 0.5
    outer_rad: float = 50.0
    # output (config.py:157-161)
    out_dir: str = "out"
    save_fields_every: int = 1
    verbose: bool = True
    write_csv: bool = True
    write_xdmf: bool = True
    # forces (config.py:164-169)
    compute_forces: bool = True
    ref_length: float | None = None
    moment_center: tuple[float, float] = (0.0, 0.0)
    # restart (config.py:172-174)
    restart_in: str | None = None
    restart_out: str = "restart.npz"
    restart_every: int = 0
    # --- device solver (additive) ---
    precision: str = "fp64"
    """"fp64": FP64 Jacobian and exact FP64 block-LU Newton step (the
    reference's MUMPS semantics).  "mixed": FP32 Jacobian, FP32 block-LU
    preconditioned FP32 GMRES, FP64 residual / state / update."""
    refactor_every: int = 1
    """mixed only: refactor the FP32 block LU every N Newton steps (0: only when
    GMRES misses gmres_rtol within gmres_max_iter iterations)."""
    gmres_rtol: float = 1.0e-6
    gmres_max_iter: int = 20
    gmres_forcing: float = 0.0
    """mixed only: > 0 enables the Eisenstat-Walker GMRES tolerance per Newton
    step, capped at this value (gmres_rtol stays the floor)."""
    gmres_forcing_target: float = 0.0
    """mixed only, with gmres_forcing: > 0 is an absolute ||R||_2 the solve is
    content with; steps close to it solve no more accurately than it needs."""

    def inv_dtau(self) -> float:
        return 0.0 if math.isinf(self.dtau) else 1.0 / self.dtau

    def inv_2dt(self) -> float:
        """1/(2 dt) for the BDF2 physical-time terms; 0 in steady mode."""
        return 0.0 if self.steady else 1.0 / (2.0 * self.dt_phys)

    def cos_alpha(self) -> float:
        return math.cos(math.radians(self.alpha_deg))

    def sin_alpha(self) -> float:
        return math.sin(math.radians(self.alpha_deg))

    def ff_bc_code(self) -> int:
        """Integer far-field selector (config.py:152-160)."""
        try:
            return _FF_CODES[self.farfield_bc.lower()]
        except KeyError:
            raise ValueError(f"unknown farfield_bc={self.farfield_bc!r}; "
                             f"choose one of {sorted(_FF_CODES)}")

    def u_rot(self, t: float) -> float:
        """Rotational wall speed at physical time t (config.py:162-166)."""
        if self.steady:
            return self.rot_speed
        return self.rot_speed if t <= self.rot_until else 0.0

    def precision_code(self) -> int:
        codes = {"fp64": 0, "mixed": 1}
        try:
            return codes[self.precision.lower()]
        except KeyError:
            raise ValueError(f"unknown precision={self.precision!r}; choose 'fp64' or 'mixed'")

    def check_linsolve(self) -> None:
        if self.linsolve.lower() not in _LINSOLVE:
            raise ValueError(f"unknown linsolve={self.linsolve!r} (choose one of {_LINSOLVE})")

    def ignored_reference_knobs(self) -> list[str]:
        """Reference settings this device solver accepts but does not use
        (the PETSc solver menu of vorti2d/dist_solver.py:127-229 and the DMDA
        switch), one message each; empty for the reference defaults."""
        out = []
        solver = ("exact FP64 block-LU Newton steps" if self.precision.lower() == "fp64" else
                  "FP32 block-LU preconditioned GMRES under FP64 Newton refinement")
        if self.linsolve.lower() != "mumps":
            out.append(f"linsolve={self.linsolve!r}: the PETSc Krylov menu is not built; the "
                       f"device path runs {solver} (precision={self.precision!r}); GMRES with "
                       "Jacobi or ILU preconditioning does not converge on the pure-Newton "
                       "system (docs/parallel.rst:61-64)")
        defaults = Config()
        for name in ("ksp_rtol", "ksp_restart", "ilu_fill", "asm_overlap"):
            if getattr(self, name) != getattr(defaults, name):
                out.append(f"{name}={getattr(self, name)!r}: a PETSc KSP/PC setting, unused by "
                           f"the device solver (its knobs are gmres_rtol, gmres_max_iter, "
                           "gmres_forcing, refactor_every)")
        if self.distributed:
            out.append("distributed=True: one GPU holds the whole mesh; the DMDA slab "
                       "decomposition is not built")
        return out

    def to_dict(self) -> dict:
        return asdict(self)
