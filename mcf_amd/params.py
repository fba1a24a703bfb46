"""Physics and control parameters of the SDPD step.

Field names follow the reference's configuration keys (ctrl.mcf and
physics_config.mcf, read by mcf/src/io_read_ctrl.F90:157-357 and
io_read_physics.F90), so a reference config maps onto this record key for
key.  Derived rules follow the reference:

* smoothing length h = cut_off / 3 for the quintic spline
  (mcf/src/kernel_new.F90:49);
* kt is forced to 0 when Brownian is false
  (mcf/src/physics_adjust_parameters.F90:415-417);
* the cell edge is L / INT(L / (1.0001 cut_off)) (mcf/src/mcf.F90:617,
  ppm.tgz:src/ppm_neighlist_clist.F90:370-376).
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import List

PRECISION_FP64 = 0
PRECISION_MIXED = 1


@dataclass
class SDPDParams:
    num_dim: int = 3
    min_phys: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    max_phys: List[float] = field(default_factory=lambda: [1.0, 1.0, 1.0])
    cut_off: float = 0.1
    # control keys
    symmetry: bool = True
    kernel_type: int = 1          # 1 quintic spline, 2 Lucy
    rhs_density_type: int = 2     # 1 mass density, 2 number density
    rhs_force_type: int = 3       # 3 Hu-Adams, 4 Hu-Adams angular
    stateEquation_type: int = 2   # 1 Morris, 2 Batchelor (Tait)
    Newtonian: bool = True
    Brownian: bool = False
    body_force_type: int = 0
    # physics keys
    c: float = 10.0
    rho: float = 1.0
    rho_ref: float = 1.0
    gamma: float = 7.0
    eta: float = 0.1
    eta_coef: float = 4.0
    kt: float = 0.0
    dt: float = 1e-4
    tau: float = 0.1
    n_p: float = 0.0
    kt_p: float = 0.0
    body_force: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    random_seed: int = 1
    # Lees-Edwards faces (boundary config): le_dim = the dimension whose two
    # faces are Lees-Edwards, 1-based like the reference's bcdef faces
    # (2 = y, 3 = z; 0 = none), le_v_lower / le_v_upper = the face
    # velocities shear_v0(:, 2 le_dim - 1) and shear_v0(:, 2 le_dim)
    le_dim: int = 0
    le_v_lower: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    le_v_upper: List[float] = field(default_factory=lambda: [0.0, 0.0, 0.0])
    # device-path options
    precision: int = PRECISION_FP64

    def __post_init__(self) -> None:
        self.min_phys = [float(v) for v in list(self.min_phys) + [0.0] * (3 - len(self.min_phys))][:3]
        self.max_phys = [float(v) for v in list(self.max_phys) + [1.0] * (3 - len(self.max_phys))][:3]
        self.body_force = [float(v) for v in list(self.body_force) + [0.0] * (3 - len(self.body_force))][:3]
        self.le_v_lower = [float(v) for v in list(self.le_v_lower) + [0.0] * (3 - len(self.le_v_lower))][:3]
        self.le_v_upper = [float(v) for v in list(self.le_v_upper) + [0.0] * (3 - len(self.le_v_upper))][:3]

    @property
    def kt_effective(self) -> float:
        return float(self.kt) if self.Brownian else 0.0

    @property
    def length(self) -> List[float]:
        return [self.max_phys[d] - self.min_phys[d] for d in range(3)]

    def cell_counts(self) -> List[int]:
        cell = self.cut_off * 1.0001
        out = []
        for d in range(3):
            if d < self.num_dim:
                n = int((self.max_phys[d] - self.min_phys[d]) / cell)
                out.append(max(n, 1))
            else:
                out.append(1)
        return out

    def replace(self, **kw) -> "SDPDParams":
        return dataclasses.replace(self, **kw)

    def validate(self) -> None:
        """Parameter checks the reference performs at start-up
        (control_check_parameters, physics_check_parameters.F90:211-218,
        stateEquation_new.F90:80-100, kernel_new.F90)."""
        errs = []
        if self.num_dim not in (2, 3):
            errs.append("num_dim must be 2 or 3")
        if self.cut_off <= 0:
            errs.append("cut_off must be positive")
        if self.kernel_type not in (1, 2):
            errs.append("kernel_type must be 1 (quintic) or 2 (Lucy)")
        if self.rhs_density_type not in (1, 2):
            errs.append("rhs_density_type must be 1 or 2")
        if self.rhs_force_type not in (3, 4):
            errs.append("rhs_force_type must be 3 (Hu-Adams) or 4 (angular)")
        if self.stateEquation_type not in (1, 2):
            errs.append("stateEquation_type must be 1 or 2")
        if self.stateEquation_type == 1 and self.c == 0.0:
            errs.append("sound speed should not be zero")
        if self.stateEquation_type == 2 and (self.rho_ref == 0.0 or self.gamma == 0.0):
            errs.append("neither rho_ref nor gamma can be zero")
        if self.Brownian and self.rhs_density_type != 2:
            errs.append("Brownian motion requires rhs_density_type = 2")
        if not self.Newtonian and self.rhs_force_type != 3:
            errs.append("non-Newtonian flow is implemented for rhs_force_type = 3")
        if not self.Newtonian and abs(self.tau) == 0.0:
            errs.append("tau should not be zero")
        if self.eta < 0:
            errs.append("dissipative force should be non-positive (eta >= 0)")
        if self.dt <= 0:
            errs.append("dt must be positive")
        for d in range(self.num_dim):
            if self.max_phys[d] <= self.min_phys[d]:
                errs.append("max_phys must exceed min_phys")
        if errs:
            raise ValueError("; ".join(errs))
