"""Run configuration mirroring the reference's module globals for the
two-fluid (flag_eqs='PIP') HLL-branch path.

The reference keeps every setting in ``module globalvar`` (Globalvar.f90:42-101)
and derives the array sizes in ``get_parameters`` (Model_rot.f90:255-333):
``margin = s_order = 2``, ``ix = ix0 + 2*margin``, ``nvar_m = 8 (+1 for psi when
flag_divb=1, 2-D only)``, ``gm = gm_d/gm_n``.  ``PIPConfig`` holds the same
quantities with the same meaning; ``to_c()`` fills the C-ABI ``pip_cfg``.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Tuple

MARGIN = 2  # s_order=2 HLL branch: margin(1)=s_order (Model_rot.f90:286)


@dataclasses.dataclass
class PIPConfig:
    ndim: int
    nx: int
    ny: int = 1
    xmin: float = 0.0
    xmax: float = 1.0
    ymin: float = 0.0
    ymax: float = 1.0
    gm: float = 5.0 / 3.0          # gm_d/gm_n (Model_rot.f90:333)
    safety: float = 0.3            # CFL safety (setting 'safety')
    ro_lim: float = 1.0e-8         # setting 'ro_lim'
    pr_lim: float = 1.0e-8         # setting 'pr_lim'
    flag_divb: int = 0             # 1: Dedner GLM psi (forced 0 in 1-D, Model_rot.f90:309)
    db_clean: float = 0.7          # setting 'db_clean'
    t_order: int = 2               # SSP-RK order (flag_time=1)
    flag_ir: int = 0               # 0 off, 3 = Popescu empirical rates (PIP_rot.f90:178)
    flag_ir_type: int = 0          # 0 new formulation (ion_pot/arb_heat)
    flag_col: int = 1              # 1: constant alpha = col; 2..7 set_collisional col_type 1..6
    col: float = 10.0              # collision coefficient alpha_c
    flag_pip_imp: int = 0          # 1: implicit collision source (PIP_rot.f90:1333-1358)
    pip_imp_factor: float = 10.0   # its dt factor (Scheme_rot.f90:308-309; setting.txt default)
    ac_field: bool = False         # per-cell alpha_c supplied by the caller
    t_ir: float = 1.0e-5           # setting 't_IR'
    T0: float = 5500.0             # setting 'T_norm'
    n0: float = 7.5e16             # setting 'n_norm'
    beta: float = 0.1              # setting 'beta'
    flag_bnd: Tuple[int, int, int, int] = (1, 1, 1, 1)  # x-,x+,y-,y+

    def __post_init__(self):
        if self.ndim not in (1, 2):
            raise ValueError("ndim must be 1 or 2")
        if self.ndim == 1:
            self.ny = 1
            self.flag_divb = 0
            self.ymin, self.ymax = 0.0, 0.0
        for f in self.flag_bnd[: 2 * self.ndim]:
            if f not in (1, 2, 3, 4, 5, 6, 7, 10, 11):
                raise ValueError(f"unsupported boundary flag {f}")
        if self.t_order not in (2, 3):
            raise ValueError("t_order must be 2 or 3 (SSP-RK)")
        if self.flag_col not in range(1, 8):
            raise ValueError("flag_col must be 1..7")
        if self.flag_col >= 2 and self.ac_field:
            raise ValueError("flag_col 2..7 computes alpha_c; ac_field needs flag_col = 1")

    # derived sizes (Model_rot.f90:309-333)
    @property
    def ix(self) -> int:
        return self.nx + 2 * MARGIN

    @property
    def jx(self) -> int:
        return self.ny + 2 * MARGIN if self.ndim == 2 else 1

    @property
    def nvar_m(self) -> int:
        return 8 + (1 if self.flag_divb == 1 else 0)

    @property
    def nvar_h(self) -> int:
        return 5

    @property
    def dx(self) -> float:
        # set_coordinate: dr(1)=(end(1)-start(1))/(ix-2*margin(1)) (Model_rot.f90:420)
        return (self.xmax - self.xmin) / float(self.nx)

    @property
    def dy(self) -> float:
        if self.ndim < 2:
            return 100.0  # dy(1)=100 in 1-D (Model_rot.f90:426-427)
        return (self.ymax - self.ymin) / float(self.ny)

    @property
    def ncells(self) -> int:
        return self.nx * self.ny

    def x_centers(self):
        """Cell centres incl. ghosts (Model_rot.f90:434-444)."""
        import numpy as np
        i = np.arange(self.ix, dtype=np.float64) - MARGIN
        return self.xmin + (i + 0.5) * self.dx

    def y_centers(self):
        import numpy as np
        if self.ndim < 2:
            return np.zeros(1)
        j = np.arange(self.jx, dtype=np.float64) - MARGIN
        return self.ymin + (j + 0.5) * self.dy

    def bytes_per_cell_update(self) -> int:
        """Algorithmic HBM bytes per cell-update of a fused SSP-RK2 step:
        stage 1 reads U^n and writes U^(1); stage 2 reads U^(1), U^n and writes
        U^(n+1) -> 5*nv*8 B, plus 2*8 B per spatially varying field read in
        each stage (per-cell alpha_c, arb_heat)."""
        nv = self.nvar_m + self.nvar_h
        extra = 0
        if self.ac_field:
            extra += 16
        if self.flag_ir >= 1 and self.flag_ir_type == 0:
            extra += 16
        return 5 * nv * 8 + extra

    def replace(self, **kw) -> "PIPConfig":
        return dataclasses.replace(self, **kw)


def split_fractions(xi_n: float):
    """Neutral/plasma split used by the reference ICs (shock_tube.f90:34-37):
    rho_n = f_n rho, rho_p = f_p rho, p_n = f_p_n p, p_p = f_p_p p."""
    f_n = xi_n
    f_p = 1.0 - xi_n
    f_p_n = f_n / (f_n + 2.0 * f_p)
    f_p_p = 2.0 * f_p / (f_n + 2.0 * f_p)
    return f_n, f_p, f_p_n, f_p_p


TWO_PI = 2.0 * math.pi
