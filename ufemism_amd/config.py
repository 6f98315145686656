"""Configuration subset used by the ice-dynamics path.

Defaults are the reference's ``*_config`` defaults
(src/configuration_module.f90:124-126 for zeta, :169-184 for the SIA/SSA
parameters).  ``use_tau_c_override`` has no reference counterpart: it lets a
caller replace the Martin et al. (2011) till yield stress by a prescribed
field, which is how a friction coefficient beta [Pa m^-1/3 yr^1/3] is
expressed under the reference's regularised-Coulomb law (tau_c = beta *
u_threshold**q, q = 0.3, u_threshold = 100 m/yr).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .fields import ZETA_DEFAULT


@dataclasses.dataclass
class DynParams:
    zeta: np.ndarray = dataclasses.field(default_factory=lambda: ZETA_DEFAULT.copy())
    m_enh_sia: float = 5.0
    m_enh_ssa: float = 0.7
    SSA_RN_tol: float = 1e-5
    SSA_max_outer_loops: int = 50
    SSA_max_residual_UV: float = 2.5
    SSA_SOR_omega: float = 1.2
    SSA_max_inner_loops: int = 10000
    use_analytical_GL_flux: bool = False
    ssa_forced_zero: bool = False
    zero_Hi_at_edges: bool = True
    use_tau_c_override: bool = False
    # ice_physical_properties: 0 A fields are inputs, 1 Arrhenius from Ti,
    # 2 the fixed benchmark values (src/general_ice_model_data_module.f90:298-462)
    physical_properties: int = 0

    @property
    def nZ(self) -> int:
        return int(len(self.zeta))


# regularised-Coulomb constants, src/ice_dynamics_module.f90:738-740
Q_PLASTIC = 0.30
U_THRESHOLD = 100.0


def tau_c_from_beta(beta):
    """Yield stress that turns tau_b = tau_c (|u|/u_th)^q into tau_b = beta |u|^q."""
    return np.asarray(beta, dtype=np.float64) * U_THRESHOLD ** Q_PLASTIC
