"""Field catalogue of the ice-dynamics path.

Each entry names a field of the reference's ``type_mesh`` / ``type_ice_model``
(src/data_types_module.f90:15-345) and gives its grid and element type.  The
product C-ABI and the CPU oracle both address fields by these names, so one
table drives host-array allocation on both sides.

Grid codes (first dimension x second dimension, Fortran column-major):

* ``Aa``    nV            ``Aa2`` nV x 2      ``Aac``  nV x nC_mem
* ``Aac1``  nV x (nC_mem+1)                  ``Aaz``  nV x nZ
* ``Ac``    nAc           ``Ac2`` nAc x 2     ``Ac4``  nAc x 4
* ``Acz``   nAc x nZ
* ``AaAc``  nVAaAc        ``AaAc2`` nVAaAc x 2
* ``AaAcc`` nVAaAc x nC_mem                  ``AaAcc1`` nVAaAc x (nC_mem+1)
* ``AaAc5`` nVAaAc x 5
"""
from __future__ import annotations

import numpy as np

F64, I32 = np.float64, np.int32

# mesh arrays produced by the set-up routines (secondary data)
MESH_SECONDARY = {
    # name: (grid, dtype)
    "Nx": ("Aac1", F64), "Ny": ("Aac1", F64), "Nxx": ("Aac1", F64),
    "Nxy": ("Aac1", F64), "Nyy": ("Aac1", F64),
    "iAci": ("Aac", I32), "Aci": ("Ac4", I32), "VAc": ("Ac2", F64),
    "Nx_Ac": ("Ac4", F64), "Ny_Ac": ("Ac4", F64), "No_Ac": ("Ac4", F64),
    "Np_Ac": ("Ac", F64), "edge_index_Ac": ("Ac", I32),
    "VAaAc": ("AaAc2", F64), "nCAaAc": ("AaAc", I32), "CAaAc": ("AaAcc", I32),
    "Nx_AaAc": ("AaAcc1", F64), "Ny_AaAc": ("AaAcc1", F64), "Nxx_AaAc": ("AaAcc1", F64),
    "Nxy_AaAc": ("AaAcc1", F64), "Nyy_AaAc": ("AaAcc1", F64),
    "colour": ("AaAc", I32), "colour_vi": ("AaAc5", I32),
}

_AA = ["Hi", "Hb", "SL", "Hs", "dHb_dt", "dHi_dt", "dHs_dt", "Hi_prev", "A_flow_mean",
       "dHi_dx", "dHi_dy", "dHs_dx", "dHs_dy", "dHs_dx_shelf", "dHs_dy_shelf",
       "U_SIA", "V_SIA", "D_SIA", "U_SSA", "V_SSA"]
_AC = ["Hi_Ac", "Hb_Ac", "SL_Ac", "Hs_Ac",
       "dHi_dx_Ac", "dHi_dy_Ac", "dHi_dp_Ac", "dHi_do_Ac",
       "dHb_dx_Ac", "dHb_dy_Ac", "dHb_dp_Ac", "dHb_do_Ac",
       "dHs_dx_Ac", "dHs_dy_Ac", "dHs_dp_Ac", "dHs_do_Ac",
       "dSL_dx_Ac", "dSL_dy_Ac", "dSL_dp_Ac", "dSL_do_Ac",
       "dHs_dx_shelf_Ac", "dHs_dy_shelf_Ac", "A_flow_mean_Ac",
       "D_SIA_Ac", "Ux_SIA_Ac", "Uy_SIA_Ac", "Up_SIA_Ac", "Uo_SIA_Ac",
       "Ux_SSA_Ac", "Uy_SSA_Ac", "Up_SSA_Ac", "Uo_SSA_Ac", "Qabs_GL_Ac", "Qp_GL_Ac"]
_AAAC = ["Hi_AaAc", "Hb_AaAc", "SL_AaAc", "dHs_dx_shelf_AaAc", "dHs_dy_shelf_AaAc",
         "A_flow_mean_AaAc", "U_SSA_AaAc", "V_SSA_AaAc",
         "dU_SSA_dx_AaAc", "dU_SSA_dy_AaAc", "dV_SSA_dx_AaAc", "dV_SSA_dy_AaAc",
         "eta_AaAc", "N_AaAc", "N_AaAc_prev", "S_AaAc", "phi_fric_AaAc", "tau_c_AaAc",
         "RHSx_AaAc", "RHSy_AaAc", "eu_i_AaAc", "ev_i_AaAc", "LHSx_AaAc", "LHSy_AaAc",
         "resU_AaAc", "resV_AaAc", "tau_c_override_AaAc"]
_MASKS = ["mask_land", "mask_ocean", "mask_lake", "mask_ice", "mask_sheet", "mask_shelf",
          "mask_coast", "mask_margin", "mask_gl", "mask_cf", "mask"]

ICE_FIELDS = {}
for _n in _AA:
    ICE_FIELDS[_n] = ("Aa", F64)
ICE_FIELDS["A_flow"] = ("Aaz", F64)
for _n in ("Ti", "Ti_pmp", "Cpi", "Ki", "U_3D", "V_3D", "W_3D"):
    ICE_FIELDS[_n] = ("Aaz", F64)
ICE_FIELDS["dVi_in"] = ("Aac", F64)
for _n in _AC:
    ICE_FIELDS[_n] = ("Ac", F64)
ICE_FIELDS["A_flow_Ac"] = ("Acz", F64)
ICE_FIELDS["D_SIA_3D_Ac"] = ("Acz", F64)
ICE_FIELDS["Ti_Ac"] = ("Acz", F64)
for _n in _AAAC:
    ICE_FIELDS[_n] = ("AaAc", F64)
for _n in _MASKS:
    ICE_FIELDS[_n] = ("Aa", I32)
for _n in _MASKS:
    ICE_FIELDS[_n + "_Ac"] = ("Ac", I32)


def grid_shape(grid: str, nV: int, nAc: int, nC_mem: int, nZ: int):
    nA = nV + nAc
    return {
        "Aa": (nV,), "Aa2": (nV, 2), "Aac": (nV, nC_mem), "Aac1": (nV, nC_mem + 1),
        "Aaz": (nV, nZ), "Ac": (nAc,), "Ac2": (nAc, 2), "Ac4": (nAc, 4), "Acz": (nAc, nZ),
        "AaAc": (nA,), "AaAc2": (nA, 2), "AaAcc": (nA, nC_mem), "AaAcc1": (nA, nC_mem + 1),
        "AaAc5": (nA, 5),
    }[grid]


def alloc(spec: dict, nV: int, nAc: int, nC_mem: int, nZ: int) -> dict:
    out = {}
    for name, (grid, dt) in spec.items():
        out[name] = np.zeros(grid_shape(grid, nV, nAc, nC_mem, nZ), dtype=dt, order="F")
    return out


# Default vertical grid (src/configuration_module.f90:124-126)
ZETA_DEFAULT = np.array([0.00, 0.10, 0.20, 0.30, 0.40, 0.50, 0.60, 0.70, 0.80, 0.90,
                         0.925, 0.95, 0.975, 0.99, 1.00])
