/*
 * ufe_oracle.h -- CPU oracle for the UFEMISM ice-dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The tests compare the HIP library with this
 * restatement, and it is the timed CPU baseline.  Only tests/,
 * the smoke check and bench.py's CPU-baseline leg may load it; the product
 * library never links or calls it.
 *
 * Every array uses the reference's own storage: Fortran column-major,
 * 1-based index values with 0 in unused slots, 2-D widths nC_mem or
 * nC_mem+1 (src/data_types_module.f90:216-345, src/data_types_module.f90:15-214).
 *
 * Fields are addressed through two pointer tables (double and int) indexed
 * by the enums below, so that the Python side can bind them by name through
 * om_field_name() without mirroring a C struct by hand.
 */
#ifndef UFE_ORACLE_H
#define UFE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mesh scalars (type_mesh, src/data_types_module.f90:216-345). */
typedef struct {
  int nV, nTri, nC_mem, nAc, nVAaAc;
  double xmin, xmax, ymin, ymax;
  int colour_nV[5];
} om_mesh_dims;

/* double-valued mesh arrays */
enum om_mesh_d {
  OM_V,          /* (nV,2) */
  OM_A,          /* (nV) */
  OM_Cw,         /* (nV,nC_mem) */
  OM_Nx, OM_Ny, OM_Nxx, OM_Nxy, OM_Nyy,          /* (nV,nC_mem+1) */
  OM_VAc,        /* (nAc,2) */
  OM_Nx_Ac, OM_Ny_Ac, OM_No_Ac,                  /* (nAc,4) */
  OM_Np_Ac,      /* (nAc) */
  OM_VAaAc,      /* (nVAaAc,2) */
  OM_Nx_AaAc, OM_Ny_AaAc, OM_Nxx_AaAc, OM_Nxy_AaAc, OM_Nyy_AaAc, /* (nVAaAc,nC_mem+1) */
  OM_MESH_D_COUNT
};
/* int-valued mesh arrays */
enum om_mesh_i {
  OM_nC, OM_C, OM_niTri, OM_iTri, OM_edge_index, OM_Tri,
  OM_iAci,       /* (nV,nC_mem) */
  OM_Aci,        /* (nAc,4) */
  OM_edge_index_Ac, /* (nAc) */
  OM_nCAaAc,     /* (nVAaAc) */
  OM_CAaAc,      /* (nVAaAc,nC_mem) */
  OM_colour,     /* (nVAaAc) */
  OM_colour_vi,  /* (nVAaAc,5) */
  OM_MESH_I_COUNT
};

typedef struct {
  om_mesh_dims d;
  double *f[OM_MESH_D_COUNT];
  int *i[OM_MESH_I_COUNT];
} om_mesh;

/* Ice-model fields (type_ice_model subset, src/data_types_module.f90:15-214). */
enum om_ice_d {
  /* Aa (nV) */
  OI_Hi, OI_Hb, OI_SL, OI_Hs, OI_dHb_dt, OI_dHi_dt, OI_dHs_dt, OI_Hi_prev,
  OI_A_flow_mean, OI_dHi_dx, OI_dHi_dy, OI_dHs_dx, OI_dHs_dy,
  OI_dHs_dx_shelf, OI_dHs_dy_shelf, OI_U_SIA, OI_V_SIA, OI_D_SIA, OI_U_SSA, OI_V_SSA,
  /* Aa (nV,nZ) */
  OI_A_flow, OI_Ti, OI_Ti_pmp, OI_Cpi, OI_Ki, OI_U_3D, OI_V_3D, OI_W_3D,
  /* Aa (nV,nC_mem) */
  OI_dVi_in,
  /* Ac (nAc) */
  OI_Hi_Ac, OI_Hb_Ac, OI_SL_Ac, OI_Hs_Ac,
  OI_dHi_dx_Ac, OI_dHi_dy_Ac, OI_dHi_dp_Ac, OI_dHi_do_Ac,
  OI_dHb_dx_Ac, OI_dHb_dy_Ac, OI_dHb_dp_Ac, OI_dHb_do_Ac,
  OI_dHs_dx_Ac, OI_dHs_dy_Ac, OI_dHs_dp_Ac, OI_dHs_do_Ac,
  OI_dSL_dx_Ac, OI_dSL_dy_Ac, OI_dSL_dp_Ac, OI_dSL_do_Ac,
  OI_dHs_dx_shelf_Ac, OI_dHs_dy_shelf_Ac, OI_A_flow_mean_Ac,
  OI_D_SIA_Ac, OI_Ux_SIA_Ac, OI_Uy_SIA_Ac, OI_Up_SIA_Ac, OI_Uo_SIA_Ac,
  OI_Ux_SSA_Ac, OI_Uy_SSA_Ac, OI_Up_SSA_Ac, OI_Uo_SSA_Ac,
  OI_Qabs_GL_Ac, OI_Qp_GL_Ac,
  /* Ac (nAc,nZ) */
  OI_A_flow_Ac, OI_D_SIA_3D_Ac, OI_Ti_Ac,
  /* AaAc (nVAaAc) */
  OI_Hi_AaAc, OI_Hb_AaAc, OI_SL_AaAc, OI_dHs_dx_shelf_AaAc, OI_dHs_dy_shelf_AaAc,
  OI_A_flow_mean_AaAc, OI_U_SSA_AaAc, OI_V_SSA_AaAc,
  OI_dU_SSA_dx_AaAc, OI_dU_SSA_dy_AaAc, OI_dV_SSA_dx_AaAc, OI_dV_SSA_dy_AaAc,
  OI_eta_AaAc, OI_N_AaAc, OI_N_AaAc_prev, OI_S_AaAc, OI_phi_fric_AaAc, OI_tau_c_AaAc,
  OI_RHSx_AaAc, OI_RHSy_AaAc, OI_eu_i_AaAc, OI_ev_i_AaAc, OI_LHSx_AaAc, OI_LHSy_AaAc,
  OI_resU_AaAc, OI_resV_AaAc,
  /* AaAc (nVAaAc): basal-friction override (no reference field) */
  OI_tau_c_override_AaAc,
  OM_ICE_D_COUNT
};
enum om_ice_i {
  OI_mask_land, OI_mask_ocean, OI_mask_lake, OI_mask_ice, OI_mask_sheet,
  OI_mask_shelf, OI_mask_coast, OI_mask_margin, OI_mask_gl, OI_mask_cf, OI_mask,
  OI_mask_land_Ac, OI_mask_ocean_Ac, OI_mask_lake_Ac, OI_mask_ice_Ac, OI_mask_sheet_Ac,
  OI_mask_shelf_Ac, OI_mask_coast_Ac, OI_mask_margin_Ac, OI_mask_gl_Ac, OI_mask_cf_Ac,
  OI_mask_Ac,
  OM_ICE_I_COUNT
};

typedef struct {
  double *f[OM_ICE_D_COUNT];
  int *i[OM_ICE_I_COUNT];
} om_ice;

/* Configuration subset (src/configuration_module.f90:124-126, 169-184). */
typedef struct {
  int nZ;
  double zeta[32];
  double m_enh_sia, m_enh_ssa;
  double SSA_RN_tol;
  int SSA_max_outer_loops;
  double SSA_max_residual_UV;
  double SSA_SOR_omega;
  int SSA_max_inner_loops;
  int use_analytical_GL_flux;   /* calculate_GL_flux (Tsai 2015) at the GL Ac points, skipped by SOR */
  int ssa_forced_zero;          /* EISMINT/Halfar/Bueler benchmarks: SSA velocities set to zero */
  int zero_Hi_at_edges;         /* calculate_ice_thickness_change edge condition */
  int use_tau_c_override;       /* replace basal_yield_stress tau_c by OI_tau_c_override_AaAc */
} om_params;

/* Solver statistics returned by om_solve_SSA. */
typedef struct {
  int picard_iterations;        /* viscosity iterations started */
  int picard_converged;         /* exited through RN < SSA_RN_tol */
  int resets;                   /* SOR instability resets */
  int total_sweeps;
  double last_RN;
  double last_max_residual;
  int sweeps[64];               /* SOR sweeps per Picard iteration (first 64) */
  int status;
} om_ssa_stats;

/* status codes (shared meaning with the product C-ABI) */
enum {
  OM_OK = 0,
  OM_SOR_NOT_CONVERGED = 1,     /* reference: warning only */
  OM_SSA_UNSTABLE = 3,          /* reference: MPI_ABORT after a second reset */
  OM_COLOURING_IDENTIFY = 4,    /* reference: identify() always aborts */
  OM_COLOURING_FAILED = 5,      /* reference: other colouring aborts */
  OM_BAD_INPUT = 6
};

const char *om_mesh_field_name_d(int id);
const char *om_mesh_field_name_i(int id);
const char *om_ice_field_name_d(int id);
const char *om_ice_field_name_i(int id);

/* mesh set-up */
int  om_count_Ac(const om_mesh *m);
int  om_make_Ac_mesh(om_mesh *m);
void om_get_neighbour_functions(om_mesh *m);
int  om_calculate_five_colouring_AaAc(om_mesh *m);
void om_get_neighbour_functions_vertex_gr(const double *V_vi, int n, const double *V_vc, int ld_vc,
                                          int is_edge, double *Nx, double *Ny, double *Nxx,
                                          double *Nxy, double *Nyy);

/* operators */
void om_get_mesh_derivatives(const om_mesh *m, const double *d, double *ddx, double *ddy);
void om_get_mesh_curvatures(const om_mesh *m, const double *d, double *ddxx, double *ddxy, double *ddyy);
void om_get_mesh_derivatives_3D(const om_mesh *m, const double *d, double *ddx, double *ddy, int nz);
void om_apply_Neumann_boundary(const om_mesh *m, double *d);
void om_apply_Neumann_boundary_3D(const om_mesh *m, double *d, int nz);
void om_get_mesh_derivatives_Ac(const om_mesh *m, const double *d, double *ddx, double *ddy,
                                double *ddp, double *ddo);
void om_get_mesh_derivatives_AaAc(const om_mesh *m, const double *d, double *ddx, double *ddy);
void om_get_mesh_curvatures_AaAc(const om_mesh *m, const double *d, double *ddxx, double *ddxy,
                                 double *ddyy);
void om_apply_Neumann_boundary_AaAc(const om_mesh *m, double *d);
void om_map_Aa_to_Ac(const om_mesh *m, const double *d_Aa, double *d_Ac);
void om_map_Aa_to_Ac_3D(const om_mesh *m, const double *d_Aa, double *d_Ac, int nz);
void om_map_Ac_to_Aa(const om_mesh *m, const double *d_Ac, double *d_Aa);
void om_map_Ac_to_Aa_3D(const om_mesh *m, const double *d_Ac, double *d_Aa, int nz);
void om_rotate_xy_to_po(const om_mesh *m, const double *x, const double *y, double *p, double *o);

/* physics */
double om_vertical_average(const om_params *p, const double *f);
void om_vertical_integrate(const om_params *p, const double *f, double *int_f);
void om_update_general_ice_model_data(const om_mesh *m, om_ice *ice);
void om_ice_physical_properties(const om_mesh *m, om_ice *ice, const om_params *p, int mode);
/* solve_SIA_3D, src/ice_dynamics_module.f90:317-405: U_3D, V_3D, W_3D */
void om_solve_SIA_3D(const om_mesh *m, om_ice *ice, const om_params *p);
/* the critical time steps of determine_timesteps_and_actions
 * (src/UFEMISM_main_model.f90:734-775), after the 0.9 correction:
 * out = {dt_D_2D_min, dt_V_2D_SSA_min, dt_V_3D_SIA_min} */
void om_critical_timesteps(const om_mesh *m, const om_ice *ice, const om_params *p, double out[3]);
/* first-order neighbour functions on the triangles (src/mesh_derivatives_module.f90:30-47) */
void om_triangle_neighbour_functions(const om_mesh *m, double *NxTri, double *NyTri);
/* get_upwind_derivative_vertex_3D (src/mesh_derivatives_module.f90:434-483) for
 * every vertex and layer of d (nV,nz); returns the number of (vertex, layer)
 * pairs whose upwind triangle was not found (their derivatives are zero) */
int om_upwind_derivatives_3D(const om_mesh *m, const double *R, const double *NxTri, const double *NyTri,
                             const double *U, const double *V, const double *d, int nz, double *ddx,
                             double *ddy);
void om_determine_masks(const om_mesh *m, om_ice *ice);
void om_solve_SIA(const om_mesh *m, om_ice *ice, const om_params *p);
void om_basal_yield_stress(const om_mesh *m, om_ice *ice);
void om_calculate_GL_flux(const om_mesh *m, om_ice *ice);
void om_SSA_effective_viscosity(const om_mesh *m, om_ice *ice, const om_params *p);
void om_SSA_sliding_term(const om_mesh *m, om_ice *ice);
int  om_solve_SSA_linearised(const om_mesh *m, om_ice *ice, const om_params *p, int *did_reset,
                             double *last_max_residual, int *sweeps);
int  om_solve_SSA(const om_mesh *m, om_ice *ice, const om_params *p, om_ssa_stats *st);
void om_calculate_ice_thickness_change(const om_mesh *m, om_ice *ice, const double *SMB_year,
                                       const double *BMB, double dt, const int *mask_noice,
                                       const om_params *p);
int  om_set_threads(int n);

/* Primary-mesh geometry, src/mesh_help_functions_module.f90 (reference layout,
 * column-major, 1-based index values). */
typedef struct {
  int nV, nTri, nC_mem;
  double xmin, xmax, ymin, ymax, tol_dist;
  const double *V;                     /* (nV,2) */
  const int *nC, *C, *niTri, *iTri, *edge_index;
  int *Tri;                            /* (nTri,3), rotated by the circumcentre retry */
  const int *Tri_edge_index;           /* (nTri) */
  double *Tricc;                       /* (nTri,2) */
} om_geom;

/* update_triangle_circumcenter over every triangle (:854-897); returns the
 * number of triangles still without a valid circumcentre, *rotated counts the retries */
int om_update_triangle_circumcenters(om_geom *g, int *rotated);
/* find_Voronoi_cell_areas (:15-50) with find_Voronoi_cell_vertices(_free,
 * _edge, _corner) and crop_circumcenter (:219-487); returns the number of
 * Voronoi vertices outside the domain (the reference stops there) */
int om_find_Voronoi_cell_areas(const om_geom *g, double *A);
/* the Voronoi cell vertices of one vertex vi (1-based) after clamping; returns nVor */
int om_find_Voronoi_cell_vertices(const om_geom *g, int vi, double *Vor, int max_vor, int *outside);
/* find_connection_widths (:101-172); returns the number of failed connections */
int om_find_connection_widths(const om_geom *g, double *Cw);
/* find_triangle_areas (:173-192) */
void om_find_triangle_areas(const om_geom *g, double *TriA);
/* determine_mesh_resolution (:193-216) */
void om_determine_mesh_resolution(const om_geom *g, double *R, double *rmin, double *rmax);

/* ---- remapping, src/mesh_mapping_module.f90 (ufe_oracle_remap.c) ---- */
void om_voronoi_geometric_centres(const om_geom *g, const double *A, double *GC);
int om_cell_polygon(const om_geom *g, int vi, double *P, int cap);
int om_clip_convex(const double *S, int n, const double *C, int k, double *out, int cap);
int om_find_containing_vertex(const om_geom *g, const double p[2], int vi);
int om_find_containing_triangle(const om_geom *g, const int *TriC, const double p[2], int t);
int om_remap_conservative(const om_geom *src, const double *A_src, const double *GC_src, const om_geom *dst,
                          const double *A_dst, int *ptr, int *vi, double *w0, double *w1x, double *w1y,
                          int max_total);
int om_remap_trilin_nearest(const om_geom *src, const int *TriC, const om_geom *dst, int *tri_vi, double *tri_w,
                            int *nn_vi);

#ifdef __cplusplus
}
#endif
#endif
