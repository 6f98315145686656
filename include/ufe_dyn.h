/*
 * ufe_dyn.h -- C-ABI of the MI355X ice-dynamics library (libufe_dyn.so).
 *
 * The library replaces the reference's ice-dynamics path
 * (src/ice_dynamics_module.f90, src/general_ice_model_data_module.f90 and the
 * mesh operators of src/mesh_derivatives_module.f90 and
 * src/mesh_ArakawaC_module.f90) with hand-written CDNA4 HIP kernels.  Every
 * entry point takes plain pointers and sizes and returns an int status.
 *
 * Layout at the boundary is the reference's own: host arrays are Fortran
 * column-major, index arrays hold 1-based indices with 0 in unused slots,
 * and 2-D widths are nC_mem or nC_mem+1 (src/data_types_module.f90:15-345).
 * The library converts once, at mesh upload, into its device layouts.
 *
 * Threading: a context is bound to one HIP device and one stream and is not
 * thread-safe; callers serialise (in the reference every routine is a
 * collective over MPI ranks; here rank 0 calls and the others wait).
 */
#ifndef UFE_DYN_H
#define UFE_DYN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (replace the reference's MPI_ABORT / warnings) -------- */
enum {
  UFE_OK = 0,
  UFE_SOR_NOT_CONVERGED = 1,   /* src/ice_dynamics_module.f90:685-689 (warning) */
  UFE_SSA_RESET = 2,           /* :679-684, one reset happened (informational) */
  UFE_SSA_UNSTABLE = 3,        /* :533-540, second reset: reference aborts */
  UFE_COLOURING_IDENTIFY = 4,  /* src/mesh_five_colour_module.f90:452: identify() aborts */
  UFE_COLOURING_FAILED = 5,    /* src/mesh_five_colour_module.f90:168-171, 270-273, 329-341 */
  UFE_BAD_INPUT = 6,
  UFE_HIP_ERROR = 7,
  UFE_NOT_SUPPORTED = 8,      /* reserved for options an entry point does not take; none is returned now */
  UFE_HOST_ERROR = 9           /* host memory or threads could not be had (no exception leaves the library) */
};

enum { UFE_FP64 = 0, UFE_MIXED = 1, UFE_FP64_KRYLOV = 2 };
/* preconditioner of the mixed-precision Krylov solves */
enum { UFE_PRECOND_SOR = 0, UFE_PRECOND_MULTIGRID = 1 };

/* Mesh in reference layout.  Primary arrays are inputs everywhere; the
 * secondary arrays are written by ufe_mesh_build and read by ufe_mesh_upload.
 * Mirrors type_mesh, src/data_types_module.f90:216-345. */
typedef struct ufe_mesh_desc {
  int32_t nV, nTri, nC_mem, nAc;      /* nAc = (sum of nC) / 2 */
  double xmin, xmax, ymin, ymax;
  /* primary */
  const double *V;                    /* (nV,2)        */
  const int32_t *nC;                  /* (nV)          */
  const int32_t *C;                   /* (nV,nC_mem)   */
  const int32_t *niTri;               /* (nV)          */
  const int32_t *iTri;                /* (nV,nC_mem)   */
  const int32_t *edge_index;          /* (nV)          */
  const int32_t *Tri;                 /* (nTri,3)      */
  const double *A;                    /* (nV)   Voronoi cell areas */
  const double *Cw;                   /* (nV,nC_mem) connection widths */
  /* secondary */
  double *Nx, *Ny, *Nxx, *Nxy, *Nyy;  /* (nV,nC_mem+1) */
  int32_t *iAci;                      /* (nV,nC_mem)   */
  int32_t *Aci;                       /* (nAc,4)       */
  double *VAc;                        /* (nAc,2)       */
  double *Nx_Ac, *Ny_Ac, *No_Ac;      /* (nAc,4)       */
  double *Np_Ac;                      /* (nAc)         */
  int32_t *edge_index_Ac;             /* (nAc)         */
  double *VAaAc;                      /* (nVAaAc,2)    */
  int32_t *nCAaAc;                    /* (nVAaAc)      */
  int32_t *CAaAc;                     /* (nVAaAc,nC_mem) */
  double *Nx_AaAc, *Ny_AaAc, *Nxx_AaAc, *Nxy_AaAc, *Nyy_AaAc; /* (nVAaAc,nC_mem+1) */
  int32_t *colour;                    /* (nVAaAc)      */
  int32_t *colour_vi;                 /* (nVAaAc,5)    */
  int32_t colour_nV[5];
} ufe_mesh_desc;

/* Configuration subset, src/configuration_module.f90:124-126, 169-184. */
typedef struct ufe_params {
  int32_t nZ;
  double zeta[32];
  double m_enh_sia, m_enh_ssa;
  double SSA_RN_tol;
  int32_t SSA_max_outer_loops;
  double SSA_max_residual_UV;
  double SSA_SOR_omega;
  int32_t SSA_max_inner_loops;
  int32_t use_analytical_GL_flux;   /* calculate_GL_flux (:848-949) prescribes the GL Ac velocities,
                                       which SOR skips (:624-631) and the Krylov modes hold fixed */
  int32_t ssa_forced_zero;          /* EISMINT/Halfar/Bueler: SSA velocities zero (:431-449) */
  int32_t zero_Hi_at_edges;         /* ice-thickness edge condition (:189-229) */
  int32_t use_tau_c_override;       /* tau_c := field "tau_c_override_AaAc" after basal_yield_stress */
  int32_t precision;                /* solve_SSA linear solves: UFE_FP64 (the reference's SOR and
                                       stopping rule), UFE_MIXED or UFE_FP64_KRYLOV (GMRES-IR below) */
  int32_t sweep_batch;              /* SOR sweeps queued per host convergence read; 0 = adaptive */
  /* UFE_MIXED / UFE_FP64_KRYLOV: each Picard iteration's linearised system is
   * solved to the FP64 relative residual SSA_linear_rel_tol by restarted
   * GMRES(SSA_gmres_restart) preconditioned by SSA_precond_sweeps SOR sweeps,
   * FP32 sweeps and Krylov vectors for UFE_MIXED, FP64 for UFE_FP64_KRYLOV,
   * with FP64 residuals and iterative refinement either way. */
  double SSA_linear_rel_tol;        /* <= 0: 1e-6 */
  int32_t SSA_gmres_restart;        /* <= 0: 40 (SOR) or 20 (multigrid) */
  int32_t SSA_precond_sweeps;       /* <= 0: 5 */
  /* ice_physical_properties (src/general_ice_model_data_module.f90:298-462),
   * run at the end of update_general_ice_model_data (:91) when non-zero:
   * UFE_PHYS_INPUT (0): A_flow, A_flow_Ac, A_flow_mean, A_flow_mean_Ac are inputs;
   * UFE_PHYS_ARRHENIUS (1): A from Ti by the Arrhenius law, Ti_pmp, Cpi, Ki (:370-460);
   * UFE_PHYS_BENCHMARK_FIXED (2): the fixed values of the EISMINT/Halfar/Bueler
   * benchmarks (:323-348). */
  int32_t physical_properties;
  /* UFE_MIXED or UFE_FP64_KRYLOV: the preconditioner of the GMRES-IR linear solves,
   * UFE_PRECOND_SOR (SSA_precond_sweeps SOR sweeps, default 5) or
   * UFE_PRECOND_MULTIGRID (aggregation multigrid, SSA_precond_sweeps SOR
   * smoothing sweeps before and after the coarse correction, default 2). */
  int32_t SSA_precond_type;
} ufe_params;

#define UFE_PHYS_INPUT 0
#define UFE_PHYS_ARRHENIUS 1
#define UFE_PHYS_BENCHMARK_FIXED 2

/* Primary-mesh geometry (src/mesh_help_functions_module.f90), the step the
 * reference takes after creating or reading a mesh before make_Ac_mesh
 * (src/mesh_creation_module.f90:1725-1731, src/restart_module.f90:89-96).
 * Host arrays in the reference layout; outputs may be NULL to skip them. */
typedef struct ufe_mesh_geometry_desc {
  int32_t nV, nTri, nC_mem;
  double xmin, xmax, ymin, ymax;
  double tol_dist;                  /* points closer than this coincide; the reference uses
                                       ((xmax-xmin)+(ymax-ymin)) * 1e-9 / 2 */
  const double *V;                  /* (nV,2) */
  const int32_t *nC, *C;            /* (nV), (nV,nC_mem) */
  const int32_t *niTri, *iTri;      /* (nV), (nV,nC_mem) */
  const int32_t *edge_index;        /* (nV) */
  int32_t *Tri;                     /* (nTri,3); update_triangle_circumcenter rotates the
                                       vertices of a triangle whose circumcentre comes out invalid */
  const int32_t *Tri_edge_index;    /* (nTri) */
  int32_t compute_Tricc;            /* 1: Tricc := update_triangle_circumcenter (:854-897);
                                       0: Tricc is an input (read from a restart file) */
  double *Tricc;                    /* (nTri,2) in/out */
  double *A;                        /* (nV)        find_Voronoi_cell_areas (:15-50, :219-487) */
  double *Cw;                       /* (nV,nC_mem) find_connection_widths (:101-172) */
  double *TriA;                     /* (nTri)      find_triangle_areas (:173-192) */
  double *R;                        /* (nV)        determine_mesh_resolution (:193-216) */
} ufe_mesh_geometry_desc;

typedef struct ufe_mesh_geometry_stats {
  double resolution_min, resolution_max;
  double area_error;                /* |1 - sum(A) / domain area|; the reference warns above 1e-4 (:41-45) */
  int32_t rotated_triangles;        /* circumcentre retries with rotated vertices */
  int32_t invalid_circumcentres;    /* still invalid after the retries (the reference prints an error) */
  int32_t voronoi_outside;          /* Voronoi vertices beyond tol_dist outside the domain (the reference stops) */
  int32_t connection_failures;      /* connections without a shared triangle, or with a single
                                       non-boundary one (the reference stops) */
} ufe_mesh_geometry_stats;

/* Remapping between two meshes, src/mesh_mapping_module.f90. */
#define UFE_REMAP_TRILIN 0              /* remap_trilin_* (:3845-3918) */
#define UFE_REMAP_NEAREST_NEIGHBOUR 1   /* remap_nearest_neighbour_* (:3919-3963) */
#define UFE_REMAP_CONS_1ST_ORDER 2      /* remap_cons_1st_order_* (:3964-4012) */
#define UFE_REMAP_CONS_2ND_ORDER 3      /* remap_cons_2nd_order_* (:4013-4124) */

typedef struct ufe_remap_stats {
  int32_t n_entries;                /* conservative entries (source cells per destination cell, summed) */
  int32_t n_fallback;               /* destination vertices given triangle weights by the conservativity check */
  int32_t trilin_missing;           /* destination vertices without a containing source triangle */
  double seconds;
} ufe_remap_stats;

/* Statistics of one solve_SSA call. */
typedef struct ufe_ssa_stats {
  int32_t picard_iterations;
  int32_t picard_converged;
  int32_t resets;
  int32_t total_sweeps;
  double last_RN;
  double last_max_residual;
  int32_t sweeps[64];               /* SOR sweeps of the first 64 Picard iterations */
  int32_t status;
  double sor_seconds;               /* device time spent in SOR sweeps */
  double total_seconds;
} ufe_ssa_stats;

/* Options of one linearised solve to a relative residual.  The reference
 * stops its SOR on max|dU| < SSA_max_residual_UV (src/ice_dynamics_module.f90:
 * 651-676); this solve instead stops on the FP64 relative residual
 * ||b - A w||_2 / ||b||_2 <= rel_tol, checked every check_interval sweeps.
 * UFE_FP64: FP64 sweeps.  UFE_MIXED: FP64 iterative refinement around FP32
 * sweeps (the correction equation A d = b - A w is swept in FP32, w += d in
 * FP64). */
typedef struct ufe_linsolve_opts {
  double rel_tol;
  int32_t max_sweeps;
  int32_t precision;                /* UFE_FP64 or UFE_MIXED */
  double omega;                     /* over-relaxation; <= 0: params.SSA_SOR_omega */
  int32_t check_interval;           /* plain SOR: sweeps between FP64 residual evaluations (<= 0: 16) */
  int32_t use_graph;                /* reserved, must be 0 */
  int32_t gmres_restart;            /* 0: plain SOR sweeps; m > 0: restarted GMRES(m) preconditioned
                                       by precond_sweeps SOR sweeps, one FP64 refinement per cycle */
  int32_t precond_sweeps;           /* SOR sweeps per preconditioner application (<= 0: 2) */
  int32_t orth_passes;              /* classical Gram-Schmidt passes, 1 or 2 (<= 0: 2) */
  int32_t method;                   /* 0: GMRES(gmres_restart) (plain SOR if that is 0);
                                       1: BiCGStab preconditioned by precond_sweeps SOR sweeps,
                                       scalars kept on the device, FP64 refinement around it */
  int32_t precond_full_sweeps;      /* 0: successive preconditioner sweeps alternate direction and
                                       skip the colour the previous sweep ended on; 1: full sweeps */
  int32_t precond_neumann_every_sweep; /* 0: Neumann update of the edge rows after the last sweep of a
                                       preconditioner application only (the vector handed to the
                                       Krylov method satisfies it); 1: after every sweep */
  int32_t precond_type;             /* UFE_PRECOND_SOR: precond_sweeps SOR sweeps from zero;
                                       UFE_PRECOND_MULTIGRID (a Krylov method): one aggregation-
                                       multigrid K-cycle in FP32, precond_sweeps (<= 0: 2) FP32 SOR
                                       sweeps before and after the coarse correction, around FP32
                                       (UFE_MIXED) or FP64 Krylov vectors; set up on the first use
                                       per mesh (ufe_multigrid_info) */
} ufe_linsolve_opts;

/* Statistics of one linearised solve. */
typedef struct ufe_linsolve_stats {
  int32_t sweeps;                   /* SOR sweeps performed (FP32 sweeps in mixed mode), counted as
                                       colour passes / 5 */
  int32_t refinements;              /* FP64 residual evaluations */
  int32_t matvecs;                  /* operator applications outside the SOR sweeps (GMRES) */
  double rel_residual;              /* ||b - A w||_2 / ||b||_2 in FP64 at exit */
  double seconds;                   /* wall time of the solve on the host */
  int32_t status;
} ufe_linsolve_stats;

/* Live kernel timing: start/stop events attached to the dispatch of every
 * 8th SOR colour-kernel launch (hipExtLaunchKernelGGL), i.e. the kernel's own
 * execution time on the library's stream. */
typedef struct ufe_kernel_profile {
  int64_t launches;                 /* colour-kernel launches timed (the sample) */
  double seconds;                   /* summed event time of those launches */
  double algorithmic_bytes;         /* summed algorithmic bytes of those launches */
  int64_t rows;                     /* summed rows (vertex-sweeps) of those launches */
  int64_t total_launches;           /* colour-kernel launches while profiling was on */
} ufe_kernel_profile;

typedef struct ufe_ctx ufe_ctx;

/* ---- context ------------------------------------------------------------ */
int ufe_ctx_create(int device, ufe_ctx **out);
int ufe_ctx_destroy(ufe_ctx *ctx);
int ufe_set_params(ufe_ctx *ctx, const ufe_params *p);
const char *ufe_last_error(ufe_ctx *ctx);
int ufe_version(void);

/* ---- mesh set-up -------------------------------------------------------- */
/* Secondary mesh data from the primary arrays, in the order of
 * src/restart_module.f90:89-102 / src/mesh_creation_module.f90:1725-1737:
 * make_Ac_mesh (src/mesh_ArakawaC_module.f90:24-285), make_combined_AaAc_mesh
 * (:286-540), get_neighbour_functions (src/mesh_derivatives_module.f90:14-75)
 * and calculate_five_colouring_AaAc (src/mesh_five_colour_module.f90:18-137).
 * Neighbour functions are computed on the device of ctx. */
int ufe_mesh_build(ufe_ctx *ctx, ufe_mesh_desc *m);
/* Upload a complete mesh (primary + secondary) and derive device layouts. */
int ufe_mesh_upload(ufe_ctx *ctx, const ufe_mesh_desc *m);
/* create_remapping_arrays (src/mesh_mapping_module.f90:673-697): trilinear,
 * nearest-neighbour and conservative (1st and 2nd order) remapping arrays from
 * the mesh of src to the mesh of dst, both uploaded on the same device.
 * gs and gd describe the two meshes with their geometry (Tricc and A as
 * ufe_mesh_geometry computes them).  The conservative weights are the
 * reference's line integrals over the overlap of each destination cell with
 * each source cell, without the domain boundary (:2348-2415); destination
 * edge vertices get zero weights, and destination vertices that fail the
 * conservativity check take triangle weights (:2416-2566).  The arrays live
 * in dst until its next mesh upload or remapping build.  UFE_BAD_INPUT where
 * the reference stops (more than 1 % of the vertices fail the check). */
int ufe_remap_build(ufe_ctx *dst, ufe_ctx *src, const ufe_mesh_geometry_desc *gs,
                    const ufe_mesh_geometry_desc *gd, ufe_remap_stats *st);
/* remap_field_dp / remap_field_dp_3D (:4125-4228) on device buffers:
 * d_src (nV_src x nz) -> d_dst (nV_dst x nz), method UFE_REMAP_*. */
int ufe_remap_field(ufe_ctx *dst, ufe_ctx *src, int method, const double *d_src, double *d_dst, int nz);
/* remap_ice_model (src/ice_dynamics_module.f90:1198-1394) for the fields of
 * this library: Hi, Hi_prev and Ti conservatively (1st order) from the fields
 * of src into those of dst; the SSA velocities of dst start from zero. */
int ufe_remap_ice_model(ufe_ctx *dst, ufe_ctx *src);
/* Copies the remapping arrays of dst to host memory (NULL skips an array):
 * nn (nV_dst), tvi/tw (nV_dst x 3), start/end (nV_dst, 0-based, end
 * exclusive) and vi (1-based)/w0/w1x/w1y (*n_entries). */
int ufe_remap_export(ufe_ctx *dst, int *n_entries, int32_t *nn, int32_t *tvi, double *tw, int32_t *start,
                     int32_t *end, int32_t *vi, double *w0, double *w1x, double *w1y);
/* Tricc, A, Cw, TriA and R of a primary mesh on the GPU
 * (src/mesh_help_functions_module.f90:15-216, 219-487, 854-897).  Returns
 * UFE_BAD_INPUT where the reference stops or reports an invalid mesh. */
int ufe_mesh_geometry(ufe_ctx *ctx, ufe_mesh_geometry_desc *g, ufe_mesh_geometry_stats *st);
/* Sizes of the uploaded mesh: nV, nAc, nVAaAc, nC_mem. */
int ufe_mesh_sizes(ufe_ctx *ctx, int32_t *out4);

/* ---- fields ------------------------------------------------------------- */
/* Ice-model fields (type_ice_model, src/data_types_module.f90:15-214) live on
 * the device.  They are addressed by id; ids, names and grids can be listed.
 * grid codes: 0 Aa(nV) 1 Aa x nZ 2 Aa x nC_mem 3 Ac(nAc) 4 Ac x nZ 5 AaAc(nVAaAc)
 * dtype codes: 0 float64 1 int32 */
int ufe_field_count(void);
const char *ufe_field_name(int id);
int ufe_field_id(const char *name);
int ufe_field_info(int id, int32_t *grid, int32_t *dtype);
int ufe_field_set(ufe_ctx *ctx, int id, const void *host);   /* reference layout */
int ufe_field_get(ufe_ctx *ctx, int id, void *host);
int ufe_field_set_device(ufe_ctx *ctx, int id, const void *dev);
int ufe_field_get_device(ufe_ctx *ctx, int id, void *dev);
int ufe_field_fill(ufe_ctx *ctx, int id, double value);     /* every element := value */

/* ---- the hot path (operate on the device-resident fields) --------------- */
/* update_general_ice_model_data, src/general_ice_model_data_module.f90:23-96;
 * it ends with ice_physical_properties when ufe_params.physical_properties is
 * not UFE_PHYS_INPUT (otherwise the A_flow fields are inputs). */
int ufe_update_general_ice_model_data(ufe_ctx *ctx);
/* ice_physical_properties, src/general_ice_model_data_module.f90:298-462, on
 * its own: mode UFE_PHYS_ARRHENIUS maps Ti to Ti_Ac (:40-43) and sets Ti_pmp,
 * A_flow, Cpi, Ki, A_flow_mean, A_flow_Ac and A_flow_mean_Ac from Ti and the
 * sheet masks; UFE_PHYS_BENCHMARK_FIXED sets the benchmark constants. */
int ufe_ice_physical_properties(ufe_ctx *ctx, int mode);
/* solve_SIA, src/ice_dynamics_module.f90:240-314 */
int ufe_solve_SIA(ufe_ctx *ctx);
/* solve_SSA, src/ice_dynamics_module.f90:408-557 (with basal_yield_stress
 * :780-844, SSA_effective_viscosity :695-726, SSA_sliding_term :727-779 and
 * solve_SSA_linearised :558-694) */
int ufe_solve_SSA(ufe_ctx *ctx, ufe_ssa_stats *st);
/* calculate_ice_thickness_change, src/ice_dynamics_module.f90:31-237.
 * SMB_year, BMB and mask_noice are the fields of the same names. */
int ufe_calculate_ice_thickness_change(ufe_ctx *ctx, double dt);
/* solve_SIA_3D, src/ice_dynamics_module.f90:317-405: fields U_3D, V_3D and
 * W_3D (Aa x nZ) from U_SSA, V_SSA, A_flow and the geometry of
 * update_general_ice_model_data (Hi, dHs_dx/dy, dHi_dx/dy, dHb_dt, masks). */
int ufe_solve_SIA_3D(ufe_ctx *ctx);
/* The critical time steps of determine_timesteps_and_actions
 * (src/UFEMISM_main_model.f90:734-780) from D_SIA_Ac, U_SSA, V_SSA, U_3D,
 * V_3D and the Voronoi areas, after its 0.9 correction factor:
 * dt_crit[3] = {dt_D_2D_min, dt_V_2D_SSA_min, dt_V_3D_SIA_min}. */
int ufe_critical_timesteps(ufe_ctx *ctx, double *dt_crit);

/* ---- linearised SSA at fixed viscosity (bench / mixed precision) -------- */
/* Recompute eta, N, S and the SOR coefficients from the current velocities
 * (one Picard step without the solve). */
int ufe_ssa_prepare(ufe_ctx *ctx);
/* Solve the linearised system from the current velocities. */
int ufe_ssa_linear_solve(ufe_ctx *ctx, const ufe_linsolve_opts *opts, ufe_linsolve_stats *st);

/* The aggregation multigrid hierarchy of the uploaded mesh (built by the first
 * UFE_PRECOND_MULTIGRID solve; levels = 0 before).  No counterpart in the
 * reference, whose linearised SSA is solved by SOR alone. */
typedef struct ufe_multigrid_info {
  int32_t levels;                   /* including the fine level */
  int32_t points[16];               /* points per level; [0]: the SOR rows */
  int64_t couplings[16];            /* off-diagonal couplings per coarse level */
  int32_t direct;                   /* the coarsest level is solved by a dense inverse */
  int32_t kstep[16];                /* level solved by two cycles (K-cycle) */
  double setup_seconds;
  int32_t tail_level;               /* first of the small levels solved in one workgroup per visit
                                       (0: none; UFE_AMG_TAIL=0 in the environment at set-up turns
                                       it off) */
  int32_t fused[16];                /* halo depth of a level whose cycle steps run as two kernels
                                       with block halos, for up to that many smoothing sweeps
                                       (0: none; the transposed hierarchy too when the level's
                                       pattern is symmetric; UFE_AMG_FUSE=0 at set-up turns them
                                       off) */
  int32_t fuse_list_max[16];        /* its longest block list (own points and halo) */
} ufe_multigrid_info;
int ufe_multigrid_info_get(ufe_ctx *ctx, ufe_multigrid_info *out);
/* FP64 relative residual of the current velocities. */
int ufe_ssa_residual(ufe_ctx *ctx, double *rel_residual);

/* No counterpart in the reference: an allocation-free variant for callers
 * that cannot allocate device memory (serving loops, harnesses that forbid
 * hipMalloc).  It is the plain FP64 SOR path of ufe_ssa_linear_solve
 * (precision UFE_FP64, gmres_restart 0, method 0: the reference's sweeps of
 * src/ice_dynamics_module.f90:601-691, stopped on the FP64 relative residual
 * every check_interval sweeps) with the same sweeps, velocities and
 * statistics.  Its residual partial sums live in a file-scope device array of
 * the library (512 KiB on each GPU it runs on, kept until the process ends)
 * instead of context buffers; it calls no allocating routine.  Meshes with
 * more than 8,388,608 SOR rows (nVAaAc) are refused with UFE_BAD_INPUT.  All
 * calls on one GPU share that array: calls that overlap on one GPU (from
 * several contexts or threads) race.  omega <= 0: params.SSA_SOR_omega;
 * check_interval <= 0: 16. */
int ufe_ssa_sor_solve_noalloc(ufe_ctx *ctx, double rel_tol, int32_t max_sweeps, double omega,
                              int32_t check_interval, ufe_linsolve_stats *st);

/* ---- adjoint of the SSA solve ------------------------------------------- */
/* No counterpart in the reference.  The forward runs exactly
 * picard_iterations Picard steps of solve_SSA (src/ice_dynamics_module.f90:
 * 498-545: viscosity :695-726, sliding :727-779, linearised system :558-694)
 * from the current velocities, solving each linearised system to the FP64
 * relative residual lin_rel_tol (GMRES-IR, as ufe_ssa_linear_solve) instead
 * of the reference's SOR stopping rule, and stores every checkpoint_every-th
 * Picard iterate.  The backward takes dJ/dU and dJ/dV of the final AaAc
 * velocities and returns dJ/dlog(tau_c) and dJ/dA on the AaAc points, the
 * exact gradient of the K-step map with exact linear solves: per step one
 * transposed solve (FP64 GMRES, transposed-SOR preconditioner) and the chain
 * through the sliding term, the right-hand side and the effective viscosity.
 * Iterates between checkpoints are recomputed in the reverse sweep.
 * With params.use_analytical_GL_flux the forward prescribes the grounding-line
 * Ac velocities once (calculate_GL_flux, :848-949) and holds them in every
 * step, as solve_SSA does; the backward keeps those rows out of the transposed
 * solves and adds their contribution to dJ/dA (the prescribed velocity is
 * proportional to A at the grounded end of each grounding-line Ac point).
 * All arrays of the backward are device pointers of nVAaAc doubles in the
 * reference AaAc order (Aa points 1..nV, then Ac points). */
typedef struct ufe_adjoint_opts {
  double lin_rel_tol;               /* relative residual of every forward and adjoint linear solve */
  int32_t picard_iterations;        /* K >= 1 Picard steps (no early exit) */
  int32_t checkpoint_every;         /* store u_0, u_c, u_2c, ...; others recomputed (>= 1) */
  int32_t precision;                /* forward linear solves: UFE_FP64 or UFE_MIXED; adjoint solves FP64 */
  int32_t gmres_restart;            /* GMRES(m), 1..64, forward and adjoint */
  int32_t precond_sweeps;           /* SOR sweeps per preconditioner application */
  double omega;                     /* <= 0: params.SSA_SOR_omega */
  int32_t max_sweeps;               /* per linear solve */
  int32_t precond_type;             /* UFE_PRECOND_SOR, or UFE_PRECOND_MULTIGRID for the forward solves
                                       and the adjoint solves (the transposed hierarchy: every coarse
                                       operator transposed, FP64 transposed SOR smoothing on the fine
                                       level); precond_sweeps then counts smoothing sweeps (<= 0: 2) */
} ufe_adjoint_opts;

typedef struct ufe_adjoint_stats {
  int32_t picard_iterations;
  int32_t checkpoints;              /* iterates stored */
  int32_t forward_sweeps;           /* SOR sweeps of the forward linear solves */
  int32_t recompute_sweeps;         /* SOR sweeps spent recomputing iterates in the backward */
  int32_t adjoint_sweeps;           /* transposed SOR sweeps of the adjoint solves */
  int32_t adjoint_matvecs;
  double max_forward_residual;      /* worst final relative residual of the forward solves */
  double max_adjoint_residual;      /* worst final relative residual of the adjoint solves */
  double forward_seconds, backward_seconds;
  int32_t status;
} ufe_adjoint_stats;

int ufe_ssa_adjoint_forward(ufe_ctx *ctx, const ufe_adjoint_opts *opts, ufe_adjoint_stats *st);
int ufe_ssa_adjoint_backward(ufe_ctx *ctx, const double *dJ_dU, const double *dJ_dV,
                             double *grad_log_tau_c, double *grad_A, ufe_adjoint_stats *st);

/* ---- autotuned launch shape of the FP32 sweep ----------------------------- */
/* No counterpart in the reference.  The FP32 colour pass has 32 launch
 * shapes: threads per block 64/128/256/512, block order (schedule 0: XCD-major
 * bands of rows, 1: round-robin) and load variant (0: plain, 1: non-temporal
 * loads for the once-streamed coefficients, 2: as 1 with eight slots' loads in
 * flight, 3: plain loads, all six slots' loads of a six-slot row in one batch).
 * All give identical results.  The autotuner times `sweeps` sweeps
 * of the uploaded mesh (single or batch) per shape and keeps the fastest for
 * every later FP32 sweep of the context (default: 256, 0, 0). */
typedef struct ufe_sweep_tuning {
  int32_t threads_per_block;
  int32_t schedule;
  int32_t loads;
  int32_t n_candidates;
  int32_t cand_threads[32];
  int32_t cand_schedule[32];
  int32_t cand_loads[32];
  double cand_us_per_sweep[32];
} ufe_sweep_tuning;

int ufe_sweep_autotune(ufe_ctx *ctx, int sweeps, ufe_sweep_tuning *out);
int ufe_sweep_tuning_set(ufe_ctx *ctx, int threads_per_block, int schedule, int loads);

/* ---- batched ensembles --------------------------------------------------- */
/* No counterpart in the reference.  Many independent SSA systems solved
 * together on one GPU: the caller joins the member meshes into one mesh whose
 * Aa vertices and Ac points come in member-major blocks (member 0's vertices,
 * member 1's, ..., then member 0's Ac points, ...; connectivity, colour lists
 * and index arrays renumbered accordingly; every member keeps its own corner
 * vertices, edge_index 2/4/6/8), declares the member sizes with
 * ufe_batch_define BEFORE ufe_mesh_upload, and uploads the union.  Each
 * colour's SOR rows of all members then form one range, so one launch sweeps
 * the colour in every member; reductions, Krylov coefficients and
 * convergence are per member.  nmembers = 0 clears the declaration. */
typedef struct ufe_batch_opts {
  double rel_tol;                   /* per-member FP64 relative residual */
  int32_t precision;                /* UFE_MIXED (FP32 sweeps and Krylov vectors) or UFE_FP64 */
  int32_t gmres_restart;            /* GMRES(m), 1..64, every member */
  int32_t precond_sweeps;           /* SOR sweeps per preconditioner application (<= 0: 4) */
  double omega;                     /* <= 0: params.SSA_SOR_omega */
  int32_t max_cycles;               /* GMRES cycles (all members in lockstep; <= 0: unbounded) */
  int32_t use_graph;                /* 1: each cycle of all members is one hipGraph launch */
  int32_t precond_type;             /* UFE_PRECOND_SOR, or UFE_PRECOND_MULTIGRID (UFE_MIXED; the union
                                       of the members is aggregated as one mesh, precond_sweeps
                                       smoothing sweeps, default 2) */
} ufe_batch_opts;

typedef struct ufe_batch_stats {
  int32_t cycles;                   /* GMRES cycles run */
  int32_t sweeps;                   /* SOR sweeps per member (colour passes / 5) */
  int32_t converged;                /* members at the tolerance */
  double max_rel_residual;          /* worst member at exit */
  double seconds;
  int32_t status;                   /* UFE_OK or UFE_SOR_NOT_CONVERGED */
} ufe_batch_stats;

int ufe_batch_define(ufe_ctx *ctx, int nmembers, const int32_t *nV, const int32_t *nAc);
/* Solve every member's current linearised system (after ufe_ssa_prepare) from
 * the current velocities; member_rel (nmembers doubles, may be NULL) receives
 * each member's final relative residual. */
int ufe_batch_linear_solve(ufe_ctx *ctx, const ufe_batch_opts *opts, double *member_rel, ufe_batch_stats *st);
/* picard_iterations Picard steps of solve_SSA (src/ice_dynamics_module.f90:
 * 498-545) for every member, each linearised system solved as above. */
int ufe_batch_solve_SSA(ufe_ctx *ctx, int picard_iterations, const ufe_batch_opts *opts, double *member_rel,
                        ufe_batch_stats *st);

/* ---- operators (device pointers, original index order) ----------------- */
/* src/mesh_derivatives_module.f90:315-432, 485-597 */
int ufe_op_derivatives_Aa(ufe_ctx *ctx, const double *d, double *ddx, double *ddy);
int ufe_op_curvatures_Aa(ufe_ctx *ctx, const double *d, double *ddxx, double *ddxy, double *ddyy);
int ufe_op_derivatives_Aa_3D(ufe_ctx *ctx, const double *d, double *ddx, double *ddy, int nz);
int ufe_op_neumann_Aa(ufe_ctx *ctx, double *d);
int ufe_op_neumann_Aa_3D(ufe_ctx *ctx, double *d, int nz);
/* src/mesh_ArakawaC_module.f90:542-844 */
/* get_upwind_derivative_vertex_3D, src/mesh_derivatives_module.f90:434-483, on
 * every vertex and layer of d (nV x nz): the derivative on the triangle
 * upwind of (U,V), or the vertex derivative where |U|,|V| < 1e-10.  *missing
 * counts the pairs without an upwind triangle (the reference prints an error
 * for each); their derivatives are zero. */
int ufe_op_upwind_derivatives_3D(ufe_ctx *ctx, const double *U, const double *V, const double *d, double *ddx,
                                 double *ddy, int nz, int *missing);
int ufe_op_derivatives_Ac(ufe_ctx *ctx, const double *d, double *ddx, double *ddy, double *ddp,
                          double *ddo);
int ufe_op_derivatives_AaAc(ufe_ctx *ctx, const double *d, double *ddx, double *ddy);
int ufe_op_curvatures_AaAc(ufe_ctx *ctx, const double *d, double *ddxx, double *ddxy,
                           double *ddyy);
int ufe_op_neumann_AaAc(ufe_ctx *ctx, double *d);
int ufe_op_map_Aa_to_Ac(ufe_ctx *ctx, const double *d_Aa, double *d_Ac);
int ufe_op_map_Aa_to_Ac_3D(ufe_ctx *ctx, const double *d_Aa, double *d_Ac, int nz);
int ufe_op_map_Ac_to_Aa(ufe_ctx *ctx, const double *d_Ac, double *d_Aa);
int ufe_op_map_Ac_to_Aa_3D(ufe_ctx *ctx, const double *d_Ac, double *d_Aa, int nz);
int ufe_op_rotate_xy_to_po(ufe_ctx *ctx, const double *x, const double *y, double *p, double *o);

/* ---- measurement -------------------------------------------------------- */
/* Enable/disable event timing of the SOR colour kernels; reading resets. */
int ufe_profile_enable(ufe_ctx *ctx, int enable);
int ufe_profile_read(ufe_ctx *ctx, ufe_kernel_profile *out);
/* Algorithmic bytes of one FP64 / FP32 sweep over all SOR rows, and the
 * mean AaAc degree of those rows: B(w) = 4 (n+1) + w (2 n + 9) per row. */
int ufe_sweep_bytes(ufe_ctx *ctx, double *fp64_bytes, double *fp32_bytes, double *mean_degree,
                    int64_t *rows);

/* ---- device helpers for callers that manage their own buffers ----------- */
int ufe_device_alloc(ufe_ctx *ctx, size_t bytes, void **dev);
int ufe_device_free(ufe_ctx *ctx, void *dev);
int ufe_memcpy_h2d(ufe_ctx *ctx, void *dev, const void *host, size_t bytes);
int ufe_memcpy_d2h(ufe_ctx *ctx, void *host, const void *dev, size_t bytes);
int ufe_synchronize(ufe_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
