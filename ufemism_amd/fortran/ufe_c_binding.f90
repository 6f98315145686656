! ufe_c_binding.f90 -- ISO_C_BINDING interfaces of libufe_dyn (include/ufe_dyn.h).
!
! The derived types below mirror the C structs field for field (BIND(C)), and
! each interface binds one C-ABI entry point by name.  Arrays cross the
! boundary as C_LOC of contiguous Fortran arrays, in the reference layout.
MODULE ufe_c_binding

  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: UFE_OK = 0, UFE_SOR_NOT_CONVERGED = 1, UFE_SSA_RESET = 2, &
                               UFE_SSA_UNSTABLE = 3, UFE_COLOURING_IDENTIFY = 4, &
                               UFE_COLOURING_FAILED = 5, UFE_BAD_INPUT = 6, UFE_HIP_ERROR = 7, &
                               UFE_NOT_SUPPORTED = 8, UFE_HOST_ERROR = 9
  INTEGER(C_INT), PARAMETER :: UFE_FP64 = 0, UFE_MIXED = 1, UFE_FP64_KRYLOV = 2

  TYPE, BIND(C) :: ufe_mesh_desc
    INTEGER(C_INT32_T) :: nV, nTri, nC_mem, nAc
    REAL(C_DOUBLE)     :: xmin, xmax, ymin, ymax
    TYPE(C_PTR) :: V, nC, C, niTri, iTri, edge_index, Tri, A, Cw
    TYPE(C_PTR) :: Nx, Ny, Nxx, Nxy, Nyy
    TYPE(C_PTR) :: iAci, Aci, VAc, Nx_Ac, Ny_Ac, No_Ac, Np_Ac, edge_index_Ac
    TYPE(C_PTR) :: VAaAc, nCAaAc, CAaAc, Nx_AaAc, Ny_AaAc, Nxx_AaAc, Nxy_AaAc, Nyy_AaAc
    TYPE(C_PTR) :: colour, colour_vi
    INTEGER(C_INT32_T) :: colour_nV(5)
  END TYPE ufe_mesh_desc

  TYPE, BIND(C) :: ufe_mesh_geometry_desc
    INTEGER(C_INT32_T) :: nV, nTri, nC_mem
    REAL(C_DOUBLE)     :: xmin, xmax, ymin, ymax, tol_dist
    TYPE(C_PTR) :: V, nC, C, niTri, iTri, edge_index, Tri, Tri_edge_index
    INTEGER(C_INT32_T) :: compute_Tricc
    TYPE(C_PTR) :: Tricc, A, Cw, TriA, R
  END TYPE ufe_mesh_geometry_desc

  TYPE, BIND(C) :: ufe_mesh_geometry_stats
    REAL(C_DOUBLE)     :: resolution_min, resolution_max, area_error
    INTEGER(C_INT32_T) :: rotated_triangles, invalid_circumcentres, voronoi_outside, connection_failures
  END TYPE ufe_mesh_geometry_stats

  TYPE, BIND(C) :: ufe_params
    INTEGER(C_INT32_T) :: nZ
    REAL(C_DOUBLE)     :: zeta(32)
    REAL(C_DOUBLE)     :: m_enh_sia, m_enh_ssa
    REAL(C_DOUBLE)     :: SSA_RN_tol
    INTEGER(C_INT32_T) :: SSA_max_outer_loops
    REAL(C_DOUBLE)     :: SSA_max_residual_UV
    REAL(C_DOUBLE)     :: SSA_SOR_omega
    INTEGER(C_INT32_T) :: SSA_max_inner_loops
    INTEGER(C_INT32_T) :: use_analytical_GL_flux, ssa_forced_zero, zero_Hi_at_edges
    INTEGER(C_INT32_T) :: use_tau_c_override, precision, sweep_batch
    REAL(C_DOUBLE)     :: SSA_linear_rel_tol
    INTEGER(C_INT32_T) :: SSA_gmres_restart, SSA_precond_sweeps
    INTEGER(C_INT32_T) :: physical_properties
    INTEGER(C_INT32_T) :: SSA_precond_type
  END TYPE ufe_params

  TYPE, BIND(C) :: ufe_ssa_stats
    INTEGER(C_INT32_T) :: picard_iterations, picard_converged, resets, total_sweeps
    REAL(C_DOUBLE)     :: last_RN, last_max_residual
    INTEGER(C_INT32_T) :: sweeps(64)
    INTEGER(C_INT32_T) :: status
    REAL(C_DOUBLE)     :: sor_seconds, total_seconds
  END TYPE ufe_ssa_stats

  INTERFACE
    FUNCTION ufe_ctx_create(device, ctx) BIND(C, name="ufe_ctx_create") RESULT(r)
      IMPORT :: C_INT, C_PTR
      INTEGER(C_INT), VALUE :: device
      TYPE(C_PTR) :: ctx
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_ctx_destroy(ctx) BIND(C, name="ufe_ctx_destroy") RESULT(r)
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_set_params(ctx, p) BIND(C, name="ufe_set_params") RESULT(r)
      IMPORT :: C_INT, C_PTR, ufe_params
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufe_params) :: p
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_last_error(ctx) BIND(C, name="ufe_last_error") RESULT(s)
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
      TYPE(C_PTR) :: s
    END FUNCTION
    FUNCTION ufe_mesh_build(ctx, m) BIND(C, name="ufe_mesh_build") RESULT(r)
      IMPORT :: C_INT, C_PTR, ufe_mesh_desc
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufe_mesh_desc) :: m
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_mesh_geometry(ctx, g, st) BIND(C, name="ufe_mesh_geometry") RESULT(r)
      IMPORT :: C_INT, C_PTR, ufe_mesh_geometry_desc, ufe_mesh_geometry_stats
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufe_mesh_geometry_desc) :: g
      TYPE(ufe_mesh_geometry_stats) :: st
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_mesh_upload(ctx, m) BIND(C, name="ufe_mesh_upload") RESULT(r)
      IMPORT :: C_INT, C_PTR, ufe_mesh_desc
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufe_mesh_desc) :: m
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_field_id(name) BIND(C, name="ufe_field_id") RESULT(r)
      IMPORT :: C_INT, C_CHAR
      CHARACTER(KIND=C_CHAR), DIMENSION(*) :: name
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_field_set(ctx, id, host) BIND(C, name="ufe_field_set") RESULT(r)
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, host
      INTEGER(C_INT), VALUE :: id
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_field_get(ctx, id, host) BIND(C, name="ufe_field_get") RESULT(r)
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, host
      INTEGER(C_INT), VALUE :: id
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_update_general_ice_model_data(ctx) BIND(C, name="ufe_update_general_ice_model_data") RESULT(r)
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_ice_physical_properties(ctx, mode) BIND(C, name="ufe_ice_physical_properties") RESULT(r)
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: mode
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_solve_SIA(ctx) BIND(C, name="ufe_solve_SIA") RESULT(r)
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_solve_SIA_3D(ctx) BIND(C, name="ufe_solve_SIA_3D") RESULT(r)
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_critical_timesteps(ctx, dt_crit) BIND(C, name="ufe_critical_timesteps") RESULT(r)
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE) :: dt_crit(3)
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_solve_SSA(ctx, st) BIND(C, name="ufe_solve_SSA") RESULT(r)
      IMPORT :: C_INT, C_PTR, ufe_ssa_stats
      TYPE(C_PTR), VALUE :: ctx
      TYPE(ufe_ssa_stats) :: st
      INTEGER(C_INT) :: r
    END FUNCTION
    FUNCTION ufe_calculate_ice_thickness_change(ctx, dt) BIND(C, name="ufe_calculate_ice_thickness_change") RESULT(r)
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt
      INTEGER(C_INT) :: r
    END FUNCTION
  END INTERFACE

CONTAINS

  ! NUL-terminated copy of a Fortran string for the C side
  FUNCTION c_string(s) RESULT(cs)
    CHARACTER(LEN=*), INTENT(IN) :: s
    CHARACTER(KIND=C_CHAR), DIMENSION(LEN_TRIM(s) + 1) :: cs
    INTEGER :: i
    DO i = 1, LEN_TRIM(s)
      cs(i) = s(i:i)
    END DO
    cs(LEN_TRIM(s) + 1) = C_NULL_CHAR
  END FUNCTION c_string

  ! Fortran copy of a C string (the library's last error message)
  FUNCTION f_string(p) RESULT(s)
    TYPE(C_PTR), INTENT(IN) :: p
    CHARACTER(LEN=512) :: s
    CHARACTER(KIND=C_CHAR), POINTER :: chars(:)
    INTEGER :: i
    s = ''
    IF (.NOT. C_ASSOCIATED(p)) RETURN
    CALL C_F_POINTER(p, chars, [512])
    DO i = 1, 512
      IF (chars(i) == C_NULL_CHAR) EXIT
      s(i:i) = chars(i)
    END DO
  END FUNCTION f_string

END MODULE ufe_c_binding
