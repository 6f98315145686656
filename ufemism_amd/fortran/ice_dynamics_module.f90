! ice_dynamics_module.f90 -- Fortran host of the MI355X ice-dynamics library.
!
! Keeps the call surface of the reference's ice_dynamics_module and
! general_ice_model_data_module:
!   update_general_ice_model_data( mesh, ice, time)          (src/general_ice_model_data_module.f90:23-30)
!   solve_SIA( mesh, ice)                                    (src/ice_dynamics_module.f90:240-249)
!   solve_SSA( mesh, ice)                                    (src/ice_dynamics_module.f90:408-417)
!   calculate_ice_thickness_change( mesh, ice, SMB, BMB, dt, mask_noice)  (src/ice_dynamics_module.f90:31-44)
! with minimal mirrors of type_mesh / type_ice_model (src/data_types_module.f90:15-345)
! in the reference layout (1-based, column-major).  Each routine copies the
! fields the reference routine reads to the device, runs the HIP kernels
! through the C-ABI and copies back what the reference routine writes.  With
! ufe_host_sync = .FALSE. the fields stay on the device between calls.
!
! Error handling follows the reference: conditions it aborts on stop the
! program (ERROR STOP), the SOR non-convergence warning is printed.
MODULE ice_dynamics_module

  USE, INTRINSIC :: ISO_C_BINDING
  USE ufe_c_binding
  IMPLICIT NONE

  INTEGER, PARAMETER :: dp = KIND(1.0D0)

  ! Configuration subset (src/configuration_module.f90:124-126, 169-184)
  TYPE type_dynamics_config
    INTEGER  :: nZ = 15
    REAL(dp) :: zeta(15) = [0.00_dp, 0.10_dp, 0.20_dp, 0.30_dp, 0.40_dp, 0.50_dp, 0.60_dp, 0.70_dp, &
                            0.80_dp, 0.90_dp, 0.925_dp, 0.95_dp, 0.975_dp, 0.99_dp, 1.00_dp]
    REAL(dp) :: m_enh_sia = 5.0_dp
    REAL(dp) :: m_enh_ssa = 0.7_dp
    REAL(dp) :: SSA_RN_tol = 1E-5_dp
    INTEGER  :: SSA_max_outer_loops = 50
    REAL(dp) :: SSA_max_residual_UV = 2.5_dp
    REAL(dp) :: SSA_SOR_omega = 1.2_dp
    INTEGER  :: SSA_max_inner_loops = 10000
    LOGICAL  :: use_analytical_GL_flux = .FALSE.
    LOGICAL  :: ssa_forced_zero = .FALSE.
    LOGICAL  :: zero_Hi_at_edges = .TRUE.
    ! linear solves of solve_SSA: UFE_FP64 runs the reference SOR; UFE_MIXED and
    ! UFE_FP64_KRYLOV solve each linearised system to SSA_linear_rel_tol by GMRES-IR
    INTEGER  :: precision = 0
    REAL(dp) :: SSA_linear_rel_tol = 1E-6_dp
    ! 0: the library defaults (40 and 5 with SOR sweeps, 20 and 2 with multigrid)
    INTEGER  :: SSA_gmres_restart = 0
    INTEGER  :: SSA_precond_sweeps = 0
    ! preconditioner of those solves: 0 SOR sweeps, 1 aggregation multigrid
    INTEGER  :: SSA_precond_type = 0
    ! ice_physical_properties at the end of update_general_ice_model_data:
    ! 0 the A_flow fields are inputs, 1 Arrhenius law from Ti, 2 fixed benchmark values
    INTEGER  :: physical_properties = 0
  END TYPE type_dynamics_config

  TYPE(type_dynamics_config), SAVE :: C
  LOGICAL, SAVE :: ufe_host_sync = .TRUE.

  TYPE type_mesh
    INTEGER  :: nV = 0, nTri = 0, nC_mem = 16, nAc = 0, nVAaAc = 0
    REAL(dp) :: xmin, xmax, ymin, ymax
    REAL(dp), ALLOCATABLE :: V(:,:), A(:), Cw(:,:)
    INTEGER,  ALLOCATABLE :: nC(:), C(:,:), niTri(:), iTri(:,:), edge_index(:), Tri(:,:)
    ! triangle data of the primary mesh; with Tri_edge_index present, initialise_mesh_dynamics
    ! computes A, Cw, TriA and R (and Tricc when absent) as the reference does
    INTEGER,  ALLOCATABLE :: Tri_edge_index(:)
    REAL(dp), ALLOCATABLE :: Tricc(:,:), TriA(:), R(:)
    REAL(dp) :: tol_dist = 0._dp, resolution_min = 0._dp, resolution_max = 0._dp
    ! secondary data (make_Ac_mesh, make_combined_AaAc_mesh, neighbour functions, colouring)
    REAL(dp), ALLOCATABLE :: Nx(:,:), Ny(:,:), Nxx(:,:), Nxy(:,:), Nyy(:,:)
    INTEGER,  ALLOCATABLE :: iAci(:,:), Aci(:,:), edge_index_Ac(:)
    REAL(dp), ALLOCATABLE :: VAc(:,:), Nx_Ac(:,:), Ny_Ac(:,:), No_Ac(:,:), Np_Ac(:)
    REAL(dp), ALLOCATABLE :: VAaAc(:,:), Nx_AaAc(:,:), Ny_AaAc(:,:), Nxx_AaAc(:,:), Nxy_AaAc(:,:), Nyy_AaAc(:,:)
    INTEGER,  ALLOCATABLE :: nCAaAc(:), CAaAc(:,:), colour(:), colour_vi(:,:)
    INTEGER :: colour_nV(5) = 0
    TYPE(C_PTR) :: ufe = C_NULL_PTR          ! device context holding this mesh
  END TYPE type_mesh

  TYPE type_ice_model
    ! Aa
    REAL(dp), ALLOCATABLE :: Hi(:), Hb(:), SL(:), Hs(:), dHb_dt(:), dHi_dt(:), dHs_dt(:), Hi_prev(:)
    REAL(dp), ALLOCATABLE :: A_flow(:,:), A_flow_mean(:)
    REAL(dp), ALLOCATABLE :: Ti(:,:), Ti_pmp(:,:), Cpi(:,:), Ki(:,:)
    REAL(dp), ALLOCATABLE :: dHi_dx(:), dHi_dy(:), dHs_dx(:), dHs_dy(:), dHs_dx_shelf(:), dHs_dy_shelf(:)
    REAL(dp), ALLOCATABLE :: U_SIA(:), V_SIA(:), D_SIA(:), U_SSA(:), V_SSA(:), dVi_in(:,:)
    REAL(dp), ALLOCATABLE :: U_3D(:,:), V_3D(:,:), W_3D(:,:)
    INTEGER,  ALLOCATABLE :: mask_land(:), mask_ocean(:), mask_lake(:), mask_ice(:), mask_sheet(:), &
                             mask_shelf(:), mask_coast(:), mask_margin(:), mask_gl(:), mask_cf(:), mask(:)
    ! Ac
    REAL(dp), ALLOCATABLE :: Hi_Ac(:), Hb_Ac(:), SL_Ac(:), Hs_Ac(:), A_flow_Ac(:,:), A_flow_mean_Ac(:)
    REAL(dp), ALLOCATABLE :: Ti_Ac(:,:)
    REAL(dp), ALLOCATABLE :: dHi_dx_Ac(:), dHi_dy_Ac(:), dHi_dp_Ac(:), dHi_do_Ac(:)
    REAL(dp), ALLOCATABLE :: dHb_dx_Ac(:), dHb_dy_Ac(:), dHb_dp_Ac(:), dHb_do_Ac(:)
    REAL(dp), ALLOCATABLE :: dHs_dx_Ac(:), dHs_dy_Ac(:), dHs_dp_Ac(:), dHs_do_Ac(:)
    REAL(dp), ALLOCATABLE :: dSL_dx_Ac(:), dSL_dy_Ac(:), dSL_dp_Ac(:), dSL_do_Ac(:)
    REAL(dp), ALLOCATABLE :: dHs_dx_shelf_Ac(:), dHs_dy_shelf_Ac(:)
    REAL(dp), ALLOCATABLE :: D_SIA_3D_Ac(:,:), D_SIA_Ac(:), Ux_SIA_Ac(:), Uy_SIA_Ac(:), Up_SIA_Ac(:), Uo_SIA_Ac(:)
    REAL(dp), ALLOCATABLE :: Ux_SSA_Ac(:), Uy_SSA_Ac(:), Up_SSA_Ac(:), Uo_SSA_Ac(:)
    INTEGER,  ALLOCATABLE :: mask_land_Ac(:), mask_ocean_Ac(:), mask_lake_Ac(:), mask_ice_Ac(:), &
                             mask_sheet_Ac(:), mask_shelf_Ac(:), mask_coast_Ac(:), mask_margin_Ac(:), &
                             mask_gl_Ac(:), mask_cf_Ac(:), mask_Ac(:)
    ! AaAc
    REAL(dp), ALLOCATABLE :: U_SSA_AaAc(:), V_SSA_AaAc(:), eta_AaAc(:), N_AaAc(:), S_AaAc(:), tau_c_AaAc(:)
    ! basal-friction override (no reference field): tau_c replaces the till yield stress when allocated
    REAL(dp), ALLOCATABLE :: tau_c_override_AaAc(:)
    ! solver statistics of the last solve_SSA
    INTEGER :: SSA_picard_iterations = 0
    INTEGER :: SSA_sweeps(64) = 0
  END TYPE type_ice_model

  TYPE type_SMB_model
    REAL(dp), ALLOCATABLE :: SMB_year(:)
  END TYPE type_SMB_model
  TYPE type_BMB_model
    REAL(dp), ALLOCATABLE :: BMB(:)
  END TYPE type_BMB_model

  INTERFACE put
    MODULE PROCEDURE put_r1, put_r2, put_i1
  END INTERFACE
  INTERFACE get
    MODULE PROCEDURE get_r1, get_r2, get_i1
  END INTERFACE

CONTAINS

  ! ---- boundary helpers ------------------------------------------------------

  SUBROUTINE check(mesh, r, what)
    TYPE(type_mesh), INTENT(IN) :: mesh
    INTEGER(C_INT),  INTENT(IN) :: r
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (r == UFE_OK) RETURN
    IF (r == UFE_SOR_NOT_CONVERGED) THEN
      WRITE(0,*) ' WARNING - SSA SOR solver doesnt converge!'
      RETURN
    END IF
    WRITE(0,*) TRIM(what), ' - ERROR ', r, ': ', TRIM(f_string(ufe_last_error(mesh%ufe)))
    ERROR STOP 1
  END SUBROUTINE check

  SUBROUTINE put_r1(mesh, name, a)
    TYPE(type_mesh), INTENT(IN) :: mesh
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(dp), TARGET, INTENT(IN) :: a(:)
    CALL check(mesh, ufe_field_set(mesh%ufe, ufe_field_id(c_string(name)), C_LOC(a)), 'put '//name)
  END SUBROUTINE put_r1
  SUBROUTINE put_r2(mesh, name, a)
    TYPE(type_mesh), INTENT(IN) :: mesh
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(dp), TARGET, INTENT(IN) :: a(:,:)
    CALL check(mesh, ufe_field_set(mesh%ufe, ufe_field_id(c_string(name)), C_LOC(a)), 'put '//name)
  END SUBROUTINE put_r2
  SUBROUTINE put_i1(mesh, name, a)
    TYPE(type_mesh), INTENT(IN) :: mesh
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, TARGET, INTENT(IN) :: a(:)
    CALL check(mesh, ufe_field_set(mesh%ufe, ufe_field_id(c_string(name)), C_LOC(a)), 'put '//name)
  END SUBROUTINE put_i1
  SUBROUTINE get_r1(mesh, name, a)
    TYPE(type_mesh), INTENT(IN) :: mesh
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(dp), TARGET, INTENT(INOUT) :: a(:)
    CALL check(mesh, ufe_field_get(mesh%ufe, ufe_field_id(c_string(name)), C_LOC(a)), 'get '//name)
  END SUBROUTINE get_r1
  SUBROUTINE get_r2(mesh, name, a)
    TYPE(type_mesh), INTENT(IN) :: mesh
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(dp), TARGET, INTENT(INOUT) :: a(:,:)
    CALL check(mesh, ufe_field_get(mesh%ufe, ufe_field_id(c_string(name)), C_LOC(a)), 'get '//name)
  END SUBROUTINE get_r2
  SUBROUTINE get_i1(mesh, name, a)
    TYPE(type_mesh), INTENT(IN) :: mesh
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, TARGET, INTENT(INOUT) :: a(:)
    CALL check(mesh, ufe_field_get(mesh%ufe, ufe_field_id(c_string(name)), C_LOC(a)), 'get '//name)
  END SUBROUTINE get_i1

  FUNCTION params_of(ncfg) RESULT(p)
    TYPE(type_dynamics_config), INTENT(IN) :: ncfg
    TYPE(ufe_params) :: p
    p%nZ = ncfg%nZ
    p%zeta = 0._dp
    p%zeta(1:ncfg%nZ) = ncfg%zeta(1:ncfg%nZ)
    p%m_enh_sia = ncfg%m_enh_sia
    p%m_enh_ssa = ncfg%m_enh_ssa
    p%SSA_RN_tol = ncfg%SSA_RN_tol
    p%SSA_max_outer_loops = ncfg%SSA_max_outer_loops
    p%SSA_max_residual_UV = ncfg%SSA_max_residual_UV
    p%SSA_SOR_omega = ncfg%SSA_SOR_omega
    p%SSA_max_inner_loops = ncfg%SSA_max_inner_loops
    p%use_analytical_GL_flux = MERGE(1, 0, ncfg%use_analytical_GL_flux)
    p%ssa_forced_zero = MERGE(1, 0, ncfg%ssa_forced_zero)
    p%zero_Hi_at_edges = MERGE(1, 0, ncfg%zero_Hi_at_edges)
    p%use_tau_c_override = 0
    p%precision = ncfg%precision
    p%sweep_batch = 0
    p%SSA_linear_rel_tol = ncfg%SSA_linear_rel_tol
    p%SSA_gmres_restart = ncfg%SSA_gmres_restart
    p%SSA_precond_sweeps = ncfg%SSA_precond_sweeps
    p%physical_properties = ncfg%physical_properties
    p%SSA_precond_type = ncfg%SSA_precond_type
  END FUNCTION params_of

  ! ---- mesh set-up -----------------------------------------------------------

  ! Secondary mesh data and the device copy of the mesh, the steps the
  ! reference takes after creating or reading a mesh
  ! (src/mesh_creation_module.f90:1725-1737, src/restart_module.f90:89-102).
  SUBROUTINE initialise_mesh_dynamics(mesh, device)
    TYPE(type_mesh), TARGET, INTENT(INOUT) :: mesh
    INTEGER, INTENT(IN), OPTIONAL :: device
    TYPE(ufe_mesh_desc) :: d
    TYPE(ufe_params) :: p
    INTEGER(C_INT) :: dev
    INTEGER :: nV, nAc, nA, w
    dev = 0
    IF (PRESENT(device)) dev = device
    IF (.NOT. C_ASSOCIATED(mesh%ufe)) THEN
      IF (ufe_ctx_create(dev, mesh%ufe) /= UFE_OK) THEN
        WRITE(0,*) 'initialise_mesh_dynamics - ERROR: no HIP device'
        ERROR STOP 1
      END IF
    END IF
    p = params_of(C)
    CALL check(mesh, ufe_set_params(mesh%ufe, p), 'ufe_set_params')
    IF (ALLOCATED(mesh%Tri_edge_index)) CALL calculate_mesh_geometry(mesh)
    nV = mesh%nV
    nAc = SUM(mesh%nC(1:nV)) / 2
    nA = nV + nAc
    mesh%nAc = nAc
    mesh%nVAaAc = nA
    w = mesh%nC_mem
    ALLOCATE(mesh%Nx(nV, w+1), mesh%Ny(nV, w+1), mesh%Nxx(nV, w+1), mesh%Nxy(nV, w+1), mesh%Nyy(nV, w+1))
    ALLOCATE(mesh%iAci(nV, w), mesh%Aci(nAc, 4), mesh%edge_index_Ac(nAc))
    ALLOCATE(mesh%VAc(nAc, 2), mesh%Nx_Ac(nAc, 4), mesh%Ny_Ac(nAc, 4), mesh%No_Ac(nAc, 4), mesh%Np_Ac(nAc))
    ALLOCATE(mesh%VAaAc(nA, 2), mesh%Nx_AaAc(nA, w+1), mesh%Ny_AaAc(nA, w+1), mesh%Nxx_AaAc(nA, w+1), &
             mesh%Nxy_AaAc(nA, w+1), mesh%Nyy_AaAc(nA, w+1))
    ALLOCATE(mesh%nCAaAc(nA), mesh%CAaAc(nA, w), mesh%colour(nA), mesh%colour_vi(nA, 5))
    d%nV = nV
    d%nTri = mesh%nTri
    d%nC_mem = w
    d%nAc = nAc
    d%xmin = mesh%xmin
    d%xmax = mesh%xmax
    d%ymin = mesh%ymin
    d%ymax = mesh%ymax
    d%V = C_LOC(mesh%V)
    d%nC = C_LOC(mesh%nC)
    d%C = C_LOC(mesh%C)
    d%niTri = C_LOC(mesh%niTri)
    d%iTri = C_LOC(mesh%iTri)
    d%edge_index = C_LOC(mesh%edge_index)
    d%Tri = C_LOC(mesh%Tri)
    d%A = C_LOC(mesh%A)
    d%Cw = C_LOC(mesh%Cw)
    d%Nx = C_LOC(mesh%Nx)
    d%Ny = C_LOC(mesh%Ny)
    d%Nxx = C_LOC(mesh%Nxx)
    d%Nxy = C_LOC(mesh%Nxy)
    d%Nyy = C_LOC(mesh%Nyy)
    d%iAci = C_LOC(mesh%iAci)
    d%Aci = C_LOC(mesh%Aci)
    d%VAc = C_LOC(mesh%VAc)
    d%Nx_Ac = C_LOC(mesh%Nx_Ac)
    d%Ny_Ac = C_LOC(mesh%Ny_Ac)
    d%No_Ac = C_LOC(mesh%No_Ac)
    d%Np_Ac = C_LOC(mesh%Np_Ac)
    d%edge_index_Ac = C_LOC(mesh%edge_index_Ac)
    d%VAaAc = C_LOC(mesh%VAaAc)
    d%nCAaAc = C_LOC(mesh%nCAaAc)
    d%CAaAc = C_LOC(mesh%CAaAc)
    d%Nx_AaAc = C_LOC(mesh%Nx_AaAc)
    d%Ny_AaAc = C_LOC(mesh%Ny_AaAc)
    d%Nxx_AaAc = C_LOC(mesh%Nxx_AaAc)
    d%Nxy_AaAc = C_LOC(mesh%Nxy_AaAc)
    d%Nyy_AaAc = C_LOC(mesh%Nyy_AaAc)
    d%colour = C_LOC(mesh%colour)
    d%colour_vi = C_LOC(mesh%colour_vi)
    d%colour_nV = 0
    CALL check(mesh, ufe_mesh_build(mesh%ufe, d), 'ufe_mesh_build')
    mesh%colour_nV = d%colour_nV
    CALL check(mesh, ufe_mesh_upload(mesh%ufe, d), 'ufe_mesh_upload')
  END SUBROUTINE initialise_mesh_dynamics

  ! Voronoi cell areas, triangle areas, connection widths and resolution, plus
  ! the circumcentres when the mesh has none: the reference's steps after
  ! creating or reading a mesh (src/mesh_creation_module.f90:1725-1731,
  ! src/restart_module.f90:89-96), on the GPU.
  SUBROUTINE calculate_mesh_geometry(mesh)
    TYPE(type_mesh), TARGET, INTENT(INOUT) :: mesh
    TYPE(ufe_mesh_geometry_desc) :: g
    TYPE(ufe_mesh_geometry_stats) :: st
    INTEGER :: nV, nT, w
    nV = mesh%nV
    nT = mesh%nTri
    w = mesh%nC_mem
    ! src/restart_module.f90:85 (tol = 1e-9)
    IF (mesh%tol_dist <= 0._dp) mesh%tol_dist = ((mesh%xmax - mesh%xmin) + (mesh%ymax - mesh%ymin)) * 1E-9_dp / 2._dp
    g%compute_Tricc = 0
    IF (.NOT. ALLOCATED(mesh%Tricc)) THEN
      ALLOCATE(mesh%Tricc(nT, 2))
      g%compute_Tricc = 1
    END IF
    IF (.NOT. ALLOCATED(mesh%A)) ALLOCATE(mesh%A(nV))
    IF (.NOT. ALLOCATED(mesh%Cw)) ALLOCATE(mesh%Cw(nV, w))
    IF (.NOT. ALLOCATED(mesh%TriA)) ALLOCATE(mesh%TriA(nT))
    IF (.NOT. ALLOCATED(mesh%R)) ALLOCATE(mesh%R(nV))
    g%nV = nV
    g%nTri = nT
    g%nC_mem = w
    g%xmin = mesh%xmin
    g%xmax = mesh%xmax
    g%ymin = mesh%ymin
    g%ymax = mesh%ymax
    g%tol_dist = mesh%tol_dist
    g%V = C_LOC(mesh%V)
    g%nC = C_LOC(mesh%nC)
    g%C = C_LOC(mesh%C)
    g%niTri = C_LOC(mesh%niTri)
    g%iTri = C_LOC(mesh%iTri)
    g%edge_index = C_LOC(mesh%edge_index)
    g%Tri = C_LOC(mesh%Tri)
    g%Tri_edge_index = C_LOC(mesh%Tri_edge_index)
    g%Tricc = C_LOC(mesh%Tricc)
    g%A = C_LOC(mesh%A)
    g%Cw = C_LOC(mesh%Cw)
    g%TriA = C_LOC(mesh%TriA)
    g%R = C_LOC(mesh%R)
    CALL check(mesh, ufe_mesh_geometry(mesh%ufe, g, st), 'ufe_mesh_geometry')
    mesh%resolution_min = st%resolution_min
    mesh%resolution_max = st%resolution_max
    ! src/mesh_help_functions_module.f90:41-45
    IF (st%area_error > 0.0001_dp) WRITE(0,*) 'calculate_mesh_geometry - WARNING: sum of Voronoi cell areas ', &
      'doesnt match square area of mesh! (error of ', st%area_error * 100._dp, ' %)'
  END SUBROUTINE calculate_mesh_geometry

  SUBROUTINE finalise_mesh_dynamics(mesh)
    TYPE(type_mesh), INTENT(INOUT) :: mesh
    INTEGER(C_INT) :: r
    IF (C_ASSOCIATED(mesh%ufe)) r = ufe_ctx_destroy(mesh%ufe)
    mesh%ufe = C_NULL_PTR
  END SUBROUTINE finalise_mesh_dynamics

  ! Zero-initialised ice model (src/ice_dynamics_module.f90:1029-1197)
  SUBROUTINE allocate_ice_model(mesh, ice)
    TYPE(type_mesh), INTENT(IN) :: mesh
    TYPE(type_ice_model), INTENT(INOUT) :: ice
    INTEGER :: nV, nAc, nA, nZ
    nV = mesh%nV
    nAc = mesh%nAc
    nA = mesh%nVAaAc
    nZ = C%nZ
    ALLOCATE(ice%Hi(nV), ice%Hb(nV), ice%SL(nV), ice%Hs(nV), ice%dHb_dt(nV), ice%dHi_dt(nV), ice%dHs_dt(nV), &
             ice%Hi_prev(nV), ice%A_flow(nV, nZ), ice%A_flow_mean(nV), ice%dHi_dx(nV), ice%dHi_dy(nV), &
             ice%dHs_dx(nV), ice%dHs_dy(nV), ice%dHs_dx_shelf(nV), ice%dHs_dy_shelf(nV), ice%U_SIA(nV), &
             ice%V_SIA(nV), ice%D_SIA(nV), ice%U_SSA(nV), ice%V_SSA(nV), ice%dVi_in(nV, mesh%nC_mem), source=0._dp)
    ALLOCATE(ice%Ti(nV, nZ), ice%Ti_pmp(nV, nZ), ice%Cpi(nV, nZ), ice%Ki(nV, nZ), ice%Ti_Ac(nAc, nZ), &
             ice%U_3D(nV, nZ), ice%V_3D(nV, nZ), ice%W_3D(nV, nZ), source=0._dp)
    ALLOCATE(ice%mask_land(nV), ice%mask_ocean(nV), ice%mask_lake(nV), ice%mask_ice(nV), ice%mask_sheet(nV), &
             ice%mask_shelf(nV), ice%mask_coast(nV), ice%mask_margin(nV), ice%mask_gl(nV), ice%mask_cf(nV), &
             ice%mask(nV), source=0)
    ALLOCATE(ice%Hi_Ac(nAc), ice%Hb_Ac(nAc), ice%SL_Ac(nAc), ice%Hs_Ac(nAc), ice%A_flow_Ac(nAc, nZ), &
             ice%A_flow_mean_Ac(nAc), ice%dHi_dx_Ac(nAc), ice%dHi_dy_Ac(nAc), ice%dHi_dp_Ac(nAc), &
             ice%dHi_do_Ac(nAc), ice%dHb_dx_Ac(nAc), ice%dHb_dy_Ac(nAc), ice%dHb_dp_Ac(nAc), ice%dHb_do_Ac(nAc), &
             ice%dHs_dx_Ac(nAc), ice%dHs_dy_Ac(nAc), ice%dHs_dp_Ac(nAc), ice%dHs_do_Ac(nAc), ice%dSL_dx_Ac(nAc), &
             ice%dSL_dy_Ac(nAc), ice%dSL_dp_Ac(nAc), ice%dSL_do_Ac(nAc), ice%dHs_dx_shelf_Ac(nAc), &
             ice%dHs_dy_shelf_Ac(nAc), ice%D_SIA_3D_Ac(nAc, nZ), ice%D_SIA_Ac(nAc), ice%Ux_SIA_Ac(nAc), &
             ice%Uy_SIA_Ac(nAc), ice%Up_SIA_Ac(nAc), ice%Uo_SIA_Ac(nAc), ice%Ux_SSA_Ac(nAc), ice%Uy_SSA_Ac(nAc), &
             ice%Up_SSA_Ac(nAc), ice%Uo_SSA_Ac(nAc), source=0._dp)
    ALLOCATE(ice%mask_land_Ac(nAc), ice%mask_ocean_Ac(nAc), ice%mask_lake_Ac(nAc), ice%mask_ice_Ac(nAc), &
             ice%mask_sheet_Ac(nAc), ice%mask_shelf_Ac(nAc), ice%mask_coast_Ac(nAc), ice%mask_margin_Ac(nAc), &
             ice%mask_gl_Ac(nAc), ice%mask_cf_Ac(nAc), ice%mask_Ac(nAc), source=0)
    ALLOCATE(ice%U_SSA_AaAc(nA), ice%V_SSA_AaAc(nA), ice%eta_AaAc(nA), ice%N_AaAc(nA), ice%S_AaAc(nA), &
             ice%tau_c_AaAc(nA), source=0._dp)
  END SUBROUTINE allocate_ice_model

  ! ---- the dynamics path -------------------------------------------------------

  SUBROUTINE update_general_ice_model_data(mesh, ice, time)
    TYPE(type_mesh), INTENT(IN) :: mesh
    TYPE(type_ice_model), INTENT(INOUT) :: ice
    REAL(dp), INTENT(IN) :: time
    IF (ufe_host_sync) THEN
      CALL put(mesh, 'Hi', ice%Hi)
      CALL put(mesh, 'Hb', ice%Hb)
      CALL put(mesh, 'SL', ice%SL)
      CALL put(mesh, 'dHb_dt', ice%dHb_dt)
      CALL put(mesh, 'dHi_dt', ice%dHi_dt)
      IF (C%physical_properties == 1) CALL put(mesh, 'Ti', ice%Ti)
    END IF
    CALL check(mesh, ufe_update_general_ice_model_data(mesh%ufe), 'update_general_ice_model_data')
    IF (.NOT. ufe_host_sync) RETURN
    IF (C%physical_properties /= 0) THEN
      ! ice_physical_properties (src/general_ice_model_data_module.f90:298-462)
      CALL get(mesh, 'A_flow', ice%A_flow)
      CALL get(mesh, 'A_flow_mean', ice%A_flow_mean)
      CALL get(mesh, 'Ti_pmp', ice%Ti_pmp)
      CALL get(mesh, 'A_flow_Ac', ice%A_flow_Ac)
      CALL get(mesh, 'A_flow_mean_Ac', ice%A_flow_mean_Ac)
      IF (C%physical_properties == 1) CALL get(mesh, 'Ti_Ac', ice%Ti_Ac)
      CALL get(mesh, 'Cpi', ice%Cpi)
      CALL get(mesh, 'Ki', ice%Ki)
    END IF
    CALL get(mesh, 'Hs', ice%Hs)
    CALL get(mesh, 'dHs_dt', ice%dHs_dt)
    CALL get(mesh, 'dHi_dx', ice%dHi_dx)
    CALL get(mesh, 'dHi_dy', ice%dHi_dy)
    CALL get(mesh, 'dHs_dx', ice%dHs_dx)
    CALL get(mesh, 'dHs_dy', ice%dHs_dy)
    CALL get(mesh, 'dHs_dx_shelf', ice%dHs_dx_shelf)
    CALL get(mesh, 'dHs_dy_shelf', ice%dHs_dy_shelf)
    CALL get(mesh, 'mask_land', ice%mask_land)
    CALL get(mesh, 'mask_ocean', ice%mask_ocean)
    CALL get(mesh, 'mask_lake', ice%mask_lake)
    CALL get(mesh, 'mask_ice', ice%mask_ice)
    CALL get(mesh, 'mask_sheet', ice%mask_sheet)
    CALL get(mesh, 'mask_shelf', ice%mask_shelf)
    CALL get(mesh, 'mask_coast', ice%mask_coast)
    CALL get(mesh, 'mask_margin', ice%mask_margin)
    CALL get(mesh, 'mask_gl', ice%mask_gl)
    CALL get(mesh, 'mask_cf', ice%mask_cf)
    CALL get(mesh, 'mask', ice%mask)
    CALL get(mesh, 'Hi_Ac', ice%Hi_Ac)
    CALL get(mesh, 'Hb_Ac', ice%Hb_Ac)
    CALL get(mesh, 'SL_Ac', ice%SL_Ac)
    CALL get(mesh, 'Hs_Ac', ice%Hs_Ac)
    CALL get(mesh, 'dHi_dx_Ac', ice%dHi_dx_Ac)
    CALL get(mesh, 'dHi_dy_Ac', ice%dHi_dy_Ac)
    CALL get(mesh, 'dHi_dp_Ac', ice%dHi_dp_Ac)
    CALL get(mesh, 'dHi_do_Ac', ice%dHi_do_Ac)
    CALL get(mesh, 'dHb_dx_Ac', ice%dHb_dx_Ac)
    CALL get(mesh, 'dHb_dy_Ac', ice%dHb_dy_Ac)
    CALL get(mesh, 'dHb_dp_Ac', ice%dHb_dp_Ac)
    CALL get(mesh, 'dHb_do_Ac', ice%dHb_do_Ac)
    CALL get(mesh, 'dHs_dx_Ac', ice%dHs_dx_Ac)
    CALL get(mesh, 'dHs_dy_Ac', ice%dHs_dy_Ac)
    CALL get(mesh, 'dHs_dp_Ac', ice%dHs_dp_Ac)
    CALL get(mesh, 'dHs_do_Ac', ice%dHs_do_Ac)
    CALL get(mesh, 'dSL_dx_Ac', ice%dSL_dx_Ac)
    CALL get(mesh, 'dSL_dy_Ac', ice%dSL_dy_Ac)
    CALL get(mesh, 'dSL_dp_Ac', ice%dSL_dp_Ac)
    CALL get(mesh, 'dSL_do_Ac', ice%dSL_do_Ac)
    CALL get(mesh, 'dHs_dx_shelf_Ac', ice%dHs_dx_shelf_Ac)
    CALL get(mesh, 'dHs_dy_shelf_Ac', ice%dHs_dy_shelf_Ac)
    CALL get(mesh, 'mask_land_Ac', ice%mask_land_Ac)
    CALL get(mesh, 'mask_ocean_Ac', ice%mask_ocean_Ac)
    CALL get(mesh, 'mask_lake_Ac', ice%mask_lake_Ac)
    CALL get(mesh, 'mask_ice_Ac', ice%mask_ice_Ac)
    CALL get(mesh, 'mask_sheet_Ac', ice%mask_sheet_Ac)
    CALL get(mesh, 'mask_shelf_Ac', ice%mask_shelf_Ac)
    CALL get(mesh, 'mask_coast_Ac', ice%mask_coast_Ac)
    CALL get(mesh, 'mask_margin_Ac', ice%mask_margin_Ac)
    CALL get(mesh, 'mask_gl_Ac', ice%mask_gl_Ac)
    CALL get(mesh, 'mask_cf_Ac', ice%mask_cf_Ac)
    CALL get(mesh, 'mask_Ac', ice%mask_Ac)
  END SUBROUTINE update_general_ice_model_data

  SUBROUTINE solve_SIA(mesh, ice)
    TYPE(type_mesh), INTENT(IN) :: mesh
    TYPE(type_ice_model), INTENT(INOUT) :: ice
    IF (ufe_host_sync) THEN
      CALL put(mesh, 'Hi_Ac', ice%Hi_Ac)
      CALL put(mesh, 'dHs_dx_Ac', ice%dHs_dx_Ac)
      CALL put(mesh, 'dHs_dy_Ac', ice%dHs_dy_Ac)
      CALL put(mesh, 'dHs_dp_Ac', ice%dHs_dp_Ac)
      CALL put(mesh, 'dHs_do_Ac', ice%dHs_do_Ac)
      CALL put(mesh, 'mask_sheet_Ac', ice%mask_sheet_Ac)
      CALL put(mesh, 'A_flow_Ac', ice%A_flow_Ac)
    END IF
    CALL check(mesh, ufe_solve_SIA(mesh%ufe), 'solve_SIA')
    IF (.NOT. ufe_host_sync) RETURN
    CALL get(mesh, 'D_SIA_3D_Ac', ice%D_SIA_3D_Ac)
    CALL get(mesh, 'D_SIA_Ac', ice%D_SIA_Ac)
    CALL get(mesh, 'Ux_SIA_Ac', ice%Ux_SIA_Ac)
    CALL get(mesh, 'Uy_SIA_Ac', ice%Uy_SIA_Ac)
    CALL get(mesh, 'Up_SIA_Ac', ice%Up_SIA_Ac)
    CALL get(mesh, 'Uo_SIA_Ac', ice%Uo_SIA_Ac)
    CALL get(mesh, 'U_SIA', ice%U_SIA)
    CALL get(mesh, 'V_SIA', ice%V_SIA)
    CALL get(mesh, 'D_SIA', ice%D_SIA)
  END SUBROUTINE solve_SIA

  ! solve_SIA_3D (src/ice_dynamics_module.f90:317-405): U_3D, V_3D, W_3D from the
  ! SSA velocities, A_flow and the geometry of update_general_ice_model_data
  SUBROUTINE solve_SIA_3D(mesh, ice)
    TYPE(type_mesh), INTENT(IN) :: mesh
    TYPE(type_ice_model), INTENT(INOUT) :: ice
    IF (ufe_host_sync) THEN
      CALL put(mesh, 'U_SSA', ice%U_SSA)
      CALL put(mesh, 'V_SSA', ice%V_SSA)
      CALL put(mesh, 'A_flow', ice%A_flow)
      CALL put(mesh, 'dHb_dt', ice%dHb_dt)
    END IF
    CALL check(mesh, ufe_solve_SIA_3D(mesh%ufe), 'solve_SIA_3D')
    IF (.NOT. ufe_host_sync) RETURN
    CALL get(mesh, 'U_3D', ice%U_3D)
    CALL get(mesh, 'V_3D', ice%V_3D)
    CALL get(mesh, 'W_3D', ice%W_3D)
  END SUBROUTINE solve_SIA_3D

  ! The critical time steps of determine_timesteps_and_actions
  ! (src/UFEMISM_main_model.f90:734-780, after the 0.9 correction): the
  ! loops and MPI_ALLREDUCEs there become this one call
  SUBROUTINE calculate_critical_timesteps(mesh, ice, dt_D_2D_min, dt_V_2D_SSA_min, dt_V_3D_SIA_min)
    TYPE(type_mesh), INTENT(IN) :: mesh
    TYPE(type_ice_model), INTENT(IN) :: ice
    REAL(dp), INTENT(OUT) :: dt_D_2D_min, dt_V_2D_SSA_min, dt_V_3D_SIA_min
    REAL(C_DOUBLE) :: dt(3)
    IF (ufe_host_sync) THEN
      CALL put(mesh, 'D_SIA_Ac', ice%D_SIA_Ac)
      CALL put(mesh, 'U_SSA', ice%U_SSA)
      CALL put(mesh, 'V_SSA', ice%V_SSA)
      CALL put(mesh, 'U_3D', ice%U_3D)
      CALL put(mesh, 'V_3D', ice%V_3D)
    END IF
    CALL check(mesh, ufe_critical_timesteps(mesh%ufe, dt), 'critical_timesteps')
    dt_D_2D_min = dt(1)
    dt_V_2D_SSA_min = dt(2)
    dt_V_3D_SIA_min = dt(3)
  END SUBROUTINE calculate_critical_timesteps

  SUBROUTINE solve_SSA(mesh, ice)
    TYPE(type_mesh), INTENT(IN) :: mesh
    TYPE(type_ice_model), INTENT(INOUT) :: ice
    TYPE(ufe_ssa_stats) :: st
    TYPE(ufe_params) :: p
    INTEGER(C_INT) :: r
    p = params_of(C)
    IF (ALLOCATED(ice%tau_c_override_AaAc)) p%use_tau_c_override = 1
    CALL check(mesh, ufe_set_params(mesh%ufe, p), 'ufe_set_params')
    IF (ufe_host_sync) THEN
      CALL put(mesh, 'mask_sheet', ice%mask_sheet)
      CALL put(mesh, 'Hi', ice%Hi)
      CALL put(mesh, 'Hb', ice%Hb)
      CALL put(mesh, 'SL', ice%SL)
      CALL put(mesh, 'Hi_Ac', ice%Hi_Ac)
      CALL put(mesh, 'Hb_Ac', ice%Hb_Ac)
      CALL put(mesh, 'SL_Ac', ice%SL_Ac)
      CALL put(mesh, 'dHs_dx_shelf', ice%dHs_dx_shelf)
      CALL put(mesh, 'dHs_dy_shelf', ice%dHs_dy_shelf)
      CALL put(mesh, 'dHs_dx_shelf_Ac', ice%dHs_dx_shelf_Ac)
      CALL put(mesh, 'dHs_dy_shelf_Ac', ice%dHs_dy_shelf_Ac)
      CALL put(mesh, 'A_flow_mean', ice%A_flow_mean)
      CALL put(mesh, 'A_flow_mean_Ac', ice%A_flow_mean_Ac)
      CALL put(mesh, 'U_SSA', ice%U_SSA)
      CALL put(mesh, 'V_SSA', ice%V_SSA)
      CALL put(mesh, 'Ux_SSA_Ac', ice%Ux_SSA_Ac)
      CALL put(mesh, 'Uy_SSA_Ac', ice%Uy_SSA_Ac)
      CALL put(mesh, 'N_AaAc', ice%N_AaAc)
      IF (ALLOCATED(ice%tau_c_override_AaAc)) CALL put(mesh, 'tau_c_override_AaAc', ice%tau_c_override_AaAc)
    END IF
    r = ufe_solve_SSA(mesh%ufe, st)
    ice%SSA_picard_iterations = st%picard_iterations
    ice%SSA_sweeps = st%sweeps
    IF (r == UFE_SSA_UNSTABLE) THEN
      WRITE(0,*) 'solve_SSA - ERROR: SSA remains unstable after resetting velocities to zero!'
      ERROR STOP 1
    END IF
    CALL check(mesh, r, 'solve_SSA')
    IF (.NOT. ufe_host_sync) RETURN
    CALL get(mesh, 'U_SSA', ice%U_SSA)
    CALL get(mesh, 'V_SSA', ice%V_SSA)
    CALL get(mesh, 'Ux_SSA_Ac', ice%Ux_SSA_Ac)
    CALL get(mesh, 'Uy_SSA_Ac', ice%Uy_SSA_Ac)
    CALL get(mesh, 'Up_SSA_Ac', ice%Up_SSA_Ac)
    CALL get(mesh, 'Uo_SSA_Ac', ice%Uo_SSA_Ac)
    CALL get(mesh, 'U_SSA_AaAc', ice%U_SSA_AaAc)
    CALL get(mesh, 'V_SSA_AaAc', ice%V_SSA_AaAc)
    CALL get(mesh, 'eta_AaAc', ice%eta_AaAc)
    CALL get(mesh, 'N_AaAc', ice%N_AaAc)
    CALL get(mesh, 'S_AaAc', ice%S_AaAc)
    CALL get(mesh, 'tau_c_AaAc', ice%tau_c_AaAc)
  END SUBROUTINE solve_SSA

  SUBROUTINE calculate_ice_thickness_change(mesh, ice, SMB, BMB, dt, mask_noice)
    TYPE(type_mesh), INTENT(IN) :: mesh
    TYPE(type_ice_model), INTENT(INOUT) :: ice
    TYPE(type_SMB_model), INTENT(IN) :: SMB
    TYPE(type_BMB_model), INTENT(IN) :: BMB
    REAL(dp), INTENT(IN) :: dt
    INTEGER, INTENT(IN) :: mask_noice(:)
    TYPE(ufe_params) :: p
    INTEGER, ALLOCATABLE :: m(:)
    p = params_of(C)
    CALL check(mesh, ufe_set_params(mesh%ufe, p), 'ufe_set_params')
    ALLOCATE(m(SIZE(mask_noice)))
    m = mask_noice
    CALL put(mesh, 'SMB_year', SMB%SMB_year)
    CALL put(mesh, 'BMB', BMB%BMB)
    CALL put(mesh, 'mask_noice', m)
    IF (ufe_host_sync) THEN
      CALL put(mesh, 'Hi', ice%Hi)
      CALL put(mesh, 'Up_SIA_Ac', ice%Up_SIA_Ac)
      CALL put(mesh, 'Up_SSA_Ac', ice%Up_SSA_Ac)
    END IF
    CALL check(mesh, ufe_calculate_ice_thickness_change(mesh%ufe, REAL(dt, C_DOUBLE)), &
               'calculate_ice_thickness_change')
    IF (.NOT. ufe_host_sync) RETURN
    CALL get(mesh, 'dVi_in', ice%dVi_in)
    CALL get(mesh, 'dHi_dt', ice%dHi_dt)
    CALL get(mesh, 'Hi', ice%Hi)
    CALL get(mesh, 'Hi_prev', ice%Hi_prev)
  END SUBROUTINE calculate_ice_thickness_change

END MODULE ice_dynamics_module
