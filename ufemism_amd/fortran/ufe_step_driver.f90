! ufe_step_driver.f90 -- one dynamics step through the Fortran host.
!
! Reads a primary mesh, parameters and an initial ice state from an
! unformatted stream file, performs the reference's per-step sequence
!   update_general_ice_model_data -> solve_SIA -> solve_SSA -> calculate_ice_thickness_change
! through ice_dynamics_module, and writes the velocities and thickness.
! Usage: ufe_step_driver <input> <output>
PROGRAM ufe_step_driver
  USE ice_dynamics_module
  IMPLICIT NONE
  TYPE(type_mesh), TARGET :: mesh
  TYPE(type_ice_model) :: ice
  TYPE(type_SMB_model) :: SMB
  TYPE(type_BMB_model) :: BMB
  CHARACTER(LEN=1024) :: fin, fout, mode
  INTEGER :: u, nV, nTri, nC_mem, nZ, k, nmax_out, nmax_in
  INTEGER, ALLOCATABLE :: mask_noice(:)
  REAL(dp) :: A_flow, tau_c, dt, hdr(4), pr(5), dt_crit(3)

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  ! optional third argument: 'mixed_multigrid' solves every linearised SSA
  ! system to 1e-10 by mixed GMRES-IR with the multigrid preconditioner
  mode = ''
  IF (COMMAND_ARGUMENT_COUNT() >= 3) CALL GET_COMMAND_ARGUMENT(3, mode)
  OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(u) nV, nTri, nC_mem, nZ
  READ(u) hdr
  mesh%nV = nV
  mesh%nTri = nTri
  mesh%nC_mem = nC_mem
  mesh%xmin = hdr(1)
  mesh%xmax = hdr(2)
  mesh%ymin = hdr(3)
  mesh%ymax = hdr(4)
  ALLOCATE(mesh%V(nV, 2), mesh%nC(nV), mesh%C(nV, nC_mem), mesh%niTri(nV), mesh%iTri(nV, nC_mem), &
           mesh%edge_index(nV), mesh%Tri(nTri, 3), mesh%A(nV), mesh%Cw(nV, nC_mem), mesh%Tri_edge_index(nTri))
  READ(u) mesh%V, mesh%nC, mesh%C, mesh%niTri, mesh%iTri, mesh%edge_index, mesh%Tri, mesh%A, mesh%Cw
  ! with Tri_edge_index the host computes Tricc, A and Cw itself, as the reference does
  READ(u) mesh%Tri_edge_index
  C%nZ = nZ
  READ(u) C%zeta(1:nZ)
  READ(u) pr, dt
  C%m_enh_sia = pr(1)
  C%m_enh_ssa = pr(2)
  C%SSA_RN_tol = pr(3)
  C%SSA_max_residual_UV = pr(4)
  C%SSA_SOR_omega = pr(5)
  READ(u) nmax_out, nmax_in
  C%SSA_max_outer_loops = nmax_out
  C%SSA_max_inner_loops = nmax_in
  IF (TRIM(mode) == 'mixed_multigrid') THEN
    C%precision = 1
    C%SSA_precond_type = 1
    C%SSA_linear_rel_tol = 1E-10_dp
  END IF

  CALL initialise_mesh_dynamics(mesh)
  CALL allocate_ice_model(mesh, ice)
  ALLOCATE(SMB%SMB_year(nV), BMB%BMB(nV), mask_noice(nV))
  READ(u) ice%Hi, ice%Hb, ice%SL, A_flow, tau_c, SMB%SMB_year, BMB%BMB, mask_noice
  CLOSE(u)
  ice%A_flow = A_flow
  ice%A_flow_mean = A_flow
  ice%A_flow_Ac = A_flow
  ice%A_flow_mean_Ac = A_flow
  ALLOCATE(ice%tau_c_override_AaAc(mesh%nVAaAc))
  ice%tau_c_override_AaAc = tau_c

  CALL update_general_ice_model_data(mesh, ice, 0._dp)
  CALL solve_SIA(mesh, ice)
  CALL solve_SSA(mesh, ice)
  CALL calculate_ice_thickness_change(mesh, ice, SMB, BMB, dt, mask_noice)
  ! the velocity part of the thermodynamics and the next time steps
  CALL solve_SIA_3D(mesh, ice)
  CALL calculate_critical_timesteps(mesh, ice, dt_crit(1), dt_crit(2), dt_crit(3))

  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) mesh%nAc, ice%SSA_picard_iterations, (ice%SSA_sweeps(k), k = 1, 64)
  WRITE(u) ice%U_SSA, ice%V_SSA, ice%Up_SSA_Ac, ice%Up_SIA_Ac, ice%Hi, ice%dHi_dt
  WRITE(u) ice%U_3D, ice%W_3D, mesh%A, dt_crit
  CLOSE(u)
  CALL finalise_mesh_dynamics(mesh)
END PROGRAM ufe_step_driver
