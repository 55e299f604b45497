! Golden-vector driver for the compiled reference.
!
! TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile.ref against the reference
! modules compiled in place from /root/reference/src (outputs in oracle/_ref/).
! It calls the reference's own module procedures in the order solve_nse uses
! them (nse.f90:23-85, initialize_nse nse.f90:89-112 minus the Silo output
! call write_data, which this driver never reaches), adds the seeded
! initial-condition noise of oracle/ic_noise.c, and dumps the module arrays as
! raw little-endian binaries for tests/golden/make_golden.py.
!
! The reference's case is fixed at compile time (parameters.f90:84-173:
! 36 x 6 x 12 cells, R_o=2, L=2, Ra=1e5, Pr=0.7, DT=0.01, tolerances 1e-6 and
! 5e-3 for p); this driver changes none of it.
!
! usage: ref_driver OUTDIR NSTEPS NOISE_AMP SEED
program ref_driver
  use iso_c_binding
  use class_general
  use class_mesh
  use class_create_geometry
  use class_create_mesh
  use class_prepare_mesh
  use class_parameters
  use class_geometry
  use class_nim
  use class_bc
  use class_nse
  implicit none
  interface
    function oracle_ic_noise(seed, field, idx) bind(C, name="oracle_ic_noise") result(r)
      import :: c_int64_t, c_int32_t, c_double
      integer(c_int64_t), value :: seed
      integer(c_int32_t), value :: field
      integer(c_int64_t), value :: idx
      real(c_double) :: r
    end function
  end interface

  character(len=512) :: outdir, arg
  integer :: nsteps, istep, i
  real(8) :: amp
  integer(8) :: seed
  real :: eu, ev, ew, ep, eT
  integer, allocatable :: cell_nodes(:,:)
  real(8), allocatable :: node_xyz(:,:)
  real(8) :: scal(8)

  call get_command_argument(1, outdir)
  call get_command_argument(2, arg); read(arg,*) nsteps
  call get_command_argument(3, arg); read(arg,*) amp
  call get_command_argument(4, arg); read(arg,*) seed

  ! ---- mesh, metrics and boundary tables (main.f90:36-45, nse.f90:31-38) ----
  call create_hollow(R_i,R_o,L,0.,0.,L/2.,NMESH_i,NMESH_j,NMESH_k)
  call prepare_mesh
  call fill_global_arrays
  TI=DT
  STEP=1
  call evaluate_all_mesh_constants
  call evaluate_pressure_bc_array
  call calculate_pressure_nim_constants
  call find_derivatives_pressure_boundary
  call apply_all_dirichlet_bcs

  allocate(cell_nodes(8,NMESH), node_xyz(3,NNODE))
  do i=1,NMESH
    cell_nodes(:,i) = MESHES_ARRAY(i)%nodes
  enddo
  do i=1,NNODE
    node_xyz(1,i) = NODES_ARRAY(i)%x
    node_xyz(2,i) = NODES_ARRAY(i)%y
    node_xyz(3,i) = NODES_ARRAY(i)%z
  enddo
  scal = 0
  scal(1) = NMESH_i; scal(2) = NMESH_j; scal(3) = NMESH_k
  scal(4) = NMESH; scal(5) = NFACE; scal(6) = tot_faces_bcs; scal(7) = NNODE
  call dumpr('geo_dims', scal, 8)
  call dumpi('geo_meshes_faces', meshes_faces, 6*NMESH)
  call dumpi('geo_meshes_neighbors', meshes_neighbors, 6*NMESH)
  call dumpi('geo_cell_nodes', cell_nodes, 8*NMESH)
  call dumpr('geo_node_xyz', node_xyz, 3*NNODE)
  call dumpi('geo_faces_shared', faces_shared_meshes, 2*NFACE)
  call dumpi('geo_faces_BC', faces_BC, NFACE)
  call dumpi('geo_bc_faces', BCs_faces, tot_faces_bcs)
  call dumpi('geo_bc_id', BCs_faces_id, tot_faces_bcs)
  call dumpi('geo_bc_position', BCs_faces_position, tot_faces_bcs)
  call dumpi('geo_bc_mesh', BCs_faces_mesh_num, tot_faces_bcs)
  call dumpi('geo_bc_points', BCs_faces_points, 5*tot_faces_bcs)
  call dumpr('geo_volume', meshes_volume, NMESH)
  call dumpr('geo_xi_x', meshes_xi_x, 9*NMESH)
  call dumpr('geo_gij', meshes_gij, 9*NMESH)
  call dumpr('geo_gij_faces', meshes_gij_faces, 54*NMESH)
  call dumpr('geo_G1', meshes_G1, NMESH)
  call dumpr('geo_G2', meshes_G2, NMESH)
  call dumpr('geo_G3', meshes_G3, NMESH)
  call dumpr('geo_faces_area', faces_area, NFACE)
  call dumpr('geo_faces_normal', faces_normal, 3*NFACE)
  call dumpr('geo_faces_tangential', faces_tangential, 3*NFACE)
  call dumpr('geo_faces_bitangential', faces_bitangential, 3*NFACE)
  call dumpr('geo_faces_mid', faces_mid_coordinates, 3*NFACE)
  call dumpr('geo_bc_a', BCs_faces_a, 20*tot_faces_bcs)
  call dumpr('geo_bc_ap', BCs_faces_ap, 8*tot_faces_bcs)
  call dumpr('geo_AA_px', AA_px, 15*NMESH)
  call dumpr('geo_AA_py', AA_py, 15*NMESH)
  call dumpr('geo_AA_pz', AA_pz, 15*NMESH)

  ! ---- seeded perturbation of the cell averages (all time levels) ----
  do i=1,NMESH
    meshes_uxyz(i) = amp*oracle_ic_noise(seed, 0, int(i-1,8))
    meshes_vxyz(i) = amp*oracle_ic_noise(seed, 1, int(i-1,8))
    meshes_wxyz(i) = amp*oracle_ic_noise(seed, 2, int(i-1,8))
    meshes_Txyz(i) = amp*oracle_ic_noise(seed, 3, int(i-1,8))
  enddo
  meshes_uxyz_old = meshes_uxyz; meshes_uxyzt = meshes_uxyz
  meshes_vxyz_old = meshes_vxyz; meshes_vxyzt = meshes_vxyz
  meshes_wxyz_old = meshes_wxyz; meshes_wxyzt = meshes_wxyz
  meshes_Txyz_old = meshes_Txyz; meshes_Txyzt = meshes_Txyz
  call dump_state('s0')

  ! ---- time steps exactly as solve_nse (nse.f90:47-75) ----
  do istep=1,nsteps
    call evaluate_time_dependent_quantities
    call calculate_velocity_nim_constants_formulation2
    call integrate_nse
    scal = 0
    scal(1) = Dilatation; scal(2) = dudt_max; scal(3) = dTdt_max
    call update_time_step
    call find_Nu2
    scal(4) = TI; scal(5) = STEP
    write(arg,'(A,I0)') 's', istep
    call dump_state(trim(arg))
    call dumpr(trim(arg)//'_scalars', scal, 8)
  enddo

  ! ---- one instrumented outer iteration from the current state ----
  call dump_old('it0')
  call calculate_velocity_nim_constants_formulation2
  call dumpr('it_M_A', M_A, 18*NMESH)
  call dumpr('it_M_F1', M_F1, 5*NMESH)
  call dumpr('it_M_F2', M_F2, 5*NMESH)
  call dumpr('it_M_F3', M_F3, 5*NMESH)
  call dumpr('it_M_F4', M_F4, 8*NMESH)
  call dumpr('it_meshes_u', meshes_u, NMESH)
  call dumpr('it_meshes_v', meshes_v, NMESH)
  call dumpr('it_meshes_w', meshes_w, NMESH)
  call update_velocity_sources
  call dumpr('it_R_B_u', meshes_R_B_u, NMESH)
  call dumpr('it_R_B_v', meshes_R_B_v, NMESH)
  call dumpr('it_R_B_w', meshes_R_B_w, NMESH)
  call dumpr('it_R_C_u', meshes_R_C_u, NMESH)
  call dumpr('it_R_C_v', meshes_R_C_v, NMESH)
  call dumpr('it_R_C_w', meshes_R_C_w, NMESH)
  call dumpr('it_R_D_u', meshes_R_D_u, NMESH)
  call dumpr('it_R_D_v', meshes_R_D_v, NMESH)
  call dumpr('it_R_D_w', meshes_R_D_w, NMESH)
  call integrate_velocity_field
  call dumpr('it_vel_faces_u', faces_u, NFACE)
  call dumpr('it_vel_faces_v', faces_v, NFACE)
  call dumpr('it_vel_faces_w', faces_w, NFACE)
  call dumpr('it_vel_uxyz', meshes_uxyz, NMESH)
  call dumpr('it_vel_vxyz', meshes_vxyz, NMESH)
  call dumpr('it_vel_wxyz', meshes_wxyz, NMESH)
  call update_pressure_sources
  call dumpr('it_R_b_p', meshes_R_b_p, NMESH)
  call dumpr('it_R_D_p', meshes_R_D_p, NMESH)
  call dumpr('it_dila', meshes_dila, NMESH)
  call apply_pressure_bc
  call dumpr('it_pbc_faces_p', faces_p, NFACE)
  call integrate_pressure_field
  call dumpr('it_p_faces_p', faces_p, NFACE)
  call calculate_temperature_nim_constants_formulation2
  call dumpr('it_MT_A', MT_A, 18*NMESH)
  call dumpr('it_MT_F1', MT_F1, 5*NMESH)
  call dumpr('it_MT_F2', MT_F2, 5*NMESH)
  call dumpr('it_MT_F3', MT_F3, 5*NMESH)
  call dumpr('it_MT_F4', MT_F4, 8*NMESH)
  call apply_zero_flux
  call dumpr('it_zf_faces_T', faces_T, NFACE)
  call update_temperature_sources
  call dumpr('it_R_D_T', meshes_R_D_T, NMESH)
  call integrate_temperature_field
  call dumpr('it_T_faces_T', faces_T, NFACE)
  call dumpr('it_T_Txyz', meshes_Txyz, NMESH)
  call find_error(eu, ev, ew, ep, eT)
  scal = 0
  scal(1) = eu; scal(2) = ev; scal(3) = ew; scal(4) = ep; scal(5) = eT
  call dumpr('it_err', scal, 8)

contains

  subroutine dumpr(name, a, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(8), intent(in) :: a(n)
    integer :: u
    open(newunit=u, file=trim(outdir)//'/'//name//'.f64', access='stream', &
         form='unformatted', status='replace')
    write(u) a
    close(u)
  end subroutine

  subroutine dumpi(name, a, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    integer, intent(in) :: a(n)
    integer :: u
    open(newunit=u, file=trim(outdir)//'/'//name//'.i32', access='stream', &
         form='unformatted', status='replace')
    write(u) a
    close(u)
  end subroutine

  subroutine dump_state(p)
    character(len=*), intent(in) :: p
    call dumpr(p//'_faces_u', faces_u, NFACE)
    call dumpr(p//'_faces_v', faces_v, NFACE)
    call dumpr(p//'_faces_w', faces_w, NFACE)
    call dumpr(p//'_faces_p', faces_p, NFACE)
    call dumpr(p//'_faces_T', faces_T, NFACE)
    call dumpr(p//'_uxyz', meshes_uxyz, NMESH)
    call dumpr(p//'_vxyz', meshes_vxyz, NMESH)
    call dumpr(p//'_wxyz', meshes_wxyz, NMESH)
    call dumpr(p//'_Txyz', meshes_Txyz, NMESH)
    call dumpr(p//'_uxyz_old', meshes_uxyz_old, NMESH)
    call dumpr(p//'_vxyz_old', meshes_vxyz_old, NMESH)
    call dumpr(p//'_wxyz_old', meshes_wxyz_old, NMESH)
    call dumpr(p//'_Txyz_old', meshes_Txyz_old, NMESH)
    call dumpr(p//'_uxyzt', meshes_uxyzt, NMESH)
    call dumpr(p//'_vxyzt', meshes_vxyzt, NMESH)
    call dumpr(p//'_wxyzt', meshes_wxyzt, NMESH)
    call dumpr(p//'_Txyzt', meshes_Txyzt, NMESH)
    call dumpr(p//'_p', meshes_p, NMESH)
    call dumpr(p//'_dila', meshes_dila, NMESH)
  end subroutine

  subroutine dump_old(p)
    character(len=*), intent(in) :: p
    call dumpr(p//'_faces_u_old', faces_u_old, NFACE)
    call dumpr(p//'_faces_v_old', faces_v_old, NFACE)
    call dumpr(p//'_faces_w_old', faces_w_old, NFACE)
    call dumpr(p//'_faces_p_old', faces_p_old, NFACE)
    call dumpr(p//'_faces_T_old', faces_T_old, NFACE)
    call dumpr(p//'_uxyz_old_con', meshes_uxyz_old_con, NMESH)
    call dumpr(p//'_vxyz_old_con', meshes_vxyz_old_con, NMESH)
    call dumpr(p//'_wxyz_old_con', meshes_wxyz_old_con, NMESH)
    call dumpr(p//'_Txyz_old_con', meshes_Txyz_old_con, NMESH)
  end subroutine
end program ref_driver
