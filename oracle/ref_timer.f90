! CPU timing driver for the compiled reference at the 1M-cell case.
!
! TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile.ref (target c2) against
! the reference's solver modules compiled in place from /root/reference/src
! with the case constants set to 128 x 256 x 32 cells (outputs in
! oracle/_ref/c2/). It sets up the mesh, metrics and boundary tables as
! solve_nse does (main.f90:36-45, nse.f90:31-38), adds the seeded
! initial-condition noise of oracle/ic_noise.c, builds the step's momentum
! constants (nse.f90:49-50) and times K outer iterations of integrate_nse's
! loop body (nse.f90:134-162: the three sweeps with their sources, the
! pressure walls, the energy constants, the adiabatic walls and find_error)
! in the reference's own order on one core: the reference has no OpenMP
! parallel regions (it calls omp_get_wtime only). Prints one line:
!   ref_iterations K seconds T cells NMESH
!
! usage: ref_timer K NOISE_AMP SEED
program ref_timer
  use iso_c_binding
  use omp_lib
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

  character(len=64) :: arg
  integer :: k, nit, i
  real(8) :: amp, t0, t1
  integer(8) :: seed
  real :: eu, ev, ew, ep, eT

  call get_command_argument(1, arg); read(arg,*) nit
  call get_command_argument(2, arg); read(arg,*) amp
  call get_command_argument(3, arg); read(arg,*) seed

  t0 = omp_get_wtime()
  call create_hollow(R_i,R_o,L,0.,0.,L/2.,NMESH_i,NMESH_j,NMESH_k)
  print '(A,F10.2)', 'setup: mesh ', omp_get_wtime() - t0
  call prepare_mesh
  print '(A,F10.2)', 'setup: connectivity ', omp_get_wtime() - t0
  call fill_global_arrays
  TI=DT
  STEP=1
  call evaluate_all_mesh_constants
  call evaluate_pressure_bc_array
  call calculate_pressure_nim_constants
  call find_derivatives_pressure_boundary
  call apply_all_dirichlet_bcs
  print '(A,F10.2)', 'setup: metrics and constants ', omp_get_wtime() - t0
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

  call evaluate_time_dependent_quantities
  call calculate_velocity_nim_constants_formulation2
  t0 = omp_get_wtime()
  do k=1,nit
    call update_velocity_sources
    call integrate_velocity_field
    call update_pressure_sources
    call apply_pressure_bc
    call integrate_pressure_field
    call calculate_temperature_nim_constants_formulation2
    call apply_zero_flux
    call update_temperature_sources
    call integrate_temperature_field
    call find_error(eu, ev, ew, ep, eT)
  enddo
  t1 = omp_get_wtime()
  print '(A,I0,A,F12.6,A,I0)', 'ref_iterations ', nit, ' seconds ', t1 - t0, ' cells ', NMESH
end program ref_timer
