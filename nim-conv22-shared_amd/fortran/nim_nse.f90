! Fortran host time loop over the MI355X library: the module a reference
! build would use in place of class_nse's solve_nse (nse.f90:23-85). Mesh and
! metrics come from nim_host_geometry, every per-step routine is a call
! through nim_hip_api, and the restart file keeps the reference's record
! layout (restart.f90:8-69).
module nim_nse
  use iso_c_binding
  use nim_host_geometry
  use nim_hip_api
  implicit none
  private
  public :: nim_run_config, nim_setup, nim_solve, nim_write_restart, nim_read_restart, nim_finish, nim_write_data

  type :: nim_run_config
    type(nim_geo_case) :: geo
    real(c_double) :: Ra = 1.0d5, Pr = 0.7d0, DT = 0.01d0, rho = 1.0d0, g_beta = 1.0d0, th_tc = 1.0d0
    real(c_double) :: relax = 0.4d0, v_limit = 1.0d-2, T_hot = 1.0d0, T_cold = 0.0d0
    real(c_double) :: noise = 1.0d-3
    integer(c_int64_t) :: seed = 20221
    real(c_double) :: tol(5) = (/ 1.0d-6, 1.0d-6, 1.0d-6, 5.0d-3, 1.0d-6 /)  ! parameters.f90:146-150
    real(c_double) :: dudt_epsilon = 1.0d-5                                  ! parameters.f90:151
    logical :: check_steady = .true.
    integer :: max_iter = 200000
    ! > 0: mixed-precision inner loop with this many FP32 passes per FP64
    ! residual (tolerances then apply to the FP64 residual measure)
    integer :: mixed_inner = 0
    ! per-cell metrics on the device (nim_device_metrics, bit for bit the
    ! host's build_geometry, about 10x faster at 8M cells); .false.: host
    logical :: device_metrics = .true.
    ! write the cell fields every step_out steps (STEP_OUT, parameters.f90:138;
    ! 0: never) as plots_dir/stepNNNNN.vtk, listed in plots.visit
    integer :: step_out = 0
    ! what tol(4) bounds: 0 max |dp| (find_error, the reference's), 1 half
    ! the range of dp (blind to p's constant level)
    integer :: p_measure = 0
    ! > 0: Anderson acceleration of the inner loop's passes (FP64) or cycles
    ! (mixed) every accel_period of them over the last `accel` steps (1..16);
    ! the same stopping test and fixed point (nim_inner_solve_accel); of the
    ! pressure faces (accel_scope 1) or of the whole iterate (0)
    integer :: accel = 0
    integer :: accel_period = 1
    integer :: accel_scope = 1
    character(len=256) :: plots_dir = 'plots'
  end type

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: N, NF, NB, nth, nrr, nzz
  real(c_double), save :: TI, DTs
  real(c_double), allocatable, save :: node_xyz(:,:)   ! (3, NNODE), for the field output
  integer, save :: STEP
  ! host copies of the state in the reference's layout
  real(c_double), allocatable, target, save :: faces_u(:), faces_v(:), faces_w(:), faces_p(:), faces_T(:)
  real(c_double), allocatable, target, save :: uxyz(:), vxyz(:), wxyz(:), Txyz(:)
  real(c_double), allocatable, target, save :: uxyz_old(:), vxyz_old(:), wxyz_old(:), Txyz_old(:)
  real(c_double), allocatable, target, save :: uxyzt(:), vxyzt(:), wxyzt(:), Txyzt(:), meshes_p(:), dila(:)

  interface
    integer(c_int32_t) function nimh_initial_state(cs, T_hot, T_cold, amp, seed, fu, fv, fw, fp, fT, &
        u, v, w, T) bind(C, name="nimh_initial_state")
      import :: c_int32_t, c_double, c_int64_t, nim_geo_case
      type(nim_geo_case), intent(in) :: cs
      real(c_double), value :: T_hot, T_cold, amp
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: fu(*), fv(*), fw(*), fp(*), fT(*), u(*), v(*), w(*), T(*)
    end function
  end interface

contains

  subroutine check(rc, what)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= NIM_OK) then
      print '(A, A, I0, A, A)', what, ' failed with code ', rc, ': ', nim_error_message()
      error stop 1
    endif
  end subroutine

  function state_ptrs() result(st)
    type(nim_state) :: st
    st%faces_u = c_loc(faces_u); st%faces_v = c_loc(faces_v); st%faces_w = c_loc(faces_w)
    st%faces_p = c_loc(faces_p); st%faces_T = c_loc(faces_T)
    st%uxyz = c_loc(uxyz); st%vxyz = c_loc(vxyz); st%wxyz = c_loc(wxyz); st%Txyz = c_loc(Txyz)
    st%uxyz_old = c_loc(uxyz_old); st%vxyz_old = c_loc(vxyz_old); st%wxyz_old = c_loc(wxyz_old)
    st%Txyz_old = c_loc(Txyz_old); st%uxyzt = c_loc(uxyzt); st%vxyzt = c_loc(vxyzt); st%wxyzt = c_loc(wxyzt)
    st%Txyzt = c_loc(Txyzt); st%meshes_p = c_loc(meshes_p); st%meshes_dila = c_loc(dila)
  end function

  ! initialize_nse (nse.f90:89-112) + apply_all_dirichlet_bcs (nse.f90:38)
  subroutine nim_setup(cfg)
    type(nim_run_config), intent(in) :: cfg
    type(nim_dims) :: dims
    type(nim_params) :: prm
    type(nim_geometry) :: g
    real(c_double), allocatable, target :: volume(:), xi_x(:,:,:), gij(:,:,:), gij_faces(:,:,:,:), G1(:), G2(:), G3(:)
    real(c_double), allocatable, target :: area(:), nrm(:,:), tan1(:,:), tan2(:,:), mid(:,:), bc_a(:,:), nodes(:,:)
    real(c_double), allocatable :: nx(:), ny(:), nzz(:)
    integer(c_int32_t), allocatable, target :: mfaces(:,:), nbrs(:,:), bcf(:), bcp(:), bcm(:), bcpts(:,:)
    integer :: nt, nr, nz, nn
    nt = cfg%geo%n_theta; nr = cfg%geo%n_r; nz = cfg%geo%n_z
    N = nt*nr*nz; NF = 3*N + nt*nz + nt*nr; NB = 2*nt*nr + 2*nt*nz; nn = nt*(nr+1)*(nz+1)
    allocate(volume(N), xi_x(3,3,N), gij(3,3,N), gij_faces(3,3,6,N), G1(N), G2(N), G3(N))
    allocate(area(NF), nrm(3,NF), tan1(3,NF), tan2(3,NF), mid(3,NF), bc_a(20,NB), nodes(3,nn))
    allocate(mfaces(6,N), nbrs(6,N), bcf(NB), bcp(NB), bcm(NB), bcpts(5,NB))
    dims = nim_dims(nt, nr, nz)
    if (cfg%device_metrics) then
      ! nodes and connectivity on the host, the metrics on the device, then
      ! the wall tables, which read the face metrics
      allocate(nx(nn), ny(nn), nzz(nn))
      call make_nodes(cfg%geo, nx, ny, nzz)
      nodes(1,:) = nx; nodes(2,:) = ny; nodes(3,:) = nzz
      call mesh_tables(cfg%geo, mfaces, nbrs, bcf, bcp, bcm)
      call check(nim_device_metrics(dims, nodes, volume, xi_x, gij, gij_faces, G1, G2, G3, area, nrm, tan1, tan2, &
                                    mid), 'nim_device_metrics')
      call bc_tables(cfg%geo, nrm, tan1, tan2, mid, mfaces, nbrs, bcf, bcp, bcm, bcpts, bc_a)
    else
      call build_geometry(cfg%geo, volume, xi_x, gij, gij_faces, G1, G2, G3, area, nrm, tan1, tan2, mid, &
                          mfaces, nbrs, bcf, bcp, bcm, bcpts, bc_a, nodes)
    endif
    ! parameters.f90:166-173
    prm%nu = 1./sqrt(cfg%Ra/(cfg%g_beta*cfg%th_tc*cfg%Pr))
    prm%Re = 1./prm%nu
    prm%alpha = prm%nu/cfg%Pr
    prm%rho = cfg%rho; prm%g_beta = cfg%g_beta; prm%DT = cfg%DT
    prm%w_u = cfg%relax; prm%w_v = cfg%relax; prm%w_w = cfg%relax; prm%w_p = cfg%relax
    prm%w_p2 = cfg%relax; prm%w_T = cfg%relax
    prm%v_limit = cfg%v_limit; prm%R_i = cfg%geo%R_i; prm%R_o = cfg%geo%R_o
    call check(nim_ctx_create(dims, prm, ctx), 'nim_ctx_create')
    g%volume = c_loc(volume); g%xi_x = c_loc(xi_x); g%gij = c_loc(gij); g%gij_faces = c_loc(gij_faces)
    g%G1 = c_loc(G1); g%G2 = c_loc(G2); g%G3 = c_loc(G3); g%faces_area = c_loc(area)
    g%faces_normal = c_loc(nrm); g%faces_mid = c_loc(mid); g%bc_points = c_loc(bcpts); g%bc_a = c_loc(bc_a)
    call check(nim_upload_geometry(ctx, g), 'nim_upload_geometry')
    nth = nt; nrr = nr; nzz = nz
    if (allocated(node_xyz)) deallocate(node_xyz)
    allocate(node_xyz(3, nn))
    node_xyz = nodes
    call check(nim_build_pressure_coeffs(ctx), 'nim_build_pressure_coeffs')
    call check(nim_set_p_measure(ctx, int(cfg%p_measure, c_int32_t)), 'nim_set_p_measure')
    if (cfg%accel > 0) then
      call check(nim_accel_options(ctx, int(cfg%accel, c_int32_t), int(cfg%accel_period, c_int32_t), &
                                   4_c_int32_t), 'nim_accel_options')
      call check(nim_accel_scope(ctx, int(cfg%accel_scope, c_int32_t)), 'nim_accel_scope')
    endif
    allocate(faces_u(NF), faces_v(NF), faces_w(NF), faces_p(NF), faces_T(NF))
    allocate(uxyz(N), vxyz(N), wxyz(N), Txyz(N), uxyz_old(N), vxyz_old(N), wxyz_old(N), Txyz_old(N))
    allocate(uxyzt(N), vxyzt(N), wxyzt(N), Txyzt(N), meshes_p(N), dila(N))
    call check(nimh_initial_state(cfg%geo, cfg%T_hot, cfg%T_cold, cfg%noise, cfg%seed, faces_u, faces_v, &
                                  faces_w, faces_p, faces_T, uxyz, vxyz, wxyz, Txyz), 'nimh_initial_state')
    uxyz_old = uxyz; vxyz_old = vxyz; wxyz_old = wxyz; Txyz_old = Txyz
    uxyzt = uxyz; vxyzt = vxyz; wxyzt = wxyz; Txyzt = Txyz
    meshes_p = 0; dila = 0
    call check(nim_upload_state(ctx, state_ptrs()), 'nim_upload_state')
    TI = cfg%DT   ! nse.f90:31-32
    STEP = 1
    DTs = cfg%DT
  end subroutine

  ! the time loop of solve_nse (nse.f90:47-75); returns the steps run
  integer function nim_solve(cfg, nsteps, restart_every, restart_prefix) result(done)
    type(nim_run_config), intent(in) :: cfg
    integer, intent(in) :: nsteps, restart_every
    character(len=*), intent(in) :: restart_prefix
    integer :: l
    integer(c_int32_t) :: iters, rc
    real(c_double) :: err(5), dudt, dTdt, dil, nu(5)
    character(len=32) :: tag
    done = 0
    do l = 1, nsteps
      call check(nim_build_momentum_coeffs(ctx), 'nim_build_momentum_coeffs')
      ! the reference's *_old arrays start at zero, so its first pass always
      ! sees the Dirichlet wall values as an increment: at least 2 passes
      if (cfg%accel > 0) then
        if (cfg%mixed_inner > 0) then
          rc = nim_inner_solve_accel(ctx, cfg%tol, int(cfg%max_iter / cfg%mixed_inner + 1, c_int32_t), 0_c_int32_t, &
                                     int(cfg%mixed_inner, c_int32_t), iters, err)
          iters = (iters - 1) * int(cfg%mixed_inner, c_int32_t)
        else
          rc = nim_inner_solve_accel(ctx, cfg%tol, int(cfg%max_iter, c_int32_t), &
                                     merge(2_c_int32_t, 0_c_int32_t, l == 1), 0_c_int32_t, iters, err)
        endif
        call check(rc, 'nim_inner_solve_accel')
      else if (cfg%mixed_inner > 0) then
        rc = nim_inner_solve_mixed(ctx, cfg%tol, int(cfg%max_iter / cfg%mixed_inner + 1, c_int32_t), 0_c_int32_t, &
                                   int(cfg%mixed_inner, c_int32_t), iters, err)
        call check(rc, 'nim_inner_solve_mixed')
        iters = (iters - 1) * int(cfg%mixed_inner, c_int32_t)
      else
        rc = nim_inner_solve(ctx, cfg%tol, int(cfg%max_iter, c_int32_t), merge(2_c_int32_t, 0_c_int32_t, l == 1), &
                             iters, err)
        call check(rc, 'nim_inner_solve')
      endif
      call check(nim_end_step(ctx, dudt, dTdt, dil), 'nim_end_step')
      ! write_data every STEP_OUT steps (nse.f90:59)
      if (cfg%step_out > 0) then
        if (mod(STEP, cfg%step_out) == 0) call nim_write_data(trim(cfg%plots_dir))
      endif
      print '(A, I6, A, I6, A, 1p, E11.4, A, E11.4, A, E11.4, A, E11.4, A, E11.4, A, E11.4)', 'step #', STEP, &
            '  | # iters. ', iters, '  | err V=', max(err(1), err(2), err(3)), '  | err p=', err(4), &
            '  | err T=', err(5), '  | D=', dil, '  | dudt=', dudt, '  | dTdt=', dTdt
      if (restart_every > 0) then
        if (mod(STEP, restart_every) == 0) then
          write(tag, '(I0)') STEP
          call nim_write_restart(trim(restart_prefix)//trim(tag))
        endif
      endif
      if (cfg%check_steady .and. dudt < cfg%dudt_epsilon .and. dTdt < cfg%dudt_epsilon) then
        print '(A, 1p, E11.4)', 'Steady state is reached, dudt = ', dudt
        if (restart_every > 0) then
          write(tag, '(I0)') STEP
          call nim_write_restart(trim(restart_prefix)//trim(tag))
        endif
        if (cfg%step_out > 0) call nim_write_data(trim(cfg%plots_dir))   ! nse.f90:66
        done = done + 1
        exit
      endif
      ! update_time_step (nse.f90:1294-1296) and find_Nu2
      STEP = STEP + 1
      TI = TI + 2*DTs
      call check(nim_advance_time(ctx), 'nim_advance_time')
      call check(nim_nusselt(ctx, nu), 'nim_nusselt')
      print '(A, 1p, E24.16, A, E24.16, A, E24.16)', 'Nu_H=', nu(1), ' |  Nu_C=', nu(2), ' | Average = ', nu(3)
      done = done + 1
    enddo
  end function

  ! write_restart (restart.f90:8-39): same records, same order
  subroutine nim_write_restart(path)
    character(len=*), intent(in) :: path
    integer :: u
    call check(nim_download_state(ctx, state_ptrs()), 'nim_download_state')
    open(newunit=u, file=path, form='unformatted', status='replace')
    write(u) TI
    write(u) STEP
    write(u) faces_u
    write(u) faces_v
    write(u) faces_w
    write(u) faces_p
    write(u) faces_T
    write(u) uxyz
    write(u) vxyz
    write(u) wxyz
    write(u) Txyz
    write(u) uxyz_old
    write(u) vxyz_old
    write(u) wxyz_old
    write(u) Txyz_old
    write(u) dila
    close(u)
  end subroutine

  ! read_restart (restart.f90:42-69) followed by update_time_step as
  ! initialize_nse does (nse.f90:104-108)
  subroutine nim_read_restart(path)
    character(len=*), intent(in) :: path
    integer :: u
    call check(nim_download_state(ctx, state_ptrs()), 'nim_download_state')
    open(newunit=u, file=path, form='unformatted', status='old')
    read(u) TI
    read(u) STEP
    read(u) faces_u
    read(u) faces_v
    read(u) faces_w
    read(u) faces_p
    read(u) faces_T
    read(u) uxyz
    read(u) vxyz
    read(u) wxyz
    read(u) Txyz
    read(u) uxyz_old
    read(u) vxyz_old
    read(u) wxyz_old
    read(u) Txyz_old
    read(u) dila
    close(u)
    call check(nim_upload_state(ctx, state_ptrs()), 'nim_upload_state')
    STEP = STEP + 1
    TI = TI + 2*DTs
    call check(nim_advance_time(ctx), 'nim_advance_time')
  end subroutine

  ! write_data (postprocess.f90:2-124) without Silo, which is not available
  ! here: the same mesh (nodes, hexahedra) and the same zone-centred single
  ! precision fields u, v, w (time-centred), p, T (time-centred) and D
  ! (dilatation), as a legacy VTK unstructured grid dir/stepNNNNN.vtk with
  ! the cycle and time of the reference's option list, and the file name
  ! appended to plots.visit
  subroutine nim_write_data(dir)
    character(len=*), intent(in) :: dir
    character(len=5) :: x1
    character(len=300) :: name
    real(c_double) :: dtime
    integer :: u, c, i, j, k, ip, nd(8)
    logical :: exist
    call check(nim_download_state(ctx, state_ptrs()), 'nim_download_state')
    dtime = STEP*2.*DTs - DTs      ! postprocess.f90:47-48
    if (STEP == 0) dtime = 0.0
    write(x1, '(I5.5)') STEP
    name = trim(dir)//'/step'//x1//'.vtk'
    call execute_command_line('mkdir -p '//trim(dir))
    print '(A, 1p, E13.6, A, I0)', 'Printing data at time = ', dtime, ' | Time step = ', STEP
    open(newunit=u, file=trim(name), status='replace', form='formatted', action='write')
    write(u, '(A)') '# vtk DataFile Version 3.0'
    write(u, '(A, I0, A, 1p, E15.8)') 'Navier Stokes Equation Solution, cycle ', STEP, ', time ', dtime
    write(u, '(A)') 'ASCII'
    write(u, '(A)') 'DATASET UNSTRUCTURED_GRID'
    write(u, '(A, I0, A)') 'POINTS ', size(node_xyz, 2), ' float'
    do i = 1, size(node_xyz, 2)
      write(u, '(1p, 3E16.8)') real(node_xyz(:, i), 4)
    enddo
    ! hexahedra: node i + nth (j + (nrr + 1) k) (0-based), bottom quad r
    ! first then theta (counter-clockwise seen from +z), then the top quad
    write(u, '(A, I0, A, I0)') 'CELLS ', N, ' ', 9*N
    do c = 0, N - 1
      i = mod(c, nth); j = mod(c / nth, nrr); k = c / (nth*nrr); ip = mod(i + 1, nth)
      nd = (/ i + nth*(j + (nrr+1)*k), i + nth*(j + 1 + (nrr+1)*k), ip + nth*(j + 1 + (nrr+1)*k), &
              ip + nth*(j + (nrr+1)*k), i + nth*(j + (nrr+1)*(k+1)), i + nth*(j + 1 + (nrr+1)*(k+1)), &
              ip + nth*(j + 1 + (nrr+1)*(k+1)), ip + nth*(j + (nrr+1)*(k+1)) /)
      write(u, '(I0, 8(1X, I0))') 8, nd
    enddo
    write(u, '(A, I0)') 'CELL_TYPES ', N
    do c = 1, N
      write(u, '(A)') '12'
    enddo
    write(u, '(A, I0)') 'CELL_DATA ', N
    call field('u', uxyzt)
    call field('v', vxyzt)
    call field('w', wxyzt)
    call field('p', meshes_p)
    call field('T', Txyzt)
    call field('D', dila)
    close(u)
    inquire(file='plots.visit', exist=exist)
    if (exist) then
      open(newunit=u, file='plots.visit', status='old', position='append', action='write')
    else
      open(newunit=u, file='plots.visit', status='new', action='write')
    endif
    write(u, '(A)') trim(name)
    close(u)
  contains
    subroutine field(fname, v)
      character(len=*), intent(in) :: fname
      real(c_double), intent(in) :: v(:)
      integer :: q
      write(u, '(A, A, A)') 'SCALARS ', fname, ' float 1'
      write(u, '(A)') 'LOOKUP_TABLE default'
      do q = 1, size(v)
        write(u, '(1p, E16.8)') real(v(q), 4)
      enddo
    end subroutine
  end subroutine

  subroutine nim_finish()
    integer(c_int32_t) :: rc
    if (c_associated(ctx)) rc = nim_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine

end module nim_nse
