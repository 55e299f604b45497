! ISO_C_BINDING interface to libnim_hip.so (include/nim_hip.h).
!
! This is the binding a Fortran host of the reference would add to call the
! MI355X time step in place of its own module subroutines: each interface
! names the reference routine it replaces.
module nim_hip_api
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: NIM_OK = 0, NIM_E_ARG = 1, NIM_E_HIP = 2, NIM_E_MAXITER = 3, &
                                   NIM_E_NONFINITE = 4, NIM_E_STATE = 5

  type, bind(C) :: nim_dims
    integer(c_int32_t) :: n_theta, n_r, n_z
  end type

  type, bind(C) :: nim_params
    real(c_double) :: nu, Re, alpha, rho, g_beta, DT
    real(c_double) :: w_u, w_v, w_w, w_p, w_p2, w_T
    real(c_double) :: v_limit, R_i, R_o
  end type

  type, bind(C) :: nim_geometry
    type(c_ptr) :: volume, xi_x, gij, gij_faces, G1, G2, G3, faces_area, faces_normal, faces_mid, bc_points, bc_a
  end type

  type, bind(C) :: nim_state
    type(c_ptr) :: faces_u, faces_v, faces_w, faces_p, faces_T
    type(c_ptr) :: uxyz, vxyz, wxyz, Txyz
    type(c_ptr) :: uxyz_old, vxyz_old, wxyz_old, Txyz_old
    type(c_ptr) :: uxyzt, vxyzt, wxyzt, Txyzt
    type(c_ptr) :: meshes_p, meshes_dila
  end type

  interface
    ! static allocation of parameters.f90:222-292
    integer(c_int32_t) function nim_ctx_create(dims, params, ctx) bind(C, name="nim_ctx_create")
      import :: c_int32_t, c_ptr, nim_dims, nim_params
      type(nim_dims), intent(in) :: dims
      type(nim_params), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int32_t) function nim_ctx_destroy(ctx) bind(C, name="nim_ctx_destroy")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! evaluate_all_mesh_constants / evaluate_pressure_bc_array results
    integer(c_int32_t) function nim_upload_geometry(ctx, geo) bind(C, name="nim_upload_geometry")
      import :: c_int32_t, c_ptr, nim_geometry
      type(c_ptr), value :: ctx
      type(nim_geometry), intent(in) :: geo
    end function
    ! the per-cell metrics of evaluate_all_mesh_constants (geometry.f90:395-632)
    ! on the device, bit for bit those of build_geometry, from the mesh nodes
    integer(c_int32_t) function nim_device_metrics(dims, node_xyz, volume, xi_x, gij, gij_faces, G1, G2, G3, &
        faces_area, faces_normal, faces_tangential, faces_bitangential, faces_mid) bind(C, name="nim_device_metrics")
      import :: c_int32_t, c_double, nim_dims
      type(nim_dims), intent(in) :: dims
      real(c_double), intent(in) :: node_xyz(*)
      real(c_double), intent(out) :: volume(*), xi_x(*), gij(*), gij_faces(*), G1(*), G2(*), G3(*)
      real(c_double), intent(out) :: faces_area(*), faces_normal(*), faces_tangential(*), faces_bitangential(*)
      real(c_double), intent(out) :: faces_mid(*)
    end function
    ! calculate_pressure_nim_constants (nim.f90:556-743)
    integer(c_int32_t) function nim_build_pressure_coeffs(ctx) bind(C, name="nim_build_pressure_coeffs")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! read_restart / initial conditions, write_restart / write_data
    integer(c_int32_t) function nim_upload_state(ctx, st) bind(C, name="nim_upload_state")
      import :: c_int32_t, c_ptr, nim_state
      type(c_ptr), value :: ctx
      type(nim_state), intent(in) :: st
    end function
    integer(c_int32_t) function nim_download_state(ctx, st) bind(C, name="nim_download_state")
      import :: c_int32_t, c_ptr, nim_state
      type(c_ptr), value :: ctx
      type(nim_state), intent(in) :: st
    end function
    ! calculate_velocity_nim_constants_formulation2 (nim.f90:160-330)
    integer(c_int32_t) function nim_build_momentum_coeffs(ctx) bind(C, name="nim_build_momentum_coeffs")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! integrate_nse loop (nse.f90:117-165)
    integer(c_int32_t) function nim_inner_solve(ctx, tol, max_iter, min_iter, iters, err) &
        bind(C, name="nim_inner_solve")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tol(5)
      integer(c_int32_t), value :: max_iter, min_iter
      integer(c_int32_t), intent(out) :: iters
      real(c_double), intent(out) :: err(5)
    end function
    ! mixed-precision form of the same loop: FP64 residuals, `inner` FP32
    ! correction passes per residual, FP64 update (no reference counterpart)
    integer(c_int32_t) function nim_inner_solve_mixed(ctx, tol, max_cycles, min_cycles, inner, cycles, err) &
        bind(C, name="nim_inner_solve_mixed")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tol(5)
      integer(c_int32_t), value :: max_cycles, min_cycles, inner
      integer(c_int32_t), intent(out) :: cycles
      real(c_double), intent(out) :: err(5)
    end function
    ! the same loop (inner = 0: FP64 passes; inner > 0: mixed cycles) with
    ! Anderson acceleration of its passes: the plain loop's stopping test and
    ! fixed point (no reference counterpart)
    integer(c_int32_t) function nim_inner_solve_accel(ctx, tol, max_iter, min_iter, inner, iters, err) &
        bind(C, name="nim_inner_solve_accel")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tol(5)
      integer(c_int32_t), value :: max_iter, min_iter, inner
      integer(c_int32_t), intent(out) :: iters
      real(c_double), intent(out) :: err(5)
    end function
    ! history depth, passes (cycles) per extrapolation, bytes per history value
    integer(c_int32_t) function nim_accel_options(ctx, depth, period, basis_bytes) &
        bind(C, name="nim_accel_options")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: depth, period, basis_bytes
    end function
    ! extrapolated part: 0 the whole iterate, 1 the pressure faces (default)
    integer(c_int32_t) function nim_accel_scope(ctx, scope) bind(C, name="nim_accel_scope")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: scope
    end function
    ! what tol(4) bounds: 0 max |dp| (find_error), 1 half the range of dp
    integer(c_int32_t) function nim_set_p_measure(ctx, mode) bind(C, name="nim_set_p_measure")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: mode
    end function
    ! FP64 residual measure of the current state
    integer(c_int32_t) function nim_residual(ctx, err) bind(C, name="nim_residual")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: err(5)
    end function
    ! check_steady_state + evaluate_D (nse.f90:167-168)
    integer(c_int32_t) function nim_end_step(ctx, dudt, dTdt, dil) bind(C, name="nim_end_step")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dudt, dTdt, dil
    end function
    ! update_time_step (nse.f90:1292-1313)
    integer(c_int32_t) function nim_advance_time(ctx) bind(C, name="nim_advance_time")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    ! find_Nu2 (nse.f90:2699-2798)
    integer(c_int32_t) function nim_nusselt(ctx, nu) bind(C, name="nim_nusselt")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: nu(5)
    end function
    function nim_last_error() bind(C, name="nim_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
  end interface

contains

  ! message of the last failed call, as a Fortran string
  function nim_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    interface
      integer(c_size_t) function c_strlen(str) bind(C, name="strlen")
        import :: c_ptr, c_size_t
        type(c_ptr), value :: str
      end function
    end interface
    p = nim_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function

end module nim_hip_api
