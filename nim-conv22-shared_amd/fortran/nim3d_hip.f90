! Drop-in driver: the reference's main program (main.f90:5-75) with the time
! step on the MI355X. Reads a namelist file (default nim3d.nml) describing
! the case; every knob that the reference fixes at compile time
! (parameters.f90:84-173) is a run-time input here.
!
! &case n_theta=36, n_r=6, n_z=12, R_i=1.0471975511966, R_o=2., height=2.,
!       blend_in=1.25, blend_out=1.75, Ra=1e5, Pr=0.7, DT=0.01, nsteps=3,
!       tol=1e-6,1e-6,1e-6,5e-3,1e-6, noise=1e-3, seed=20221,
!       restart_every=0, restart_prefix='restart/rest.step', restart_in='',
!       final_out='', mixed_inner=0 /
! step_out > 0 writes the cell fields every step_out steps (the reference's
! STEP_OUT, parameters.f90:138) to plots_dir/stepNNNNN.vtk (legacy VTK in
! place of Silo) and lists them in plots.visit.
! p_measure=1 makes tol(4) bound half the range of the pressure increments
! instead of their largest magnitude (which counts p's constant-level drift).
! mixed_inner > 0 selects the mixed-precision inner loop (FP64 residuals,
! that many FP32 correction passes per residual). accel=m > 0 extrapolates
! the inner loop's passes (or mixed cycles) every accel_period of them over
! the last m steps (Anderson acceleration; the same stopping test and fixed
! point, far fewer passes for p at tight tolerances) over the pressure faces
! (accel_scope=1) or the whole iterate (0). device_metrics=.false.
! computes the per-cell metrics on the host instead of the device (the same
! bits either way).
program nim3d_hip
  use iso_c_binding
  use nim_host_geometry
  use nim_nse
  implicit none
  type(nim_run_config) :: cfg
  integer :: n_theta = 36, n_r = 6, n_z = 12, nsteps = 3, restart_every = 0, done, u, ios, mixed_inner = 0
  integer :: step_out = 0, p_measure = 0, accel = 0, accel_period = 1, accel_scope = 1
  real(c_double) :: R_i = 1.0d0*(3.14159265359d0/3.0d0), R_o = 2.0d0, height = 2.0d0
  real(c_double) :: blend_in = 1.25d0, blend_out = 1.75d0, grading = 1.5d0, hex_order = -1.0d0
  real(c_double) :: Ra = 1.0d5, Pr = 0.7d0, DT = 0.01d0, noise = 1.0d-3, tol(5) = (/1d-6, 1d-6, 1d-6, 5d-3, 1d-6/)
  real(c_double) :: T_hot = 1.0d0, T_cold = 0.0d0
  integer(c_int64_t) :: seed = 20221
  logical :: check_steady = .true., device_metrics = .true.
  character(len=256) :: nml = 'nim3d.nml', restart_prefix = 'restart/rest.step', restart_in = '', final_out = ''
  character(len=256) :: plots_dir = 'plots'
  namelist /case/ n_theta, n_r, n_z, R_i, R_o, height, blend_in, blend_out, grading, hex_order, Ra, Pr, DT, &
                  nsteps, tol, noise, seed, T_hot, T_cold, restart_every, restart_prefix, restart_in, &
                  check_steady, final_out, mixed_inner, device_metrics, step_out, plots_dir, &
                  p_measure, accel, accel_period, accel_scope

  if (command_argument_count() >= 1) call get_command_argument(1, nml)
  open(newunit=u, file=trim(nml), status='old', iostat=ios)
  if (ios == 0) then
    read(u, nml=case)
    close(u)
  endif
  cfg%geo = nim_geo_case(n_theta, n_r, n_z, 0, R_i, R_o, height, blend_in, blend_out, grading, hex_order, &
                         3.14159265359d0)
  cfg%Ra = Ra; cfg%Pr = Pr; cfg%DT = DT; cfg%tol = tol; cfg%noise = noise; cfg%seed = seed
  cfg%T_hot = T_hot; cfg%T_cold = T_cold; cfg%check_steady = check_steady; cfg%mixed_inner = mixed_inner
  cfg%device_metrics = device_metrics
  cfg%step_out = step_out; cfg%plots_dir = plots_dir; cfg%p_measure = p_measure
  cfg%accel = accel; cfg%accel_period = accel_period; cfg%accel_scope = accel_scope
  call nim_setup(cfg)
  if (len_trim(restart_in) > 0) call nim_read_restart(trim(restart_in))
  done = nim_solve(cfg, nsteps, restart_every, trim(restart_prefix))
  if (len_trim(final_out) > 0) call nim_write_restart(trim(final_out))
  call nim_finish()
  print '(A, I0, A)', 'completed ', done, ' steps'
end program nim3d_hip
