! C-callable entry points of the Fortran host library (libnim_host.so):
! case geometry and initial state in the reference's layout, ready for
! nim_upload_geometry / nim_upload_state of libnim_hip.so. Python test and
! benchmark drivers call these through ctypes; a Fortran host uses the module
! procedures directly.
module nim_host_capi
  use iso_c_binding
  use nim_host_geometry
  implicit none

  interface
    ! seeded uniform value in [-1, 1) for (seed, field, cell); nim_host_noise.c
    function nimh_noise(seed, field, idx) bind(C, name="nimh_noise") result(r)
      import :: c_int64_t, c_int32_t, c_double
      integer(c_int64_t), value :: seed
      integer(c_int32_t), value :: field
      integer(c_int64_t), value :: idx
      real(c_double) :: r
    end function
  end interface

contains

  integer(c_int32_t) function nimh_build_geometry(cs, volume, xi_x, gij, gij_faces, G1, G2, G3, faces_area, &
      faces_normal, faces_tangential, faces_bitangential, faces_mid, meshes_faces, meshes_neighbors, &
      bc_faces, bc_position, bc_mesh, bc_points, bc_a, node_xyz) bind(C, name="nimh_build_geometry") result(rc)
    type(nim_geo_case), intent(in) :: cs
    real(c_double), intent(out) :: volume(*), xi_x(*), gij(*), gij_faces(*), G1(*), G2(*), G3(*)
    real(c_double), intent(out) :: faces_area(*), faces_normal(*), faces_tangential(*), faces_bitangential(*)
    real(c_double), intent(out) :: faces_mid(*), bc_a(*), node_xyz(*)
    integer(c_int32_t), intent(out) :: meshes_faces(*), meshes_neighbors(*), bc_faces(*), bc_position(*)
    integer(c_int32_t), intent(out) :: bc_mesh(*), bc_points(*)
    integer :: N, NF, NB, NN
    rc = 1
    if (cs%n_theta < 4 .or. mod(cs%n_theta, 2) /= 0 .or. cs%n_r < 3 .or. cs%n_z < 3) return
    if (.not. (cs%R_o > cs%R_i .and. cs%R_i > 0 .and. cs%height > 0)) return
    N = cs%n_theta*cs%n_r*cs%n_z
    NF = 3*N + cs%n_theta*cs%n_z + cs%n_theta*cs%n_r
    NB = 2*cs%n_theta*cs%n_r + 2*cs%n_theta*cs%n_z
    NN = cs%n_theta*(cs%n_r+1)*(cs%n_z+1)
    call build_geometry(cs, volume(1:N), xi_x3(xi_x, N), xi_x3(gij, N), gf4(gij_faces, N), G1(1:N), G2(1:N), &
                        G3(1:N), faces_area(1:NF), v2(faces_normal, 3, NF), v2(faces_tangential, 3, NF), &
                        v2(faces_bitangential, 3, NF), v2(faces_mid, 3, NF), i2(meshes_faces, 6, N), &
                        i2(meshes_neighbors, 6, N), bc_faces(1:NB), bc_position(1:NB), bc_mesh(1:NB), &
                        i2(bc_points, 5, NB), v2(bc_a, 20, NB), v2(node_xyz, 3, NN))
    rc = 0
  contains
    function xi_x3(a, n) result(p)
      real(c_double), target, intent(inout) :: a(3,3,*)
      integer, intent(in) :: n
      real(c_double), pointer :: p(:,:,:)
      p => a(:,:,1:n)
    end function
    function gf4(a, n) result(p)
      real(c_double), target, intent(inout) :: a(3,3,6,*)
      integer, intent(in) :: n
      real(c_double), pointer :: p(:,:,:,:)
      p => a(:,:,:,1:n)
    end function
    function v2(a, m, n) result(p)
      integer, intent(in) :: m, n
      real(c_double), target, intent(inout) :: a(m,*)
      real(c_double), pointer :: p(:,:)
      p => a(:,1:n)
    end function
    function i2(a, m, n) result(p)
      integer, intent(in) :: m, n
      integer(c_int32_t), target, intent(inout) :: a(m,*)
      integer(c_int32_t), pointer :: p(:,:)
      p => a(:,1:n)
    end function
  end function

  ! The parts of nimh_build_geometry around the per-cell metrics, for a host
  ! that evaluates those on the device (nim_device_metrics of libnim_hip.so):
  ! the deformed nodes (node_xyz(3,NNODE)), the cell/boundary tables, and
  ! the pressure-wall tables that need the face vectors.
  integer(c_int32_t) function nimh_nodes(cs, node_xyz) bind(C, name="nimh_nodes") result(rc)
    type(nim_geo_case), intent(in) :: cs
    real(c_double), intent(out) :: node_xyz(3, *)
    real(c_double), allocatable :: nx(:), ny(:), nz(:)
    integer :: NN
    NN = cs%n_theta*(cs%n_r+1)*(cs%n_z+1)
    allocate(nx(NN), ny(NN), nz(NN))
    call make_nodes(cs, nx, ny, nz)
    node_xyz(1, 1:NN) = nx; node_xyz(2, 1:NN) = ny; node_xyz(3, 1:NN) = nz
    rc = 0
  end function

  integer(c_int32_t) function nimh_mesh_tables(cs, meshes_faces, meshes_neighbors, bc_faces, bc_position, &
      bc_mesh) bind(C, name="nimh_mesh_tables") result(rc)
    type(nim_geo_case), intent(in) :: cs
    integer(c_int32_t), intent(out) :: meshes_faces(6, *), meshes_neighbors(6, *), bc_faces(*), bc_position(*)
    integer(c_int32_t), intent(out) :: bc_mesh(*)
    integer :: N, NB
    N = cs%n_theta*cs%n_r*cs%n_z
    NB = 2*cs%n_theta*cs%n_r + 2*cs%n_theta*cs%n_z
    call mesh_tables(cs, meshes_faces(:, 1:N), meshes_neighbors(:, 1:N), bc_faces(1:NB), bc_position(1:NB), &
                     bc_mesh(1:NB))
    rc = 0
  end function

  integer(c_int32_t) function nimh_bc_tables(cs, faces_normal, faces_tangential, faces_bitangential, faces_mid, &
      meshes_faces, meshes_neighbors, bc_faces, bc_position, bc_mesh, bc_points, bc_a) &
      bind(C, name="nimh_bc_tables") result(rc)
    type(nim_geo_case), intent(in) :: cs
    real(c_double), intent(in) :: faces_normal(3, *), faces_tangential(3, *), faces_bitangential(3, *)
    real(c_double), intent(in) :: faces_mid(3, *)
    integer(c_int32_t), intent(in) :: meshes_faces(6, *), meshes_neighbors(6, *), bc_faces(*), bc_position(*)
    integer(c_int32_t), intent(in) :: bc_mesh(*)
    integer(c_int32_t), intent(out) :: bc_points(5, *)
    real(c_double), intent(out) :: bc_a(20, *)
    integer :: N, NF, NB
    N = cs%n_theta*cs%n_r*cs%n_z
    NF = 3*N + cs%n_theta*cs%n_z + cs%n_theta*cs%n_r
    NB = 2*cs%n_theta*cs%n_r + 2*cs%n_theta*cs%n_z
    call bc_tables(cs, faces_normal(:, 1:NF), faces_tangential(:, 1:NF), faces_bitangential(:, 1:NF), &
                   faces_mid(:, 1:NF), meshes_faces(:, 1:N), meshes_neighbors(:, 1:N), bc_faces(1:NB), &
                   bc_position(1:NB), bc_mesh(1:NB), bc_points(:, 1:NB), bc_a(:, 1:NB))
    rc = 0
  end function

  ! Initial state: velocities and pressure zero on every face, the Dirichlet
  ! wall temperatures on the hexagon (T_hot) and the cylinder (T_cold)
  ! (boundary_condition_4/5, general.f90:85-117; apply_dirichlet_bc_i returns
  ! the constant exactly), top/bottom faces 0 (adiabatic, set by the first
  ! sweep), and cell averages of u, v, w, T perturbed by amp*noise on every
  ! time level (the reference starts from rest, general.f90:16-31).
  integer(c_int32_t) function nimh_initial_state(cs, T_hot, T_cold, amp, seed, faces_u, faces_v, faces_w, &
      faces_p, faces_T, uxyz, vxyz, wxyz, Txyz) bind(C, name="nimh_initial_state") result(rc)
    type(nim_geo_case), intent(in) :: cs
    real(c_double), value :: T_hot, T_cold, amp
    integer(c_int64_t), value :: seed
    real(c_double), intent(out) :: faces_u(*), faces_v(*), faces_w(*), faces_p(*), faces_T(*)
    real(c_double), intent(out) :: uxyz(*), vxyz(*), wxyz(*), Txyz(*)
    integer :: N, NF, nt, nr, nz, o5, o6, i, k
    integer(c_int64_t) :: c
    nt = cs%n_theta; nr = cs%n_r; nz = cs%n_z
    N = nt*nr*nz
    NF = 3*N + nt*nz + nt*nr
    faces_u(1:NF) = 0; faces_v(1:NF) = 0; faces_w(1:NF) = 0; faces_p(1:NF) = 0; faces_T(1:NF) = 0
    ! hexagon faces follow bottom and top in the face numbering
    o5 = 3*N + nt*nz + nt*nr - 2*nt*nz
    o6 = o5 + nt*nz
    do k = 0, nz-1
      do i = 0, nt-1
        faces_T(o5 + i + nt*k + 1) = T_hot
        faces_T(o6 + i + nt*k + 1) = T_cold
      enddo
    enddo
    do c = 0, N-1
      uxyz(c+1) = amp*nimh_noise(seed, 0, c)
      vxyz(c+1) = amp*nimh_noise(seed, 1, c)
      wxyz(c+1) = amp*nimh_noise(seed, 2, c)
      Txyz(c+1) = amp*nimh_noise(seed, 3, c)
    enddo
    rc = 0
  end function

end module nim_host_capi
