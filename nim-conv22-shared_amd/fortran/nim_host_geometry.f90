! Host-side mesh and metric setup of the NIM case, with run-time sizes.
!
! Restates the reference's one-time setup so that the hot path receives the
! same inputs the reference computes for itself:
!   create_hollow            create_geometry.f90:321-622  (structured O-grid)
!   deform_nodes             general.f90:428-484, 700-788 (radial grading,
!                                                          circle->hexagon)
!   prepare_mesh / link_mesh prepare_mesh.f90:11-183      (local frames; on
!                            this O-grid they are the structured frames below)
!   mapping coefficients     create_mesh.f90:254-303
!   evaluate_curvilinear     curvilinear.f90:9-160
!   volume_calculate         create_mesh.f90:306-322
!   find_normal              geometry.f90:76-190
!   calculate_diffusion_tensor_faces_i   geometry.f90:395-497
!   convection_diffusion_constants_mesh_i geometry.f90:500-632
!   evaluate_pressure_bc_array           geometry.f90:879-975
!   fill_global_arrays (BC tables)       geometry.f90:978-1032
! The reference sizes everything at compile time (parameters.f90:84-123) and
! hard-codes the case; here the O-grid size, radii, height, blend radii and
! grading exponent are arguments. Expression order follows the reference so
! that on the reference's own case the arrays agree with it to rounding.
!
! Local frame of cell (i,j,k) (0-based theta, r, z indices): nodes
! 1..8 = (i,j,k) (i,j+1,k) (i+1,j+1,k) (i+1,j,k) and the same at k+1, so the
! map coordinate xi1 runs outward in r, xi2 along +theta, xi3 along +z and the
! faces 1..6 are inner-r, outer-r, -theta, +theta, bottom, top
! (mesh.f90:65-69). Where two cells compute a value of a shared face, the
! higher-numbered cell's value is kept, as the reference's serial loop
! (OMP_NUM_THREADS=1, run.sh:2) leaves it.
module nim_host_geometry
  use iso_c_binding
  implicit none
  private
  public :: nim_geo_case, build_geometry, gauss_points, make_nodes, mesh_tables, bc_tables

  integer, parameter :: dp = c_double

  ! Five-point Gauss rule in the reference's order (parameters.f90:18-22)
  integer, parameter :: NG = 5

  type, bind(C) :: nim_geo_case
    integer(c_int32_t) :: n_theta, n_r, n_z, pad
    real(c_double) :: R_i         ! hexagon circumradius (inner annulus radius before deformation)
    real(c_double) :: R_o         ! cylinder radius
    real(c_double) :: height      ! axial length, z in [0, height]
    real(c_double) :: blend_in    ! radius up to which nodes become fully hexagonal (general.f90:433: 1.25)
    real(c_double) :: blend_out   ! radius beyond which nodes stay circular (general.f90:433: 1.75)
    real(c_double) :: grading     ! wall refinement exponent (general.f90:430: 1.5)
    real(c_double) :: hex_order   ! blending exponent (general.f90:433: -1)
    real(c_double) :: pi_const    ! the reference's MATH_PI (parameters.f90:31)
  end type

  type :: curvi
    real(dp) :: x, y, z, detJ
    real(dp) :: jac(3,3), jinv(3,3), g_ij(3,3), gij(3,3), gam(3,3,3)
    real(dp) :: gco(3,3)   ! covariant base vectors g_k = rows of jac
    real(dp) :: gcon(3,3)  ! contravariant base vectors g^k = columns of jinv
  end type

contains

  subroutine gauss_points(x, w)
    real(dp), intent(out) :: x(NG), w(NG)
    x = (/ -1._dp/3._dp*sqrt(5._dp-2._dp*sqrt(10._dp/7._dp)), 1._dp/3._dp*sqrt(5._dp-2._dp*sqrt(10._dp/7._dp)), 0._dp, &
           -1._dp/3._dp*sqrt(5._dp+2._dp*sqrt(10._dp/7._dp)), 1._dp/3._dp*sqrt(5._dp+2._dp*sqrt(10._dp/7._dp)) /)
    w = (/ (322._dp+13._dp*sqrt(70._dp))/900._dp, (322._dp+13._dp*sqrt(70._dp))/900._dp, 128._dp/225._dp, &
           (322._dp-13._dp*sqrt(70._dp))/900._dp, (322._dp-13._dp*sqrt(70._dp))/900._dp /)
  end subroutine

  ! ------------------------------------------------------------------
  ! node generation and deformation
  ! ------------------------------------------------------------------
  subroutine make_nodes(cs, nx, ny, nz)
    type(nim_geo_case), intent(in) :: cs
    real(dp), intent(out) :: nx(:), ny(:), nz(:)
    integer :: i, j, k, n, nt, nr, nzc
    real(dp) :: PI, dr, dtheta, dz, r, th, x, y, rr, Roi, rn
    real(dp) :: p1, p2, Rin, Ron
    nt = cs%n_theta; nr = cs%n_r; nzc = cs%n_z
    PI = 4.D0*DATAN(1.D0)
    dr = (cs%R_o - cs%R_i)/nr
    dtheta = 2*PI/(nt)
    dz = cs%height/nzc
    n = 0
    do k = 1, nzc+1
      do j = 1, nr+1
        do i = 1, nt
          n = n + 1
          r = cs%R_i + (j-1)*dr
          th = (i-1)*dtheta
          nx(n) = r*cos(th)
          ny(n) = r*sin(th)
          ! create_hollow shifts by (cx, cy, cz) = (0, 0, L/2) (main.f90:36)
          nz(n) = ((k-1)*dz - cs%height/2._dp) + cs%height/2._dp
          nx(n) = nx(n) + 0._dp
          ny(n) = ny(n) + 0._dp
        enddo
      enddo
    enddo
    ! refine_walls_radial_order (general.f90:463-484)
    Roi = cs%R_o - cs%R_i
    do n = 1, size(nx)
      x = nx(n); y = ny(n)
      rr = sqrt(x*x + y*y) - cs%R_i
      th = atan(y/(x))
      if (x < 0) th = th + cs%pi_const
      rr = rr/Roi
      if (rr <= 0.5_dp .and. rr > 0._dp) rr = ((2*rr)**cs%grading)/2._dp
      if (rr > 0.5_dp .and. rr < 1._dp) rr = 1._dp - ((-2._dp*(rr-1))**cs%grading)/2._dp
      rr = (rr*Roi + cs%R_i)
      nx(n) = rr*cos(th)
      ny(n) = rr*sin(th)
    enddo
    ! create_hexagon_from_annulus_order_n_flipped_three_regions (general.f90:700-722)
    Rin = cs%blend_in; Ron = cs%blend_out; rn = cs%hex_order
    do n = 1, size(nx)
      x = nx(n); y = ny(n)
      rr = sqrt(x*x + y*y)
      call circle_to_hexagon(cs%pi_const, x, y, p1, p2)
      if (rr <= Rin) then
        nx(n) = p1
        ny(n) = p2
      elseif (rr > Rin .and. rr < Ron) then
        nx(n) = -(x-p1)/(rpow(Rin,rn)-rpow(Ron,rn))*rpow(rr,rn) - (p1*rpow(Ron,rn)-x*rpow(Rin,rn))/(rpow(Rin,rn)-rpow(Ron,rn))
        ny(n) = -(y-p2)/(rpow(Rin,rn)-rpow(Ron,rn))*rpow(rr,rn) - (p2*rpow(Ron,rn)-y*rpow(Rin,rn))/(rpow(Rin,rn)-rpow(Ron,rn))
      endif
    enddo
  end subroutine

  ! real power as the reference's x**n with a real exponent; n = -1 reduces
  ! to a reciprocal exactly
  pure real(dp) function rpow(x, n)
    real(dp), intent(in) :: x, n
    if (n == -1._dp) then
      rpow = 1._dp/x
    else
      rpow = x**n
    endif
  end function

  ! convert_circle_to_hexagon_mine (general.f90:726-788): radial projection
  ! of a point on the circle of radius R onto the hexagon of circumradius R
  ! with vertices at theta = 0, 60, ... degrees
  subroutine circle_to_hexagon(MP, x1, y1, x2, y2)
    real(dp), intent(in) :: MP, x1, y1
    real(dp), intent(out) :: x2, y2
    real(dp) :: theta, R, a1, b1, c1, a2, b2, c2, pp1, pp2
    theta = atan(y1/(x1))
    if (x1 < 0) theta = theta + MP
    R = sqrt(x1*x1 + y1*y1)
    x2 = x1
    y2 = y1
    if (theta > 0._dp .and. theta < MP/3._dp) then
      pp1 = R; pp2 = 0._dp
      a1 = (sin(0._dp)-sin(MP/3._dp))/(cos(0._dp)-cos(MP/3._dp))
      b1 = -1._dp; c1 = pp2 - a1*pp1; a2 = y1/x1; b2 = -1._dp; c2 = 0._dp
      x2 = (b1*c2-b2*c1)/(a1*b2-a2*b1)
      y2 = (c1*a2-c2*a1)/(a1*b2-a2*b1)
    elseif (theta > MP/3._dp .and. theta < 2*MP/3._dp) then
      y2 = R*sin(MP/3._dp)
      x2 = x1*y2/y1
    elseif (theta > 2*MP/3._dp .and. theta < 3*MP/3._dp) then
      pp1 = -R; pp2 = 0._dp
      a1 = (sin(2*MP/3._dp)-sin(3*MP/3._dp))/(cos(2*MP/3._dp)-cos(3*MP/3._dp))
      b1 = -1._dp; c1 = pp2 - a1*pp1; a2 = y1/x1; b2 = -1._dp; c2 = 0._dp
      x2 = (b1*c2-b2*c1)/(a1*b2-a2*b1)
      y2 = (c1*a2-c2*a1)/(a1*b2-a2*b1)
    elseif (theta > 3*MP/3._dp .and. theta < 4*MP/3._dp) then
      pp1 = -R; pp2 = 0._dp
      a1 = (sin(3*MP/3._dp)-sin(4*MP/3._dp))/(cos(3*MP/3._dp)-cos(4*MP/3._dp))
      b1 = -1._dp; c1 = pp2 - a1*pp1; a2 = y1/x1; b2 = -1._dp; c2 = 0._dp
      x2 = (b1*c2-b2*c1)/(a1*b2-a2*b1)
      y2 = (c1*a2-c2*a1)/(a1*b2-a2*b1)
    elseif ((theta > 4*MP/3._dp .and. theta < 5*MP/3._dp) .or. (theta > -2*MP/3._dp .and. theta < -MP/3._dp)) then
      y2 = R*sin(4*MP/3._dp)
      x2 = x1*y2/y1
    elseif (theta > -MP/3._dp .and. theta < 0._dp) then
      pp1 = R; pp2 = 0._dp
      a1 = (sin(-MP/3._dp)-sin(0._dp))/(cos(-MP/3._dp)-cos(0._dp))
      b1 = -1._dp; c1 = pp2 - a1*pp1; a2 = y1/x1; b2 = -1._dp; c2 = 0._dp
      x2 = (b1*c2-b2*c1)/(a1*b2-a2*b1)
      y2 = (c1*a2-c2*a1)/(a1*b2-a2*b1)
    endif
  end subroutine

  ! ------------------------------------------------------------------
  ! trilinear map of one cell (create_mesh.f90:254-303, curvilinear.f90)
  ! ------------------------------------------------------------------
  subroutine map_coeffs(x, y, z, C, D, E)
    real(dp), intent(in) :: x(8), y(8), z(8)
    real(dp), intent(out) :: C(0:7), D(0:7), E(0:7)
    call one(x, C)
    call one(y, D)
    call one(z, E)
  contains
    subroutine one(v, A)
      real(dp), intent(in) :: v(8)
      real(dp), intent(out) :: A(0:7)
      A(0)=(v(7)+v(8)+v(5)+v(6)+v(3)+v(4)+v(1)+v(2))/8._dp
      A(1)=(v(7)-v(8)-v(5)+v(6)+v(3)-v(4)-v(1)+v(2))/8._dp
      A(2)=(v(7)+v(8)-v(5)-v(6)+v(3)+v(4)-v(1)-v(2))/8._dp
      A(3)=(v(7)+v(8)+v(5)+v(6)-v(3)-v(4)-v(1)-v(2))/8._dp
      A(4)=(v(7)-v(8)+v(5)-v(6)+v(3)-v(4)+v(1)-v(2))/8._dp
      A(5)=(v(7)-v(8)-v(5)+v(6)-v(3)+v(4)+v(1)-v(2))/8._dp
      A(6)=(v(7)+v(8)-v(5)-v(6)-v(3)-v(4)+v(1)+v(2))/8._dp
      A(7)=(v(7)-v(8)+v(5)-v(6)-v(3)+v(4)-v(1)+v(2))/8._dp
    end subroutine
  end subroutine

  pure real(dp) function det3(a)
    real(dp), intent(in) :: a(0:2,0:2)
    det3 = a(0,0)*a(1,1)*a(2,2) + a(0,1)*a(1,2)*a(2,0) + a(0,2)*a(1,0)*a(2,1) &
         - a(0,2)*a(1,1)*a(2,0) - a(0,1)*a(1,0)*a(2,2) - a(0,0)*a(1,2)*a(2,1)
  end function

  pure real(dp) function dot3(a, b)
    real(dp), intent(in) :: a(3), b(3)
    dot3 = a(1)*b(1) + a(2)*b(2) + a(3)*b(3)
  end function

  subroutine curvilinear(C, D, E, x, y, z, cv, want_gamma)
    real(dp), intent(in) :: C(0:7), D(0:7), E(0:7), x, y, z
    type(curvi), intent(out) :: cv
    logical, intent(in) :: want_gamma
    real(dp) :: a(0:2,0:2), inv(0:2,0:2), r
    real(dp) :: g12(3), g13(3), g23(3)
    integer :: p, q
    cv%x = C(0)+C(1)*x+C(2)*y+C(3)*z+C(4)*x*y+C(5)*x*z+C(6)*y*z+C(7)*x*y*z
    cv%y = D(0)+D(1)*x+D(2)*y+D(3)*z+D(4)*x*y+D(5)*x*z+D(6)*y*z+D(7)*x*y*z
    cv%z = E(0)+E(1)*x+E(2)*y+E(3)*z+E(4)*x*y+E(5)*x*z+E(6)*y*z+E(7)*x*y*z
    cv%jac(1,1)=C(1)+C(4)*y+C(5)*z+C(7)*y*z
    cv%jac(1,2)=D(1)+D(4)*y+D(5)*z+D(7)*y*z
    cv%jac(1,3)=E(1)+E(4)*y+E(5)*z+E(7)*y*z
    cv%jac(2,1)=C(2)+C(4)*x+C(6)*z+C(7)*x*z
    cv%jac(2,2)=D(2)+D(4)*x+D(6)*z+D(7)*x*z
    cv%jac(2,3)=E(2)+E(4)*x+E(6)*z+E(7)*x*z
    cv%jac(3,1)=C(3)+C(5)*x+C(6)*y+C(7)*x*y
    cv%jac(3,2)=D(3)+D(5)*x+D(6)*y+D(7)*x*y
    cv%jac(3,3)=E(3)+E(5)*x+E(6)*y+E(7)*x*y
    a = cv%jac
    ! inv3x3 (general.f90:849-857): reciprocal determinant times adjugate
    r = 1/det3(a)
    inv(0,0) = r*(a(1,1)*a(2,2)-a(2,1)*a(1,2)); inv(0,1) = r*(a(0,2)*a(2,1)-a(0,1)*a(2,2))
    inv(0,2) = r*(a(0,1)*a(1,2)-a(0,2)*a(1,1)); inv(1,0) = r*(a(1,2)*a(2,0)-a(1,0)*a(2,2))
    inv(1,1) = r*(a(0,0)*a(2,2)-a(0,2)*a(2,0)); inv(1,2) = r*(a(0,2)*a(1,0)-a(0,0)*a(1,2))
    inv(2,0) = r*(a(1,0)*a(2,1)-a(1,1)*a(2,0)); inv(2,1) = r*(a(0,1)*a(2,0)-a(0,0)*a(2,1))
    inv(2,2) = r*(a(0,0)*a(1,1)-a(0,1)*a(1,0))
    cv%jinv = inv
    do p = 1, 3
      cv%gco(:,p) = cv%jac(p,:)
      cv%gcon(:,p) = cv%jinv(:,p)
    enddo
    do p = 1, 3
      do q = 1, 3
        cv%g_ij(p,q) = dot3(cv%gco(:,p), cv%gco(:,q))
        cv%gij(p,q) = dot3(cv%gcon(:,p), cv%gcon(:,q))
      enddo
    enddo
    cv%detJ = det3(a)
    if (want_gamma) then
      ! find_gamma_tensor (curvilinear.f90:109-160)
      g12 = (/ C(4)+C(7)*z, D(4)+D(7)*z, E(4)+E(7)*z /)
      g13 = (/ C(5)+C(7)*y, D(5)+D(7)*y, E(5)+E(7)*y /)
      g23 = (/ C(6)+C(7)*x, D(6)+D(7)*x, E(6)+E(7)*x /)
      cv%gam = 0._dp
      do p = 1, 3
        cv%gam(p,1,2) = dot3(cv%gcon(:,p), g12)
        cv%gam(p,1,3) = dot3(cv%gcon(:,p), g13)
        cv%gam(p,2,3) = dot3(cv%gcon(:,p), g23)
        cv%gam(p,2,1) = cv%gam(p,1,2)
        cv%gam(p,3,1) = cv%gam(p,1,3)
        cv%gam(p,3,2) = cv%gam(p,2,3)
      enddo
    endif
  end subroutine

  ! ------------------------------------------------------------------
  ! all per-cell metrics and per-face quantities
  ! ------------------------------------------------------------------
  ! Arrays are in the reference's layout and numbering; face numbers are
  ! computed with face_number (create_geometry.f90:626-664).
  integer function face_number(nt, nr, nz, pos, i, j, k) result(f)
    integer, intent(in) :: nt, nr, nz, pos, i, j, k  ! 0-based cell indices
    integer :: N, o1, o2, o3, o4, o5, o6, jj, kk, ii
    N = nt*nr*nz
    o1 = N; o2 = o1 + nt*(nr-1)*nz; o3 = o2 + nt*nr*(nz-1); o4 = o3 + nt*nr
    o5 = o4 + nt*nr; o6 = o5 + nt*nz
    select case (pos)
    case (1, 2)
      jj = j + (pos-1)
      if (jj == 0) then
        f = o5 + i + nt*k
      elseif (jj == nr) then
        f = o6 + i + nt*k
      else
        f = o1 + i + nt*((jj-1) + (nr-1)*k)
      endif
    case (3, 4)
      ii = i + (pos-3)
      if (ii == nt) ii = 0
      f = ii + nt*(j + nr*k)
    case default
      kk = k + (pos-5)
      if (kk == 0) then
        f = o3 + i + nt*j
      elseif (kk == nz) then
        f = o4 + i + nt*j
      else
        f = o2 + i + nt*(j + nr*(kk-1))
      endif
    end select
    f = f + 1  ! 1-based
  end function

  ! Builds every geometry array of the case. Output arrays (reference layout):
  ! volume(N), xi_x(3,3,N), gij(3,3,N), gij_faces(3,3,6,N), G1..G3(N),
  ! faces_area(NF), faces_normal(3,NF), faces_tangential(3,NF),
  ! faces_bitangential(3,NF), faces_mid(3,NF), meshes_faces(6,N),
  ! meshes_neighbors(6,N), bc_faces(NB), bc_position(NB), bc_mesh(NB),
  ! bc_points(5,NB), bc_a(20,NB), node_xyz(3,NNODE)
  subroutine build_geometry(cs, volume, xi_x, gij, gij_faces, G1, G2, G3, faces_area, faces_normal, &
                            faces_tangential, faces_bitangential, faces_mid, meshes_faces, meshes_neighbors, &
                            bc_faces, bc_position, bc_mesh, bc_points, bc_a, node_xyz)
    type(nim_geo_case), intent(in) :: cs
    real(dp), intent(out) :: volume(:), xi_x(:,:,:), gij(:,:,:), gij_faces(:,:,:,:), G1(:), G2(:), G3(:)
    real(dp), intent(out) :: faces_area(:), faces_normal(:,:), faces_tangential(:,:), faces_bitangential(:,:)
    real(dp), intent(out) :: faces_mid(:,:), bc_a(:,:), node_xyz(:,:)
    integer(c_int32_t), intent(out) :: meshes_faces(:,:), meshes_neighbors(:,:), bc_faces(:), bc_position(:)
    integer(c_int32_t), intent(out) :: bc_mesh(:), bc_points(:,:)
    integer :: nt, nr, nz, N, NF, NB, NNODE, c, i, j, k, p, b
    real(dp), allocatable :: nx(:), ny(:), nzz(:)
    ! per cell-face results, reduced to faces afterwards (last writer wins)
    real(dp), allocatable :: cf_area(:,:), cf_n(:,:,:), cf_t(:,:,:), cf_s(:,:,:), cf_mid(:,:,:)
    real(dp) :: gx(NG), gw(NG)

    nt = cs%n_theta; nr = cs%n_r; nz = cs%n_z
    N = nt*nr*nz
    NF = 3*N + nt*nz + nt*nr
    NB = 2*nt*nr + 2*nt*nz
    NNODE = nt*(nr+1)*(nz+1)
    call gauss_points(gx, gw)
    allocate(nx(NNODE), ny(NNODE), nzz(NNODE))
    call make_nodes(cs, nx, ny, nzz)
    node_xyz(1,:) = nx; node_xyz(2,:) = ny; node_xyz(3,:) = nzz

    allocate(cf_area(6,N), cf_n(3,6,N), cf_t(3,6,N), cf_s(3,6,N), cf_mid(3,6,N))

    call mesh_tables(cs, meshes_faces, meshes_neighbors, bc_faces, bc_position, bc_mesh)
    !$omp parallel do private(c, i, j, k) schedule(static)
    do c = 1, N
      i = mod(c-1, nt); j = mod((c-1)/nt, nr); k = (c-1)/(nt*nr)
      call cell_metrics(c, i, j, k)
    enddo
    !$omp end parallel do

    ! faces written by cells in natural order: the highest cell number wins
    do c = 1, N
      do p = 1, 6
        b = meshes_faces(p, c)
        faces_area(b) = cf_area(p, c)
        faces_normal(:, b) = cf_n(:, p, c)
        faces_tangential(:, b) = cf_t(:, p, c)
        faces_bitangential(:, b) = cf_s(:, p, c)
        faces_mid(:, b) = cf_mid(:, p, c)
      enddo
    enddo

    call bc_tables(cs, faces_normal, faces_tangential, faces_bitangential, faces_mid, meshes_faces, &
                   meshes_neighbors, bc_faces, bc_position, bc_mesh, bc_points, bc_a)

  contains

    subroutine cell_metrics(c, i, j, k)
      integer, intent(in) :: c, i, j, k
      real(dp) :: x(8), y(8), z(8), Cc(0:7), Dc(0:7), Ec(0:7)
      integer :: nd(8), ip, a, bb, ii, jj, kk, f
      real(dp) :: vol, s, dS(NG,NG), Kf(3,3,NG,NG), area, pts(3)
      real(dp) :: acc_G(3), acc_K(3,3), acc_J(3,3), G(3), wgt
      type(curvi) :: cv
      ip = mod(i+1, nt)
      nd = (/ i + nt*(j + (nr+1)*k), i + nt*(j+1 + (nr+1)*k), ip + nt*(j+1 + (nr+1)*k), ip + nt*(j + (nr+1)*k), &
              i + nt*(j + (nr+1)*(k+1)), i + nt*(j+1 + (nr+1)*(k+1)), ip + nt*(j+1 + (nr+1)*(k+1)), &
              ip + nt*(j + (nr+1)*(k+1)) /) + 1
      do a = 1, 8
        x(a) = nx(nd(a)); y(a) = ny(nd(a)); z(a) = nzz(nd(a))
      enddo
      call map_coeffs(x, y, z, Cc, Dc, Ec)
      ! volume_calculate (create_mesh.f90:306-322)
      vol = 0._dp
      do ii = 1, NG
        do jj = 1, NG
          do kk = 1, NG
            call curvilinear(Cc, Dc, Ec, gx(ii), gx(jj), gx(kk), cv, .false.)
            vol = vol + gw(ii)*gw(jj)*gw(kk)*(1._dp*cv%detJ)
          enddo
        enddo
      enddo
      volume(c) = vol
      ! find_normal (geometry.f90:76-190): values at the six face centres
      do f = 1, 6
        pts = 0._dp
        pts((f+1)/2) = merge(1._dp, -1._dp, mod(f,2) == 0)
        call curvilinear(Cc, Dc, Ec, pts(1), pts(2), pts(3), cv, .false.)
        a = (f+1)/2
        cf_n(:, f, c) = cv%gcon(:, a)/sqrt(cv%gij(a,a))
        if (a == 1) then
          cf_t(:, f, c) = cv%gco(:, 2)/sqrt(cv%g_ij(2,2)); cf_s(:, f, c) = cv%gco(:, 3)/sqrt(cv%g_ij(3,3))
        elseif (a == 2) then
          cf_t(:, f, c) = cv%gco(:, 1)/sqrt(cv%g_ij(1,1)); cf_s(:, f, c) = cv%gco(:, 3)/sqrt(cv%g_ij(3,3))
        else
          cf_t(:, f, c) = cv%gco(:, 1)/sqrt(cv%g_ij(1,1)); cf_s(:, f, c) = cv%gco(:, 2)/sqrt(cv%g_ij(2,2))
        endif
        cf_mid(:, f, c) = (/ cv%x, cv%y, cv%z /)
      enddo
      ! calculate_diffusion_tensor_faces_i (geometry.f90:415-497): face
      ! area and face-averaged g^ab/sqrt(g^aa) on the 5x5 face points
      do f = 1, 6
        a = (f+1)/2
        do ii = 1, NG
          do jj = 1, NG
            pts(a) = merge(1._dp, -1._dp, mod(f,2) == 0)
            if (a == 1) then
              pts(2) = gx(ii); pts(3) = gx(jj)
            elseif (a == 2) then
              pts(1) = gx(ii); pts(3) = gx(jj)
            else
              pts(1) = gx(ii); pts(2) = gx(jj)
            endif
            call curvilinear(Cc, Dc, Ec, pts(1), pts(2), pts(3), cv, .false.)
            dS(ii,jj) = cv%detJ*sqrt(cv%gij(a,a))
            do bb = 1, 3
              do kk = 1, 3
                Kf(bb,kk,ii,jj) = cv%gij(bb,kk)/sqrt(cv%gij(bb,bb))*dS(ii,jj)
              enddo
            enddo
          enddo
        enddo
        area = 0._dp
        do ii = 1, NG
          do jj = 1, NG
            area = area + gw(ii)*gw(jj)*dS(ii,jj)
          enddo
        enddo
        cf_area(f, c) = area
        do bb = 1, 3
          do kk = 1, 3
            s = 0._dp
            do ii = 1, NG
              do jj = 1, NG
                s = s + gw(ii)*gw(jj)*Kf(bb,kk,ii,jj)
              enddo
            enddo
            gij_faces(bb, kk, f, c) = s/area
          enddo
        enddo
      enddo
      ! convection_diffusion_constants_mesh_i (geometry.f90:500-632)
      acc_G = 0._dp; acc_K = 0._dp; acc_J = 0._dp
      do ii = 1, NG
        do jj = 1, NG
          do kk = 1, NG
            call curvilinear(Cc, Dc, Ec, gx(ii), gx(jj), gx(kk), cv, .true.)
            do a = 1, 3
              G(a) = (cv%gam(a,1,2)*cv%gij(1,2) + cv%gam(a,2,1)*cv%gij(2,1) &
                    + cv%gam(a,1,3)*cv%gij(1,3) + cv%gam(a,3,1)*cv%gij(3,1) &
                    + cv%gam(a,2,3)*cv%gij(2,3) + cv%gam(a,3,2)*cv%gij(3,2))
            enddo
            wgt = gw(ii)*gw(jj)*gw(kk)
            do a = 1, 3
              acc_G(a) = acc_G(a) + wgt*(G(a)*cv%detJ)
              do bb = 1, 3
                acc_K(a,bb) = acc_K(a,bb) + wgt*(cv%gij(a,bb)*cv%detJ)
                acc_J(a,bb) = acc_J(a,bb) + wgt*(cv%jinv(a,bb)*cv%detJ)
              enddo
            enddo
          enddo
        enddo
      enddo
      G1(c) = acc_G(1)/vol; G2(c) = acc_G(2)/vol; G3(c) = acc_G(3)/vol
      do a = 1, 3
        do bb = 1, 3
          gij(a, bb, c) = acc_K(a,bb)/vol
          xi_x(a, bb, c) = acc_J(bb,a)/vol
        enddo
      enddo
    end subroutine

  end subroutine build_geometry

  ! mesh_tables: the face numbers and neighbours of every cell (create_hollow,
  ! create_geometry.f90:449-664; neighbour 0 = wall) and the boundary-face
  ! tables of fill_global_arrays (geometry.f90:978-1032) in the boundary
  ! group order bottom, top, hexagon, cylinder
  subroutine mesh_tables(cs, meshes_faces, meshes_neighbors, bc_faces, bc_position, bc_mesh)
    type(nim_geo_case), intent(in) :: cs
    integer(c_int32_t), intent(out) :: meshes_faces(:,:), meshes_neighbors(:,:), bc_faces(:), bc_position(:)
    integer(c_int32_t), intent(out) :: bc_mesh(:)
    integer :: nt, nr, nz, N, NB, c, i, j, k, p, b
    nt = cs%n_theta; nr = cs%n_r; nz = cs%n_z
    N = nt*nr*nz
    NB = 2*nt*nr + 2*nt*nz
    do c = 1, N
      i = mod(c-1, nt); j = mod((c-1)/nt, nr); k = (c-1)/(nt*nr)
      do p = 1, 6
        meshes_faces(p, c) = face_number(nt, nr, nz, p, i, j, k)
      enddo
      meshes_neighbors(1, c) = merge(c - nt, 0, j > 0)
      meshes_neighbors(2, c) = merge(c + nt, 0, j < nr-1)
      meshes_neighbors(3, c) = mod(i-1+nt, nt) + nt*(j + nr*k) + 1
      meshes_neighbors(4, c) = mod(i+1, nt) + nt*(j + nr*k) + 1
      meshes_neighbors(5, c) = merge(c - nt*nr, 0, k > 0)
      meshes_neighbors(6, c) = merge(c + nt*nr, 0, k < nz-1)
    enddo
    b = 0
    do j = 0, nr-1
      do i = 0, nt-1
        b = b + 1; bc_position(b) = 5; bc_mesh(b) = i + nt*j + 1
      enddo
    enddo
    do j = 0, nr-1
      do i = 0, nt-1
        b = b + 1; bc_position(b) = 6; bc_mesh(b) = i + nt*(j + nr*(nz-1)) + 1
      enddo
    enddo
    do k = 0, nz-1
      do i = 0, nt-1
        b = b + 1; bc_position(b) = 1; bc_mesh(b) = i + nt*nr*k + 1
      enddo
    enddo
    do k = 0, nz-1
      do i = 0, nt-1
        b = b + 1; bc_position(b) = 2; bc_mesh(b) = i + nt*(nr-1 + nr*k) + 1
      enddo
    enddo
    do b = 1, NB
      bc_faces(b) = meshes_faces(bc_position(b), bc_mesh(b))
    enddo
  end subroutine mesh_tables

  ! evaluate_pressure_bc_array (geometry.f90:879-975): 5-point stencil
  ! normal to each wall and its finite-difference weights
  subroutine bc_tables(cs, faces_normal, faces_tangential, faces_bitangential, faces_mid, meshes_faces, &
                       meshes_neighbors, bc_faces, bc_position, bc_mesh, bc_points, bc_a)
    type(nim_geo_case), intent(in) :: cs
    real(dp), intent(in) :: faces_normal(:,:), faces_tangential(:,:), faces_bitangential(:,:), faces_mid(:,:)
    integer(c_int32_t), intent(in) :: meshes_faces(:,:), meshes_neighbors(:,:), bc_faces(:), bc_position(:)
    integer(c_int32_t), intent(in) :: bc_mesh(:)
    integer(c_int32_t), intent(out) :: bc_points(:,:)
    real(dp), intent(out) :: bc_a(:,:)
    integer :: bb, M1, M2, M3, M4, fn, q, NB
    real(dp) :: n(5), t(5), s(5), nv(3), tv(3), sv(3), dd(3), D
    real(dp), allocatable :: A(:)
    allocate(A(20))
    NB = 2*cs%n_theta*cs%n_r + 2*cs%n_theta*cs%n_z
    do bb = 1, NB
      nv = faces_normal(:, bc_faces(bb)); tv = faces_tangential(:, bc_faces(bb)); sv = faces_bitangential(:, bc_faces(bb))
      M1 = bc_mesh(bb)
      bc_points(1, bb) = bc_faces(bb)
      select case (bc_position(bb))
      case (1, 2)
        fn = 3 - bc_position(bb)
        M2 = meshes_neighbors(fn, M1)
        M3 = meshes_neighbors(4, M1); if (M3 == 0) M3 = meshes_neighbors(3, M1)
        M4 = meshes_neighbors(6, M1); if (M4 == 0) M4 = meshes_neighbors(5, M1)
      case (3, 4)
        fn = 7 - bc_position(bb)
        M2 = meshes_neighbors(fn, M1)
        M3 = meshes_neighbors(2, M1); if (M3 == 0) M3 = meshes_neighbors(1, M1)
        M4 = meshes_neighbors(6, M1); if (M4 == 0) M4 = meshes_neighbors(5, M1)
      case default
        fn = 11 - bc_position(bb)
        M2 = meshes_neighbors(fn, M1)
        M3 = meshes_neighbors(2, M1); if (M3 == 0) M3 = meshes_neighbors(1, M1)
        M4 = meshes_neighbors(4, M1); if (M4 == 0) M4 = meshes_neighbors(3, M1)
      end select
      bc_points(2, bb) = meshes_faces(fn, M1)
      bc_points(3, bb) = meshes_faces(fn, M2)
      bc_points(4, bb) = meshes_faces(fn, M3)
      bc_points(5, bb) = meshes_faces(fn, M4)
      do q = 1, 5
        dd = faces_mid(:, bc_points(q, bb)) - faces_mid(:, bc_points(1, bb))
        n(q) = dot3(dd, nv); t(q) = dot3(dd, tv); s(q) = dot3(dd, sv)
      enddo
      D = (n(3)*s(4)*t(2) - n(3)*s(5)*t(2) - n(2)*s(4)*t(3) + n(2)*s(5)*t(3) - n(3)*s(2)*t(4) + n(2)*s(3)*t(4) &
           - n(2)*s(5)*t(4) + n(3)*s(5)*t(4) + n(5)*s(3)*t(2) - n(5)*s(4)*t(2) - n(5)*s(2)*t(3) + n(5)*s(4)*t(3) &
           + n(5)*s(2)*t(4) - n(5)*s(3)*t(4) + n(3)*s(2)*t(5) - n(2)*s(3)*t(5) + n(2)*s(4)*t(5) - n(3)*s(4)*t(5) &
           + n(4)*(s(5)*t(2) + s(2)*t(3) - s(5)*t(3) - s(2)*t(5) + s(3)*(-t(2) + t(5))))
      A(2)=2*(-n(5)*s(4)*t(3) + n(4)*s(5)*t(3) + n(5)*s(3)*t(4) - n(3)*s(5)*t(4) - n(4)*s(3)*t(5) + n(3)*s(4)*t(5))/(n(2)*n(2)*D)
      A(3)=2*( n(5)*s(4)*t(2) - n(4)*s(5)*t(2) - n(5)*s(2)*t(4) + n(2)*s(5)*t(4) + n(4)*s(2)*t(5) - n(2)*s(4)*t(5))/(n(3)*n(3)*D)
      A(4)=2*( n(5)*s(3)*t(2) - n(3)*s(5)*t(2) - n(5)*s(2)*t(3) + n(2)*s(5)*t(3) + n(3)*s(2)*t(5) - n(2)*s(3)*t(5))/(n(4)*n(4)*D)
      A(5)=2*( n(4)*s(3)*t(2) - n(3)*s(4)*t(2) - n(4)*s(2)*t(3) + n(2)*s(4)*t(3) + n(3)*s(2)*t(4) - n(2)*s(3)*t(4))/(n(5)*n(5)*D)
      A(1)=-A(2)-A(3)-A(4)-A(5)
      A(6)=-A(2)*n(2)-A(3)*n(3)-A(4)*n(4)-A(5)*n(5)
      A(7)=-A(2)*t(2)-A(3)*t(3)-A(4)*t(4)-A(5)*t(5)
      A(8)=-A(2)*s(2)-A(3)*s(3)-A(4)*s(4)-A(5)*s(5)
      A(9)=-A(2)*n(2)*t(2)-A(3)*n(3)*t(3)-A(4)*n(4)*t(4)-A(5)*n(5)*t(5)
      A(10)=-A(2)*n(2)*s(2)-A(3)*n(3)*s(3)-A(4)*n(4)*s(4)-A(5)*n(5)*s(5)
      A(11)=-A(2)*s(2)*t(2)-A(3)*s(3)*t(3)-A(4)*s(4)*t(4)-A(5)*s(5)*t(5)
      A(12)=-A(2)*(t(2)*t(2))/2._dp-A(3)*(t(3)*t(3))/2._dp-A(4)*(t(4)*t(4))/2._dp-A(5)*(t(5)*t(5))/2._dp
      A(13)=-A(2)*(s(2)*s(2))/2._dp-A(3)*(s(3)*s(3))/2._dp-A(4)*(s(4)*s(4))/2._dp-A(5)*(s(5)*s(5))/2._dp
      A(14)=-A(2)*n(2)*(t(2)*t(2))/2._dp-A(3)*n(3)*(t(3)*t(3))/2._dp-A(4)*n(4)*(t(4)*t(4))/2._dp-A(5)*n(5)*(t(5)*t(5))/2._dp
      A(15)=-A(2)*n(2)*(s(2)*s(2))/2._dp-A(3)*n(3)*(s(3)*s(3))/2._dp-A(4)*n(4)*(s(4)*s(4))/2._dp-A(5)*n(5)*(s(5)*s(5))/2._dp
      A(16)=-A(2)*s(2)*(t(2)*t(2))/2._dp-A(3)*s(3)*(t(3)*t(3))/2._dp-A(4)*s(4)*(t(4)*t(4))/2._dp-A(5)*s(5)*(t(5)*t(5))/2._dp
      A(17)=-A(2)*t(2)*(s(2)*s(2))/2._dp-A(3)*t(3)*(s(3)*s(3))/2._dp-A(4)*t(4)*(s(4)*s(4))/2._dp-A(5)*t(5)*(s(5)*s(5))/2._dp
      A(18)=-A(2)*n(2)*t(2)*s(2)-A(3)*n(3)*t(3)*s(3)-A(4)*n(4)*t(4)*s(4)-A(5)*n(5)*t(5)*s(5)
      A(19)=-A(2)*t(2)**3._dp/6._dp-A(3)*t(3)**3._dp/6._dp-A(4)*t(4)**3._dp/6._dp-A(5)*t(5)**3._dp/6._dp
      A(20)=-A(2)*s(2)**3._dp/6._dp-A(3)*s(3)**3._dp/6._dp-A(4)*s(4)**3._dp/6._dp-A(5)*s(5)**3._dp/6._dp
      bc_a(:, bb) = A
    enddo
  end subroutine bc_tables

end module nim_host_geometry
