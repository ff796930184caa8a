! ref_bind.f90 -- C-callable wrappers around the reference's own Fortran CPU routines.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with the
! reference's bia/find_hyperbola.f90, find_ellipse.f90, find_parabola.f90,
! find_dXYUV_RKF.f90, Invariance_CPU.f90, Invariance_QPU.f90 and sincos.c, read where
! they lie under the reference checkout, into oracle/_ref/libbiaref.so.  Tests use it
! to pin the C restatement (oracle/bia_oracle.c); bench.py times ref_cpu_rows as the
! CPU baseline.  Nothing here is part of the product library.

module ref_bind
  use iso_c_binding
  implicit none
contains

  ! reference hyperbola solver, bia/find_hyperbola.f90:1-166
  subroutine ref_hyperbola(eps, dt, itr, yz, yn, err, ir, ier) bind(C, name='ref_hyperbola')
    real(c_double), value :: eps, dt
    integer(c_int), value :: itr
    real(c_double) :: yz(4), yn(4), err
    integer(c_int) :: ir, ier
    real(8) :: a(4), b(4), e, tol
    integer :: irr, ierr, it
    a = yz; b = 0; e = 0; tol = 1d-12; it = itr; irr = 0; ierr = 0
    call find_hyperbola(eps, dt, a, b, e, tol, it, irr, ierr)
    yn = b; err = e; ir = irr; ier = ierr
  end subroutine ref_hyperbola

  ! reference ellipse solver, bia/find_ellipse.f90:1-157
  subroutine ref_ellipse(eps, dt, itr, yz, yn, err, ir, ier) bind(C, name='ref_ellipse')
    real(c_double), value :: eps, dt
    integer(c_int), value :: itr
    real(c_double) :: yz(4), yn(4), err
    integer(c_int) :: ir, ier
    real(8) :: a(4), b(4), e, tol
    integer :: irr, ierr, it
    a = yz; b = 0; e = 0; tol = 1d-12; it = itr; irr = 0; ierr = 0
    call find_ellipse(eps, dt, a, b, e, tol, it, irr, ierr)
    yn = b; err = e; ir = irr; ier = ierr
  end subroutine ref_ellipse

  ! reference parabola solver, bia/find_parabola.f90:1-135 (prints its iterations)
  subroutine ref_parabola(eps, dt, yz, yn, err) bind(C, name='ref_parabola')
    real(c_double), value :: eps, dt
    real(c_double) :: yz(4), yn(4), err
    real(8) :: a(4), b(4), e
    a = yz; b = 0; e = 0
    call find_parabola(eps, dt, a, b, e)
    yn = b; err = e
  end subroutine ref_parabola

  ! reference per-pair RKF45, bia/find_dXYUV_RKF.f90:1-170.  Its field routine reads the
  ! coupling from common /const/ (find_dXYUV_RKF.f90:161), set here before the call.
  subroutine ref_rkf(eps, dt, yz, yn, err, ier) bind(C, name='ref_rkf')
    real(c_double), value :: eps, dt
    real(c_double) :: yz(4), yn(4), err
    integer(c_int) :: ier
    real(8) :: eps_c
    common / const / eps_c
    real(8) :: a(4), b(4), e
    integer :: ierr
    eps_c = eps
    a = yz; b = 0; e = 0; ierr = 0
    call find_dXYUV_RKF(eps, dt, a, b, e, ierr)
    yn = b; err = e; ier = ierr
  end subroutine ref_rkf

  ! reference invariance in double (Invariance_CPU.f90) and quad (Invariance_QPU.f90);
  ! out = KE, PE, Lx, Ly, Lz, Ua, Va, Wa, Xa, Ya, Za
  subroutine ref_invariance(n, eps, quad, x, y, z, u, v, w, out) bind(C, name='ref_invariance')
    integer(c_int), value :: n, quad
    real(c_double), value :: eps
    real(c_double) :: x(n), y(n), z(n), u(n), v(n), w(n), out(11)
    real(8) :: KE, PE, Lx, Ly, Lz, Ua, Va, Wa, Xa, Ya, Za
    integer :: np
    np = n
    if (quad /= 0) then
      call Invariance_QPU(np, eps, x, y, z, u, v, w, KE, PE, Lx, Ly, Lz, Ua, Va, Wa, Xa, Ya, Za)
    else
      call Invariance_CPU(np, eps, x, y, z, u, v, w, KE, PE, Lx, Ly, Lz, Ua, Va, Wa, Xa, Ya, Za)
    end if
    out = (/ KE, PE, Lx, Ly, Lz, Ua, Va, Wa, Xa, Ya, Za /)
  end subroutine ref_invariance

  ! The reference CPU pair loop of bia.f90:142-196 over rows i0..i1-1 (0-based), calling
  ! the reference's find_hyperbola / find_parabola / find_ellipse (method 1, 'ALG') or
  ! find_dXYUV_RKF (method 0, 'RKF').  State and deltas are SoA (x,y,z,u,v,w).  The frame
  ! of bia.f90:284-314 is restated in frame3 below (it lives in bia.f90's program file).
  subroutine ref_cpu_rows(method, n, eps, dt, itr, st, i0, i1, delta) bind(C, name='ref_cpu_rows')
    integer(c_int), value :: method, n, itr, i0, i1
    real(c_double), value :: eps, dt
    real(c_double) :: st(n, 6), delta(n, 6)
    real(8) :: r(3), g(3), e_xi(3), e_eta(3), yz(4), yn(4), err, tol, rij, eij, eps_c
    real(8) :: dr_xi, dr_eta, dg_xi, dg_eta, dxh(3), dvh(3)
    integer :: i, j, ir, ier, it
    common / const / eps_c
    eps_c = eps; tol = 1d-12; it = itr
    do i = i0 + 1, i1
      do j = i + 1, n
        r = st(i, 1:3) - st(j, 1:3)
        g = st(i, 4:6) - st(j, 4:6)
        call frame3(r, g, e_xi, e_eta, yz)
        if (method == 0) then
          call find_dXYUV_RKF(eps, dt, yz, yn, err, ier)
        else
          rij = sqrt(r(1)**2 + r(2)**2 + r(3)**2)
          eij = g(1)**2 + g(2)**2 + g(3)**2 + 2 * eps / rij
          if (eij > 0) then
            call find_hyperbola(eps, dt, yz, yn, err, tol, it, ir, ier)
          else if (eij == 0) then
            call find_parabola(eps, dt, yz, yn, err)
          else
            call find_ellipse(eps, dt, yz, yn, err, tol, it, ir, ier)
          end if
        end if
        dr_xi = yn(1) - yz(1); dr_eta = yn(2) - yz(2)
        dg_xi = yn(3) - yz(3); dg_eta = yn(4) - yz(4)
        dxh = (dr_xi * e_xi + dr_eta * e_eta - g * dt) / 2
        dvh = (dg_xi * e_xi + dg_eta * e_eta) / 2
        delta(i, 1:3) = delta(i, 1:3) + dxh; delta(j, 1:3) = delta(j, 1:3) - dxh
        delta(i, 4:6) = delta(i, 4:6) + dvh; delta(j, 4:6) = delta(j, 4:6) - dvh
      end do
      delta(i, 1:3) = delta(i, 1:3) + st(i, 4:6) * dt
    end do
  end subroutine ref_cpu_rows

  ! orbital-plane frame of bia.f90:284-314: e_eta = g/|g|, e_zeta = (r x g)/|r x g|,
  ! e_xi = e_eta x e_zeta, yz = (r.e_xi, r.e_eta, 0, |g|)
  subroutine frame3(r, g, e_xi, e_eta, yz)
    real(8) :: r(3), g(3), e_xi(3), e_eta(3), yz(4), L(3), e_zeta(3), gm
    gm = sqrt(g(1)**2 + g(2)**2 + g(3)**2)
    L(1) = r(2) * g(3) - r(3) * g(2)
    L(2) = r(3) * g(1) - r(1) * g(3)
    L(3) = r(1) * g(2) - r(2) * g(1)
    e_eta = g / gm
    e_zeta = L / sqrt(sum(L**2))
    e_xi(1) = e_eta(2) * e_zeta(3) - e_eta(3) * e_zeta(2)
    e_xi(2) = e_eta(3) * e_zeta(1) - e_eta(1) * e_zeta(3)
    e_xi(3) = e_eta(1) * e_zeta(2) - e_eta(2) * e_zeta(1)
    yz(1) = r(1) * e_xi(1) + r(2) * e_xi(2) + r(3) * e_xi(3); yz(3) = 0
    yz(2) = r(1) * e_eta(1) + r(2) * e_eta(2) + r(3) * e_eta(3); yz(4) = gm
  end subroutine frame3
end module ref_bind
