! bia_hip.f90 -- ISO_C_BINDING interface to libbia_hip.so's new entry points
! (include/bia_hip.h).  The reference-compatible symbols (find_dxyuv_kernel_, invariance_gpu_,
! rkf_c_) need no interface: the reference's implicit %VAL calls bind to them directly.
module bia_hip
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: BIA_OK = 0, BIA_ERR_IO = 6, BIA_METHOD_RKF = 0, BIA_METHOD_ALG = 1
  integer(c_int), parameter :: BIA_MODE_PLAIN = 0, BIA_MODE_LEARN = 1, BIA_MODE_GUESS = 2, BIA_MODE_PLAIN_LEARN = 3
  integer(c_int), parameter :: BIA_STATS_LEN = 16
  integer(c_int), parameter :: BIA_COUPLING_GRAVITY = 1, BIA_COUPLING_COULOMB = 2

  interface
    integer(c_int) function bia_create(ctx, device) bind(C, name='bia_create')
      import :: c_int, c_ptr
      type(c_ptr) :: ctx
      integer(c_int), value :: device
    end function

    integer(c_int) function bia_destroy(ctx) bind(C, name='bia_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! which of the context's two learned initialisers the bia_mlp_* calls act on (0: the Kepler
    ! drift's start, 1: the reference solvers' starts of BIA_MODE_PLAIN_LEARN)
    integer(c_int) function bia_mlp_select(ctx, net) bind(C, name='bia_mlp_select')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: net
    end function

    ! per-pair RKF45 tolerances of the following steps (the reference's 0 / 1e-16 by default)
    integer(c_int) function bia_set_rkf_tolerance(ctx, relerr, abserr) bind(C, name='bia_set_rkf_tolerance')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: relerr, abserr
    end function

    ! largest tile-partial workspace of a step in bytes (0 = no limit; default 1 GiB)
    integer(c_int) function bia_set_workspace_limit(ctx, bytes) bind(C, name='bia_set_workspace_limit')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), value :: bytes
    end function

    ! stages and tile-partial workspace bytes of the last step's plan
    integer(c_int) function bia_plan_info(ctx, nstages, workspace_bytes) bind(C, name='bia_plan_info')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: nstages
      integer(c_long_long), intent(out) :: workspace_bytes
    end function

    integer(c_int) function bia_step_host(ctx, method, mode, n, eps, dt, itr, X, Y, Z, U, V, W, &
                                          dX, dY, dZ, dU, dV, dW, stats) bind(C, name='bia_step_host')
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: method, mode, n, itr
      real(c_double), value :: eps, dt
      real(c_double), intent(in) :: X(*), Y(*), Z(*), U(*), V(*), W(*)
      real(c_double), intent(out) :: dX(*), dY(*), dZ(*), dU(*), dV(*), dW(*)
      integer(c_long_long), intent(out) :: stats(*)
    end function

    integer(c_int) function bia_invariants_host(ctx, n, eps, X, Y, Z, U, V, W, out) &
        bind(C, name='bia_invariants_host')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      real(c_double), value :: eps
      real(c_double), intent(in) :: X(*), Y(*), Z(*), U(*), V(*), W(*)
      real(c_double), intent(out) :: out(11)
    end function

    integer(c_int) function bia_dim_rkf45_host(ctx, np, eps, x0, xe, y, init, relerr, abserr, h, ese, &
                                               nde, naccept, nreject, ier) bind(C, name='bia_dim_rkf45_host')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: np
      real(c_double), value :: eps, xe, relerr, abserr
      real(c_double) :: x0, y(*), h, ese(*)
      integer(c_int) :: init, nde, naccept, nreject, ier
    end function

    ! mass / charge: pass c_null_ptr for an array the coupling does not use, else c_loc(array)
    integer(c_int) function bia_dim_rkf45_coupled_host(ctx, coupling, coup, np, mass, charge, x0, xe, y, init, &
                                                       relerr, abserr, h, ese, nde, naccept, nreject, ier) &
        bind(C, name='bia_dim_rkf45_coupled_host')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx, mass, charge
      integer(c_int), value :: coupling, np
      real(c_double), value :: coup, xe, relerr, abserr
      real(c_double) :: x0, y(*), h, ese(*)
      integer(c_int) :: init, nde, naccept, nreject, ier
    end function

    ! the reference's record files (.bbb: nhead 0; .DIM, .BIA: nhead 1), for callers that do not
    ! read them themselves: count the records, read one into head(nhead) and soa(ld, 6), write one
    ! (append 0 truncates, 1 appends).  path: a NUL-terminated string, trim(name)//c_null_char.
    integer(c_int) function bia_records_scan(path, nhead, nrec, nmax) bind(C, name='bia_records_scan')
      import :: c_int, c_char, c_long_long
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), value :: nhead
      integer(c_long_long) :: nrec, nmax
    end function

    integer(c_int) function bia_record_read(path, nhead, rec, head, soa, ld, n) bind(C, name='bia_record_read')
      import :: c_int, c_char, c_long_long, c_double
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), value :: nhead
      integer(c_long_long), value :: rec, ld
      real(c_double) :: head(*), soa(*)
      integer(c_long_long) :: n
    end function

    integer(c_int) function bia_record_write(path, append, nhead, head, soa, ld, n) bind(C, name='bia_record_write')
      import :: c_int, c_char, c_long_long, c_double
      character(kind=c_char), dimension(*), intent(in) :: path
      integer(c_int), value :: append, nhead
      real(c_double), intent(in) :: head(*), soa(*)
      integer(c_long_long), value :: ld, n
    end function
  end interface
end module bia_hip
