! bia_run.f90 -- Fortran host driver over the bind(C) entry points: the flow of the
! reference's bia/bia.f90 (read .bbb, invariants, one BIA step, write .BIA deltas and the
! particle-1 text line, compare with the .DIM result when present) with the per-particle loop
! replaced by one device-resident whole step.
!
! usage: bia_run <prefix> <eps> <dt> [method 1=ALG|0=RKF] [mode 0=plain|2=Kepler drift]
!   reads <prefix>.bbb (and <prefix>.DIM if it exists), writes <prefix>.run.BIA; mode 1 (learned
!   start) needs a network trained and tabulated on the context first (bia_mlp_*), which this
!   driver does not do, so the step call reports BIA_ERR_STATE for it
program bia_run
  use iso_c_binding
  use bia_hip
  implicit none
  character(len=512) :: prefix, arg
  real(c_double) :: eps, dt
  integer(c_int) :: method, mode, n, s, k, nbytes
  real(c_double), allocatable :: st(:,:), X(:), Y(:), Z(:), U(:), V(:), W(:), d(:,:), fin(:,:)
  real(c_double) :: inv0(11), inv1(11), tdim
  integer(c_long_long) :: stats(16)
  type(c_ptr) :: ctx
  logical :: have_dim
  integer :: u_in, u_out

  call get_command_argument(1, prefix)
  call get_command_argument(2, arg); read(arg, *) eps
  call get_command_argument(3, arg); read(arg, *) dt
  method = 1; mode = 0
  if (command_argument_count() >= 4) then; call get_command_argument(4, arg); read(arg, *) method; end if
  if (command_argument_count() >= 5) then; call get_command_argument(5, arg); read(arg, *) mode; end if

  open(newunit=u_in, file=trim(prefix)//'.bbb', form='unformatted', access='stream', action='read')
  read(u_in) nbytes
  n = nbytes / 48
  allocate(st(6, n), X(n), Y(n), Z(n), U(n), V(n), W(n), d(n, 6), fin(6, n))
  read(u_in) st
  close(u_in)
  X = st(1,:); Y = st(2,:); Z = st(3,:); U = st(4,:); V = st(5,:); W = st(6,:)
  print '(a,i9)', 'N_p =', n

  s = bia_create(ctx, 0_c_int); if (s /= BIA_OK) stop 'bia_create failed'
  s = bia_invariants_host(ctx, n, eps, X, Y, Z, U, V, W, inv0); if (s /= BIA_OK) stop 'invariants failed'
  s = bia_step_host(ctx, method, mode, n, eps, dt, 18_c_int, X, Y, Z, U, V, W, &
                    d(:,1), d(:,2), d(:,3), d(:,4), d(:,5), d(:,6), stats)
  if (s /= BIA_OK) stop 'bia_step_host failed'
  X = X + d(:,1); Y = Y + d(:,2); Z = Z + d(:,3); U = U + d(:,4); V = V + d(:,5); W = W + d(:,6)
  s = bia_invariants_host(ctx, n, eps, X, Y, Z, U, V, W, inv1); if (s /= BIA_OK) stop 'invariants failed'
  print '(a,3es24.16)', 'E ini/end, dE/E0:', inv0(1) + inv0(2), inv1(1) + inv1(2), &
        ((inv1(1) - inv0(1)) + (inv1(2) - inv0(2))) / (inv0(1) + inv0(2))
  print '(a,4i14)', 'pairs hyp/par/ell/rkf:', stats(1:4)

  open(newunit=u_out, file=trim(prefix)//'.run.BIA', form='unformatted', action='write')
  write(u_out) dt, (d(k,1), d(k,2), d(k,3), d(k,4), d(k,5), d(k,6), k = 1, n)
  close(u_out)
  print '(es23.16,6es24.16)', dt, X(1), Y(1), Z(1), U(1), V(1), W(1)

  inquire(file=trim(prefix)//'.DIM', exist=have_dim)
  if (have_dim) then
    open(newunit=u_in, file=trim(prefix)//'.DIM', form='unformatted', action='read')
    read(u_in) tdim, fin
    close(u_in)
    print '(a,2es12.4)', 'max |r - r_DIM|, |v - v_DIM|:', &
      max(maxval(abs(X - fin(1,:))), maxval(abs(Y - fin(2,:))), maxval(abs(Z - fin(3,:)))), &
      max(maxval(abs(U - fin(4,:))), maxval(abs(V - fin(5,:))), maxval(abs(W - fin(6,:))))
  end if
  s = bia_destroy(ctx)
end program bia_run
