!> wql_run -- Fortran host for the GPU path: a two-block Cartesian
!> slip-weakening rupture (TPV5-style prestress, planar fault at x = 0,
!> free surface at y = 0), stepped through libwql.so.
!>   usage: wql_run nq nr ns nsteps [fp32]
program wql_run
  use iso_c_binding
  use wql_gpu
  implicit none
  integer :: nq, nr, ns, nsteps, ib, i, j, k, narg
  character(len=32) :: arg
  real(c_double), parameter :: h = 0.1d0, rho = 2.67d0, cs = 3.464d0, cp = 6.0d0
  real(c_double), allocatable, target :: X1(:,:,:,:), X2(:,:,:,:), S(:,:,:)
  type(wql_block_desc) :: blocks(2)
  type(wql_problem) :: prob
  type(c_ptr) :: ctx
  real(c_double) :: Lx, Ly, Lz
  integer(c_int) :: prec
  integer(8) :: c0, c1, rate

  narg = command_argument_count()
  if (narg < 4) stop 'usage: wql_run nq nr ns nsteps [fp32]'
  call get_command_argument(1, arg); read(arg, *) nq
  call get_command_argument(2, arg); read(arg, *) nr
  call get_command_argument(3, arg); read(arg, *) ns
  call get_command_argument(4, arg); read(arg, *) nsteps
  prec = WQL_FP64
  if (narg >= 5) then
     call get_command_argument(5, arg)
     if (trim(arg) == 'fp32') prec = WQL_FP32
  end if

  Lx = (nq - 1) * h; Ly = (nr - 1) * h; Lz = (ns - 1) * h
  allocate(X1(nq, nr, ns, 3), X2(nq, nr, ns, 3), S(nr, ns, 4))
  do k = 1, ns
     do j = 1, nr
        do i = 1, nq
           X1(i, j, k, 1) = -Lx + (i - 1) * h
           X2(i, j, k, 1) = (i - 1) * h
           X1(i, j, k, 2) = (j - 1) * h
           X2(i, j, k, 2) = (j - 1) * h
           X1(i, j, k, 3) = -Lz / 2 + (k - 1) * h
           X2(i, j, k, 3) = -Lz / 2 + (k - 1) * h
        end do
     end do
  end do

  do ib = 1, 2
     blocks(ib)%nq = nq; blocks(ib)%nr = nr; blocks(ib)%ns = ns
     blocks(ib)%pml = 0; blocks(ib)%npml = 0; blocks(ib)%cartesian = 1
     blocks(ib)%rho_s_p = [rho, cs, cp]
  end do
  blocks(1)%bc = [1, 0, 2, 1, 1, 1]     ! fault on the block's +q face
  blocks(2)%bc = [0, 1, 2, 1, 1, 1]     ! fault on the block's -q face
  blocks(1)%X = c_loc(X1)
  blocks(2)%X = c_loc(X2)

  prob%precision = prec
  prob%dt = 0.5d0 * h / sqrt(cs**2 + cp**2)      ! domain.f90:222-235 with CFL 0.5
  prob%t0 = 0.0d0
  prob%plastic = 0
  prob%s0_kind = 0
  prob%fault%law = WQL_SLIP_WEAKENING
  prob%fault%stress_mode = 0
  prob%fault%T0 = [-120.0d0, 0.0d0, 70.0d0]
  prob%fault%s0 = 0.0d0
  prob%fault%npatch = 1
  prob%fault%patch = 0.0d0
  prob%fault%patch(:, 1) = [Ly / 2, 0.0d0, 1.5d0, 1.5d0, 11.6d0]
  prob%fault%nuc = 0
  prob%fault%nucp = 0.0d0
  prob%fault%fs = 0.677d0; prob%fault%fd = 0.525d0; prob%fault%Dc = 0.4d0
  prob%fault%bar_wy = 0.8d0 * Ly; prob%fault%bar_wz = 0.4d0 * Lz; prob%fault%bar_big = 1.0d10
  prob%fault%rs_a = 0; prob%fault%rs_da = 0; prob%fault%rs_w = 0; prob%fault%rs_wr = 0
  prob%fault%rs_b = 0; prob%fault%rs_L = 0; prob%fault%rs_V0 = 0; prob%fault%rs_f0 = 0; prob%fault%rs_Vin = 0

  call wql_check(wql_create(prob, blocks, ctx), 'wql_create')
  call system_clock(c0, rate)
  call wql_check(wql_step(ctx, int(nsteps, c_int)), 'wql_step')
  call wql_check(wql_sync(ctx), 'wql_sync')
  call system_clock(c1)
  call wql_check(wql_get_fault(ctx, 0_c_int, c_loc(S), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                 c_null_ptr), 'wql_get_fault')
  write(*, '(a,i0,a,es22.15,a,es22.15,a,es22.15,a,f10.3,a)') 'wql_run: steps ', nsteps, ' t ', wql_time(ctx), &
       ' max_slip ', maxval(S(:, :, 4)), ' sum_slip ', sum(S(:, :, 4)), ' wall_ms ', &
       1.0d3 * real(c1 - c0, 8) / real(rate, 8), ''
  call wql_check(wql_destroy(ctx), 'wql_destroy')
end program wql_run
