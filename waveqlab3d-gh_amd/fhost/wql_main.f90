!> wql_main -- the reference's main program (main.f90:58-96) on the GPU path:
!> read the input file named on the command line, init_domain, then
!> `do n = 1, D%nt; call time_step_RK(D, D%dt, RK, n); end do`.
!>   usage: wql_main input.in [nsteps] [fp32] [dump.bin]
!>          wql_main input.in check
!> nsteps (> 0) overrides D%nt; dump.bin receives, as raw FP64 in Fortran
!> order, both blocks' F(q,r,s,1:9) and block copies' S(nr,ns,4), W, trup,
!> Svel(nr,ns,3) after the last step (for comparisons).  `check` prints the
!> namelist values the host read and exits without touching the device.
program wql_main
  use iso_c_binding
  use wql_gpu
  use wql_domain
  use wql_input
  implicit none
  type(domain_type), target :: D
  type(RK_type) :: RK
  character(len=1024) :: ifname, arg, dump
  integer :: n, nsteps, narg, stat, ib, u, k
  integer(c_int) :: prec
  logical :: grid_only = .false.
  integer(8) :: c0, c1, rate
  type(problem_params) :: p
  real(kind = wp), allocatable :: F(:,:,:,:), S(:,:,:), W(:,:,:), tr(:,:,:), Sv(:,:,:)

  narg = command_argument_count()
  if (narg < 1) stop 'usage: wql_main input.in [nsteps] [fp32] [dump.bin] | check | grid dump.bin'
  call get_command_argument(1, ifname, status = stat)
  if (stat /= 0) stop ':: Problem reading input filename from commandline.'
  nsteps = 0
  prec = WQL_FP64
  dump = ''
  do k = 2, narg
     call get_command_argument(k, arg)
     if (trim(arg) == 'check') then
        call read_input(trim(ifname), p)
        write(*, '(a,1x,a,1x,a,1x,a,1x,a)') 'problem', trim(p%problem), trim(p%coupling), trim(p%response), &
             trim(p%type_of_mesh)
        write(*, '(a,es25.17,1x,es25.17,1x,i0)') 'time', p%t_final, p%CFL, p%w_stride
        write(*, '(a,1x,a,1x,i0)') 'operator', trim(p%fd_type), p%order
        do ib = 1, 2
           write(*, '(a,i0,3(1x,i0),6(1x,es25.17),2(1x,es25.17),6(1x,i0),6(1x,l1),1x,i0,1x,a)') 'block', ib, &
                p%btp(ib)%nqrs, p%btp(ib)%aqrs, p%btp(ib)%bqrs, p%btp(ib)%lc, p%btp(ib)%rc, &
                p%btp(ib)%lqrs, p%btp(ib)%rqrs, p%btp(ib)%pml_lqrs, p%btp(ib)%pml_rqrs, p%btp(ib)%npml, &
                trim(p%btp(ib)%profile_type)
        end do
        stop
     else if (trim(arg) == 'grid') then
        grid_only = .true.
     else if (trim(arg) == 'fp32') then
        prec = WQL_FP32
     else if (index(arg, '.bin') > 0) then
        dump = arg
     else
        read(arg, *) nsteps
     end if
  end do

  call init_RK(RK)
  call init_domain(D, trim(ifname), prec)
  if (grid_only) then
     ! the node coordinates X(q, r, s, 1:3) of both blocks, then stop
     if (len_trim(dump) == 0) stop 'usage: wql_main input.in grid dump.bin'
     open(newunit = u, file = trim(dump), access = 'stream', form = 'unformatted', status = 'replace')
     do ib = 1, 2
        write(u) D%B(ib)%X
     end do
     close(u)
     call close_domain(D)
     stop
  end if
  if (nsteps > 0) D%nt = nsteps
  call system_clock(c0, rate)
  do n = 1, D%nt
     call time_step_RK(D, D%dt, RK, n)
  end do
  call wql_check(wql_sync(D%ctx), 'wql_sync')
  call system_clock(c1)
  write(*, '(a,i0,a,es22.15,a,es22.15,a,f10.3)') 'wql_main: steps ', D%nt, ' t ', D%t, ' dt ', D%dt, &
       ' wall_ms ', 1.0d3 * real(c1 - c0, 8) / real(rate, 8)
  if (len_trim(dump) > 0) then
     open(newunit = u, file = trim(dump), access = 'stream', form = 'unformatted', status = 'replace')
     do ib = 1, 2
        allocate(F(D%B(ib)%nq, D%B(ib)%nr, D%B(ib)%ns, 9))
        call get_fields(D, ib, F)
        write(u) F
        deallocate(F)
     end do
     do ib = 1, 2
        allocate(S(D%B(ib)%nr, D%B(ib)%ns, 4), W(D%B(ib)%nr, D%B(ib)%ns, 1), tr(D%B(ib)%nr, D%B(ib)%ns, 1), &
                 Sv(D%B(ib)%nr, D%B(ib)%ns, 3))
        call get_fault(D, ib, S, W, tr, Sv)
        write(u) S, W, tr, Sv
        deallocate(S, W, tr, Sv)
     end do
     close(u)
  end if
  call close_domain(D)
end program wql_main
