!> Golden-vector driver linked against the reference modules (built only in
!> the development container, see oracle/Makefile target `ref`).
!>
!> usage (always 2 MPI ranks, one block per rank, as the reference's fault
!> accumulation requires -- domain.f90:600-609):
!>
!>   mpiexec -n 2 _ref/golden <input.in> rhs  <out_prefix> <t>
!>   mpiexec -n 2 _ref/golden <input.in> traj <out_prefix> <nsteps> [x0 y0 z0 sigma amp]
!>
!> rhs : fills F (and PML Q, interface S/W) with the deterministic pattern of
!>       fill_pattern below, then dumps DF after set_rates (domain.f90:908) and
!>       again after enforce_bound_iface_conditions (domain.f90:586) at time t,
!>       stage 2 (no output side effects).
!> traj: runs nsteps of time_step_RK (time_step.f90:68) from the reference
!>       initial condition (optionally plus a Gaussian velocity pulse on both
!>       blocks) and dumps the final state.
!>
!> Every dump is raw little-endian float64 in Fortran order of the owned
!> range, file <out_prefix>_b<block>_<name>.bin.
program golden_driver

  use mpi
  use mpi3dbasic, only : start_mpi, finish_mpi
  use common, only : wp
  use datatypes, only : domain_type
  use domain, only : init_domain, in_block_comm, set_rates, exchange_fields_interface, &
                     enforce_bound_iface_conditions, scale_rates
  use time_step, only : RK_type, init_RK, time_step_RK

  implicit none

  type(domain_type) :: D
  type(RK_type) :: RK
  character(256) :: infile, mode, prefix, arg
  integer :: ierr, rank, nsteps, n, ib, unit_in, nargs
  real(kind = wp) :: t, x0, y0, z0, sig, amp
  logical :: pulse

  call start_mpi
  call MPI_Comm_rank(MPI_COMM_WORLD, rank, ierr)

  call get_command_argument(1, infile)
  call get_command_argument(2, mode)
  call get_command_argument(3, prefix)
  call get_command_argument(4, arg)
  nargs = command_argument_count()

  open(newunit = unit_in, file = trim(infile), status = 'old', action = 'read')
  call init_RK(RK)
  call init_domain(D, unit_in, trim(infile))

  ib = 1
  if (in_block_comm(2) .and. .not. in_block_comm(1)) ib = 2

  call dump_geometry(D, ib, prefix)
  ! the material of init_material (material.f90:179-690; with the anelastic
  ! response the unrelaxed moduli of init_anelastic_properties, 11-145, and
  ! Q^-1)
  call dump3(D%B(ib)%M%M, D, ib, prefix, 'Mat')
  if (D%B(ib)%M%anelastic) then
     call dump3(reshape(D%B(ib)%M%Qs_inv, [shape(D%B(ib)%M%Qs_inv), 1]), D, ib, prefix, 'Qsinv')
  end if

  select case (trim(mode))
  case ('rhs')
     read(arg, *) t
     D%t = t
     call fill_pattern(D, ib)
     call set_rates(D)
     call dump3(D%B(ib)%F%DF, D, ib, prefix, 'DFvol')
     call dump_anelastic(D, ib, prefix, 'Deta')
     call dump_pml(D, ib, prefix, 'DQvol')
     call exchange_fields_interface(D)
     call enforce_bound_iface_conditions(D, 2)
     call dump3(D%B(ib)%F%DF, D, ib, prefix, 'DFfull')
     call dump_pml(D, ib, prefix, 'DQfull')
     call dump_face(D%I(1)%DS, D, ib, prefix, 'DS')
     call dump_face(D%I(1)%DW, D, ib, prefix, 'DW')
  case ('traj')
     read(arg, *) nsteps
     pulse = (nargs >= 9)
     if (pulse) then
        call get_command_argument(5, arg); read(arg, *) x0
        call get_command_argument(6, arg); read(arg, *) y0
        call get_command_argument(7, arg); read(arg, *) z0
        call get_command_argument(8, arg); read(arg, *) sig
        call get_command_argument(9, arg); read(arg, *) amp
        call add_pulse(D, ib, x0, y0, z0, sig, amp)
     end if
     call dump3(D%B(ib)%F%F, D, ib, prefix, 'F0')
     call dump_face(D%I(1)%W, D, ib, prefix, 'W0')
     do n = 1, nsteps
        call time_step_RK(D, D%dt, RK, n)
     end do
     call dump3(D%B(ib)%F%F, D, ib, prefix, 'F')
     call dump_face(D%I(1)%S, D, ib, prefix, 'S')
     call dump_face(D%I(1)%W, D, ib, prefix, 'W')
     call dump_face(D%I(1)%trup, D, ib, prefix, 'trup')
     call dump_face(D%I(1)%Svel, D, ib, prefix, 'Svel')
     call dump_pml(D, ib, prefix, 'Q')
     if (allocated(D%B(ib)%P%P)) call dump3(D%B(ib)%P%P, D, ib, prefix, 'P')
     call dump_anelastic(D, ib, prefix, 'eta')
     if (rank == 0) then
        open(newunit = unit_in, file = trim(prefix)//'_meta.txt', status = 'replace')
        write(unit_in, '(a,es25.17)') 'dt ', D%dt
        write(unit_in, '(a,es25.17)') 't ', D%t
        write(unit_in, '(a,i0)') 'nsteps ', nsteps
        close(unit_in)
     end if
  case default
     stop 'golden_driver: unknown mode'
  end select

  call finish_mpi

contains

  !> pattern value in [-1,1]: an integer hash of (i,j,k,f,block) divided in
  !> double precision, reproducible bit-for-bit in numpy
  real(kind = wp) function pat(i, j, k, f, blk)
    integer, intent(in) :: i, j, k, f, blk
    integer(8) :: idx, h
    idx = int(i, 8) + 1000_8*int(j, 8) + 1000000_8*int(k, 8) + 1000000000_8*int(f, 8)
    h = mod(idx*7919_8 + 104729_8*int(blk, 8), 1000003_8)
    pat = 2.0_wp*(real(h, wp)/1000003.0_wp) - 1.0_wp
  end function pat

  subroutine fill_pattern(D, ib)
    type(domain_type), intent(inout) :: D
    integer, intent(in) :: ib
    integer :: i, j, k, f, side
    real(kind = wp) :: scale
    associate(C => D%B(ib)%G%C)
      do f = 1, 9
         scale = 1.0_wp
         if (f >= 4) scale = 10.0_wp
         do k = C%ms, C%ps
            do j = C%mr, C%pr
               do i = C%mq, C%pq
                  D%B(ib)%F%F(i, j, k, f) = scale*pat(i, j, k, f, ib)
               end do
            end do
         end do
      end do
      D%B(ib)%F%DF = 0.0_wp
      do side = 1, 6
         if (.not. D%B(ib)%PMLB(side)%pml) cycle
         if (.not. allocated(D%B(ib)%PMLB(side)%Q)) cycle
         do f = 1, 9
            do k = lbound(D%B(ib)%PMLB(side)%Q, 3), ubound(D%B(ib)%PMLB(side)%Q, 3)
               do j = lbound(D%B(ib)%PMLB(side)%Q, 2), ubound(D%B(ib)%PMLB(side)%Q, 2)
                  do i = lbound(D%B(ib)%PMLB(side)%Q, 1), ubound(D%B(ib)%PMLB(side)%Q, 1)
                     D%B(ib)%PMLB(side)%Q(i, j, k, f) = 0.1_wp*pat(i, j, k, f + 10*side, ib + 2)
                  end do
               end do
            end do
         end do
         D%B(ib)%PMLB(side)%DQ = 0.0_wp
      end do
      do k = C%ms, C%ps
         do j = C%mr, C%pr
            do f = 1, 4
               D%I(1)%S(j, k, f) = 0.25_wp*(pat(0, j, k, f, 7) + 1.0_wp)
            end do
            D%I(1)%W(j, k, 1) = 0.6_wp + 0.05_wp*pat(0, j, k, 0, 8)
         end do
      end do
      D%I(1)%DS = 0.0_wp
      D%I(1)%DW = 0.0_wp
      ! anelastic memory variables: component c (4..9), mechanism l (1..4)
      if (D%B(ib)%M%anelastic) then
         do k = C%ms, C%ps
            do j = C%mr, C%pr
               do i = C%mq, C%pq
                  do f = 1, 4
                     D%B(ib)%M%eta4(i, j, k, f) = 0.05_wp*pat(i, j, k, 30 + f, ib + 4)
                     D%B(ib)%M%eta5(i, j, k, f) = 0.05_wp*pat(i, j, k, 34 + f, ib + 4)
                     D%B(ib)%M%eta6(i, j, k, f) = 0.05_wp*pat(i, j, k, 38 + f, ib + 4)
                     D%B(ib)%M%eta7(i, j, k, f) = 0.05_wp*pat(i, j, k, 42 + f, ib + 4)
                     D%B(ib)%M%eta8(i, j, k, f) = 0.05_wp*pat(i, j, k, 46 + f, ib + 4)
                     D%B(ib)%M%eta9(i, j, k, f) = 0.05_wp*pat(i, j, k, 50 + f, ib + 4)
                  end do
               end do
            end do
         end do
         D%B(ib)%M%Deta4 = 0.0_wp; D%B(ib)%M%Deta5 = 0.0_wp; D%B(ib)%M%Deta6 = 0.0_wp
         D%B(ib)%M%Deta7 = 0.0_wp; D%B(ib)%M%Deta8 = 0.0_wp; D%B(ib)%M%Deta9 = 0.0_wp
      end if
    end associate
  end subroutine fill_pattern

  !> the anelastic memory variables (name 'eta') or their rates ('Deta') of
  !> block ib as six dumps of 4 mechanisms each: <name>4 .. <name>9
  subroutine dump_anelastic(D, ib, prefix, name)
    type(domain_type), intent(in) :: D
    integer, intent(in) :: ib
    character(*), intent(in) :: prefix, name
    if (.not. D%B(ib)%M%anelastic) return
    if (name == 'eta') then
       call dump3(D%B(ib)%M%eta4, D, ib, prefix, 'eta4')
       call dump3(D%B(ib)%M%eta5, D, ib, prefix, 'eta5')
       call dump3(D%B(ib)%M%eta6, D, ib, prefix, 'eta6')
       call dump3(D%B(ib)%M%eta7, D, ib, prefix, 'eta7')
       call dump3(D%B(ib)%M%eta8, D, ib, prefix, 'eta8')
       call dump3(D%B(ib)%M%eta9, D, ib, prefix, 'eta9')
    else
       call dump3(D%B(ib)%M%Deta4, D, ib, prefix, 'Deta4')
       call dump3(D%B(ib)%M%Deta5, D, ib, prefix, 'Deta5')
       call dump3(D%B(ib)%M%Deta6, D, ib, prefix, 'Deta6')
       call dump3(D%B(ib)%M%Deta7, D, ib, prefix, 'Deta7')
       call dump3(D%B(ib)%M%Deta8, D, ib, prefix, 'Deta8')
       call dump3(D%B(ib)%M%Deta9, D, ib, prefix, 'Deta9')
    end if
  end subroutine dump_anelastic

  subroutine add_pulse(D, ib, x0, y0, z0, sig, amp)
    type(domain_type), intent(inout) :: D
    integer, intent(in) :: ib
    real(kind = wp), intent(in) :: x0, y0, z0, sig, amp
    integer :: i, j, k
    real(kind = wp) :: r2
    associate(C => D%B(ib)%G%C, X => D%B(ib)%G%X)
      do k = C%ms, C%ps
         do j = C%mr, C%pr
            do i = C%mq, C%pq
               r2 = (X(i, j, k, 1) - x0)**2 + (X(i, j, k, 2) - y0)**2 + (X(i, j, k, 3) - z0)**2
               D%B(ib)%F%F(i, j, k, 1:3) = D%B(ib)%F%F(i, j, k, 1:3) + amp*exp(-r2/(2.0_wp*sig**2))
            end do
         end do
      end do
    end associate
  end subroutine add_pulse

  subroutine write_raw(fname, a, n)
    character(*), intent(in) :: fname
    integer, intent(in) :: n
    real(kind = wp), intent(in) :: a(n)
    integer :: u
    open(newunit = u, file = fname, access = 'stream', form = 'unformatted', status = 'replace')
    write(u) a
    close(u)
  end subroutine write_raw

  subroutine dump3(A, D, ib, prefix, name)
    real(kind = wp), intent(in) :: A(:,:,:,:)
    type(domain_type), intent(in) :: D
    integer, intent(in) :: ib
    character(*), intent(in) :: prefix, name
    real(kind = wp), allocatable :: tmp(:,:,:,:)
    integer :: lo(3)
    associate(C => D%B(ib)%G%C)
      if (size(A, 1) == C%pq - C%mq + 1 + 2*D%B(ib)%nb) then
         lo = [1 + D%B(ib)%nb, 1 + D%B(ib)%nb, 1 + D%B(ib)%nb]
      else
         lo = [1, 1, 1]
      end if
      allocate(tmp(C%nq, C%nr, C%ns, size(A, 4)))
      tmp = A(lo(1):lo(1)+C%nq-1, lo(2):lo(2)+C%nr-1, lo(3):lo(3)+C%ns-1, :)
      call write_raw(trim(prefix)//'_b'//achar(48+ib)//'_'//name//'.bin', tmp, size(tmp))
    end associate
  end subroutine dump3

  subroutine dump_face(A, D, ib, prefix, name)
    real(kind = wp), intent(in) :: A(:,:,:)
    type(domain_type), intent(in) :: D
    integer, intent(in) :: ib
    character(*), intent(in) :: prefix, name
    real(kind = wp), allocatable :: tmp(:,:,:)
    integer :: nb
    nb = D%B(ib)%nb
    associate(C => D%B(ib)%G%C)
      allocate(tmp(C%nr, C%ns, size(A, 3)))
      tmp = A(1+nb:nb+C%nr, 1+nb:nb+C%ns, :)
      call write_raw(trim(prefix)//'_b'//achar(48+ib)//'_'//name//'.bin', tmp, size(tmp))
    end associate
  end subroutine dump_face

  subroutine dump_pml(D, ib, prefix, name)
    type(domain_type), intent(in) :: D
    integer, intent(in) :: ib
    character(*), intent(in) :: prefix, name
    integer :: side
    real(kind = wp), allocatable :: tmp(:)
    do side = 1, 6
       if (.not. D%B(ib)%PMLB(side)%pml) cycle
       if (.not. allocated(D%B(ib)%PMLB(side)%Q)) cycle
       if (name == 'Q') then
          tmp = reshape(D%B(ib)%PMLB(side)%Q, [size(D%B(ib)%PMLB(side)%Q)])
       else
          tmp = reshape(D%B(ib)%PMLB(side)%DQ, [size(D%B(ib)%PMLB(side)%DQ)])
       end if
       call write_raw(trim(prefix)//'_b'//achar(48+ib)//'_'//name//achar(48+side)//'.bin', tmp, size(tmp))
    end do
  end subroutine dump_pml

  subroutine dump_geometry(D, ib, prefix)
    type(domain_type), intent(in) :: D
    integer, intent(in) :: ib
    character(*), intent(in) :: prefix
    integer :: side
    call dump3(D%B(ib)%G%X, D, ib, prefix, 'X')
    call dump3(D%B(ib)%G%metricx, D, ib, prefix, 'metricx')
    call dump3(D%B(ib)%G%metricy, D, ib, prefix, 'metricy')
    call dump3(D%B(ib)%G%metricz, D, ib, prefix, 'metricz')
    call dump3(reshape(D%B(ib)%G%J, [shape(D%B(ib)%G%J), 1]), D, ib, prefix, 'J')
    do side = 1, 6
       call dump_nrm(D, ib, side, prefix)
    end do
  end subroutine dump_geometry

  subroutine dump_nrm(D, ib, side, prefix)
    type(domain_type), intent(in) :: D
    integer, intent(in) :: ib, side
    character(*), intent(in) :: prefix
    real(kind = wp), allocatable :: tmp(:)
    tmp = [reshape(D%B(ib)%B(side)%n_n, [size(D%B(ib)%B(side)%n_n)]), &
           reshape(D%B(ib)%B(side)%n_m, [size(D%B(ib)%B(side)%n_m)]), &
           reshape(D%B(ib)%B(side)%n_l, [size(D%B(ib)%B(side)%n_l)])]
    call write_raw(trim(prefix)//'_b'//achar(48+ib)//'_nml'//achar(48+side)//'.bin', tmp, size(tmp))
  end subroutine dump_nrm

end program golden_driver
