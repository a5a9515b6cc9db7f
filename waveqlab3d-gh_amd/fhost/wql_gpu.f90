!> wql_gpu -- ISO_C_BINDING interface to libwql.so (include/wql.h) for a
!> Fortran host.  A WaveQLab3D-style driver keeps its own domain/block data
!> and input handling and replaces its step loop
!>     do n = 1, D%nt; call time_step_RK(D, D%dt, RK, n); end do      (main.f90:83-96)
!> with wql_create once and wql_step(ctx, nt).  Arrays cross the boundary in
!> the reference's column-major (q, r, s, field) order on the owned nodes.
module wql_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: WQL_FP64 = 0, WQL_FP32 = 1
  integer(c_int), parameter :: WQL_LOCKED = 0, WQL_SLIP_WEAKENING = 1, WQL_RATE_STATE = 2

  !> wql_block_desc (block_type fields the path reads, datatypes.f90:118-135)
  type, bind(C) :: wql_block_desc
     integer(c_int) :: nq, nr, ns
     integer(c_int) :: bc(6)
     integer(c_int) :: pml(6)
     integer(c_int) :: npml
     integer(c_int) :: cartesian
     type(c_ptr) :: X = c_null_ptr          ! real(c_double) X(nq, nr, ns, 3)
     type(c_ptr) :: M = c_null_ptr          ! real(c_double) M(nq, nr, ns, 3) or null
     real(c_double) :: rho_s_p(3)
  end type wql_block_desc

  !> wql_fault_params (Interface_Condition.f90 keyed tables made explicit)
  type, bind(C) :: wql_fault_params
     integer(c_int) :: law, stress_mode
     real(c_double) :: T0(3), s0(6)
     integer(c_int) :: npatch
     real(c_double) :: patch(5, 4)          ! patch(:, k) = (yc, zc, wy, wz, amplitude)
     integer(c_int) :: nuc
     real(c_double) :: nucp(5)
     real(c_double) :: fs, fd, Dc, bar_wy, bar_wz, bar_big
     real(c_double) :: rs_a, rs_da, rs_w, rs_wr, rs_b, rs_L, rs_V0, rs_f0, rs_Vin
     integer(c_int) :: rs_law = 0           ! 0 aging, 1 slip law + flash weakening, 2 KDN
     real(c_double) :: rs_fw = 0, rs_Vw = 0, rs_dVw = 0
     integer(c_int) :: rs_init = 0          ! 0 steady state from the prestress, 1 constant rs_W0
     real(c_double) :: rs_W0 = 0
     integer(c_int) :: sw_law = 0           ! 0 linear slip weakening, 1 TPV26/27
     real(c_double) :: coh(3) = 0, frc(5) = 0
     integer(c_int) :: rs_solver = 0        ! 0 regula falsi, 1 bracketed Newton
  end type wql_fault_params

  type, bind(C) :: wql_problem
     integer(c_int) :: precision
     real(c_double) :: dt, t0
     integer(c_int) :: plastic
     real(c_double) :: pl_c, pl_nu, pl_Tv, pl_mu, pl_beta, pl_eta
     integer(c_int) :: s0_kind
     type(wql_fault_params) :: fault
  end type wql_problem

  interface
     integer(c_int) function wql_create(prob, blocks, ctx) bind(C, name='wql_create')
       import :: c_int, c_ptr, wql_problem, wql_block_desc
       type(wql_problem), intent(in) :: prob
       type(wql_block_desc), intent(in) :: blocks(2)
       type(c_ptr), intent(out) :: ctx
     end function wql_create
     integer(c_int) function wql_destroy(ctx) bind(C, name='wql_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wql_destroy
     type(c_ptr) function wql_last_error() bind(C, name='wql_last_error')
       import :: c_ptr
     end function wql_last_error
     integer(c_int) function wql_init_state(ctx) bind(C, name='wql_init_state')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wql_init_state
     integer(c_int) function wql_set_fields(ctx, block, F) bind(C, name='wql_set_fields')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(in) :: F(*)
     end function wql_set_fields
     integer(c_int) function wql_get_fields(ctx, block, F) bind(C, name='wql_get_fields')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(out) :: F(*)
     end function wql_get_fields
     integer(c_int) function wql_get_fault(ctx, block, S, W, trup, Svel, Uhat, Vhat) bind(C, name='wql_get_fault')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       type(c_ptr), value :: S, W, trup, Svel, Uhat, Vhat
     end function wql_get_fault
     ! fields of both blocks on the fault, [9][ns][nr] each (fault output
     ! records Uface / Vface, domain.f90:741-779)
     integer(c_int) function wql_get_fault_faces(ctx, Uface, Vface) bind(C, name='wql_get_fault_faces')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), value :: Uface, Vface
     end function wql_get_fault_faces
     integer(c_int) function wql_step(ctx, nsteps) bind(C, name='wql_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function wql_step
     integer(c_int) function wql_sync(ctx) bind(C, name='wql_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wql_sync
     real(c_double) function wql_time(ctx) bind(C, name='wql_time')
       import :: c_double, c_ptr
       type(c_ptr), value :: ctx
     end function wql_time
  end interface

contains

  !> the reference stops on errors (stop 'msg'); so does this wrapper
  subroutine wql_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(wql_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
    end do
    write(*, '(a,a,i0,a,1024a)') what, ': wql error ', rc, ': ', msg(1:n)
    error stop 'wql'
  end subroutine wql_check

end module wql_gpu
