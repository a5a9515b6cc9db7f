!> wql_gpu -- ISO_C_BINDING interface to libwql.so (include/wql.h) for a
!> Fortran host: bind(C) mirrors of the descriptors and an interface for
!> every entry point.  Arrays cross the boundary in the reference's
!> column-major (q, r, s, field) order on the owned nodes.  wql_check turns
!> a non-zero return into `error stop` with the library's message, the
!> reference's `stop 'msg'` behaviour.
module wql_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: WQL_FP64 = 0, WQL_FP32 = 1
  integer(c_int), parameter :: WQL_LOCKED = 0, WQL_SLIP_WEAKENING = 1, WQL_RATE_STATE = 2
  integer(c_int), parameter :: WQL_FD_TRADITIONAL = 0, WQL_FD_UPWIND = 1, WQL_FD_UPWIND_DRP = 2
  integer(c_int), parameter :: WQL_OK = 0, WQL_E_ARG = -1, WQL_E_NOMEM = -2, WQL_E_HIP = -3, &
                               WQL_E_NONFINITE = -4, WQL_E_SOLVER = -5, WQL_E_STATE = -6

  !> wql_block_desc (block_type fields the path reads, datatypes.f90:118-135)
  type, bind(C) :: wql_block_desc
     integer(c_int) :: nq = 0, nr = 0, ns = 0
     integer(c_int) :: bc(6) = 0
     integer(c_int) :: pml(6) = 0
     integer(c_int) :: npml = 0
     integer(c_int) :: cartesian = 1
     type(c_ptr) :: X = c_null_ptr          ! real(c_double) X(nq, nr, ns, 3)
     type(c_ptr) :: M = c_null_ptr          ! real(c_double) M(nq, nr, ns, 3) or null
     real(c_double) :: rho_s_p(3) = 0
  end type wql_block_desc

  !> wql_fault_params (Interface_Condition.f90 keyed tables made explicit)
  type, bind(C) :: wql_fault_params
     integer(c_int) :: law = 0, stress_mode = 0
     real(c_double) :: T0(3) = 0, s0(6) = 0
     integer(c_int) :: npatch = 0
     real(c_double) :: patch(5, 4) = 0      ! patch(:, k) = (yc, zc, wy, wz, amplitude)
     integer(c_int) :: nuc = 0
     real(c_double) :: nucp(5) = 0
     real(c_double) :: fs = 0, fd = 0, Dc = 1, bar_wy = 1d30, bar_wz = 1d30, bar_big = 0
     real(c_double) :: rs_a = 0, rs_da = 0, rs_w = 0, rs_wr = 1, rs_b = 0, rs_L = 1, rs_V0 = 1, rs_f0 = 0, rs_Vin = 0
     integer(c_int) :: rs_law = 0           ! 0 aging, 1 slip law + flash weakening, 2 KDN
     real(c_double) :: rs_fw = 0, rs_Vw = 0, rs_dVw = 0
     integer(c_int) :: rs_init = 0          ! 0 steady state from the prestress, 1 constant rs_W0
     real(c_double) :: rs_W0 = 0
     integer(c_int) :: sw_law = 0           ! 0 linear slip weakening, 1 TPV26/27
     real(c_double) :: coh(3) = 0, frc(5) = [0d0, 0d0, 1d0, 1d0, 1d0]
     integer(c_int) :: rs_solver = 0        ! 0 regula falsi, 1 bracketed Newton
     integer(c_int) :: key = 0              ! WQL_KEY_*: per-node fault fields of a keyed problem
  end type wql_fault_params

  type, bind(C) :: wql_problem
     integer(c_int) :: precision = WQL_FP64
     real(c_double) :: dt = 0, t0 = 0
     integer(c_int) :: plastic = 0
     real(c_double) :: pl_c = 1.36d0, pl_nu = 0.1934d0, pl_Tv = 0.03d0, pl_mu = 0, pl_beta = 0, pl_eta = 0
     integer(c_int) :: s0_kind = 0
     type(wql_fault_params) :: fault
     integer(c_int) :: fd_type = WQL_FD_TRADITIONAL  ! &problem_list fd_type
     integer(c_int) :: fd_order = 5                  ! &problem_list order
     integer(c_int) :: anelastic = 0                 ! response = 'anelastic'
     real(c_double) :: an_c = 1, an_weight_exp = 0, an_fref = 1  ! &anelastic_list
  end type wql_problem

  !> wql_grid_desc (init_grid_cartesian / init_grid_curve, grid.f90:13-144)
  type, bind(C) :: wql_grid_desc
     integer(c_int) :: nq = 0, nr = 0, ns = 0
     real(c_double) :: a(3) = 0, b(3) = 0
     integer(c_int) :: mesh = 0              ! 0 uniform, 1 transfinite interpolation
     integer(c_int) :: profile = 0           ! 0 planar q faces, 1 'analytical_test_problem',
                                             ! 2 'analytical_tpv28', 3 'analytical_tpv10', 4 'analytical_tpv36'
     real(c_double) :: lc = 0, rc = 0
     type(c_ptr) :: fault_x = c_null_ptr     ! optional fault_x(nr, ns)
     integer(c_int) :: fault_side = 0
     type(c_ptr) :: top_y = c_null_ptr       ! optional top_y(nq, ns)
  end type wql_grid_desc

  !> wql_source: one moment-tensor point source (moment_tensor.f90:82-270)
  type, bind(C) :: wql_source
     integer(c_int) :: type = 0               ! WQL_STF_*: 0 gaussian ... 6 sine_moment_time_history
     real(c_double) :: duration = 0, t_init = 0
     real(c_double) :: m(6) = 0               ! mXX mYY mZZ mXY mXZ mYZ
     real(c_double) :: x(3) = 0
     integer(c_int) :: node(3) = -2           ! 0-based nearest node, -2 where the block does not hold it
     real(c_double) :: node_x(3) = -1
  end type wql_source

  interface
     ! ---- set-up ----------------------------------------------------------
     integer(c_int) function wql_problem_keys(problem, coupling, response, plastic_model, mu_beta_eta, prob) &
          bind(C, name='wql_problem_keys')
       import :: c_int, c_char, c_double, wql_problem
       character(kind=c_char), intent(in) :: problem(*), coupling(*), response(*), plastic_model(*)
       real(c_double), intent(in) :: mu_beta_eta(3)
       type(wql_problem), intent(inout) :: prob
     end function wql_problem_keys
     integer(c_int) function wql_grid_block(grid, X) bind(C, name='wql_grid_block')
       import :: c_int, c_double, wql_grid_desc
       type(wql_grid_desc), intent(in) :: grid
       real(c_double), intent(out) :: X(*)
     end function wql_grid_block
     integer(c_int) function wql_grid_faces(nq, nr, ns, faces, X) bind(C, name='wql_grid_faces')
       import :: c_int, c_double
       integer(c_int), value :: nq, nr, ns
       real(c_double), intent(in) :: faces(*)
       real(c_double), intent(out) :: X(*)
     end function wql_grid_faces
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
     integer(c_int) function wql_problem_material(problem, nq, nr, ns, X, M) bind(C, name='wql_problem_material')
       import :: c_int, c_char, c_double
       character(kind=c_char), intent(in) :: problem(*)
       integer(c_int), value :: nq, nr, ns
       real(c_double), intent(in) :: X(*)
       real(c_double), intent(out) :: M(*)
     end function wql_problem_material
     integer(c_int) function wql_abi_version() bind(C, name='wql_abi_version')
       import :: c_int
     end function wql_abi_version
     integer(c_size_t) function wql_device_bytes(ctx) bind(C, name='wql_device_bytes')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: ctx
     end function wql_device_bytes
     integer(c_size_t) function wql_estimate_bytes(prob, blocks) bind(C, name='wql_estimate_bytes')
       import :: c_size_t, wql_problem, wql_block_desc
       type(wql_problem), intent(in) :: prob
       type(wql_block_desc), intent(in) :: blocks(2)
     end function wql_estimate_bytes
     integer(c_int) function wql_init_state(ctx) bind(C, name='wql_init_state')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wql_init_state
     ! ---- state ------------------------------------------------------------
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
     integer(c_int) function wql_get_fields_noalloc(ctx, block, F) bind(C, name='wql_get_fields_noalloc')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(out) :: F(*)
     end function wql_get_fields_noalloc
     integer(c_int) function wql_get_rates(ctx, block, DF) bind(C, name='wql_get_rates')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(out) :: DF(*)
     end function wql_get_rates
     integer(c_int) function wql_set_pml(ctx, block, side, Q) bind(C, name='wql_set_pml')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block, side
       real(c_double), intent(in) :: Q(*)
     end function wql_set_pml
     integer(c_int) function wql_get_pml(ctx, block, side, Q) bind(C, name='wql_get_pml')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block, side
       real(c_double), intent(out) :: Q(*)
     end function wql_get_pml
     integer(c_int) function wql_get_pml_rates(ctx, block, side, DQ) bind(C, name='wql_get_pml_rates')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block, side
       real(c_double), intent(out) :: DQ(*)
     end function wql_get_pml_rates
     integer(c_int) function wql_get_metrics(ctx, block, metric) bind(C, name='wql_get_metrics')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(out) :: metric(*)
     end function wql_get_metrics
     integer(c_int) function wql_get_frame(ctx, block, side, nml) bind(C, name='wql_get_frame')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block, side
       real(c_double), intent(out) :: nml(*)
     end function wql_get_frame
     integer(c_int) function wql_get_fault_faces(ctx, Uface, Vface) bind(C, name='wql_get_fault_faces')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), value :: Uface, Vface
     end function wql_get_fault_faces
     integer(c_int) function wql_find_stations(X, nq, nr, ns, n, xyz, ijk) bind(C, name='wql_find_stations')
       import :: c_int, c_double
       real(c_double), intent(in) :: X(*)
       integer(c_int), value :: nq, nr, ns, n
       real(c_double), intent(in) :: xyz(*)
       integer(c_int), intent(out) :: ijk(*)
     end function wql_find_stations
     integer(c_int) function wql_locate_sources(X, nq, nr, ns, n, xyz, src) bind(C, name='wql_locate_sources')
       import :: c_int, c_double, wql_source
       real(c_double), intent(in) :: X(*)
       integer(c_int), value :: nq, nr, ns, n
       real(c_double), intent(in) :: xyz(*)
       type(wql_source), intent(inout) :: src(*)
     end function wql_locate_sources
     integer(c_int) function wql_set_sources(ctx, block, n, src, order, alpha) bind(C, name='wql_set_sources')
       import :: c_int, c_ptr, wql_source
       type(c_ptr), value :: ctx
       integer(c_int), value :: block, n
       type(wql_source), intent(in) :: src(*)
       integer(c_int), value :: order, alpha
     end function wql_set_sources
     integer(c_int) function wql_set_sources_smooth(ctx, block, n, src, X) bind(C, name='wql_set_sources_smooth')
       import :: c_int, c_ptr, c_double, wql_source
       type(c_ptr), value :: ctx
       integer(c_int), value :: block, n
       type(wql_source), intent(in) :: src(*)
       real(c_double), intent(in) :: X(*)
     end function wql_set_sources_smooth
     integer(c_int) function wql_set_stations(ctx, block, n, ijk) bind(C, name='wql_set_stations')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: block, n
       integer(c_int), intent(in) :: ijk(*)
     end function wql_set_stations
     integer(c_int) function wql_get_stations(ctx, block, v) bind(C, name='wql_get_stations')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(out) :: v(*)
     end function wql_get_stations
     integer(c_int) function wql_set_fault(ctx, block, S, W) bind(C, name='wql_set_fault')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       type(c_ptr), value :: S, W
     end function wql_set_fault
     integer(c_int) function wql_get_fault(ctx, block, S, W, trup, Svel, Uhat, Vhat) bind(C, name='wql_get_fault')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       type(c_ptr), value :: S, W, trup, Svel, Uhat, Vhat
     end function wql_get_fault
     integer(c_int) function wql_get_fault_rates(ctx, block, DS, DW) bind(C, name='wql_get_fault_rates')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       type(c_ptr), value :: DS, DW
     end function wql_get_fault_rates
     integer(c_int) function wql_get_plastic(ctx, block, P) bind(C, name='wql_get_plastic')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(out) :: P(*)
     end function wql_get_plastic
     integer(c_int) function wql_set_anelastic(ctx, block, eta) bind(C, name='wql_set_anelastic')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(in) :: eta(*)
     end function wql_set_anelastic
     integer(c_int) function wql_get_anelastic(ctx, block, what, out) bind(C, name='wql_get_anelastic')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block, what
       real(c_double), intent(out) :: out(*)
     end function wql_get_anelastic
     integer(c_int) function wql_get_fault_peak(ctx, block, psr) bind(C, name='wql_get_fault_peak')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(out) :: psr(*)
     end function wql_get_fault_peak
     integer(c_int) function wql_fault_stats(ctx, block, stats) bind(C, name='wql_fault_stats')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: block
       real(c_double), intent(out) :: stats(6)
     end function wql_fault_stats
     ! ---- stepping ---------------------------------------------------------
     integer(c_int) function wql_eval_rates(ctx, t, what, stage) bind(C, name='wql_eval_rates')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: t
       integer(c_int), value :: what, stage
     end function wql_eval_rates
     integer(c_int) function wql_step(ctx, nsteps) bind(C, name='wql_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function wql_step
     integer(c_int) function wql_stage(ctx, stage) bind(C, name='wql_stage')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: stage
     end function wql_stage
     integer(c_int) function wql_sync(ctx) bind(C, name='wql_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function wql_sync
     real(c_double) function wql_time(ctx) bind(C, name='wql_time')
       import :: c_double, c_ptr
       type(c_ptr), value :: ctx
     end function wql_time
     ! ---- launch shape, graphs, profiling ------------------------------------
     integer(c_int) function wql_set_tile(ctx, variant, zchunk) bind(C, name='wql_set_tile')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: variant, zchunk
     end function wql_set_tile
     integer(c_int) function wql_get_tile(ctx, variant, zchunk) bind(C, name='wql_get_tile')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: variant, zchunk
     end function wql_get_tile
     integer(c_int) function wql_set_tile_order(ctx, order, group) bind(C, name='wql_set_tile_order')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: order, group
     end function wql_set_tile_order
     integer(c_int) function wql_get_tile_order(ctx, order, group) bind(C, name='wql_get_tile_order')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: order, group
     end function wql_get_tile_order
     integer(c_int) function wql_tile_variants() bind(C, name='wql_tile_variants')
       import :: c_int
     end function wql_tile_variants
     integer(c_int) function wql_use_graph(ctx, on) bind(C, name='wql_use_graph')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function wql_use_graph
     integer(c_int) function wql_profile(ctx, on) bind(C, name='wql_profile')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function wql_profile
     real(c_double) function wql_volume_kernel_ms(ctx) bind(C, name='wql_volume_kernel_ms')
       import :: c_double, c_ptr
       type(c_ptr), value :: ctx
     end function wql_volume_kernel_ms
     ! ---- serving ----------------------------------------------------------
     integer(c_int) function wql_server_create(max_active, hbm_budget, steps_per_launch, srv) &
          bind(C, name='wql_server_create')
       import :: c_int, c_size_t, c_ptr
       integer(c_int), value :: max_active, steps_per_launch
       integer(c_size_t), value :: hbm_budget
       type(c_ptr), intent(out) :: srv
     end function wql_server_create
     integer(c_int) function wql_server_destroy(srv) bind(C, name='wql_server_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: srv
     end function wql_server_destroy
     integer(c_int) function wql_server_submit(srv, prob, blocks, nsteps, id) bind(C, name='wql_server_submit')
       import :: c_int, c_ptr, wql_problem, wql_block_desc
       type(c_ptr), value :: srv
       type(wql_problem), intent(in) :: prob
       type(wql_block_desc), intent(in) :: blocks(2)
       integer(c_int), value :: nsteps
       integer(c_int), intent(out) :: id
     end function wql_server_submit
     integer(c_int) function wql_server_poll(srv, done, max_done) bind(C, name='wql_server_poll')
       import :: c_int, c_ptr
       type(c_ptr), value :: srv
       integer(c_int), intent(out) :: done(*)
       integer(c_int), value :: max_done
     end function wql_server_poll
     integer(c_int) function wql_server_result(srv, id, S, trup, psr, stats) bind(C, name='wql_server_result')
       import :: c_int, c_ptr
       type(c_ptr), value :: srv
       integer(c_int), value :: id
       type(c_ptr), value :: S, trup, psr, stats
     end function wql_server_result
     integer(c_int) function wql_server_status(srv, out) bind(C, name='wql_server_status')
       import :: c_int, c_ptr, c_long
       type(c_ptr), value :: srv
       integer(c_long), intent(out) :: out(4)
     end function wql_server_status
     integer(c_int) function wql_server_set_policy(srv, policy) bind(C, name='wql_server_set_policy')
       import :: c_int, c_ptr
       type(c_ptr), value :: srv
       integer(c_int), value :: policy
     end function wql_server_set_policy
  end interface

contains

  !> the library's last error message as a Fortran string
  function wql_error_message() result(text)
    character(len=:), allocatable :: text
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    call c_f_pointer(wql_last_error(), msg, [1024])
    n = 0
    do while (n < 1024)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: text)
    text = transfer(msg(1:n), text)
  end function wql_error_message

  !> the reference stops on errors (stop 'msg'); so does this wrapper
  subroutine wql_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    if (rc == 0) return
    write(*, '(a,a,i0,a,a)') what, ': wql error ', rc, ': ', wql_error_message()
    error stop 'wql'
  end subroutine wql_check

  !> a NUL-terminated copy of a Fortran string for the C ABI
  function c_string(text) result(cs)
    character(*), intent(in) :: text
    character(kind=c_char, len=len_trim(text) + 1) :: cs
    cs = trim(text) // c_null_char
  end function c_string

end module wql_gpu
