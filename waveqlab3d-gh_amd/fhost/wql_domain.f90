!> wql_domain -- the reference's domain / time-stepping API for the Fortran
!> host, over libwql: `init_domain(D, infile)`, `time_step_RK(D, dt, RK, n)`
!> and `close_domain(D)` keep the call shape of domain.f90:31-584 and
!> time_step.f90:20-150, so the reference's main program (main.f90:58-96)
!> runs unchanged on top of them.  The whole block state lives on the GPU;
!> `D%B(ib)%X` keeps the node coordinates on the host, and the fields reach
!> the host only through `get_fields`.
module wql_domain
  use iso_c_binding
  use wql_gpu
  use wql_input
  implicit none
  private
  public :: domain_type, block_type, station_set, RK_type, init_RK, init_domain, time_step_RK, close_domain, &
            get_fields, get_fault, wp

  integer, parameter :: wp = c_double

  !> Carpenter-Kennedy LSRK(5,4) (time_step.f90:20-65); the library holds
  !> the same table and uses its own copy
  type :: RK_type
     integer :: nstage = 5
     real(kind = wp) :: A(5), B(5), C(0:5)
  end type RK_type

  type :: block_type
     integer :: nq = 0, nr = 0, ns = 0
     real(kind = wp), allocatable :: X(:,:,:,:)   ! X(q, r, s, 1:3)
     real(kind = wp), allocatable :: M(:,:,:,:)   ! M(q, r, s, 1:3) of a keyed medium (material.f90:179-690)
  end type block_type

  !> one block's stations (seismogram_type, datatypes.f90): the located
  !> stations' output units
  type :: station_set
     integer :: n = 0
     integer, allocatable :: units(:)
  end type station_set

  type :: domain_type
     type(problem_params) :: p                   ! the input namelists
     character(len=256) :: name, problem, coupling, response
     real(kind = wp) :: t = 0, dt = 0, t_final = 0, CFL = 0.5_wp
     integer :: nt = 0, w_stride = 1
     logical :: w_fault = .false.
     type(block_type) :: B(2)
     type(wql_problem) :: prob
     type(wql_block_desc) :: desc(2)
     type(c_ptr) :: ctx = c_null_ptr
     integer :: fault_units(8) = -1              ! the eight fault output files
     type(output_params) :: out                  ! &output_list
     type(station_set) :: st(2)                  ! station seismograms per block
  end type domain_type

  character(len=5), parameter :: EXTS(8) = ['Uface', 'Vface', 'S    ', 'Uhat ', 'Vhat ', 'Svel ', 'trup ', 'state']

contains

  subroutine init_RK(RK)
    type(RK_type), intent(out) :: RK
    RK%nstage = 5
    RK%A = [0.0_wp, -567301805773.0_wp/1357537059087.0_wp, -2404267990393.0_wp/2016746695238.0_wp, &
            -3550918686646.0_wp/2091501179385.0_wp, -1275806237668.0_wp/842570457699.0_wp]
    RK%B = [1432997174477.0_wp/9575080441755.0_wp, 5161836677717.0_wp/13612068292357.0_wp, &
            1720146321549.0_wp/2090206949498.0_wp, 3134564353537.0_wp/4481467310338.0_wp, &
            2277821191437.0_wp/14882151754819.0_wp]
    RK%C = [0.0_wp, 1432997174477.0_wp/9575080441755.0_wp, 2526269341429.0_wp/6820363962896.0_wp, &
            2006345519317.0_wp/3224310063776.0_wp, 2802321613138.0_wp/2924317926251.0_wp, 1.0_wp]
  end subroutine init_RK

  !> block_time_step (block.f90:437-452)
  real(kind = wp) function block_time_step(b, CFL)
    type(block_temp_parameters), intent(in) :: b
    real(kind = wp), intent(in) :: CFL
    real(kind = wp) :: h(3)
    integer :: d
    do d = 1, 3
       h(d) = (b%bqrs(d) - b%aqrs(d)) / real(b%nqrs(d) - 1, wp)
    end do
    block_time_step = CFL * minval(h) / sqrt(b%rho_s_p(2)**2 + b%rho_s_p(3)**2)
  end function block_time_step

  !> init_domain (domain.f90:31-564): namelists, time step, grids, the
  !> device context; precision is WQL_FP64 (default) or WQL_FP32
  subroutine init_domain(D, ifname, precision)
    type(domain_type), intent(inout), target :: D
    character(*), intent(in) :: ifname
    integer(c_int), intent(in), optional :: precision
    type(wql_grid_desc) :: g
    integer :: ib, k, u, stat
    real(kind = wp) :: dt(2)
    real(kind = wp), allocatable, target :: fx(:,:)

    call read_input(ifname, D%p)
    D%name = D%p%name; D%problem = D%p%problem; D%coupling = D%p%coupling; D%response = D%p%response
    D%t_final = D%p%t_final; D%CFL = D%p%CFL; D%w_stride = D%p%w_stride; D%w_fault = D%p%w_fault
    if (D%p%nblocks /= 2) error stop 'init_domain: the path runs two blocks joined by one interface'
    D%t = 0.0_wp
    do ib = 1, 2
       dt(ib) = block_time_step(D%p%btp(ib), D%CFL)
    end do
    D%dt = min(dt(1), dt(2))
    D%nt = floor(D%t_final / D%dt)                       ! domain.f90:494

    do ib = 1, 2
       associate(b => D%p%btp(ib))
         D%B(ib)%nq = b%nqrs(1); D%B(ib)%nr = b%nqrs(2); D%B(ib)%ns = b%nqrs(3)
         allocate(D%B(ib)%X(b%nqrs(1), b%nqrs(2), b%nqrs(3), 3))
         g%nq = b%nqrs(1); g%nr = b%nqrs(2); g%ns = b%nqrs(3)
         g%a = b%aqrs; g%b = b%bqrs; g%lc = b%lc; g%rc = b%rc
         if (trim(D%p%type_of_mesh) == 'cartesian') then
            g%mesh = 0
         else if (trim(D%p%type_of_mesh) == 'curvilinear') then
            g%mesh = 1
         else
            error stop 'init_domain: type_of_mesh must be cartesian or curvilinear'
         end if
         ! fault profiles of Curve_Grid_3D3 (grid.f90:625-925); other names
         ! keep the q faces planar, as the reference does
         g%profile = 0
         g%fault_x = c_null_ptr
         select case (trim(b%profile_type))
         case ('analytical_test_problem')
            g%profile = 1
         case ('analytical_tpv28')
            g%profile = 2
         case ('analytical_tpv10')
            g%profile = 3
         case ('analytical_tpv36')
            g%profile = 4
         case ('read_from_memomry_fractal')
            ! the x column of the profile file times the face weight, as an
            ! offset field of the weighted q face (read_2darray, grid.f90:2457)
            if (g%mesh == 1 .and. (b%lc /= 0.0_wp .or. b%rc /= 0.0_wp)) then
               if (b%lc /= 0.0_wp .and. b%rc /= 0.0_wp) &
                    error stop 'init_domain: a fault profile file on both q faces of a block is not supported'
               call read_fault_profile(trim(b%profile_path), g%nr, g%ns, fx)
               if (b%rc /= 0.0_wp) then
                  fx = b%rc * fx
                  g%fault_side = 0
               else
                  fx = b%lc * fx
                  g%fault_side = 1
               end if
               g%fault_x = c_loc(fx)
            end if
         end select
         if (D%p%use_topography .and. g%mesh == 1) then
            ! the topography branch of Curve_Grid_3D3: faces here, volume on
            ! the device
            if (allocated(fx)) deallocate(fx)
            call topography_grid(b, D%p%topo, D%B(ib)%X)
         else
            call wql_check(wql_grid_block(g, D%B(ib)%X), 'wql_grid_block')
         end if
         if (allocated(fx)) deallocate(fx)
         D%desc(ib)%nq = g%nq; D%desc(ib)%nr = g%nr; D%desc(ib)%ns = g%ns
         ! boundary codes Lx Rx Ly Ry Lz Rz and PML sides (block.f90:27-269)
         D%desc(ib)%bc = [b%lqrs(1), b%rqrs(1), b%lqrs(2), b%rqrs(2), b%lqrs(3), b%rqrs(3)]
         do k = 1, 3
            D%desc(ib)%pml(2 * k - 1) = merge(1, 0, b%pml_lqrs(k))
            D%desc(ib)%pml(2 * k) = merge(1, 0, b%pml_rqrs(k))
         end do
         D%desc(ib)%npml = b%npml
         D%desc(ib)%cartesian = merge(1, 0, g%mesh == 0)
         D%desc(ib)%X = c_loc(D%B(ib)%X)
         D%desc(ib)%M = c_null_ptr
         ! the medium init_material keys by problem name, if any, or
         ! init_material_from_file (material.f90:706-747): lambda, mu, rho
         ! from the block's three material_path files, raw FP64 with q fastest
         allocate(D%B(ib)%M(g%nq, g%nr, g%ns, 3))
         if (trim(D%p%material_source) == 'file') then
            do k = 1, 3
               open(newunit=u, file=trim(b%material_path(k)), access='stream', form='unformatted', &
                    status='old', action='read', iostat=stat)
               if (stat /= 0) error stop 'init_domain: cannot open a material_path file'
               read(u, iostat=stat) D%B(ib)%M(:, :, :, k)
               if (stat /= 0) error stop 'init_domain: a material_path file is shorter than its block'
               close(u)
            end do
            k = 1
         else
            k = wql_problem_material(c_string(D%problem), g%nq, g%nr, g%ns, D%B(ib)%X, D%B(ib)%M)
         end if
         if (k < 0) call wql_check(k, 'wql_problem_material')
         if (k == 1) then
            D%desc(ib)%M = c_loc(D%B(ib)%M)
         else
            deallocate(D%B(ib)%M)
         end if
         D%desc(ib)%rho_s_p = b%rho_s_p
       end associate
    end do
    if (any(D%p%btp(1)%mu_beta_eta /= D%p%btp(2)%mu_beta_eta)) &
         error stop 'init_domain: both blocks must give the same mu_beta_eta'

    D%prob%precision = WQL_FP64
    if (present(precision)) D%prob%precision = precision
    D%prob%dt = D%dt
    D%prob%t0 = D%t
    call wql_check(wql_problem_keys(c_string(D%problem), c_string(D%coupling), c_string(D%p%response), &
                                    c_string(D%p%plastic_model), D%p%btp(1)%mu_beta_eta, D%prob), &
                   'wql_problem_keys')
    ! spatial operator (&problem_list fd_type / order, block.f90:101-255)
    select case (trim(D%p%fd_type))
    case ('upwind')
       D%prob%fd_type = WQL_FD_UPWIND
    case ('upwind_drp')
       D%prob%fd_type = WQL_FD_UPWIND_DRP
    case default
       D%prob%fd_type = WQL_FD_TRADITIONAL
    end select
    D%prob%fd_order = D%p%order
    ! anelastic response: &anelastic_list (init_anelastic_properties, material.f90:11-145)
    if (trim(D%p%response) == 'anelastic') &
         call read_anelastic_list(ifname, D%prob%an_c, D%prob%an_weight_exp, D%prob%an_fref)
    call wql_check(wql_create(D%prob, D%desc, D%ctx), 'wql_create')
    ! Init_Vel_State (domain.f90:526-529)
    call wql_check(wql_init_state(D%ctx), 'wql_init_state')
    if (D%w_fault) call open_fault_output(D)
    call init_sources(D, ifname)
    call read_output_list(ifname, D%out)
    if (D%out%output_seismograms) then
       do ib = 1, 2
          call init_stations(D, ib, ifname)
       end do
    end if
  end subroutine init_domain

  !> time_step_RK (time_step.f90:68-150): one LSRK(5,4) step on the device;
  !> the fault output of write_output (time_step.f90:137) every w_stride
  !> steps: the start-of-step records before the step, the records of its
  !> first stage after it (the library keeps those on the device)
  subroutine time_step_RK(D, dt, RK, n)
    type(domain_type), intent(inout) :: D
    real(kind = wp), intent(in) :: dt
    type(RK_type), intent(in) :: RK
    integer, intent(in) :: n
    logical :: out
    if (abs(dt - D%dt) > 1e-15_wp * D%dt .or. RK%nstage /= 5) error stop 'time_step_RK: dt and RK are fixed at init_domain'
    ! write_output (time_step.f90:137): stations at the start of the step
    if (mod(n - 1, D%w_stride) == 0) call write_stations(D)
    out = D%w_fault .and. mod(n - 1, D%w_stride) == 0
    if (out) call write_fault_start(D)
    call wql_check(wql_step(D%ctx, 1_c_int), 'wql_step')
    if (out) call write_fault_stage1(D)
    D%t = wql_time(D%ctx)
  end subroutine time_step_RK

  !> close_domain (domain.f90:567-584)
  subroutine close_domain(D)
    type(domain_type), intent(inout) :: D
    integer :: k, ib
    if (c_associated(D%ctx)) then
       call wql_check(wql_sync(D%ctx), 'wql_sync')
       call wql_check(wql_destroy(D%ctx), 'wql_destroy')
       D%ctx = c_null_ptr
    end if
    do k = 1, 8
       if (D%fault_units(k) >= 0) close(D%fault_units(k))
       D%fault_units(k) = -1
    end do
    ! destroy_seismogram (seismogram.f90:491-527)
    do ib = 1, 2
       do k = 1, D%st(ib)%n
          close(D%st(ib)%units(k))
       end do
       D%st(ib)%n = 0
    end do
  end subroutine close_domain

  !> the fields of block ib (1 or 2), F(q, r, s, 1:9)
  subroutine get_fields(D, ib, F)
    type(domain_type), intent(inout) :: D
    integer, intent(in) :: ib
    real(kind = wp), intent(out) :: F(:,:,:,:)
    call wql_check(wql_get_fields(D%ctx, int(ib - 1, c_int), F), 'wql_get_fields')
  end subroutine get_fields

  !> block ib's copy of the interface state: S(nr, ns, 4), W, trup (nr, ns, 1), Svel(nr, ns, 3)
  subroutine get_fault(D, ib, S, W, trup, Svel)
    type(domain_type), intent(inout) :: D
    integer, intent(in) :: ib
    real(kind = wp), intent(out), target :: S(:,:,:), W(:,:,:), trup(:,:,:), Svel(:,:,:)
    call wql_check(wql_get_fault(D%ctx, int(ib - 1, c_int), c_loc(S), c_loc(W), c_loc(trup), c_loc(Svel), &
                                 c_null_ptr, c_null_ptr), 'wql_get_fault')
  end subroutine get_fault

  ! ---- moment-tensor sources (moment_tensor.f90:82-1080) --------------------
  !> init_moment_tensor for both blocks: each block's source list, located
  !> on its grid; with the unified list only the preferred block holds the
  !> sources (moment_tensor.f90:346-358)
  subroutine init_sources(D, ifname)
    type(domain_type), intent(inout) :: D
    character(*), intent(in) :: ifname
    type(moment_params) :: mp
    character(len=64), allocatable :: kind(:)
    real(kind = wp), allocatable :: dur(:), t0(:), mt(:,:), xyz(:,:)
    integer, allocatable :: alpha(:)
    type(wql_source), allocatable :: src(:)
    integer :: ib, n, s, k
    logical :: union
    character(len=24), parameter :: names(0:6) = [character(len=24) :: 'gaussian', 'LOH_discontinuity', 'Brune', &
         'Cos-Bell', 'cosine', 'LOH_moment_time_history', 'sine_moment_time_history']
    call read_moment_list(ifname, mp)
    if (.not. mp%use_moment_tensor) return
    do ib = 1, 2
       call read_tensor_list(ifname, ib, kind, dur, t0, mt, xyz, alpha, n, union)
       allocate(src(max(n, 1)))
       do s = 1, n
          src(s)%type = -1
          do k = 0, 6
             if (trim(kind(s)) == trim(names(k))) src(s)%type = k
          end do
          if (src(s)%type < 0) error stop 'invalid source type'
          src(s)%duration = dur(s); src(s)%t_init = t0(s); src(s)%m = mt(:, s); src(s)%x = xyz(:, s)
       end do
       if (D%p%mollify_source) then
          ! set_moment_tensor_smooth: every source of the block's list spreads
          ! over the block's nodes (domain.f90:926-953)
          call wql_check(wql_set_sources_smooth(D%ctx, int(ib - 1, c_int), int(n, c_int), src, D%B(ib)%X), &
                         'wql_set_sources_smooth')
          deallocate(src)
          cycle
       end if
       call wql_check(wql_locate_sources(D%B(ib)%X, int(D%B(ib)%nq, c_int), int(D%B(ib)%nr, c_int), &
                                         int(D%B(ib)%ns, c_int), int(n, c_int), xyz, src), 'wql_locate_sources')
       if (union .and. ib /= mp%tensor_block_preference) then
          do s = 1, n
             src(s)%node = -2
             src(s)%node_x = -1.0_wp
          end do
       end if
       k = 0
       if (n > 0) k = alpha(1)
       call wql_check(wql_set_sources(D%ctx, int(ib - 1, c_int), int(n, c_int), src, int(mp%order, c_int), &
                                      int(k, c_int)), 'wql_set_sources')
       deallocate(src)
    end do
  end subroutine init_sources

  ! ---- station seismograms (seismogram.f90:12-489) -------------------------
  !> init_seismogram for block ib: the stations' nodes (Find_Coordinates,
  !> seismogram.f90:531-650, in the library), one file per located station
  !> named <name>_<i>_<j>_<k>_block<b>.dat or, with station_xyz_index,
  !> <name>_<x>_<y>_<z>.dat, under <station_file_directory>/block<b>
  subroutine init_stations(D, ib, ifname)
    type(domain_type), intent(inout) :: D
    integer, intent(in) :: ib
    character(*), intent(in) :: ifname
    real(kind = wp), allocatable :: xyz(:,:)
    integer(c_int), allocatable :: ijk(:,:), nodes(:,:)
    integer :: n, s, m, stat
    character(len=32) :: xs, ys, zs, bs
    character(len=512) :: dir, fn
    call read_station_list(ifname, D%out, ib, xyz, n)
    allocate(ijk(3, max(n, 1)), nodes(3, max(n, 1)))
    call wql_check(wql_find_stations(D%B(ib)%X, int(D%B(ib)%nq, c_int), int(D%B(ib)%nr, c_int), &
                                     int(D%B(ib)%ns, c_int), int(n, c_int), xyz, ijk), 'wql_find_stations')
    write(bs, '(i0)') ib
    if (len_trim(D%out%station_file_directory) > 0 .and. trim(adjustl(D%out%station_file_directory)) /= '.') then
       dir = trim(adjustl(D%out%station_file_directory)) // '/block' // trim(bs)
    else
       dir = 'block' // trim(bs)
    end if
    call execute_command_line('mkdir -p "' // trim(dir) // '"')
    allocate(D%st(ib)%units(max(n, 1)))
    m = 0
    do s = 1, n
       if (any(ijk(:, s) < 0)) cycle
       m = m + 1
       nodes(:, m) = ijk(:, s)
       if (D%out%station_xyz_index) then
          write(xs, '(f20.3)') xyz(1, s)
          write(ys, '(f20.3)') xyz(2, s)
          write(zs, '(f20.3)') xyz(3, s)
          fn = trim(adjustl(D%name)) // '_' // trim(adjustl(xs)) // '_' // trim(adjustl(ys)) // '_' // &
               trim(adjustl(zs)) // '.dat'
       else
          write(fn, '(a,i0,a,i0,a,i0,a,i0,a)') trim(adjustl(D%name)) // '_', ijk(1, s) + 1, '_', ijk(2, s) + 1, '_', &
               ijk(3, s) + 1, '_block', ib, '.dat'
       end if
       ! the problem name may carry a directory of its own (name = 'data/...')
       if (index(fn, '/', back = .true.) > 0) &
            call execute_command_line('mkdir -p "' // trim(dir) // '/' // fn(1:index(fn, '/', back = .true.) - 1) // '"')
       open(newunit = D%st(ib)%units(m), file = trim(dir) // '/' // trim(fn), status = 'replace', action = 'write', &
            iostat = stat)
       if (stat /= 0) error stop 'init_seismogram: cannot open a station file'
    end do
    D%st(ib)%n = m
    call wql_check(wql_set_stations(D%ctx, int(ib - 1, c_int), int(m, c_int), nodes), 'wql_set_stations')
  end subroutine init_stations

  !> write_seismogram (seismogram.f90:460-478): t, vx, vy, vz per station
  subroutine write_stations(D)
    type(domain_type), intent(inout) :: D
    real(kind = wp), allocatable :: v(:,:)
    integer :: ib, s
    do ib = 1, 2
       if (D%st(ib)%n == 0) cycle
       allocate(v(3, D%st(ib)%n))
       call wql_check(wql_get_stations(D%ctx, int(ib - 1, c_int), v), 'wql_get_stations')
       do s = 1, D%st(ib)%n
          write(D%st(ib)%units(s), '(4f15.10)') D%t, v(:, s)
       end do
       deallocate(v)
    end do
  end subroutine write_stations

  ! ---- fault output (fault_output.f90:13-135): raw FP32 records ------------
  subroutine open_fault_output(D)
    type(domain_type), intent(inout) :: D
    integer :: k, stat
    do k = 1, 8
       open(newunit = D%fault_units(k), file = trim(D%name) // '_interface.' // trim(EXTS(k)), &
            access = 'stream', form = 'unformatted', status = 'replace', action = 'write', iostat = stat)
       if (stat /= 0) error stop 'init_fault_output: cannot open the fault output files'
    end do
  end subroutine open_fault_output

  subroutine write_fault_start(D)
    type(domain_type), intent(inout) :: D
    real(kind = wp), allocatable, target :: U(:,:,:), V(:,:,:), S(:,:,:), W(:,:,:)
    integer :: nr, ns
    nr = D%B(1)%nr; ns = D%B(1)%ns
    allocate(U(nr, ns, 9), V(nr, ns, 9), S(nr, ns, 4), W(nr, ns, 1))
    call wql_check(wql_get_fault_faces(D%ctx, c_loc(U), c_loc(V)), 'wql_get_fault_faces')
    call wql_check(wql_get_fault(D%ctx, 0_c_int, c_loc(S), c_loc(W), c_null_ptr, c_null_ptr, c_null_ptr, &
                                 c_null_ptr), 'wql_get_fault')
    write(D%fault_units(1)) real(U, 4)
    write(D%fault_units(2)) real(V, 4)
    write(D%fault_units(3)) real(S, 4)
    write(D%fault_units(8)) real(W, 4)
  end subroutine write_fault_start

  subroutine write_fault_stage1(D)
    type(domain_type), intent(inout) :: D
    real(kind = wp), allocatable, target :: Uh(:,:,:), Vh(:,:,:), tr(:,:,:)
    integer :: nr, ns
    nr = D%B(1)%nr; ns = D%B(1)%ns
    allocate(Uh(nr, ns, 6), Vh(nr, ns, 6), tr(nr, ns, 1))
    call wql_check(wql_get_fault(D%ctx, 0_c_int, c_null_ptr, c_null_ptr, c_loc(tr), c_null_ptr, c_loc(Uh), &
                                 c_loc(Vh)), 'wql_get_fault')
    write(D%fault_units(4)) real(Uh, 4)
    write(D%fault_units(5)) real(Vh, 4)
    write(D%fault_units(6)) real(Uh(:, :, 1:3), 4)
    write(D%fault_units(7)) real(tr, 4)
  end subroutine write_fault_stage1

  !> x offsets fx(nr, ns) of a fault profile file: nr*ns rows 'x y z', y
  !> fastest (read_2darray / reshape_2darray, grid.f90:2457-2505)
  !> Gen_Surf_Mesh (grid.f90:2230-2288): S(i, j) of the 2-D transfinite
  !> interpolation of the edges XL, XR (along j) and XB, XT (along i)
  pure function surf(mx, my, XL, XR, XB, XT) result(S)
    integer, intent(in) :: mx, my
    real(kind = wp), intent(in) :: XL(my), XR(my), XB(mx), XT(mx)
    real(kind = wp) :: S(mx, my), q, r
    integer :: i, j
    do j = 1, my
       r = real(j - 1, wp) * (1.0_wp / real(my - 1, wp))
       do i = 1, mx
          q = real(i - 1, wp) * (1.0_wp / real(mx - 1, wp))
          S(i, j) = (((((((1.0_wp - q) * XL(j) + q * XR(j)) + (1.0_wp - r) * XB(i)) + r * XT(i)) &
                    - ((1.0_wp - q) * (1.0_wp - r)) * XL(1)) - (q * (1.0_wp - r)) * XR(1)) &
                    - (r * (1.0_wp - q)) * XT(1)) - (r * q) * XT(mx)
       end do
    end do
  end function surf

  !> the topography branch of Curve_Grid_3D3 (use_topography, grid.f90:
  !> 1349-1860): the top face from 'analytical_topography' or a
  !> 'read_topo_from_memory' file (rows x y z, z fastest), the q faces between
  !> it and y = b_y (or x, y, z of a 'read_from_memomry_fractal' file of
  !> faultsize rows on a weighted face), the s faces and the bottom face from
  !> their edges; wql_grid_faces interpolates the volume on the device
  subroutine topography_grid(b, topo, X)
    type(block_temp_parameters), intent(in) :: b
    real(kind = wp), intent(in) :: topo
    real(kind = wp), intent(out) :: X(:,:,:,:)
    real(kind = wp), parameter :: pi = 3.141592653589793_wp
    real(kind = wp), allocatable :: Xt(:,:), Yt(:,:), Zt(:,:), Xl(:,:), Yl(:,:), Zl(:,:), Xr(:,:), Yr(:,:), Zr(:,:)
    real(kind = wp), allocatable :: Xf(:,:), Yf(:,:), Xk(:,:), Yk(:,:), Xb(:,:), Zb(:,:), d(:,:), F(:), zc(:), rr(:), qq(:)
    real(kind = wp) :: row(3), r1, hill, hx, hz
    integer :: mx, my, mz, i, k, u, stat, ny, nz, o
    mx = b%nqrs(1); my = b%nqrs(2); mz = b%nqrs(3)
    allocate(Xt(mx, mz), Yt(mx, mz), Zt(mx, mz))
    select case (trim(b%topography_type))
    case ('analytical_topography')
       hx = (b%bqrs(1) - b%aqrs(1)) / real(mx - 1, wp)
       hz = (b%bqrs(3) - b%aqrs(3)) / real(mz - 1, wp)
       do k = 1, mz
          do i = 1, mx
             Xt(i, k) = b%aqrs(1) + real(i - 1, wp) * hx
             Zt(i, k) = b%aqrs(3) + real(k - 1, wp) * hz
             r1 = sqrt((Xt(i, k) - 10.0_wp) * (Xt(i, k) - 10.0_wp) + (Zt(i, k) - 10.0_wp) * (Zt(i, k) - 10.0_wp))
             hill = 0.0_wp
             if (r1 <= 3.0_wp) hill = -0.3_wp * (1.0_wp + cos(pi * r1 / 3.0_wp))
             Yt(i, k) = b%aqrs(2) - (topo * 1.5_wp) * hill
          end do
       end do
    case ('read_topo_from_memory')
       open(newunit = u, file = trim(b%topography_path), status = 'old', action = 'read', iostat = stat)
       if (stat /= 0) error stop 'init_domain: cannot open the topography file'
       do i = 1, mx
          do k = 1, mz
             read(u, *, iostat = stat) row
             if (stat /= 0) error stop 'init_domain: short topography file'
             Xt(i, k) = row(1); Yt(i, k) = b%aqrs(2) + topo * row(2); Zt(i, k) = row(3)
          end do
       end do
       close(u)
    case default
       error stop 'init_domain: specify topography profile'
    end select
    allocate(zc(mz), rr(my), qq(mx))
    do k = 1, mz
       zc(k) = b%aqrs(3) + (b%bqrs(3) - b%aqrs(3)) * (real(k - 1, wp) * (1.0_wp / real(mz - 1, wp)))
    end do
    do i = 1, my
       rr(i) = real(i - 1, wp) * (1.0_wp / real(my - 1, wp))
    end do
    do i = 1, mx
       qq(i) = real(i - 1, wp) * (1.0_wp / real(mx - 1, wp))
    end do
    allocate(Xl(my, mz), Yl(my, mz), Zl(my, mz), Xr(my, mz), Yr(my, mz), Zr(my, mz))
    Yl = surf(my, mz, Yt(1, :), spread(b%bqrs(2), 1, mz), Yt(1, 1) + (b%bqrs(2) - Yt(1, 1)) * rr, &
              Yt(1, mz) + (b%bqrs(2) - Yt(1, mz)) * rr)
    Yr = surf(my, mz, Yt(mx, :), spread(b%bqrs(2), 1, mz), Yt(mx, 1) + (b%bqrs(2) - Yt(mx, 1)) * rr, &
              Yt(mx, mz) + (b%bqrs(2) - Yt(mx, mz)) * rr)
    Zl = surf(my, mz, zc, zc, spread(b%aqrs(3), 1, my), spread(b%bqrs(3), 1, my))
    Zr = Zl
    Xl = b%aqrs(1); Xr = b%bqrs(1)
    select case (trim(b%profile_type))
    case ('read_from_memomry_fractal')
       ny = my; nz = mz
       if (all(b%faultsize > 0)) then
          ny = b%faultsize(1); nz = b%faultsize(2)
       end if
       if (ny < my .or. nz < mz) error stop 'init_domain: faultsize smaller than the fault face'
       allocate(d(3, ny * nz))
       open(newunit = u, file = trim(b%profile_path), status = 'old', action = 'read', iostat = stat)
       if (stat /= 0) error stop 'init_domain: cannot open the fault profile file'
       do i = 1, ny * nz
          read(u, *, iostat = stat) d(:, i)
          if (stat /= 0) error stop 'init_domain: short fault profile file'
       end do
       close(u)
       do k = 1, mz
          do i = 1, my
             o = i + ny * (k - 1)
             if (b%rc /= 0.0_wp) then
                Xl(i, k) = b%aqrs(1) + d(1, o); Yl(i, k) = d(2, o); Zl(i, k) = d(3, o)
             end if
             if (b%lc /= 0.0_wp) then
                Xr(i, k) = b%bqrs(1) + d(1, o); Yr(i, k) = d(2, o); Zr(i, k) = d(3, o)
             end if
          end do
       end do
    case ('no_path')
    case default
       error stop "init_domain: the topography branch takes the fault profile 'no_path' or 'read_from_memomry_fractal'"
    end select
    allocate(Xf(mx, my), Yf(mx, my), Xk(mx, my), Yk(mx, my))
    Xf = surf(mx, my, Xl(:, 1), Xr(:, 1), Xt(:, 1), Xl(my, 1) + (Xr(my, 1) - Xl(my, 1)) * qq)
    Yf = surf(mx, my, Yl(:, 1), Yr(:, 1), Yt(:, 1), spread(b%bqrs(2), 1, mx))
    Xk = surf(mx, my, Xl(:, mz), Xr(:, mz), Xt(:, mz), Xl(my, mz) + (Xr(my, mz) - Xl(my, mz)) * qq)
    Yk = surf(mx, my, Yl(:, mz), Yr(:, mz), Yt(:, mz), spread(b%bqrs(2), 1, mx))
    allocate(Xb(mx, mz), Zb(mx, mz))
    Xb = surf(mx, mz, Xl(my, :), Xr(my, :), Xf(:, my), Xk(:, my))
    Zb = surf(mx, mz, Zl(my, :), Zr(my, :), spread(b%aqrs(3), 1, mx), spread(b%bqrs(3), 1, mx))
    ! wql_grid_faces' layout: per component the q = 0, q = nq-1, r = 0,
    ! r = nr-1, s = 0, s = ns-1 faces, each in Fortran order
    F = [reshape(Xl, [my * mz]), reshape(Xr, [my * mz]), reshape(Xt, [mx * mz]), reshape(Xb, [mx * mz]), &
         reshape(Xf, [mx * my]), reshape(Xk, [mx * my]), &
         reshape(Yl, [my * mz]), reshape(Yr, [my * mz]), reshape(Yt, [mx * mz]), &
         spread(b%bqrs(2), 1, mx * mz), reshape(Yf, [mx * my]), reshape(Yk, [mx * my]), &
         reshape(Zl, [my * mz]), reshape(Zr, [my * mz]), reshape(Zt, [mx * mz]), reshape(Zb, [mx * mz]), &
         spread(b%aqrs(3), 1, mx * my), spread(b%bqrs(3), 1, mx * my)]
    call wql_check(wql_grid_faces(int(mx, c_int), int(my, c_int), int(mz, c_int), F, X), 'wql_grid_faces')
  end subroutine topography_grid

  subroutine read_fault_profile(path, nr, ns, fx)
    character(*), intent(in) :: path
    integer(c_int), intent(in) :: nr, ns
    real(kind = wp), allocatable, intent(out) :: fx(:,:)
    integer :: u, stat, j, k
    real(kind = wp) :: row(3)
    allocate(fx(nr, ns))
    open(newunit = u, file = path, status = 'old', action = 'read', iostat = stat)
    if (stat /= 0) error stop 'init_domain: cannot open the fault profile file'
    do k = 1, ns
       do j = 1, nr
          read(u, *, iostat = stat) row
          if (stat /= 0) error stop 'init_domain: short fault profile file'
          fx(j, k) = row(1)
       end do
    end do
    close(u)
  end subroutine read_fault_profile

end module wql_domain
