!> wql_input -- the reference's input-file API for the Fortran host: the
!> &problem_list and &block_list namelists with the variables, types and
!> defaults of init_domain (domain.f90:99-211, datatypes.f90:255-266),
!> including the `&block_list = 'file'` indirection, which reads the block
!> list from a second file named relative to the input file
!> (input_utils.f90:9-94).
module wql_input
  use iso_c_binding, only : c_double
  implicit none
  private
  public :: problem_params, block_temp_parameters, read_input, output_params, read_output_list, read_station_list, &
            moment_params, read_moment_list, read_tensor_list, read_anelastic_list

  integer, parameter :: wp = c_double

  !> block parameters of &block_list, component for component the
  !> reference's block_temp_parameters (datatypes.f90:255-266)
  type :: block_temp_parameters
     integer :: nqrs(3) = 0
     real(kind = wp) :: aqrs(3) = 0, bqrs(3) = 0
     real(kind = wp) :: lc = 0, rc = 0
     real(kind = wp) :: rho_s_p(3) = [2.67_wp, 3.464_wp, 6.0_wp]
     real(kind = wp) :: mu_beta_eta(3) = 0
     integer :: lqrs(3) = 1, rqrs(3) = 1
     character(len=256) :: profile_type = 'none', profile_path = 'No-path', material_path(3) = 'none'
     character(len=256) :: topography_type = 'none', topography_path = 'No-path'
     logical :: pml_lqrs(3) = .false., pml_rqrs(3) = .false.
     integer :: npml = 0, faultsize(2) = 0
  end type block_temp_parameters

  !> the &problem_list variables with init_domain's defaults (domain.f90:119-141)
  type :: problem_params
     character(len=256) :: name = 'default', problem = 'TPV5', response = 'elastic', plastic_model = 'default'
     character(len=256) :: coupling = 'locked', fd_type = 'traditional', mesh_source = 'compute'
     character(len=256) :: type_of_mesh = 'cartesian', material_source = 'hardcode'
     integer :: nblocks = 2, nt = 0, order = 5, w_stride = 1
     real(kind = wp) :: CFL = 0.5_wp, t_final = 0.0_wp, topo = 1.0_wp
     logical :: interpol = .false., w_fault = .true., use_topography = .false., mollify_source = .false.
     type(block_temp_parameters) :: btp(2)
  end type problem_params

  !> the station-output variables of &output_list with init_seismogram's
  !> defaults (seismogram.f90:54-95)
  type :: output_params
     logical :: output_seismograms = .false., station_xyz_index = .false.
     character(len=256) :: station_list = 'infile', station_list_file = '', station_file_directory = 'seismogram'
  end type output_params

  !> &moment_list with init_block's defaults (block.f90:63-85)
  type :: moment_params
     logical :: use_moment_tensor = .false.
     integer :: order = 2, tensor_block_preference = 1
  end type moment_params

contains

  !> &moment_list of the file `path` (block.f90:63-85)
  subroutine read_moment_list(path, m)
    character(*), intent(in) :: path
    type(moment_params), intent(out) :: m
    logical :: use_moment_tensor, show_moment_tensor_info, show_moment_tesnsor_info
    integer :: order, tensor_block_preference, u, stat
    character(len=512) :: msg
    namelist /moment_list/ use_moment_tensor, order, tensor_block_preference, show_moment_tensor_info, &
                           show_moment_tesnsor_info
    use_moment_tensor = .false.; order = 2; tensor_block_preference = 1
    show_moment_tensor_info = .false.; show_moment_tesnsor_info = .false.
    open(newunit = u, file = trim(path), status = 'old', action = 'read', iostat = stat)
    if (stat /= 0) error stop ':: Cannot open input file'
    read(u, nml = moment_list, iostat = stat, iomsg = msg)
    close(u)
    if (stat > 0) then
       write(*, '(a)') trim(msg)
       error stop 'error reading namelist moment_list'
    end if
    m%use_moment_tensor = use_moment_tensor
    m%order = order
    m%tensor_block_preference = merge(2, 1, tensor_block_preference == 2)
  end subroutine read_moment_list

  !> the moment-tensor source lines block `block` reads (init_moment_tensor,
  !> moment_tensor.f90:82-270): the unified '!---begin:tensor_list---' list
  !> (union = .true.) or the block's own '..._listU---' / '..._listV---'
  !> list; per source type, duration, t_init, mXX mYY mZZ mXY mXZ mYZ, x y z
  !> and alpha, list-directed
  subroutine read_tensor_list(path, block, kind, dur, t0, mt, xyz, alpha, n, union)
    character(*), intent(in) :: path
    integer, intent(in) :: block
    character(len=64), allocatable, intent(out) :: kind(:)
    real(kind = wp), allocatable, intent(out) :: dur(:), t0(:), mt(:,:), xyz(:,:)
    integer, allocatable, intent(out) :: alpha(:)
    integer, intent(out) :: n
    logical, intent(out) :: union
    character(len=256) :: begin, end, line, t
    integer :: u, stat, k
    logical :: found
    call marker_present(path, '!---begin:tensor_list---', union)
    if (union) then
       begin = '!---begin:tensor_list---'; end = '!---end:tensor_list---'
    else if (block == 1) then
       begin = '!---begin:tensor_listU---'; end = '!---end:tensor_listU---'
    else
       begin = '!---begin:tensor_listV---'; end = '!---end:tensor_listV---'
    end if
    open(newunit = u, file = trim(path), status = 'old', action = 'read', iostat = stat)
    if (stat /= 0) error stop ':: Cannot open input file'
    found = .false.
    do
       read(u, '(a)', iostat = stat) line
       if (stat /= 0) exit
       if (trim(line) == trim(begin)) then
          found = .true.
          exit
       end if
    end do
    if (.not. found) error stop 'error reading the moment tensor list'
    n = 0
    do
       read(u, '(a)', iostat = stat) line
       if (stat /= 0) error stop 'error reading the moment tensor list'
       if (trim(line) == trim(end)) exit
       t = adjustl(line)
       if (len_trim(t) == 0 .or. t(1:1) == '!') cycle
       n = n + 1
    end do
    allocate(kind(n), dur(n), t0(n), mt(6, n), xyz(3, n), alpha(n))
    rewind(u)
    do
       read(u, '(a)') line
       if (trim(line) == trim(begin)) exit
    end do
    k = 0
    do while (k < n)
       read(u, '(a)') line
       t = adjustl(line)
       if (len_trim(t) == 0 .or. t(1:1) == '!') cycle
       k = k + 1
       read(line, *) kind(k), dur(k), t0(k), mt(:, k), xyz(:, k), alpha(k)
    end do
    close(u)
  end subroutine read_tensor_list

  subroutine marker_present(path, marker, found)
    character(*), intent(in) :: path, marker
    logical, intent(out) :: found
    character(len=1024) :: line
    integer :: u, stat
    found = .false.
    open(newunit = u, file = trim(path), status = 'old', action = 'read', iostat = stat)
    if (stat /= 0) return
    do
       read(u, '(a)', iostat = stat) line
       if (stat /= 0) exit
       if (trim(line) == marker) then
          found = .true.
          exit
       end if
    end do
    close(u)
  end subroutine marker_present

  !> &output_list of the file `path` (seismogram.f90:54-104); the group's
  !> other variables are read and not used by this host
  subroutine read_output_list(path, o)
    character(*), intent(in) :: path
    type(output_params), intent(out) :: o
    logical :: output_exact_moment, output_seismograms, output_station_info, output_fault_topo, &
               output_fields_block1, output_fields_block2, station_xyz_index, debug_station_info, save_debug_station
    integer :: stride_fields, u, stat
    character(len=256) :: station_list, station_list_file, station_file_directory
    character(len=512) :: msg
    namelist /output_list/ output_exact_moment, output_seismograms, output_station_info, output_fault_topo, &
                           output_fields_block1, output_fields_block2, stride_fields, station_xyz_index, &
                           debug_station_info, save_debug_station, station_list, station_list_file, &
                           station_file_directory
    output_exact_moment = .false.; output_seismograms = o%output_seismograms; output_station_info = .true.
    output_fault_topo = .false.; output_fields_block1 = .false.; output_fields_block2 = .false.
    stride_fields = 1; station_xyz_index = o%station_xyz_index; debug_station_info = .false.
    save_debug_station = .false.; station_list = o%station_list; station_list_file = o%station_list_file
    station_file_directory = o%station_file_directory
    open(newunit = u, file = trim(path), status = 'old', action = 'read', iostat = stat)
    if (stat /= 0) error stop ':: Cannot open input file'
    read(u, nml = output_list, iostat = stat, iomsg = msg)
    close(u)
    if (stat > 0) then
       write(*, '(a)') trim(msg)
       error stop 'error reading namelist output_list'
    end if
    o%output_seismograms = output_seismograms
    o%station_xyz_index = station_xyz_index
    o%station_list = station_list
    o%station_list_file = station_list_file
    o%station_file_directory = station_file_directory
  end subroutine read_output_list

  !> the stations block `block` looks for (seismogram.f90:128-326):
  !> xyz(3, n) from the lines between '!---begin:station_list---' and
  !> '!---end:station_list---' of the input file or of station_list_file
  !> (station_list = 'extfile'), or the block's own list
  !> ('..._listU---' block 1, '..._listV---' block 2) in the input file
  subroutine read_station_list(path, o, block, xyz, n)
    character(*), intent(in) :: path
    type(output_params), intent(in) :: o
    integer, intent(in) :: block
    real(kind = wp), allocatable, intent(out) :: xyz(:,:)
    integer, intent(out) :: n
    character(len=1024) :: src
    character(len=1) :: tag
    logical :: ext, found
    ext = trim(o%station_list) == 'extfile' .or. trim(o%station_list) == 'exfile'
    src = path
    if (ext) then
       if (len_trim(o%station_list_file) == 0) error stop 'error: station_list_file must be specified when station_list=extfile'
       src = o%station_list_file
    end if
    call list_between(trim(src), '!---begin:station_list---', '!---end:station_list---', xyz, n, found)
    if (found) return
    if (ext) error stop 'error: !---begin:station_list--- not found in external station file'
    tag = merge('U', 'V', block == 1)
    call list_between(trim(src), '!---begin:station_list' // tag // '---', '!---end:station_list' // tag // '---', &
                      xyz, n, found)
    if (.not. found) error stop 'error reading seismogram station list'
  end subroutine read_station_list

  subroutine list_between(path, begin, end, xyz, n, found)
    character(*), intent(in) :: path, begin, end
    real(kind = wp), allocatable, intent(out) :: xyz(:,:)
    integer, intent(out) :: n
    logical, intent(out) :: found
    character(len=1024) :: line
    integer :: u, stat, k
    n = 0
    found = .false.
    open(newunit = u, file = path, status = 'old', action = 'read', iostat = stat)
    if (stat /= 0) then
       write(*, '(a,a)') 'error: cannot open station list file: ', trim(path)
       error stop
    end if
    do
       read(u, '(a)', iostat = stat) line
       if (stat /= 0) exit
       if (trim(line) == begin) then
          found = .true.
          exit
       end if
    end do
    if (.not. found) then
       close(u)
       allocate(xyz(3, 0))
       return
    end if
    do
       read(u, '(a)', iostat = stat) line
       if (stat /= 0) error stop 'error reading seismogram station list'
       if (trim(line) == end) exit
       n = n + 1
    end do
    allocate(xyz(3, n))
    rewind(u)
    do
       read(u, '(a)') line
       if (trim(line) == begin) exit
    end do
    do k = 1, n
       read(u, *) xyz(:, k)
    end do
    close(u)
  end subroutine list_between

  !> &anelastic_list of the file `path` with init_anelastic_properties'
  !> defaults (material.f90:30-39): Q_s = c v_s, the weight set, fref
  subroutine read_anelastic_list(path, c, weight_exp, fref)
    character(*), intent(in) :: path
    real(kind = wp), intent(out) :: c, weight_exp, fref
    namelist /anelastic_list/ c, weight_exp, fref
    integer :: u, stat
    c = 1.0_wp
    weight_exp = 0.0_wp
    fref = 1.0_wp
    open(newunit = u, file = trim(path), status = 'old', action = 'read', iostat = stat)
    if (stat /= 0) error stop ':: Cannot open input file'
    read(u, nml = anelastic_list, iostat = stat)
    close(u)
    if (stat > 0) error stop 'error reading namelist anelastic_list'
  end subroutine read_anelastic_list

  !> read &problem_list and &block_list from the file `path`
  subroutine read_input(path, p)
    character(*), intent(in) :: path
    type(problem_params), intent(out) :: p
    character(len=256) :: name, problem, response, plastic_model, coupling, fd_type, mesh_source, &
                          type_of_mesh, material_source
    integer :: nblocks, nt, order, w_stride
    real(kind = wp) :: CFL, t_final, topo
    logical :: interpol, w_fault, use_topography, mollify_source
    type(block_temp_parameters) :: btp(2)
    namelist /problem_list/ name, problem, response, plastic_model, nblocks, &
                            nt, CFL, coupling, fd_type, order, t_final, &
                            mesh_source, type_of_mesh, material_source, &
                            interpol, w_stride, w_fault, use_topography, topo, mollify_source
    namelist /block_list/ btp
    integer :: u, stat
    character(len=512) :: msg
    character(len=1024) :: block_file
    logical :: has_file

    name = p%name; problem = p%problem; response = p%response; plastic_model = p%plastic_model
    coupling = p%coupling; fd_type = p%fd_type; mesh_source = p%mesh_source; type_of_mesh = p%type_of_mesh
    material_source = p%material_source; nblocks = p%nblocks; nt = p%nt; order = p%order
    w_stride = p%w_stride; CFL = p%CFL; t_final = p%t_final; topo = p%topo; interpol = p%interpol
    w_fault = p%w_fault; use_topography = p%use_topography; mollify_source = p%mollify_source
    btp = p%btp

    ! scanned before the input is opened for the namelist reads: a file is
    ! connected to one unit at a time
    call find_block_list_file(path, block_file, has_file)
    open(newunit = u, file = trim(path), status = 'old', action = 'read', iostat = stat)
    if (stat /= 0) error stop ':: Cannot open input file'
    read(u, nml = problem_list, iostat = stat, iomsg = msg)
    if (stat > 0) then
       write(*, '(a)') trim(msg)
       error stop 'error reading namelist problem_list'
    end if
    if (has_file) then
       close(u)
       open(newunit = u, file = trim(block_file), status = 'old', action = 'read', iostat = stat)
       if (stat /= 0) error stop 'error opening block list file'
    else
       rewind(u)
    end if
    read(u, nml = block_list, iostat = stat, iomsg = msg)
    if (stat > 0) then
       write(*, '(a)') trim(msg)
       error stop 'error reading namelist block_list'
    end if
    close(u)

    p%name = name; p%problem = problem; p%response = response; p%plastic_model = plastic_model
    p%coupling = coupling; p%fd_type = fd_type; p%mesh_source = mesh_source; p%type_of_mesh = type_of_mesh
    p%material_source = material_source; p%nblocks = nblocks; p%nt = nt; p%order = order
    p%w_stride = w_stride; p%CFL = CFL; p%t_final = t_final; p%topo = topo; p%interpol = interpol
    p%w_fault = w_fault; p%use_topography = use_topography; p%mollify_source = mollify_source
    p%btp = btp
  end subroutine read_input

  !> the first `&block_list` line of the input: `&block_list = 'file'`
  !> names the block-list file, relative to the input file's directory
  !> unless absolute (input_utils.f90:45-94)
  subroutine find_block_list_file(path, block_file, has_file)
    character(*), intent(in) :: path
    character(len=*), intent(out) :: block_file
    logical, intent(out) :: has_file
    character(len=1024) :: line, low, rest
    integer :: u, stat, pos, peq, q1, q2, k
    character :: quote

    has_file = .false.
    block_file = ''
    open(newunit = u, file = trim(path), status = 'old', action = 'read', iostat = stat)
    if (stat /= 0) return
    do
       read(u, '(a)', iostat = stat) line
       if (stat /= 0) exit
       k = index(line, '!')
       if (k > 0) line = line(1:k - 1)
       line = adjustl(line)
       if (len_trim(line) == 0) cycle
       low = lower(line)
       pos = index(low, '&block_list')
       if (pos == 0) cycle
       rest = line(pos + len('&block_list'):)
       peq = index(rest, '=')
       if (peq == 0) exit
       ! an assignment right after the group name (not btp(...)%... = )
       if (len_trim(rest(1:peq - 1)) /= 0) exit
       rest = adjustl(rest(peq + 1:))
       quote = rest(1:1)
       if (quote /= "'" .and. quote /= '"') exit
       q1 = 2
       q2 = index(rest(q1:), quote)
       if (q2 == 0) exit
       block_file = rest(q1:q1 + q2 - 2)
       if (block_file(1:1) /= '/') block_file = trim(parent_dir(path)) // trim(block_file)
       has_file = len_trim(block_file) > 0
       exit
    end do
    close(u)
  end subroutine find_block_list_file

  function parent_dir(path) result(dir)
    character(*), intent(in) :: path
    character(len=len(path)) :: dir
    integer :: k
    k = index(path, '/', back = .true.)
    dir = ''
    if (k > 0) dir = path(1:k)
  end function parent_dir

  function lower(s) result(t)
    character(*), intent(in) :: s
    character(len=len(s)) :: t
    integer :: i, c
    t = s
    do i = 1, len(s)
       c = iachar(s(i:i))
       if (c >= iachar('A') .and. c <= iachar('Z')) t(i:i) = achar(c + 32)
    end do
  end function lower

end module wql_input
