!> DOUG.ctl reading with the reference's rules (src/utils/DOUG_utils.F90:629-1149):
!> one `key value` per line, fields split on blanks only, key lower-cased,
!> `#` in column 1 starts a comment, the first occurrence of a key wins
!> (a later one is reported and ignored), integers read with (i10), reals
!> with (f10.5), logicals with (l1).
module doug_ctl
  implicit none

  integer, parameter :: rk = kind(1.0d0)
  integer, parameter :: L = 150

  type ctl_values
    integer :: solver = -1, method = -1, fine_method = -1, coarse_method = -1
    integer :: num_subdomains = -1, levels = -1, overlap = -1, smoothers = -1
    integer :: input_type = -1, radius1 = -1, radius2 = -1
    integer :: minasize1 = -1, minasize2 = -1, maxasize1 = -1, maxasize2 = -1
    integer :: debug = -1, verbose = -1, plotting = -1, solve_maxiters = -1, num_iters = -1
    real(rk) :: strong1 = -1.0_rk, strong2 = -1.0_rk, solve_tolerance = -1.0_rk
    logical :: symmstruct = .false., symmnumeric = .false.
    character(L) :: assembled_mtx_file = '', assembled_rhs_file = '', solution_file = ''
    ! elemental input (input_type 1)
    character(L) :: info_file = '', freedom_lists_file = '', elemmat_rhs_file = ''
    character(L) :: freemap_file = '', dump_matrix_file = ''
    integer :: number_of_blocks = -1
    logical :: dump_matrix_only = .false.
    integer :: assembled_mtx_format = -1, assembled_rhs_format = -1, solution_format = -1
    ! extension of this build: a non-overlapping partition of the unknowns
    ! (text: nparts, then nparts+1 starts, then the node list, 1-based)
    character(L) :: partition_file = ''
    ! extensions of this build: factor precision of the subdomain solves
    ! (0 FP64, 1 FP32 factors inside FP64 PCG) and the launch-shape autotuner
    integer :: factor_precision = 0
    logical :: autotune = .false.
    ! extension of this build: the decomposition of a run on this many MPI
    ! processes (virtual ranks), with an optional rank map file
    integer :: virtual_ranks = 0
    character(L) :: rank_map_file = ''
  end type ctl_values

  ! every word the reference's control file accepts (src/utils/controls.F90)
  character(len=24), parameter :: known_words(57) = [character(len=24) :: &
    'solver', 'method', 'levels', 'overlap', 'smoothers', 'input_type', 'matrix_type', &
    'info_file', 'freedom_lists_file', 'elemmat_rhs_file', 'coords_file', 'freemap_file', &
    'freedom_mask_file', 'number_of_blocks', 'strong1', 'strong2', 'solve_tolerance', &
    'solution_format', 'solution_file', 'radius1', 'radius2', 'minasize1', 'minasize2', &
    'maxasize1', 'maxasize2', 'debug', 'verbose', 'plotting', 'initial_guess', &
    'start_vec_type', 'start_vec_file', 'symmstruct', 'symmnumeric', 'solve_maxiters', &
    'maxcie', 'maxnd', 'cutbal', 'center_type', 'hanging_nodes', 'interpolation_type', &
    'assembled_mtx_file', 'assembled_mtx_format', 'assembled_rhs_file', &
    'assembled_rhs_format', 'dump_matrix_only', 'dump_matrix_file', 'num_subdomains', &
    'fine_method', 'num_iters', 'grid_size', 'coarse_method', 'partition_file', &
    'factor_precision', 'autotune', 'virtual_ranks', 'rank_map_file', 'dummy']

contains

  !> getword (DOUG_utils.F90:1518-1544): the num-th blank-separated word.
  subroutine getword(string, num, wordx)
    character(*), intent(in) :: string
    integer, intent(in) :: num
    character(*), intent(out) :: wordx
    integer :: i, j, k, n
    n = len_trim(string)
    j = 1
    k = 1
    do i = 1, num
      do while (j <= n)
        if (string(j:j) /= ' ') exit
        j = j + 1
      end do
      if (j > n) then
        wordx = ' '
        return
      end if
      k = j
      do while (j <= len(string))
        if (string(j:j) == ' ') exit
        j = j + 1
      end do
    end do
    wordx = string(k:min(j, len(string)))
  end subroutine getword

  function tolower(s) result(r)
    character(*), intent(in) :: s
    character(len(s)) :: r
    integer :: i, c
    r = s
    do i = 1, len(s)
      c = ichar(s(i:i))
      if (c >= ichar('A') .and. c <= ichar('Z')) r(i:i) = char(c + 32)
    end do
  end function tolower

  subroutine repeated(w1, w2)
    character(*), intent(in) :: w1, w2
    write(*, '(a,a,a,a)') ' WARNING: control ', trim(w1), ' repeated; ignoring value ', trim(w2)
  end subroutine repeated

  subroutine set_int(v, w1, w2)
    integer, intent(inout) :: v
    character(*), intent(in) :: w1, w2
    if (v /= -1) then
      call repeated(w1, w2)
    else
      read(w2, '(i10)') v
    end if
  end subroutine set_int

  !> like set_int, for a key whose unset value is 0
  subroutine set_ranks_key(v, w1, w2)
    integer, intent(inout) :: v
    character(*), intent(in) :: w1, w2
    if (v /= 0) then
      call repeated(w1, w2)
    else
      read(w2, '(i10)') v
    end if
  end subroutine set_ranks_key

  subroutine set_real(v, w1, w2)
    real(rk), intent(inout) :: v
    character(*), intent(in) :: w1, w2
    if (v /= -1.0_rk) then
      call repeated(w1, w2)
    else
      read(w2, '(f10.5)') v
    end if
  end subroutine set_real

  subroutine set_logical(v, w1, w2)
    logical, intent(inout) :: v
    character(*), intent(in) :: w1, w2
    if (v) then
      call repeated(w1, w2)
    else
      read(w2, '(l1)') v
    end if
  end subroutine set_logical

  subroutine set_string(v, w1, w2)
    character(*), intent(inout) :: v
    character(*), intent(in) :: w1, w2
    if (len_trim(v) > 0) then
      call repeated(w1, w2)
    else
      v = trim(w2)
    end if
  end subroutine set_string

  !> CtrlData_initFromFile + util_actionOnCtrlArg for the keys on the path.
  subroutine ctl_read(fname, c, ok)
    character(*), intent(in) :: fname
    type(ctl_values), intent(out) :: c
    logical, intent(out) :: ok
    character(300) :: line, w1, w2
    integer :: u, ios, k
    logical :: known
    ok = .false.
    open(newunit=u, file=trim(fname), status='old', form='formatted', iostat=ios)
    if (ios /= 0) return
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '#') cycle
      call getword(line, 1, w1)
      w1 = tolower(w1)
      call getword(line, 2, w2)
      known = .false.
      do k = 1, size(known_words)
        if (trim(w1) == trim(known_words(k))) known = .true.
      end do
      if (.not. known) then
        write(*, '(a,a)') ' WARNING: unknown control ', trim(w1)
        cycle
      end if
      select case (trim(w1))
      case ('solver');               call set_int(c%solver, w1, w2)
      case ('method');               call set_int(c%method, w1, w2)
      case ('fine_method');          call set_int(c%fine_method, w1, w2)
      case ('num_iters');            call set_int(c%num_iters, w1, w2)
      case ('coarse_method');        call set_int(c%coarse_method, w1, w2)
      case ('num_subdomains');       call set_int(c%num_subdomains, w1, w2)
      case ('levels');               call set_int(c%levels, w1, w2)
      case ('overlap');              call set_int(c%overlap, w1, w2)
      case ('smoothers');            call set_int(c%smoothers, w1, w2)
      case ('input_type');           call set_int(c%input_type, w1, w2)
      case ('radius1');              call set_int(c%radius1, w1, w2)
      case ('radius2');              call set_int(c%radius2, w1, w2)
      case ('minasize1');            call set_int(c%minasize1, w1, w2)
      case ('minasize2');            call set_int(c%minasize2, w1, w2)
      case ('maxasize1');            call set_int(c%maxasize1, w1, w2)
      case ('maxasize2');            call set_int(c%maxasize2, w1, w2)
      case ('debug');                call set_int(c%debug, w1, w2)
      case ('verbose');              call set_int(c%verbose, w1, w2)
      case ('plotting');             call set_int(c%plotting, w1, w2)
      case ('solve_maxiters');       call set_int(c%solve_maxiters, w1, w2)
      case ('strong1');              call set_real(c%strong1, w1, w2)
      case ('strong2');              call set_real(c%strong2, w1, w2)
      case ('solve_tolerance');      call set_real(c%solve_tolerance, w1, w2)
      case ('symmstruct');           call set_logical(c%symmstruct, w1, w2)
      case ('symmnumeric');          call set_logical(c%symmnumeric, w1, w2)
      case ('assembled_mtx_file');   call set_string(c%assembled_mtx_file, w1, w2)
      case ('assembled_rhs_file');   call set_string(c%assembled_rhs_file, w1, w2)
      case ('solution_file');        call set_string(c%solution_file, w1, w2)
      case ('partition_file');       call set_string(c%partition_file, w1, w2)
      case ('factor_precision');     call set_int(c%factor_precision, w1, w2)
      case ('autotune');             call set_logical(c%autotune, w1, w2)
      case ('virtual_ranks');        call set_ranks_key(c%virtual_ranks, w1, w2)
      case ('rank_map_file');        call set_string(c%rank_map_file, w1, w2)
      case ('assembled_mtx_format'); call set_int(c%assembled_mtx_format, w1, w2)
      case ('assembled_rhs_format'); call set_int(c%assembled_rhs_format, w1, w2)
      case ('info_file');            call set_string(c%info_file, w1, w2)
      case ('freedom_lists_file');   call set_string(c%freedom_lists_file, w1, w2)
      case ('elemmat_rhs_file');     call set_string(c%elemmat_rhs_file, w1, w2)
      case ('freemap_file');         call set_string(c%freemap_file, w1, w2)
      case ('number_of_blocks');     call set_int(c%number_of_blocks, w1, w2)
      case ('dump_matrix_only');     call set_logical(c%dump_matrix_only, w1, w2)
      case ('dump_matrix_file');     call set_string(c%dump_matrix_file, w1, w2)
      case ('solution_format');      call set_int(c%solution_format, w1, w2)
      case default
        ! known to DOUG but not used on this path
      end select
    end do
    close(u)
    ! the reference's default of 100 for solve_maxiters sits after its read
    ! loop but the END= branch jumps past it (DOUG_utils.F90:673-678), so an
    ! unset value stays -1 and pcg_weigs turns it into 100000
    ok = .true.
  end subroutine ctl_read

  !> CtrlData_print (DOUG_utils.F90:1155-1400): every value on the path as
  !> `name = value`, integers (i5), reals (1p,e12.3), logicals (l5), so the
  !> values the (i10)/(f10.5)/(l1) reads produced can be checked.
  subroutine ctl_print(c)
    type(ctl_values), intent(in) :: c
    write(*, *)
    write(*, *) 'Shared control parameters:'
    call pi('solver', c%solver)
    call pi('method', c%method)
    call pi('fine_method', c%fine_method)
    call pi('num_iters', c%num_iters)
    call pi('coarse_method', c%coarse_method)
    call pi('levels', c%levels)
    call pi('overlap', c%overlap)
    call pi('smoothers', c%smoothers)
    call pi('input_type', c%input_type)
    call pi('radius1', c%radius1)
    call pi('radius2', c%radius2)
    call pi('minasize1', c%minasize1)
    call pi('minasize2', c%minasize2)
    call pi('maxasize1', c%maxasize1)
    call pi('maxasize2', c%maxasize2)
    call pr('strong1', c%strong1)
    call pr('strong2', c%strong2)
    call pi('solve_maxiters', c%solve_maxiters)
    call pr('solve_tolerance', c%solve_tolerance)
    call pl('symmstruct', c%symmstruct)
    call pl('symmnumeric', c%symmnumeric)
    call pi('num_subdomains', c%num_subdomains)
    call pi('virtual_ranks', c%virtual_ranks)
    call pi('factor_precision', c%factor_precision)
    write(*, *) 'Master control parameters:'
    call pi('assembled_mtx_format', c%assembled_mtx_format)
    call ps('assembled_mtx_file', c%assembled_mtx_file)
    call pi('assembled_rhs_format', c%assembled_rhs_format)
    call ps('assembled_rhs_file', c%assembled_rhs_file)
    call pi('solution_format', c%solution_format)
    call ps('solution_file', c%solution_file)
    call pl('dump_matrix_only', c%dump_matrix_only)
  contains
    subroutine pi(w, v)
      character(*), intent(in) :: w
      integer, intent(in) :: v
      write(*, '(a20,'' = '',i5)') w, v
    end subroutine pi
    subroutine pr(w, v)
      character(*), intent(in) :: w
      real(rk), intent(in) :: v
      write(*, '(a20,'' = '',1p,e12.3)') w, v
    end subroutine pr
    subroutine pl(w, v)
      character(*), intent(in) :: w
      logical, intent(in) :: v
      write(*, '(a20,'' = '',l5)') w, v
    end subroutine pl
    subroutine ps(w, v)
      character(*), intent(in) :: w, v
      write(*, '(a20,'' = '',a)') w, trim(v)
    end subroutine ps
  end subroutine ctl_print

end module doug_ctl
