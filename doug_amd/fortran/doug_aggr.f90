!> doug_aggr: the DOUG driver (src/main/aggr.F90:67-295) for one process with
!> all subdomains on one MI355X.  Reads DOUG.ctl and the assembled input,
!> builds the two-level Schwarz preconditioner and runs PCG through
!> libdoug_hip.so, then writes the solution in DOUG's formats.
!>
!> usage: doug_aggr [-f DOUG.ctl] [-p [prefix]] [-q [file]] [-h]
program doug_aggr
  use iso_c_binding
  use doug_capi
  use doug_ctl
  use doug_io
  implicit none

  character(len=300) :: ctlfile, arg, profprefix, msg
  type(ctl_values) :: c
  type(doug_params) :: p
  type(doug_info) :: info
  type(doug_tuning) :: tuning
  type(c_ptr) :: ctx
  integer(c_int32_t), allocatable :: indi(:), indj(:), pstarts(:), pnodes(:)
  integer(c_int32_t), allocatable, target :: ranks(:)
  real(c_double), allocatable :: val(:), b(:), x(:), hist(:), xg(:)
  integer :: n, nnz, k, nargs, maxit, pu
  integer(c_int32_t) :: it
  real(c_double) :: cond, tol, t0, t1
  logical :: ok, prof
  integer(c_int) :: rc

  ctlfile = 'DOUG.ctl'
  prof = .false.
  profprefix = 'prof'
  nargs = command_argument_count()
  k = 1
  do while (k <= nargs)
    call get_command_argument(k, arg)
    select case (trim(arg))
    case ('-f')
      k = k + 1
      call get_command_argument(k, ctlfile)
      if (len_trim(ctlfile) == 0) call fatal('No DOUG control file specified')
    case ('-p')
      prof = .true.
      if (k < nargs) then
        call get_command_argument(k + 1, arg)
        if (arg(1:1) /= '-') then
          profprefix = arg
          k = k + 1
        end if
      end if
    case ('-q')
      if (k < nargs) then
        call get_command_argument(k + 1, arg)
        if (arg(1:1) /= '-') k = k + 1
      end if
    case ('-h')
      write(*, '(a)') 'usage: doug_aggr [-f DOUG.ctl] [-p [prefix]] [-q [file]] [-h]'
      stop
    case default
      call fatal('Unrecognised command line argument: '//trim(arg))
    end select
    k = k + 1
  end do

  write(*, '(a)') ' Parsing DOUG control file: '//trim(ctlfile)
  call ctl_read(ctlfile, c, ok)
  if (.not. ok) call fatal('Unable to open DOUG control file: '//trim(ctlfile))
  call ctl_print(c)
  if (c%solver /= 2) call fatal('[DOUG main] : Wrong solution method specified')
  if (c%input_type /= 1 .and. c%input_type /= 2 .and. c%input_type /= -1) &
    call fatal('[DOUG main] : Unrecognised input type.')
  if (c%coarse_method > 1) call fatal('Wrong coarse method')

  if (c%input_type == 1) then
    ! elemental input (Distribution_elem.F90:27-163), one partition
    if (c%number_of_blocks > 1) call fatal('multi-variable elemental input (number_of_blocks > 1) is not supported')
    write(*, '(a,a)') ' ##### Elemental input: ##### ', trim(c%elemmat_rhs_file)
    call read_elemental(c%info_file, c%freedom_lists_file, c%freemap_file, c%elemmat_rhs_file, &
                        n, nnz, indi, indj, val, b, msg, ok)
    if (.not. ok) call fatal(trim(msg))
    write(*, *) 'n,nnz:', n, nnz
    allocate(x(n))
    ! a separate global RHS replaces the elements' ones when its file exists
    ! (DOUG_utils.F90:685-691, Vect.F90:603-632)
    inquire(file=trim(c%assembled_rhs_file), exist=ok)
    if (len_trim(c%assembled_rhs_file) > 0 .and. ok) then
      write(*, '(a,a)') ' ##### Assembled RHS file: ##### ', trim(c%assembled_rhs_file)
      call info_ngf(c%info_file, k)
      allocate(xg(k))
      call read_vector(c%assembled_rhs_file, max(c%assembled_rhs_format, 0), xg, ok)
      if (.not. ok) call fatal('Unable to read vector file: '//trim(c%assembled_rhs_file))
      call remap_global_rhs(c%freedom_lists_file, c%info_file, xg, b, ok)
      if (.not. ok) call fatal('RHS file does not match the freedoms')
    end if
    ! conversion of elemental input to an assembled matrix file
    ! (main/geom.F90:105-110)
    if (c%dump_matrix_only) then
      call dump_matrix(c%dump_matrix_file, n, nnz, indi, indj, val, ok)
      if (.not. ok) call fatal('[SpMtx_dumpMatrix] : Could not open file.')
      stop
    end if
  else
    ! matrix and right-hand side (Distribution_assm.F90:65-76)
    write(*, '(a,a)') ' ##### Assembled input file: ##### ', trim(c%assembled_mtx_file)
    if (c%assembled_mtx_format > 2) &
      call fatal('[ReadInSparseAssembledHeader] Data format not recognized. 0=text, 1=binary, 2=XDR')
    call read_assembled(c%assembled_mtx_file, max(c%assembled_mtx_format, 0), n, nnz, &
                        indi, indj, val, ok)
    if (.not. ok) call fatal('Unable to read assembled file: '//trim(c%assembled_mtx_file))
    write(*, *) 'n,nnz:', n, nnz
    ! the matrix as read, in the assembled text form (main/geom.F90:105-110
    ! dumps D%A whatever the input type)
    if (c%dump_matrix_only) then
      call dump_matrix(c%dump_matrix_file, n, nnz, indi, indj, val, ok)
      if (.not. ok) call fatal('[SpMtx_dumpMatrix] : Could not open file.')
      stop
    end if
    allocate(b(n), x(n))
    if (len_trim(c%assembled_rhs_file) > 0) then
      write(*, '(a,a)') ' ##### Assembled RHS file: ##### ', trim(c%assembled_rhs_file)
      call read_vector(c%assembled_rhs_file, c%assembled_rhs_format, b, ok)
      if (.not. ok) call fatal('Unable to read vector file: '//trim(c%assembled_rhs_file))
    else
      b = 1.0_c_double
    end if
  end if

  rc = doug_ctx_create(0_c_int, ctx)
  if (rc /= DOUG_OK) call fatal('no GPU available for libdoug_hip')
  call doug_params_default(p)
  p%radius1 = c%radius1
  p%radius2 = c%radius2
  p%minasize1 = c%minasize1
  p%minasize2 = c%minasize2
  p%maxasize1 = c%maxasize1
  p%maxasize2 = c%maxasize2
  p%strong1 = c%strong1
  p%strong2 = c%strong2
  p%smoothers = c%smoothers
  p%overlap = c%overlap
  if (c%levels >= 1) p%levels = c%levels
  if (c%method >= 0) p%method = c%method
  if (c%fine_method >= 0) p%fine_method = c%fine_method
  p%num_iters = c%num_iters
  p%symmstruct = merge(1, 0, c%symmstruct)
  p%symmnumeric = merge(1, 0, c%symmnumeric)
  call check(doug_set_params(ctx, p))
  call check(doug_set_matrix_coo(ctx, int(n, c_int32_t), int(nnz, c_int64_t), indi, indj, val, &
                                 1_c_int32_t))
  ! DOUG reads num_subdomains only with several processes; a value > 1 selects
  ! that path's subdomain source (parts of the fine-aggregate graph)
  if (len_trim(c%partition_file) == 0 .and. c%num_subdomains > 1 .and. c%virtual_ranks <= 0) &
    call check(doug_set_num_subdomains(ctx, int(c%num_subdomains, c_int32_t)))
  if (len_trim(c%partition_file) > 0) then
    call read_partition(c%partition_file, n, pstarts, pnodes, ok)
    if (.not. ok) call fatal('Unable to read partition file: '//trim(c%partition_file))
    call check(doug_set_partition(ctx, int(size(pstarts) - 1, c_int32_t), pstarts, pnodes, &
                                  1_c_int32_t))
  end if

  ! the decomposition of a run on several processes (virtual ranks); the
  ! rank map file takes the place of METIS's partition
  if (c%virtual_ranks > 0) then
    if (len_trim(c%rank_map_file) > 0) then
      call read_rank_map(c%rank_map_file, n, ranks, ok)
      if (.not. ok) call fatal('Unable to read rank map file: '//trim(c%rank_map_file))
      call check(doug_set_ranks(ctx, int(c%virtual_ranks, c_int32_t), c_loc(ranks), 1_c_int32_t))
    else
      call check(doug_set_ranks(ctx, int(c%virtual_ranks, c_int32_t), c_null_ptr, 1_c_int32_t))
    end if
    write(*, '(a,i0)') ' virtual ranks: ', c%virtual_ranks
    if (c%num_subdomains > 1) call check(doug_set_num_subdomains(ctx, int(c%num_subdomains, c_int32_t)))
  end if

  call cpu_time(t0)
  if (c%factor_precision /= 0 .and. c%factor_precision /= 1) &
    call fatal('factor_precision must be 0 or 1')
  call check(doug_setup(ctx, int(c%factor_precision, c_int32_t)))
  call check(doug_get_info(ctx, info))
  if (c%autotune) then
    call check(doug_autotune(ctx, 20_c_int32_t, tuning))
    write(*, '(a,i0,a,i0,a,i0)') ' autotune: SpMV rows/workgroup ', tuning%spmv_rows, &
      ', XCD order ', tuning%spmv_xcd_order, ', subdomain order ', tuning%trsv_order
  end if
  write(*, *) '# fine aggregates:', info%nagr_fine
  write(*, *) '# coarse aggregates:', info%nagr_coarse
  write(*, *) '# subdomains:', info%nsub, ' overlap:', info%overlap

  ! pcg_weigs (pcg.F90:289-462)
  tol = c%solve_tolerance
  maxit = c%solve_maxiters
  if (maxit == -1) maxit = 100000
  allocate(hist(maxit))
  x = 0.0_c_double
  write(*, '(/a)') 'Preconditioned conjugate gradient:'
  write(*, '(a,i6,a,e8.2)') 'maxit = ', maxit, ', tol = ', tol
  call check(doug_pcg(ctx, b, x, tol, int(maxit, c_int32_t), it, cond, hist))
  call cpu_time(t1)
  do k = 1, it
    write(*, '(i5,a,e22.15)') k, ': res_norm=', hist(k)
  end do
  write(*, '(a,i3,a,e10.4)') '#it:', it, ' Cond#: ', cond
  write(*, *) 'time spent in pcg():', t1 - t0
  write(*, *) '    ...of which setup:', info%t_total
  write(*, *) '       ...of which factorisation:', info%t_factor + info%t_coarse_factor

  if (prof) then
    open(newunit=pu, file=trim(profprefix)//'.00', status='replace', form='formatted')
    write(pu, "(I0,':fine aggregates:',I0)") 0, info%nagr_fine
    write(pu, "(I0,':coarse aggregates:',I0)") 0, info%nagr_coarse
    write(pu, "(I0,':pcg iterations:',I0)") 0, it
    write(pu, "(I0,':setup time:',F0.3)") 0, info%t_total
    close(pu)
  end if

  if (len_trim(c%solution_file) > 0 .and. c%solution_format >= 0) then
    call write_solution(c%solution_file, c%solution_format, x, ok)
    if (.not. ok) write(*, *) 'Solution not written.'
  end if
  rc = doug_ctx_destroy(ctx)

contains

  subroutine fatal(msg)
    character(*), intent(in) :: msg
    write(*, '(a)') ' DOUG_abort: '//trim(msg)
    error stop 1
  end subroutine fatal

  subroutine check(code)
    integer(c_int), intent(in) :: code
    if (code /= DOUG_OK) call fatal(trim(doug_error_string(ctx)))
  end subroutine check

end program doug_aggr
