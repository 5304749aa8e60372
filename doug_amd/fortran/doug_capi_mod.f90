!> ISO_C_BINDING interface to libdoug_hip.so (include/doug.h).
module doug_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: DOUG_OK = 0, DOUG_PREC_FP64 = 0, DOUG_PREC_FP32_FACTORS = 1
  integer(c_int32_t), parameter :: DOUG_OPT_ITERATION_GRAPH = 1, DOUG_OPT_SPMV_SHAPE = 2, &
    DOUG_OPT_TRSV_ORDER = 3, DOUG_OPT_LOCAL_ORDER = 4, DOUG_OPT_DEVICE_COARSE = 5, &
    DOUG_OPT_BATCH_TILES = 6, DOUG_OPT_DEVICE_BLOCKS = 7, DOUG_OPT_KEEP_SETUP_TEMPS = 8, &
    DOUG_OPT_BATCH_COLUMN_GROUPS = 9

  !> doug_params (field order and meaning as in include/doug.h)
  type, bind(C) :: doug_params
    integer(c_int32_t) :: radius1, radius2, minasize1, minasize2, maxasize1, maxasize2
    real(c_double) :: strong1, strong2
    integer(c_int32_t) :: smoothers, overlap, levels, method, fine_method
    integer(c_int32_t) :: symmstruct, symmnumeric
    real(c_double) :: omega
    integer(c_int32_t) :: num_iters, reserved
  end type doug_params

  type, bind(C) :: doug_info
    integer(c_int32_t) :: n, nagr_fine, nagr_coarse, nsub, ncoarse, overlap, smoothers
    integer(c_int64_t) :: nnz, sub_total, factor_bytes, coarse_factor_bytes
    real(c_double) :: t_aggregate, t_coarse_build, t_subdomains, t_upload, t_factor
    real(c_double) :: t_coarse_factor, t_total
    integer(c_int64_t) :: factor_tiles
    real(c_double) :: factor_flops
    integer(c_int64_t) :: temp_bytes
  end type doug_info

  type, bind(C) :: doug_tuning
    integer(c_int32_t) :: spmv_shape, spmv_rows, spmv_xcd_order, trsv_order
    real(c_double) :: spmv_ms(6), trsv_ms(2)
    integer(c_int32_t) :: trsm_tiles_per_step, trsm_col_groups
    real(c_double) :: trsm_ms(2)
  end type doug_tuning

  interface
    function doug_ctx_create(device, ctx) bind(C, name="doug_ctx_create") result(rc)
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
      integer(c_int) :: rc
    end function

    function doug_ctx_destroy(ctx) bind(C, name="doug_ctx_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function doug_last_error(ctx) bind(C, name="doug_last_error") result(msg)
      import :: c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr) :: msg
    end function

    subroutine doug_params_default(p) bind(C, name="doug_params_default")
      import :: doug_params
      type(doug_params) :: p
    end subroutine

    function doug_set_params(ctx, p) bind(C, name="doug_set_params") result(rc)
      import :: c_int, c_ptr, doug_params
      type(c_ptr), value :: ctx
      type(doug_params) :: p
      integer(c_int) :: rc
    end function

    function doug_set_matrix_coo(ctx, n, nnz, indi, indj, val, base) &
        bind(C, name="doug_set_matrix_coo") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n, base
      integer(c_int64_t), value :: nnz
      integer(c_int32_t) :: indi(*), indj(*)
      real(c_double) :: val(*)
      integer(c_int) :: rc
    end function

    function doug_set_partition(ctx, nparts, starts, nodes, base) &
        bind(C, name="doug_set_partition") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nparts, base
      integer(c_int32_t) :: starts(*), nodes(*)
      integer(c_int) :: rc
    end function

    function doug_set_num_subdomains(ctx, nsub) bind(C, name="doug_set_num_subdomains") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsub
      integer(c_int) :: rc
    end function

    function doug_set_ranks(ctx, nranks, rank_of_node, base) bind(C, name="doug_set_ranks") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nranks, base
      type(c_ptr), value :: rank_of_node
      integer(c_int) :: rc
    end function

    function doug_setup(ctx, precision) bind(C, name="doug_setup") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: precision
      integer(c_int) :: rc
    end function

    function doug_get_info(ctx, info) bind(C, name="doug_get_info") result(rc)
      import :: c_int, c_ptr, doug_info
      type(c_ptr), value :: ctx
      type(doug_info) :: info
      integer(c_int) :: rc
    end function

    function doug_autotune(ctx, reps, tu) bind(C, name="doug_autotune") result(rc)
      import :: c_int, c_int32_t, c_ptr, doug_tuning
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: reps
      type(doug_tuning) :: tu
      integer(c_int) :: rc
    end function

    function doug_pcg(ctx, b, x, tol, maxit, it, cond, res_hist) &
        bind(C, name="doug_pcg") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: b(*), x(*), res_hist(*)
      real(c_double), value :: tol
      integer(c_int32_t), value :: maxit
      integer(c_int32_t) :: it
      real(c_double) :: cond
      integer(c_int) :: rc
    end function

    function doug_set_aggregates(ctx, nagr, num) bind(C, name="doug_set_aggregates") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nagr
      integer(c_int32_t) :: num(*)
      integer(c_int) :: rc
    end function

    function doug_set_strong_mask(ctx, nnz, strong) bind(C, name="doug_set_strong_mask") result(rc)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nnz
      type(c_ptr), value :: strong        ! uint8_t flags, c_null_ptr clears them
      integer(c_int) :: rc
    end function

    function doug_set_restriction_coo(ctx, nagr, nnz, ri, rj, rv, base) &
        bind(C, name="doug_set_restriction_coo") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nagr, base
      integer(c_int64_t), value :: nnz
      integer(c_int32_t) :: ri(*), rj(*)
      real(c_double) :: rv(*)
      integer(c_int) :: rc
    end function

    function doug_set_subdomains(ctx, nsub, starts, inds, base) &
        bind(C, name="doug_set_subdomains") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsub, base
      integer(c_int32_t) :: starts(*), inds(*)
      integer(c_int) :: rc
    end function

    function doug_get_rank_map(ctx, rank, ninner, nlocal, lg) bind(C, name="doug_get_rank_map") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: rank
      integer(c_int32_t) :: ninner, nlocal
      type(c_ptr), value :: lg            ! int32_t(nlocal) or c_null_ptr to ask the size
      integer(c_int) :: rc
    end function

    function doug_partition_graph(n, xadj, adjncy, vwgt, nparts, base, part) &
        bind(C, name="doug_partition_graph") result(rc)
      import :: c_int, c_int32_t, c_ptr
      integer(c_int32_t), value :: n, nparts, base
      integer(c_int32_t) :: xadj(*), adjncy(*), part(*)
      type(c_ptr), value :: vwgt          ! int32_t(n) or c_null_ptr
      integer(c_int) :: rc
    end function

    function doug_setup_host(ctx) bind(C, name="doug_setup_host") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function doug_get_fine_aggregates(ctx, num, starts, nodes) &
        bind(C, name="doug_get_fine_aggregates") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: num(*), starts(*), nodes(*)
      integer(c_int) :: rc
    end function

    function doug_get_coarse_aggregates(ctx, num, starts, nodes) &
        bind(C, name="doug_get_coarse_aggregates") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: num(*), starts(*), nodes(*)
      integer(c_int) :: rc
    end function

    function doug_get_subdomains(ctx, starts, inds) bind(C, name="doug_get_subdomains") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: starts(*), inds(*)
      integer(c_int) :: rc
    end function

    function doug_get_restriction(ctx, rowptr, cols, vals) bind(C, name="doug_get_restriction") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: rowptr(*), cols(*)
      real(c_double) :: vals(*)
      integer(c_int) :: rc
    end function

    function doug_get_coarse_matrix(ctx, rowptr, cols, vals) &
        bind(C, name="doug_get_coarse_matrix") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: rowptr(*), cols(*)
      real(c_double) :: vals(*)
      integer(c_int) :: rc
    end function

    function doug_restriction_nnz(ctx) bind(C, name="doug_restriction_nnz") result(nnz)
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: nnz
    end function

    function doug_coarse_matrix_nnz(ctx) bind(C, name="doug_coarse_matrix_nnz") result(nnz)
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t) :: nnz
    end function

    function doug_spmv(ctx, x, y) bind(C, name="doug_spmv") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: x(*), y(*)
      integer(c_int) :: rc
    end function

    function doug_spmv_device(ctx, dx, dy) bind(C, name="doug_spmv_device") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dx, dy
      integer(c_int) :: rc
    end function

    function doug_prec_apply(ctx, r, z) bind(C, name="doug_prec_apply") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double) :: r(*), z(*)
      integer(c_int) :: rc
    end function

    function doug_prec_apply_device(ctx, dr, dz) bind(C, name="doug_prec_apply_device") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dr, dz
      integer(c_int) :: rc
    end function

    function doug_prec_apply_device_noalloc(ctx, dr, dz) &
        bind(C, name="doug_prec_apply_device_noalloc") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dr, dz
      integer(c_int) :: rc
    end function

    function doug_pcg_device(ctx, db, dx, tol, maxit, it, cond, res_hist) &
        bind(C, name="doug_pcg_device") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx, db, dx
      real(c_double), value :: tol
      integer(c_int32_t), value :: maxit
      integer(c_int32_t) :: it
      real(c_double) :: cond
      type(c_ptr), value :: res_hist      ! host double(maxit) or c_null_ptr
      integer(c_int) :: rc
    end function

    function doug_time_operator(ctx, which, reps, ms) bind(C, name="doug_time_operator") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which, reps
      real(c_double) :: ms
      integer(c_int) :: rc
    end function

    function doug_get_tuning(ctx, tu) bind(C, name="doug_get_tuning") result(rc)
      import :: c_int, c_ptr, doug_tuning
      type(c_ptr), value :: ctx
      type(doug_tuning) :: tu
      integer(c_int) :: rc
    end function

    function doug_set_option(ctx, option, val) bind(C, name="doug_set_option") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: option
      integer(c_int64_t), value :: val
      integer(c_int) :: rc
    end function

    function doug_value_grad_device(ctx, dlam, dx, dgrad) &
        bind(C, name="doug_value_grad_device") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dlam, dx, dgrad
      integer(c_int) :: rc
    end function

    function doug_prepare_batch(ctx) bind(C, name="doug_prepare_batch") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function doug_prec_apply_batch_device(ctx, nrhs, dr, dz) &
        bind(C, name="doug_prec_apply_batch_device") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, dr, dz
      integer(c_int32_t), value :: nrhs
      integer(c_int) :: rc
    end function

    !> B and X node-major: entry (i, r) of the n x nrhs multi-vector at
    !> position i*nrhs + r (0-based), i.e. a Fortran array(nrhs, n)
    function doug_pcg_batch(ctx, nrhs, b, x, tol, maxit, its, res_hist) &
        bind(C, name="doug_pcg_batch") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nrhs, maxit
      real(c_double) :: b(*), x(*)
      real(c_double), value :: tol
      integer(c_int32_t) :: its(*)
      type(c_ptr), value :: res_hist      ! host double(maxit*nrhs) or c_null_ptr
      integer(c_int) :: rc
    end function

    function doug_pcg_batch_device(ctx, nrhs, db, dx, tol, maxit, its, res_hist) &
        bind(C, name="doug_pcg_batch_device") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: ctx, db, dx
      integer(c_int32_t), value :: nrhs, maxit
      real(c_double), value :: tol
      integer(c_int32_t) :: its(*)
      type(c_ptr), value :: res_hist
      integer(c_int) :: rc
    end function
  end interface

contains

  !> The library's last error message as a Fortran string.
  function doug_error_string(ctx) result(s)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: s
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: k
    s = ' '
    p = doug_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do k = 1, 512
      if (cs(k) == c_null_char) exit
      s(k:k) = cs(k)
    end do
  end function doug_error_string

end module doug_capi
