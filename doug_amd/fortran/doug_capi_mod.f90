!> ISO_C_BINDING interface to libdoug_hip.so (include/doug.h).
module doug_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: DOUG_OK = 0, DOUG_PREC_FP64 = 0, DOUG_PREC_FP32_FACTORS = 1

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
