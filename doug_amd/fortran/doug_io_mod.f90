!> DOUG's assembled-input and solution file formats.
!>  matrix: ReadInSparseAssembled (src/datatypes/SpMtx/SpMtx_class.F90:410-597)
!>  vector: Vect_ReadFromFile (src/datatypes/Vect.F90:840-951)
!>  solution: WriteSolutionToFile (src/utils/DOUG_utils.F90:1598-1700)
!> XDR (format 2) is read and written directly as big-endian int32 length +
!> big-endian doubles, the layout fxdr produces.
module doug_io
  use iso_c_binding, only: c_int32_t, c_double
  implicit none

  integer, parameter :: rk = kind(1.0d0)

  type triple
    integer(c_int32_t) :: i, j
    real(c_double) :: val
  end type triple

contains

  function swap4(x) result(y)
    integer(c_int32_t), intent(in) :: x
    integer(c_int32_t) :: y
    character(len=4) :: a, b
    integer :: k
    a = transfer(x, a)
    do k = 1, 4
      b(k:k) = a(5 - k:5 - k)
    end do
    y = transfer(b, y)
  end function swap4

  function swap8(x) result(y)
    real(c_double), intent(in) :: x
    real(c_double) :: y
    character(len=8) :: a, b
    integer :: k
    a = transfer(x, a)
    do k = 1, 8
      b(k:k) = a(9 - k:9 - k)
    end do
    y = transfer(b, y)
  end function swap8

  logical function little_endian()
    integer(c_int32_t) :: one = 1
    character(len=4) :: a
    a = transfer(one, a)
    little_endian = ichar(a(1:1)) == 1
  end function little_endian

  !> Assembled matrix: text (0) `n nnz` then `i j v`, binary (1) two
  !> unformatted records.  n and nnz are swapped when n > nnz; 0-based
  !> indices are detected from the first entry.
  subroutine read_assembled(fname, fmt, n, nnz, indi, indj, val, ok)
    character(*), intent(in) :: fname
    integer, intent(in) :: fmt
    integer, intent(out) :: n, nnz
    integer(c_int32_t), allocatable, intent(out) :: indi(:), indj(:)
    real(c_double), allocatable, intent(out) :: val(:)
    logical, intent(out) :: ok
    integer :: u, ios, k, t
    type(triple), allocatable :: tr(:)
    ok = .false.
    if (fmt == 1) then
      open(newunit=u, file=trim(fname), status='old', form='unformatted', iostat=ios)
      if (ios /= 0) return
      read(u, iostat=ios) n, nnz
    else
      open(newunit=u, file=trim(fname), status='old', form='formatted', iostat=ios)
      if (ios /= 0) return
      read(u, *, iostat=ios) n, nnz
    end if
    if (ios /= 0) return
    if (n > nnz) then
      t = nnz
      nnz = n
      n = t
    end if
    allocate(indi(nnz), indj(nnz), val(nnz))
    if (fmt == 1) then
      allocate(tr(nnz))
      read(u, iostat=ios) (tr(k), k = 1, nnz)
      if (ios /= 0) return
      do k = 1, nnz
        indi(k) = tr(k)%i
        indj(k) = tr(k)%j
        val(k) = tr(k)%val
      end do
    else
      do k = 1, nnz
        read(u, *, iostat=ios) indi(k), indj(k), val(k)
        if (ios /= 0) return
      end do
    end if
    close(u)
    if (indi(1) == 0) then
      indi = indi + 1
      indj = indj + 1
    end if
    ok = .true.
  end subroutine read_assembled

  !> Right-hand side: text (0) `(i9)` count then `(e23.16)` values; binary (1)
  !> one unformatted record of n doubles without a length; XDR (2).
  subroutine read_vector(fname, fmt, x, ok)
    character(*), intent(in) :: fname
    integer, intent(in) :: fmt
    real(c_double), intent(inout) :: x(:)
    logical, intent(out) :: ok
    integer :: u, ios, k, m
    integer(c_int32_t) :: m4
    ok = .false.
    select case (fmt)
    case (0)
      open(newunit=u, file=trim(fname), status='old', form='formatted', iostat=ios)
      if (ios /= 0) return
      read(u, '(i9)', iostat=ios) m
      if (ios /= 0 .or. m /= size(x)) return
      do k = 1, size(x)
        read(u, '(e23.16)', iostat=ios) x(k)
        if (ios /= 0) return
      end do
    case (1)
      open(newunit=u, file=trim(fname), status='old', form='unformatted', iostat=ios)
      if (ios /= 0) return
      read(u, iostat=ios) (x(k), k = 1, size(x))
      if (ios /= 0) return
    case (2)
      open(newunit=u, file=trim(fname), status='old', access='stream', form='unformatted', &
           iostat=ios)
      if (ios /= 0) return
      read(u, iostat=ios) m4
      if (ios /= 0) return
      if (little_endian()) m4 = swap4(m4)
      if (m4 /= size(x)) return
      read(u, iostat=ios) x
      if (ios /= 0) return
      if (little_endian()) then
        do k = 1, size(x)
          x(k) = swap8(x(k))
        end do
      end if
    case default
      return
    end select
    close(u)
    ok = .true.
  end subroutine read_vector

  !> WriteSolutionToFile: text (0) `(i6)` + `(e21.14)`, binary (1) one
  !> record, XDR (2) length + big-endian doubles.
  subroutine write_solution(fname, fmt, x, ok)
    character(*), intent(in) :: fname
    integer, intent(in) :: fmt
    real(c_double), intent(in) :: x(:)
    logical, intent(out) :: ok
    integer :: u, ios, k
    integer(c_int32_t) :: m4
    ok = .false.
    select case (fmt)
    case (0)
      open(newunit=u, file=trim(fname), form='formatted', status='replace', iostat=ios)
      if (ios /= 0) return
      write(u, '(i6)') size(x)
      do k = 1, size(x)
        write(u, '(e21.14)') x(k)
      end do
    case (1)
      open(newunit=u, file=trim(fname), form='unformatted', status='replace', iostat=ios)
      if (ios /= 0) return
      write(u) (x(k), k = 1, size(x))
    case (2)
      open(newunit=u, file=trim(fname), access='stream', form='unformatted', &
           status='replace', iostat=ios)
      if (ios /= 0) return
      m4 = size(x)
      if (little_endian()) then
        write(u) swap4(m4), (swap8(x(k)), k = 1, size(x))
      else
        write(u) m4, x
      end if
    case default
      return
    end select
    close(u)
    ok = .true.
  end subroutine write_solution

  !> Partition file (extension): nparts, nparts+1 starts, node list (1-based).
  subroutine read_partition(fname, n, starts, nodes, ok)
    character(*), intent(in) :: fname
    integer, intent(in) :: n
    integer(c_int32_t), allocatable, intent(out) :: starts(:), nodes(:)
    logical, intent(out) :: ok
    integer :: u, ios, np
    ok = .false.
    open(newunit=u, file=trim(fname), status='old', form='formatted', iostat=ios)
    if (ios /= 0) return
    read(u, *, iostat=ios) np
    if (ios /= 0 .or. np <= 0) return
    allocate(starts(np + 1), nodes(n))
    read(u, *, iostat=ios) starts
    if (ios /= 0) return
    read(u, *, iostat=ios) nodes
    if (ios /= 0) return
    close(u)
    ok = .true.
  end subroutine read_partition

end module doug_io
