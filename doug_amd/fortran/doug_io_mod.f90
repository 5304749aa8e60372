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
  !> unformatted records (n nnz; nnz triples of int32 i, int32 j, real8 v),
  !> XDR (2) big-endian int32 n, nnz then int32 i, int32 j, double v per
  !> entry.  n and nnz are swapped when n > nnz; 0-based indices are
  !> detected from the first entry.
  subroutine read_assembled(fname, fmt, n, nnz, indi, indj, val, ok)
    character(*), intent(in) :: fname
    integer, intent(in) :: fmt
    integer, intent(out) :: n, nnz
    integer(c_int32_t), allocatable, intent(out) :: indi(:), indj(:)
    real(c_double), allocatable, intent(out) :: val(:)
    logical, intent(out) :: ok
    integer :: u, ios, k, t
    integer(c_int32_t) :: h(2), ij(2)
    real(c_double) :: v
    type(triple), allocatable :: tr(:)
    ok = .false.
    if (fmt == 1) then
      open(newunit=u, file=trim(fname), status='old', form='unformatted', iostat=ios)
      if (ios /= 0) return
      read(u, iostat=ios) n, nnz
    else if (fmt == 2) then
      open(newunit=u, file=trim(fname), status='old', access='stream', form='unformatted', &
           iostat=ios)
      if (ios /= 0) return
      read(u, iostat=ios) h
      if (little_endian()) h = [swap4(h(1)), swap4(h(2))]
      n = h(1)
      nnz = h(2)
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
    else if (fmt == 2) then
      do k = 1, nnz
        read(u, iostat=ios) ij, v
        if (ios /= 0) return
        if (little_endian()) then
          ij = [swap4(ij(1)), swap4(ij(2))]
          v = swap8(v)
        end if
        indi(k) = ij(1)
        indj(k) = ij(2)
        val(k) = v
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

  !> Rank of each unknown, one integer per line (1-based), the layout of the
  !> reference's domnums.txt (Distribution_assm.F90:202-216).
  subroutine read_rank_map(fname, n, ranks, ok)
    character(*), intent(in) :: fname
    integer, intent(in) :: n
    integer(c_int32_t), allocatable, intent(out) :: ranks(:)
    logical, intent(out) :: ok
    integer :: u, ios, i
    ok = .false.
    open(newunit=u, file=trim(fname), status='old', form='formatted', iostat=ios)
    if (ios /= 0) return
    allocate(ranks(n))
    do i = 1, n
      read(u, *, iostat=ios) ranks(i)
      if (ios /= 0) then
        close(u)
        return
      end if
    end do
    close(u)
    ok = .true.
  end subroutine read_rank_map

  !> Elemental input (input_type 1) for one process: the files
  !> Mesh_initFromFile / Mesh_readFileFreelists / Mesh_readFileFreemap
  !> (Mesh.F90:293-322,427-517) read, assembled as
  !> ElemMtxs_readAndDistribute + ElemMtxsAssembleContext (ElemMtxs_distribute.F90:509-631,
  !> ElemMtxs_assemble.F90:82-289) assemble them with one partition:
  !>  - every freedom that some element lists is local, numbered in ascending
  !>    global order (Mesh_MapsAndNghbrs, Mesh.F90:1098-1195);
  !>  - element e's matrix arrives as nfrelt(e) records, record j holding
  !>    column j; entry (i, j) with a nonzero value is added, in element
  !>    order, to the slot (lf(i), lf(j)), and the element right-hand side to
  !>    b(lf(i));
  !>  - slots come out by ascending first index, then ascending second index,
  !>    stored with row and column swapped (ElemMtxs_assemble.F90:256-262).
  !> Returns the COO (1-based) in that order and b.
  subroutine read_elemental(finfo, ffree, fmap, felem, n, nnz, indi, indj, val, b, msg, ok)
    character(*), intent(in) :: finfo, ffree, fmap, felem
    integer, intent(out) :: n, nnz
    integer(c_int32_t), allocatable, intent(out) :: indi(:), indj(:)
    real(c_double), allocatable, intent(out) :: val(:), b(:)
    character(*), intent(out) :: msg
    logical, intent(out) :: ok
    integer(c_int32_t) :: tmp(5), nf4
    integer :: nell, ngf, mfrelt, u, ios, e, i, j, k, nf, nt, t, lf
    integer(c_int32_t), allocatable :: nfrelt(:), mhead(:, :), freemap(:)
    integer, allocatable :: gl(:), ti(:), tj(:), ord(:), ord2(:), cnt(:)
    real(c_double), allocatable :: elem(:, :), erhs(:), tv(:)
    real(c_double) :: acc
    ok = .false.
    msg = ''
    open(newunit=u, file=trim(finfo), status='old', form='unformatted', iostat=ios)
    if (ios /= 0) then
      msg = 'Unable to open info file: '//trim(finfo)
      return
    end if
    read(u, iostat=ios) tmp
    close(u)
    if (ios /= 0) then
      msg = 'Unable to read info file: '//trim(finfo)
      return
    end if
    nell = tmp(1)
    ngf = tmp(2)
    mfrelt = tmp(4)
    if (nell <= 0 .or. ngf <= 0 .or. mfrelt <= 0) then
      msg = 'info file holds no elements or freedoms'
      return
    end if
    ! element freedom lists
    allocate(nfrelt(nell), mhead(mfrelt, nell))
    mhead = 0
    open(newunit=u, file=trim(ffree), status='old', form='unformatted', iostat=ios)
    if (ios /= 0) then
      msg = 'Unable to open freedom lists file: '//trim(ffree)
      return
    end if
    do e = 1, nell
      read(u, iostat=ios) nf4, (mhead(k, e), k = 1, min(int(nf4), mfrelt))
      if (ios /= 0 .or. nf4 < 0 .or. nf4 > mfrelt) then
        msg = 'Unable to read freedom lists file: '//trim(ffree)
        close(u)
        return
      end if
      nfrelt(e) = nf4
      if (maxval(mhead(:, e)) > ngf .or. any(mhead(1:nf4, e) < 1)) then
        msg = '[Mesh_readFileFreelists] - freedom index out of range'
        close(u)
        return
      end if
    end do
    close(u)
    ! freedom map (read as the reference does; not used by the solve)
    if (len_trim(fmap) > 0) then
      allocate(freemap(ngf))
      open(newunit=u, file=trim(fmap), status='old', form='unformatted', iostat=ios)
      if (ios == 0) read(u, iostat=ios) freemap
      if (ios /= 0) then
        msg = 'Unable to read freemap file: '//trim(fmap)
        return
      end if
      close(u)
    end if
    ! local numbering: freedoms that appear, ascending
    allocate(gl(ngf))
    gl = 0
    do e = 1, nell
      do k = 1, nfrelt(e)
        gl(mhead(k, e)) = 1
      end do
    end do
    n = 0
    do k = 1, ngf
      if (gl(k) /= 0) then
        n = n + 1
        gl(k) = n
      end if
    end do
    ! element matrices and right-hand sides
    nt = 0
    do e = 1, nell
      nt = nt + nfrelt(e) * nfrelt(e)
    end do
    allocate(ti(nt), tj(nt), tv(nt), b(n), elem(mfrelt, mfrelt), erhs(mfrelt))
    b = 0.0_c_double
    open(newunit=u, file=trim(felem), status='old', form='unformatted', iostat=ios)
    if (ios /= 0) then
      msg = 'Unable to open element matrices file: '//trim(felem)
      return
    end if
    t = 0
    do e = 1, nell
      nf = nfrelt(e)
      do j = 1, nf
        read(u, iostat=ios) (elem(k, j), k = 1, nf)
        if (ios /= 0) exit
      end do
      if (ios == 0) read(u, iostat=ios) (erhs(k), k = 1, nf)
      if (ios /= 0) then
        msg = 'Unable to read element matrices file: '//trim(felem)
        close(u)
        return
      end if
      do i = 1, nf
        lf = gl(mhead(i, e))
        do j = 1, nf
          if (elem(i, j) == 0.0_c_double) cycle
          t = t + 1
          ti(t) = lf
          tj(t) = gl(mhead(j, e))
          tv(t) = elem(i, j)
        end do
        b(lf) = b(lf) + erhs(i)
      end do
    end do
    close(u)
    ! stable sort of the contributions by (first, second) index: by the
    ! second, then by the first; equal slots keep element order
    allocate(ord(t), ord2(t), cnt(n + 1))
    cnt = 0
    do k = 1, t
      cnt(tj(k) + 1) = cnt(tj(k) + 1) + 1
    end do
    cnt(1) = 1
    do k = 2, n + 1
      cnt(k) = cnt(k) + cnt(k - 1)
    end do
    do k = 1, t
      ord2(cnt(tj(k))) = k
      cnt(tj(k)) = cnt(tj(k)) + 1
    end do
    cnt = 0
    do k = 1, t
      cnt(ti(k) + 1) = cnt(ti(k) + 1) + 1
    end do
    cnt(1) = 1
    do k = 2, n + 1
      cnt(k) = cnt(k) + cnt(k - 1)
    end do
    do k = 1, t
      i = ord2(k)
      ord(cnt(ti(i))) = i
      cnt(ti(i)) = cnt(ti(i)) + 1
    end do
    ! one slot per (first, second), summed from zero in element order
    nnz = 0
    do k = 1, t
      if (k == 1) then
        nnz = 1
      else if (ti(ord(k)) /= ti(ord(k - 1)) .or. tj(ord(k)) /= tj(ord(k - 1))) then
        nnz = nnz + 1
      end if
    end do
    allocate(indi(nnz), indj(nnz), val(nnz))
    nnz = 0
    acc = 0.0_c_double
    do k = 1, t
      if (k > 1) then
        if (ti(ord(k)) /= ti(ord(k - 1)) .or. tj(ord(k)) /= tj(ord(k - 1))) then
          nnz = nnz + 1
          indi(nnz) = tj(ord(k - 1))
          indj(nnz) = ti(ord(k - 1))
          val(nnz) = acc
          acc = 0.0_c_double
        end if
      end if
      acc = acc + tv(ord(k))
    end do
    if (t > 0) then
      nnz = nnz + 1
      indi(nnz) = tj(ord(t))
      indj(nnz) = ti(ord(t))
      val(nnz) = acc
    end if
    ok = .true.
  end subroutine read_elemental

  !> Global right-hand side of ngf values remapped to the local freedoms
  !> (Vect_readAndBroadcastRHS, Vect.F90:603-632): b(lf) = x(lg(lf)).
  subroutine remap_global_rhs(ffree, finfo, xg, b, ok)
    character(*), intent(in) :: ffree, finfo
    real(c_double), intent(in) :: xg(:)
    real(c_double), intent(out) :: b(:)
    logical, intent(out) :: ok
    integer(c_int32_t) :: tmp(5), nf4
    integer(c_int32_t), allocatable :: mh(:)
    logical, allocatable :: seen(:)
    integer :: u, ios, e, k, lf
    ok = .false.
    open(newunit=u, file=trim(finfo), status='old', form='unformatted', iostat=ios)
    if (ios /= 0) return
    read(u, iostat=ios) tmp
    close(u)
    if (ios /= 0 .or. size(xg) /= tmp(2)) return
    allocate(seen(tmp(2)), mh(tmp(4)))
    seen = .false.
    open(newunit=u, file=trim(ffree), status='old', form='unformatted', iostat=ios)
    if (ios /= 0) return
    do e = 1, tmp(1)
      read(u, iostat=ios) nf4, (mh(k), k = 1, min(int(nf4), int(tmp(4))))
      if (ios /= 0) return
      do k = 1, nf4
        seen(mh(k)) = .true.
      end do
    end do
    close(u)
    lf = 0
    do k = 1, tmp(2)
      if (seen(k)) then
        lf = lf + 1
        if (lf > size(b)) return
        b(lf) = xg(k)
      end if
    end do
    ok = lf == size(b)
  end subroutine remap_global_rhs

  !> SpMtx_writeMatrix (SpMtx_util.f90:486-507): `ncols nnz`, then one
  !> `(I10,I10,E24.16)` line per entry, rows ascending in stored order (the
  !> matrix as arranged by rows when the reference dumps it).
  subroutine dump_matrix(fname, n, nnz, indi, indj, val, ok)
    character(*), intent(in) :: fname
    integer, intent(in) :: n, nnz
    integer(c_int32_t), intent(in) :: indi(:), indj(:)
    real(c_double), intent(in) :: val(:)
    logical, intent(out) :: ok
    integer, allocatable :: cnt(:), ord(:)
    integer :: u, ios, k
    ok = .false.
    allocate(cnt(n + 1), ord(nnz))
    cnt = 0
    do k = 1, nnz
      cnt(indi(k) + 1) = cnt(indi(k) + 1) + 1
    end do
    cnt(1) = 1
    do k = 2, n + 1
      cnt(k) = cnt(k) + cnt(k - 1)
    end do
    do k = 1, nnz
      ord(cnt(indi(k))) = k
      cnt(indi(k)) = cnt(indi(k)) + 1
    end do
    open(newunit=u, file=trim(fname), status='replace', form='formatted', iostat=ios)
    if (ios /= 0) return
    write(u, *) n, nnz
    do k = 1, nnz
      write(u, '(I10,I10,E24.16)') indi(ord(k)), indj(ord(k)), val(ord(k))
    end do
    close(u)
    ok = .true.
  end subroutine dump_matrix

  !> ngf, the number of global freedoms, from the info file
  subroutine info_ngf(finfo, ngf)
    character(*), intent(in) :: finfo
    integer, intent(out) :: ngf
    integer(c_int32_t) :: tmp(5)
    integer :: u, ios
    ngf = 0
    open(newunit=u, file=trim(finfo), status='old', form='unformatted', iostat=ios)
    if (ios /= 0) return
    read(u, iostat=ios) tmp
    close(u)
    if (ios == 0) ngf = tmp(2)
  end subroutine info_ngf

end module doug_io
