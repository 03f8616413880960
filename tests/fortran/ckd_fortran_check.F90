! ckd_fortran_check.F90 -- drives the Fortran module cosmokdtree_gpu the way
! the reference's Fortran demos drive module cosmokdtree, for
! tests/test_gpu_fortran.py, which writes the inputs and checks the outputs
! against the oracle.
!
! Usage: ckd_fortran_check <dir>.  <dir>/meta.txt holds
! "n nq k radius nb nsingle leaf" and, for periodic builds, a second line
! with the box sides.  Inputs (raw, Fortran order): pts.bin (n, ndim),
! q.bin (ndim, nq), box.bin (2*ndim, nb).  Outputs (raw int64 / prec):
! knn_d.bin, knn_i.bin (k, nq) from knn_search_many; knn1_d.bin,
! knn1_i.bin (k, nsingle) from knn_search one target at a time; ball_off,
! ball_d, ball_i from ball_search_many (sorted); ball1_n, ball1_d, ball1_i
! from ball_search one target at a time (unsorted); box_off, box_i from
! box_search_many; box1_n, box1_i from box_search one box at a time.
program ckd_fortran_check
    use iso_c_binding, only: c_int64_t
    use cosmokdtree_gpu
    implicit none
#if defined(dimen)
    integer, parameter :: ndim = dimen
#else
    integer, parameter :: ndim = 3
#endif
#if doubleprecision == 1
    integer, parameter :: prec = 8
#else
    integer, parameter :: prec = 4
#endif
    character(len=512) :: dir
    integer :: n, nq, k, nb, nsingle, leaf, j, u
    real(kind=8) :: radius8
    real(kind=prec) :: radius
    real(kind=prec), allocatable :: pts(:, :), q(:, :), boxes(:, :), d(:, :), bd(:)
    type(KDTreeNode), pointer :: tree
    type(KDTreeResult) :: res
    integer(kind=8), allocatable :: i8(:, :)
    integer(c_int64_t), allocatable :: off(:)
    integer(kind=8), allocatable :: cat_i(:), counts(:)
    real(kind=prec), allocatable :: cat_d(:)
#if longint == 1
    integer(kind=8), allocatable :: ii(:, :), bi(:)
#else
    integer(kind=4), allocatable :: ii(:, :), bi(:)
#endif
#if periodic == 1
    real(kind=8) :: L8(ndim)
#endif

    call get_command_argument(1, dir)
    open(newunit=u, file=trim(dir) // '/meta.txt', status='old', action='read')
    read(u, *) n, nq, k, radius8, nb, nsingle, leaf
#if periodic == 1
    read(u, *) L8
#endif
    close(u)
    radius = real(radius8, prec)
    allocate(pts(n, ndim), q(ndim, nq), boxes(2 * ndim, nb))
    call read_raw(trim(dir) // '/pts.bin', pts)
    call read_raw(trim(dir) // '/q.bin', q)
    call read_raw(trim(dir) // '/box.bin', boxes)

#if periodic == 1
    tree => build_kdtree(pts, real(L8, prec), leaf)
#else
    tree => build_kdtree(pts, leaf)
#endif

    ! kNN, many targets in one call, then one at a time
    allocate(d(k, nq), ii(k, nq))
    call knn_search_many(tree, q, k, d, ii)
    call write_real(trim(dir) // '/knn_d.bin', reshape(d, [size(d)]))
    call write_int(trim(dir) // '/knn_i.bin', int(reshape(ii, [size(ii)]), 8))
    deallocate(d)
    allocate(d(k, nsingle), i8(k, nsingle))
    do j = 1, nsingle
        res = knn_search(tree, q(:, j), k)
        d(:, j) = res%dist
        i8(:, j) = int(res%idx, 8)
        call deallocate_query(res)
    end do
    call write_real(trim(dir) // '/knn1_d.bin', reshape(d, [size(d)]))
    call write_int(trim(dir) // '/knn1_i.bin', reshape(i8, [size(i8)]))

    ! ball
    call ball_search_many(tree, q, radius, off, bd, bi, sorted=.true.)
    call write_int(trim(dir) // '/ball_off.bin', int(off, 8))
    call write_real(trim(dir) // '/ball_d.bin', bd)
    call write_int(trim(dir) // '/ball_i.bin', int(bi, 8))
    allocate(counts(nsingle), cat_i(0), cat_d(0))
    do j = 1, nsingle
        res = ball_search(tree, q(:, j), radius)
        counts(j) = size(res%idx)
        cat_i = [cat_i, int(res%idx, 8)]
        cat_d = [cat_d, res%dist]
        call deallocate_query(res)
    end do
    call write_int(trim(dir) // '/ball1_n.bin', counts)
    call write_real(trim(dir) // '/ball1_d.bin', cat_d)
    call write_int(trim(dir) // '/ball1_i.bin', cat_i)

    ! box
    call box_search_many(tree, boxes, off, bi)
    call write_int(trim(dir) // '/box_off.bin', int(off, 8))
    call write_int(trim(dir) // '/box_i.bin', int(bi, 8))
    deallocate(cat_i)
    allocate(cat_i(0))
    do j = 1, min(nb, nsingle)
        res = box_search(tree, boxes(:, j))
        counts(j) = size(res%idx)
        cat_i = [cat_i, int(res%idx, 8)]
        if (allocated(res%dist)) stop 'box_search must leave dist unallocated'
        call deallocate_query(res)
    end do
    call write_int(trim(dir) // '/box1_n.bin', counts(1:min(nb, nsingle)))
    call write_int(trim(dir) // '/box1_i.bin', cat_i)

    call deallocate_kdtree(tree)
    if (associated(tree)) stop 'deallocate_kdtree must nullify the tree'
    write(*, '(a)') 'ckd_fortran_check: done'

contains

    subroutine read_raw(path, a)
        character(len=*), intent(in) :: path
        real(kind=prec), intent(out) :: a(..)
        integer :: v
        open(newunit=v, file=path, access='stream', form='unformatted', status='old', action='read')
        select rank (a)
        rank (2)
            read(v) a
        end select
        close(v)
    end subroutine read_raw

    subroutine write_real(path, a)
        character(len=*), intent(in) :: path
        real(kind=prec), intent(in) :: a(:)
        integer :: v
        open(newunit=v, file=path, access='stream', form='unformatted', status='replace', action='write')
        write(v) a
        close(v)
    end subroutine write_real

    subroutine write_int(path, a)
        character(len=*), intent(in) :: path
        integer(kind=8), intent(in) :: a(:)
        integer :: v
        open(newunit=v, file=path, access='stream', form='unformatted', status='replace', action='write')
        write(v) a
        close(v)
    end subroutine write_int

end program ckd_fortran_check
