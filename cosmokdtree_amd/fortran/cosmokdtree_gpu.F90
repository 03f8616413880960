! cosmokdtree_gpu.F90 -- the Fortran interface of the reference library
! (module cosmokdtree, src/cosmokdtree.f90:1-123) over the GPU library's
! C-ABI (include/cosmokdtree_amd.h).
!
! A Fortran code switches from `use cosmokdtree` to `use cosmokdtree_gpu`
! and keeps its calls: the same public names, arguments and results, chosen
! at compile time by the same preprocessor macros (periodic, dimen,
! doubleprecision, longint; src/Makefile:6-31):
!
!   type(KDTreeNode), pointer :: tree
!   type(KDTreeResult) :: query
!   tree => build_kdtree(points)              ! points(n, ndim)
!   tree => build_kdtree(points, L, leaf)     ! periodic: box sides L(ndim)
!   query = knn_search(tree, target, k, sorted)
!   query = ball_search(tree, target, radius, sorted)
!   query = box_search(tree, box)             ! (xmin, xmax, ymin, ...)
!   call deallocate_query(query)
!   call deallocate_kdtree(tree)
!
! Results follow the reference: 1-based indices of type intkind, distances
! of kind prec, an empty kNN slot (k > n) holds idx -1 and dist
! sqrt(huge(0.0)) (cosmokdtree.f90:517-518), kNN rows sorted unless
! sorted = .false., ball rows sorted only when sorted = .true.
! (cosmokdtree.f90:529-535,1062-1066).  Errors the reference stops on (a
! periodic point outside [-L/2, L/2], a wrong number of columns) stop with
! the library's message.
!
! The tree lives in HBM: KDTreeNode holds the library's handle, not the
! reference's node fields (point, axis, left, right, ...).  Every call
! copies its inputs to the device and its results back.  One query per call
! pays that round trip each time; knn_search_many, ball_search_many and
! box_search_many search many targets in one call (no reference
! counterpart; same results as the single calls).
!
! Build: compile this file with the macros of the wanted variant and link
! with -lcosmokdtree_amd -lamdhip64 (the Makefile in this folder).
module cosmokdtree_gpu
    use iso_c_binding
    implicit none
    private
    public :: build_kdtree, deallocate_kdtree, knn_search, ball_search, box_search, &
              KDTreeNode, KDTreeResult, deallocate_query, &
              knn_search_many, ball_search_many, box_search_many

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
#if longint == 1
    integer, parameter :: intkind = 8
#else
    integer, parameter :: intkind = 4
#endif

    integer(c_int), parameter :: H2D = 1, D2H = 2  ! hipMemcpyKind

    type :: KDTreeNode
        type(c_ptr) :: handle = c_null_ptr  ! ckd_tree*
        integer(c_int64_t) :: n = 0
    end type KDTreeNode

    type :: KDTreeResult
        integer(kind=intkind), allocatable :: idx(:)
        real(kind=prec), allocatable :: dist(:)
    end type KDTreeResult

    interface
        integer(c_int) function hipMalloc(p, bytes) bind(C, name="hipMalloc")
            import :: c_ptr, c_int, c_size_t
            type(c_ptr) :: p
            integer(c_size_t), value :: bytes
        end function
        integer(c_int) function hipFree(p) bind(C, name="hipFree")
            import :: c_ptr, c_int
            type(c_ptr), value :: p
        end function
        integer(c_int) function hipMemcpy(dst, src, bytes, kind) bind(C, name="hipMemcpy")
            import :: c_ptr, c_int, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
            integer(c_int), value :: kind
        end function
        integer(c_int) function ckd_build(points, n, nd, pr, leaf, boxsize, idx_bits, stream, tree) &
                bind(C, name="ckd_build")
            import :: c_ptr, c_int, c_int32_t, c_int64_t
            type(c_ptr), value :: points, boxsize, stream
            integer(c_int64_t), value :: n
            integer(c_int32_t), value :: nd, pr, leaf, idx_bits
            type(c_ptr) :: tree
        end function
        integer(c_int) function ckd_free(tree) bind(C, name="ckd_free")
            import :: c_ptr, c_int
            type(c_ptr), value :: tree
        end function
        integer(c_int) function ckd_knn(tree, q, nq, k, sorted, dist, idx, seed, stats, stream) &
                bind(C, name="ckd_knn")
            import :: c_ptr, c_int, c_int32_t, c_int64_t
            type(c_ptr), value :: tree, q, dist, idx, seed, stats, stream
            integer(c_int64_t), value :: nq
            integer(c_int32_t), value :: k, sorted
        end function
        integer(c_int) function ckd_ball_count(tree, q, nq, r, offsets, total, stream) &
                bind(C, name="ckd_ball_count")
            import :: c_ptr, c_int, c_int64_t, c_double
            type(c_ptr), value :: tree, q, offsets, stream
            integer(c_int64_t), value :: nq
            real(c_double), value :: r
            integer(c_int64_t) :: total
        end function
        integer(c_int) function ckd_ball_fill(tree, q, nq, r, sorted, offsets, dist, idx, stream) &
                bind(C, name="ckd_ball_fill")
            import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: tree, q, offsets, dist, idx, stream
            integer(c_int64_t), value :: nq
            real(c_double), value :: r
            integer(c_int32_t), value :: sorted
        end function
        integer(c_int) function ckd_box_count(tree, b, nb, offsets, total, stream) bind(C, name="ckd_box_count")
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: tree, b, offsets, stream
            integer(c_int64_t), value :: nb
            integer(c_int64_t) :: total
        end function
        integer(c_int) function ckd_box_fill(tree, b, nb, offsets, idx, stream) bind(C, name="ckd_box_fill")
            import :: c_ptr, c_int, c_int64_t
            type(c_ptr), value :: tree, b, offsets, idx, stream
            integer(c_int64_t), value :: nb
        end function
        type(c_ptr) function ckd_last_error() bind(C, name="ckd_last_error")
            import :: c_ptr
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
        end function
    end interface

contains

    ! ------------------------------------------------------------ helpers
    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        character(len=:), allocatable :: text
        type(c_ptr) :: p
        integer(c_size_t) :: i, m
        if (rc == 0) return
        p = ckd_last_error()
        m = c_strlen(p)
        call c_f_pointer(p, msg, [max(m, 1_c_size_t)])
        allocate(character(len=m) :: text)
        do i = 1, m
            text(i:i) = msg(i)
        end do
        write(*, '(a, a, a)') 'cosmokdtree_gpu: ', what, ': ' // text
        error stop 1
    end subroutine check

    subroutine hip_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc /= 0) then
            write(*, '(a, a, a, i0)') 'cosmokdtree_gpu: ', what, ': HIP error ', rc
            error stop 1
        end if
    end subroutine hip_check

    function dev_alloc(bytes) result(p)
        integer(c_size_t), intent(in) :: bytes
        type(c_ptr) :: p
        p = c_null_ptr
        call hip_check(hipMalloc(p, max(bytes, 8_c_size_t)), 'hipMalloc')
    end function dev_alloc

    subroutine dev_free(p)
        type(c_ptr), intent(inout) :: p
        if (c_associated(p)) call hip_check(hipFree(p), 'hipFree')
        p = c_null_ptr
    end subroutine dev_free

    subroutine to_dev(dst, src, bytes)
        type(c_ptr), intent(in) :: dst, src
        integer(c_size_t), intent(in) :: bytes
        if (bytes > 0) call hip_check(hipMemcpy(dst, src, bytes, H2D), 'hipMemcpy')
    end subroutine to_dev

    subroutine to_host(dst, src, bytes)
        type(c_ptr), intent(in) :: dst, src
        integer(c_size_t), intent(in) :: bytes
        if (bytes > 0) call hip_check(hipMemcpy(dst, src, bytes, D2H), 'hipMemcpy')
    end subroutine to_host

    ! 0-based library index -> the reference's 1-based index (-1 when empty)
    elemental function ref_index(i) result(j)
        integer(c_int64_t), intent(in) :: i
        integer(kind=intkind) :: j
        if (i >= 0) then
            j = int(i + 1, intkind)
        else
            j = -1
        end if
    end function ref_index

    ! ------------------------------------------------------------- build
    ! build_kdtree (cosmokdtree.f90:124-238): points_in(n, ndim); periodic
    ! builds take the box sides L_in(ndim) (domain [-L/2, L/2]).
#if periodic == 1
    function build_kdtree(points_in, L_in, leaf) result(tree)
        real(kind=prec), intent(in) :: L_in(:)
#else
    function build_kdtree(points_in, leaf) result(tree)
#endif
        real(kind=prec), intent(in) :: points_in(:, :)
        integer, intent(in), optional :: leaf
        type(KDTreeNode), pointer :: tree
        real(kind=prec), allocatable, target :: rows(:, :)
        real(c_double), target :: box(ndim)
        type(c_ptr) :: dpts, bptr
        integer(c_int64_t) :: n
        integer(c_int32_t) :: leafsize
        integer(c_size_t) :: bytes
        if (size(points_in, 2) /= ndim) then
            write(*, '(a)') 'cosmokdtree_gpu: build_kdtree: points must have ndim columns'
            error stop 1
        end if
        n = size(points_in, 1, kind=c_int64_t)
        leafsize = 16
        if (present(leaf)) leafsize = int(leaf, c_int32_t)
        bptr = c_null_ptr
#if periodic == 1
        if (size(L_in) /= ndim) then
            write(*, '(a)') 'cosmokdtree_gpu: build_kdtree: L must have ndim entries'
            error stop 1
        end if
        box = real(L_in, c_double)
        bptr = c_loc(box)
#endif
        allocate(rows(ndim, n))  ! one point per column = row-major (n, ndim)
        rows = transpose(points_in)
        bytes = int(n, c_size_t) * ndim * prec
        dpts = dev_alloc(bytes)
        call to_dev(dpts, c_loc(rows), bytes)
        allocate(tree)
        tree%n = n
        call check(ckd_build(dpts, n, int(ndim, c_int32_t), int(prec, c_int32_t), leafsize, bptr, &
                             int(merge(64, 0, intkind == 8), c_int32_t), c_null_ptr, tree%handle), 'build_kdtree')
        call dev_free(dpts)  ! the tree keeps its own copy
    end function build_kdtree

    ! deallocate_kdtree (cosmokdtree.f90:245-267)
    subroutine deallocate_kdtree(node)
        type(KDTreeNode), pointer, intent(inout) :: node
        if (.not. associated(node)) return
        if (c_associated(node%handle)) call check(ckd_free(node%handle), 'deallocate_kdtree')
        deallocate(node)
        nullify(node)
    end subroutine deallocate_kdtree

    ! deallocate_query (cosmokdtree.f90:1427-1434)
    subroutine deallocate_query(query)
        type(KDTreeResult), intent(inout) :: query
        if (allocated(query%idx)) deallocate(query%idx)
        if (allocated(query%dist)) deallocate(query%dist)
    end subroutine deallocate_query

    ! --------------------------------------------------------------- kNN
    ! knn_search_many: the k nearest neighbours of targets(ndim, nq), rows
    ! dist(k, nq), idx(k, nq).
    subroutine knn_search_many(node, targets, k, dist, idx, sorted)
        type(KDTreeNode), pointer, intent(in) :: node
        real(kind=prec), intent(in), target :: targets(:, :)
        integer, intent(in) :: k
        real(kind=prec), intent(out), target :: dist(:, :)
        integer(kind=intkind), intent(out) :: idx(:, :)
        logical, intent(in), optional :: sorted
        integer(c_int64_t), allocatable, target :: raw(:, :)
        type(c_ptr) :: dq, dd, di
        integer(c_int64_t) :: nq
        integer(c_int32_t) :: srt
        nq = size(targets, 2, kind=c_int64_t)
        srt = 1
        if (present(sorted)) then
            if (.not. sorted) srt = 0
        end if
        if (nq == 0) return
        allocate(raw(k, nq))
        dq = dev_alloc(int(nq, c_size_t) * ndim * prec)
        dd = dev_alloc(int(nq, c_size_t) * k * prec)
        di = dev_alloc(int(nq, c_size_t) * k * 8)
        call to_dev(dq, c_loc(targets), int(nq, c_size_t) * ndim * prec)
        call check(ckd_knn(node%handle, dq, nq, int(k, c_int32_t), srt, dd, di, c_null_ptr, c_null_ptr, &
                           c_null_ptr), 'knn_search')
        call to_host(c_loc(dist), dd, int(nq, c_size_t) * k * prec)
        call to_host(c_loc(raw), di, int(nq, c_size_t) * k * 8)
        idx = ref_index(raw)
        call dev_free(dq)
        call dev_free(dd)
        call dev_free(di)
    end subroutine knn_search_many

    ! knn_search (cosmokdtree.f90:504-540)
    function knn_search(node, targett, k, sorted) result(query)
        integer, intent(in) :: k
        type(KDTreeNode), pointer, intent(in) :: node
        real(kind=prec), intent(in) :: targett(ndim)
        logical, intent(in), optional :: sorted
        type(KDTreeResult) :: query
        real(kind=prec) :: t(ndim, 1), d(k, 1)
        integer(kind=intkind) :: i(k, 1)
        t(:, 1) = targett
        call knn_search_many(node, t, k, d, i, sorted)
        allocate(query%idx(k), query%dist(k))
        query%idx = i(:, 1)
        query%dist = d(:, 1)
    end function knn_search

    ! -------------------------------------------------------------- ball
    ! ball_search_many: all points within radius of each of targets(ndim,
    ! nq), as CSR: the points of target j are entries offsets(j)+1 ..
    ! offsets(j+1) of dist and idx (offsets(1) = 0).
    subroutine ball_search_many(node, targets, radius, offsets, dist, idx, sorted)
        type(KDTreeNode), pointer, intent(in) :: node
        real(kind=prec), intent(in), target :: targets(:, :)
        real(kind=prec), intent(in) :: radius
        integer(c_int64_t), allocatable, intent(out), target :: offsets(:)
        real(kind=prec), allocatable, intent(out), target :: dist(:)
        integer(kind=intkind), allocatable, intent(out) :: idx(:)
        logical, intent(in), optional :: sorted
        integer(c_int64_t), allocatable, target :: raw(:)
        type(c_ptr) :: dq, doff, dd, di
        integer(c_int64_t) :: nq, total
        integer(c_int32_t) :: srt
        nq = size(targets, 2, kind=c_int64_t)
        srt = 0
        if (present(sorted)) then
            if (sorted) srt = 1
        end if
        allocate(offsets(nq + 1))
        offsets = 0
        if (nq == 0) then
            allocate(dist(0), idx(0))
            return
        end if
        dq = dev_alloc(int(nq, c_size_t) * ndim * prec)
        doff = dev_alloc(int(nq + 1, c_size_t) * 8)
        call to_dev(dq, c_loc(targets), int(nq, c_size_t) * ndim * prec)
        total = 0
        call check(ckd_ball_count(node%handle, dq, nq, real(radius, c_double), doff, total, c_null_ptr), &
                   'ball_search')
        dd = dev_alloc(int(total, c_size_t) * prec)
        di = dev_alloc(int(total, c_size_t) * 8)
        call check(ckd_ball_fill(node%handle, dq, nq, real(radius, c_double), srt, doff, dd, di, c_null_ptr), &
                   'ball_search')
        allocate(dist(total), idx(total), raw(total))
        call to_host(c_loc(offsets), doff, int(nq + 1, c_size_t) * 8)
        call to_host(c_loc(dist), dd, int(total, c_size_t) * prec)
        call to_host(c_loc(raw), di, int(total, c_size_t) * 8)
        idx = ref_index(raw)
        call dev_free(dq)
        call dev_free(doff)
        call dev_free(dd)
        call dev_free(di)
    end subroutine ball_search_many

    ! ball_search (cosmokdtree.f90:1015-1072)
    function ball_search(node, targett, radius, sorted) result(query)
        real(kind=prec) :: radius
        type(KDTreeNode), pointer, intent(in) :: node
        real(kind=prec), intent(in) :: targett(ndim)
        logical, intent(in), optional :: sorted
        type(KDTreeResult) :: query
        real(kind=prec) :: t(ndim, 1)
        integer(c_int64_t), allocatable :: off(:)
        t(:, 1) = targett
        call ball_search_many(node, t, radius, off, query%dist, query%idx, sorted)
    end function ball_search

    ! --------------------------------------------------------------- box
    ! box_search_many: boxes(2*ndim, nb), each (xmin, xmax, ymin, ...); CSR
    ! as for ball_search_many.
    subroutine box_search_many(node, boxes, offsets, idx)
        type(KDTreeNode), pointer, intent(in) :: node
        real(kind=prec), intent(in), target :: boxes(:, :)
        integer(c_int64_t), allocatable, intent(out), target :: offsets(:)
        integer(kind=intkind), allocatable, intent(out) :: idx(:)
        integer(c_int64_t), allocatable, target :: raw(:)
        type(c_ptr) :: db, doff, di
        integer(c_int64_t) :: nb, total
        nb = size(boxes, 2, kind=c_int64_t)
        allocate(offsets(nb + 1))
        offsets = 0
        if (nb == 0) then
            allocate(idx(0))
            return
        end if
        if (size(boxes, 1) /= 2 * ndim) then
            write(*, '(a)') 'cosmokdtree_gpu: box_search: a box holds 2*ndim values'
            error stop 1
        end if
        db = dev_alloc(int(nb, c_size_t) * 2 * ndim * prec)
        doff = dev_alloc(int(nb + 1, c_size_t) * 8)
        call to_dev(db, c_loc(boxes), int(nb, c_size_t) * 2 * ndim * prec)
        total = 0
        call check(ckd_box_count(node%handle, db, nb, doff, total, c_null_ptr), 'box_search')
        di = dev_alloc(int(total, c_size_t) * 8)
        call check(ckd_box_fill(node%handle, db, nb, doff, di, c_null_ptr), 'box_search')
        allocate(idx(total), raw(total))
        call to_host(c_loc(offsets), doff, int(nb + 1, c_size_t) * 8)
        call to_host(c_loc(raw), di, int(total, c_size_t) * 8)
        idx = ref_index(raw)
        call dev_free(db)
        call dev_free(doff)
        call dev_free(di)
    end subroutine box_search_many

    ! box_search (cosmokdtree.f90:1220-1253)
    function box_search(node, box) result(query)
        real(kind=prec) :: box(2 * ndim)
        type(KDTreeNode), pointer, intent(in) :: node
        type(KDTreeResult) :: query
        real(kind=prec) :: b(2 * ndim, 1)
        integer(c_int64_t), allocatable :: off(:)
        b(:, 1) = box
        call box_search_many(node, b, off, query%idx)  ! dist stays unallocated, as in the reference
    end function box_search

end module cosmokdtree_gpu
