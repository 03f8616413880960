! ref_bridge.f90 -- C-ABI harness around the reference Fortran module.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's own
! src/cosmokdtree.f90 (read in place, never copied) together with this file
! into oracle/_ref/libckdref_*.so, one library per compile-time variant
! (periodic, dimen, doubleprecision; longint=1 always).  Tests use it to pin
! the C oracle; bench.py times it as the host-core CPU baseline.
!
! The loops below mirror the reference's f2py bridge
! (src/cosmokdtree2py.f90): an OpenMP parallel do over query rows for kNN
! (:132-141) and ball search (:184-197), serial single-box calls for box
! search (:252-315).  Arrays are Fortran-ordered exactly as f2py hands them
! over: points(n, ndim), targets(ntar, ndim), dist(ntar, k).
module ckdref_bridge
    use cosmokdtree
    use iso_c_binding
    implicit none

#if doubleprecision == 1
    integer, parameter :: rp = 8
#else
    integer, parameter :: rp = 4
#endif
#if longint == 1
    integer, parameter :: ik = 8
#else
    integer, parameter :: ik = 4
#endif
#if defined(dimen)
    integer, parameter :: nd = dimen
#else
    integer, parameter :: nd = 3
#endif

    type(KDTreeResult), allocatable, save :: res(:)
    integer(c_int64_t), save :: dump_pos, dump_leafpos

contains

    function ckdr_ndim() result(d) bind(C, name="ckdr_ndim")
        integer(c_int) :: d
        d = nd
    end function

    function ckdr_prec() result(p) bind(C, name="ckdr_prec")
        integer(c_int) :: p
        p = rp
    end function

    function ckdr_periodic() result(p) bind(C, name="ckdr_periodic")
        integer(c_int) :: p
#if periodic == 1
        p = 1
#else
        p = 0
#endif
    end function

    ! pybuild_kdtree (cosmokdtree2py.f90:29-74)
    function ckdr_build(pts, n, leaf, boxsize) result(h) bind(C, name="ckdr_build")
        type(c_ptr), value :: pts, boxsize
        integer(c_int64_t), value :: n
        integer(c_int), value :: leaf
        type(c_ptr) :: h
        real(rp), pointer :: p(:, :), lb(:)
        type(KDTreeNode), pointer :: tree
        call c_f_pointer(pts, p, [n, int(nd, c_int64_t)])
#if periodic == 1
        call c_f_pointer(boxsize, lb, [nd])
        tree => build_kdtree(p, lb, int(leaf))
#else
        tree => build_kdtree(p, int(leaf))
#endif
        h = c_loc(tree)
    end function

    ! pydeallocate_tree (cosmokdtree2py.f90:80-97)
    subroutine ckdr_free(h) bind(C, name="ckdr_free")
        type(c_ptr), value :: h
        type(KDTreeNode), pointer :: tree
        call c_f_pointer(h, tree)
        call deallocate_kdtree(tree)
    end subroutine

    ! pyknn_search (cosmokdtree2py.f90:103-143); indices stay 1-based here.
    subroutine ckdr_knn(h, tq, ntar, k, sorted, dptr, iptr) bind(C, name="ckdr_knn")
        type(c_ptr), value :: h, tq, dptr, iptr
        integer(c_int64_t), value :: ntar
        integer(c_int), value :: k, sorted
        type(KDTreeNode), pointer :: tree
        real(rp), pointer :: targett(:, :), dist(:, :)
        integer(ik), pointer :: idx(:, :)
        type(KDTreeResult) :: query
        logical :: s
        integer(c_int64_t) :: i
        call c_f_pointer(h, tree)
        call c_f_pointer(tq, targett, [ntar, int(nd, c_int64_t)])
        call c_f_pointer(dptr, dist, [ntar, int(k, c_int64_t)])
        call c_f_pointer(iptr, idx, [ntar, int(k, c_int64_t)])
        s = sorted /= 0
        !$omp parallel shared(tree, targett, k, dist, idx, s) private(i, query)
        !$omp do
        do i = 1, ntar
            query = knn_search(tree, targett(i, :), int(k), s)
            dist(i, :) = query%dist
            idx(i, :) = query%idx
            call deallocate_query(query)
        end do
        !$omp end do
        !$omp end parallel
    end subroutine

    ! pyball_search_call_1 (cosmokdtree2py.f90:153-201): runs the searches,
    ! keeps the results, returns the per-target counts.
    subroutine ckdr_ball_run(h, tq, ntar, radius, sorted, counts) bind(C, name="ckdr_ball_run")
        type(c_ptr), value :: h, tq, counts
        integer(c_int64_t), value :: ntar
        real(c_double), value :: radius
        integer(c_int), value :: sorted
        type(KDTreeNode), pointer :: tree
        real(rp), pointer :: targett(:, :)
        integer(c_int64_t), pointer :: nball(:)
        logical :: s
        real(rp) :: r
        integer(c_int64_t) :: i
        call c_f_pointer(h, tree)
        call c_f_pointer(tq, targett, [ntar, int(nd, c_int64_t)])
        call c_f_pointer(counts, nball, [ntar])
        s = sorted /= 0
        r = real(radius, rp)
        if (allocated(res)) deallocate(res)
        allocate(res(ntar))
        !$omp parallel shared(ntar, tree, targett, r, s, nball, res) private(i)
        !$omp do
        do i = 1, ntar
            res(i) = ball_search(tree, targett(i, :), r, s)
            nball(i) = size(res(i)%idx)
        end do
        !$omp end do
        !$omp end parallel
    end subroutine

    ! pybox_search_call_1 (cosmokdtree2py.f90:252-287), one call per box,
    ! boxes(2*ndim, nbox) Fortran order = C-order (nbox, 2*ndim).
    subroutine ckdr_box_run(h, bptr, nbox, counts) bind(C, name="ckdr_box_run")
        type(c_ptr), value :: h, bptr, counts
        integer(c_int64_t), value :: nbox
        type(KDTreeNode), pointer :: tree
        real(rp), pointer :: boxes(:, :)
        integer(c_int64_t), pointer :: nb(:)
        integer(c_int64_t) :: i
        call c_f_pointer(h, tree)
        call c_f_pointer(bptr, boxes, [int(2 * nd, c_int64_t), nbox])
        call c_f_pointer(counts, nb, [nbox])
        if (allocated(res)) deallocate(res)
        allocate(res(nbox))
        do i = 1, nbox
            res(i) = box_search(tree, boxes(:, i))
            nb(i) = size(res(i)%idx)
        end do
    end subroutine

    ! pyball_search_call_2 / pybox_search_call_2 (cosmokdtree2py.f90:203-246,
    ! :289-315), into CSR rows; indices stay 1-based here.
    subroutine ckdr_res_fetch(nres, offsets, dptr, iptr) bind(C, name="ckdr_res_fetch")
        integer(c_int64_t), value :: nres
        type(c_ptr), value :: offsets, dptr, iptr
        integer(c_int64_t), pointer :: off(:)
        real(rp), pointer :: dist(:)
        integer(ik), pointer :: idx(:)
        integer(c_int64_t) :: i, c, tot
        call c_f_pointer(offsets, off, [nres + 1])
        tot = off(nres + 1)
        call c_f_pointer(iptr, idx, [max(tot, 1_c_int64_t)])
        if (c_associated(dptr)) call c_f_pointer(dptr, dist, [max(tot, 1_c_int64_t)])
        do i = 1, nres
            c = size(res(i)%idx, kind=c_int64_t)
            if (c > 0) then
                idx(off(i) + 1:off(i) + c) = res(i)%idx
                if (c_associated(dptr)) dist(off(i) + 1:off(i) + c) = res(i)%dist
            end if
        end do
        deallocate(res)
    end subroutine

    recursive subroutine count_nodes(node, c)
        type(KDTreeNode), pointer :: node
        integer(c_int64_t) :: c
        if (.not. associated(node)) return
        c = c + 1
        if (.not. node%is_leaf) then
            call count_nodes(node%left, c)
            call count_nodes(node%right, c)
        end if
    end subroutine

    function ckdr_nnodes(h) result(c) bind(C, name="ckdr_nnodes")
        type(c_ptr), value :: h
        integer(c_int64_t) :: c
        type(KDTreeNode), pointer :: tree
        call c_f_pointer(h, tree)
        c = 0
        call count_nodes(tree, c)
    end function

    recursive subroutine dump_rec(node, is_leaf, axis, split, minb, maxb, lft, rgt, start, cnt, lidx, myid)
        type(KDTreeNode), pointer :: node
        integer(c_int8_t) :: is_leaf(:)
        integer(c_int32_t) :: axis(:)
        real(rp) :: split(:), minb(:, :), maxb(:, :)
        integer(c_int64_t) :: lft(:), rgt(:), start(:), cnt(:), lidx(:)
        integer(c_int64_t), intent(out) :: myid
        integer(c_int64_t) :: id, m, j, cid
        id = dump_pos
        dump_pos = dump_pos + 1
        myid = id - 1
        minb(:, id) = node%minbounds
        maxb(:, id) = node%maxbounds
        lft(id) = -1
        rgt(id) = -1
        if (node%is_leaf) then
            is_leaf(id) = 1
            axis(id) = -1
            split(id) = 0
            m = size(node%leaf_indices, kind=c_int64_t)
            start(id) = dump_leafpos
            cnt(id) = m
            do j = 1, m
                lidx(dump_leafpos + j) = node%leaf_indices(j) - 1
            end do
            dump_leafpos = dump_leafpos + m
        else
            is_leaf(id) = 0
            axis(id) = node%axis - 1
            split(id) = node%point
            start(id) = dump_leafpos
            call dump_rec(node%left, is_leaf, axis, split, minb, maxb, lft, rgt, start, cnt, lidx, cid)
            lft(id) = cid
            call dump_rec(node%right, is_leaf, axis, split, minb, maxb, lft, rgt, start, cnt, lidx, cid)
            rgt(id) = cid
            cnt(id) = dump_leafpos - start(id)
        end if
    end subroutine

    ! Preorder dump: node arrays of length nnodes, leaf indices (0-based) in
    ! leaf storage order.  minb/maxb are (ndim, nnodes) Fortran order.
    subroutine ckdr_dump(h, nn, n, p_is_leaf, p_axis, p_split, p_minb, p_maxb, p_left, p_right, p_start, p_count, &
                         p_lidx) bind(C, name="ckdr_dump")
        type(c_ptr), value :: h, p_is_leaf, p_axis, p_split, p_minb, p_maxb, p_left, p_right, p_start, p_count, p_lidx
        integer(c_int64_t), value :: nn, n
        type(KDTreeNode), pointer :: tree
        integer(c_int8_t), pointer :: is_leaf(:)
        integer(c_int32_t), pointer :: axis(:)
        real(rp), pointer :: split(:), minb(:, :), maxb(:, :)
        integer(c_int64_t), pointer :: lft(:), rgt(:), start(:), cnt(:), lidx(:)
        integer(c_int64_t) :: rid
        call c_f_pointer(h, tree)
        call c_f_pointer(p_is_leaf, is_leaf, [nn])
        call c_f_pointer(p_axis, axis, [nn])
        call c_f_pointer(p_split, split, [nn])
        call c_f_pointer(p_minb, minb, [int(nd, c_int64_t), nn])
        call c_f_pointer(p_maxb, maxb, [int(nd, c_int64_t), nn])
        call c_f_pointer(p_left, lft, [nn])
        call c_f_pointer(p_right, rgt, [nn])
        call c_f_pointer(p_start, start, [nn])
        call c_f_pointer(p_count, cnt, [nn])
        call c_f_pointer(p_lidx, lidx, [n])
        dump_pos = 1
        dump_leafpos = 0
        call dump_rec(tree, is_leaf, axis, split, minb, maxb, lft, rgt, start, cnt, lidx, rid)
    end subroutine

end module ckdref_bridge
