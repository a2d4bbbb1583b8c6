! The library handle of the drop-in path.  The Fortran driver keeps its YYTree
! (3_6/ModYinYangTree.f90:55-81) and BlockType data; libfalcon_hip.so keeps every
! block resident in HBM.  Host primitive arrays are stale between
! falcon_sync_to_host calls, which the driver makes before ModCheck/ModSave.
!
! The reference driver writes the initial state on the host after
! ModCommunication_SetGCAll (3_6/test1.f90:133,140), so attaching at SetGCAll
! only marks the host state as not yet on the device; the first ghost fill or
! step after it copies the state up (falcon_push_pending).
module ModFalconShim
    use iso_c_binding
    use falcon_c
    use ModParameters,  only: ni, nj, nk, ng, nvar, DivB_option
    use ModYinYangTree, only: YYTree, TreeNode
    use ModBlock,       only: BlockType, GC_target
    use ModAllocation,  only: ModAllocation_GetRank
    implicit none
    private
    public :: falcon_h, falcon_check, falcon_attach, falcon_push_state, falcon_push_pending, falcon_sync_to_host
    public :: falcon_solver

    type(c_ptr), save   :: falcon_h = c_null_ptr
    ! the host state is newer than the device state (set by falcon_attach)
    logical, save       :: state_pending = .false.
    ! the implicit solver the handle holds: 0 none, 1 heat conduction
    ! (ModHeatConduction), 2 PFSS (ModPFSS); each set-up replaces the other
    integer, save       :: falcon_solver = 0

contains

    ! The reference's only failure mode is `stop 1` (3_6/ModCheck.f90:38).
    subroutine falcon_check(rc)
        integer(c_int), intent(in)      :: rc
        character(kind=c_char)          :: buf(1024)
        character(len=1024)             :: msg
        integer                         :: i, ierr
        if (rc == 0) return
        ierr = falcon_last_error(buf, 1024)
        msg = ''
        do i = 1, 1024
            if (buf(i) == c_null_char) exit
            msg(i:i) = buf(i)
        end do
        write(*, '(a,a)') 'falcon: ', trim(msg)
        stop 1
    end subroutine falcon_check

    recursive subroutine count_nodes(node, n)
        type(TreeNode), intent(in)      :: node
        integer, intent(inout)          :: n
        integer                         :: c
        n = n + 1
        if (.not. node%if_leaf) then
            do c = 1, size(node%children)
                call count_nodes(node%children(c), n)
            end do
        end if
    end subroutine count_nodes

    ! depth-first pre-order, children in their own order: the library's node list
    recursive subroutine flatten(node, par, k, parent, is_yin, rng6)
        type(TreeNode), intent(in)      :: node
        integer, intent(in)             :: par
        integer, intent(inout)          :: k
        integer(c_int), intent(inout)   :: parent(:), is_yin(:)
        real(c_double), intent(inout)   :: rng6(:,:)
        integer                         :: me, c
        k = k + 1
        me = k
        parent(me) = par - 1                    ! 0-based, -1 for a root
        is_yin(me) = merge(1, 0, node%if_yin)
        rng6(1:3, me) = node%rtp_range(:, 1)
        rng6(4:6, me) = node%rtp_range(:, 2)
        if (.not. node%if_leaf) then
            do c = 1, size(node%children)
                call flatten(node%children(c), me, k, parent, is_yin, rng6)
            end do
        end if
    end subroutine flatten

    ! Register the driver's tree, 1-D background and state with the library, and
    ! build every table (in place of ModCommunication_SetGCAll, 3_6/ModCommunication.f90:22-64).
    subroutine falcon_attach(Tree)
        type(YYTree), intent(inout), target :: Tree
        integer(c_int), allocatable     :: parent(:), is_yin(:)
        real(c_double), allocatable     :: rng6(:,:), bg(:,:)
        integer(c_int)                  :: nb, ny
        integer                         :: n, k, ib
        type(BlockType), pointer        :: B

        if (c_associated(falcon_h)) call falcon_check(falcon_destroy(falcon_h))
        call falcon_check(falcon_create(falcon_h, ni, nj, nk, ng, Tree%r_range_CGS(1), Tree%r_range_CGS(2)))
        n = 0
        call count_nodes(Tree%Yin, n)
        call count_nodes(Tree%Yang, n)
        allocate(parent(n), is_yin(n), rng6(6, n))
        k = 0
        call flatten(Tree%Yin, 0, k, parent, is_yin, rng6)
        call flatten(Tree%Yang, 0, k, parent, is_yin, rng6)
        call falcon_check(falcon_tree_set_nodes(falcon_h, int(n, c_int), parent, is_yin, rng6))
        call falcon_check(falcon_nblocks(falcon_h, nb, ny))
        if (nb /= Tree%nLocalBlocks) then
            write(*, '(a)') 'falcon: one process owns every block on its GPU'
            stop 1
        end if
        ! 1-D profiles as ModBlock_InitSSM reads them (3_6/ModBlock.f90:232-249)
        allocate(bg(-ng+1:ni+ng, 8))
        do ib = 1, Tree%nLocalBlocks
            B => Tree%LocalBlocks(ib)
            bg(:, 1) = B%g_I
            bg(:, 2) = B%rho0_I
            bg(:, 3) = B%p0_I
            bg(:, 4) = B%te0_I
            bg(:, 5) = B%gamma1_I
            bg(:, 6) = B%gamma3_I
            bg(:, 7) = B%diffusion_I + B%cooling_I
            bg(:, 8) = B%Xi_rsst_I
            call falcon_check(falcon_set_background(falcon_h, int(B%iBlock - 1, c_int), bg))
        end do
        call falcon_check(falcon_set_gc_all(falcon_h))
        ! GLM divergence cleaning (3_6/ModAdvance.f90:82-84): psi is stepped on the device too
        if (DivB_option == 1) call falcon_check(falcon_set_divb(falcon_h, 1_c_int))
        call falcon_host_tables(Tree)
        state_pending = .true.
    end subroutine falcon_attach

    ! The host-side tables that ModGC_FindGC_Single and ModGC_GetGC_Targets_single
    ! leave in every block (3_6/ModGC.f90:25-147), from the library's own (the
    ! same integers): GC_iBlocks_III and GC_targets with their ghost indices and
    ! receiver-frame cell centres.  The device exchange needs no GC_sources.
    subroutine falcon_host_tables(Tree)
        type(YYTree), intent(inout), target :: Tree
        type(BlockType), pointer        :: B
        type(GC_target), pointer        :: T
        integer(c_int32_t), allocatable :: gcib(:)
        integer(c_int32_t), allocatable, target :: ijk(:)
        integer(c_int)                  :: ntg, donor, ngc
        integer                         :: ib, it, g
        allocate(gcib((ni + 2 * ng) * (nj + 2 * ng) * (nk + 2 * ng)))
        do ib = 1, Tree%nLocalBlocks
            B => Tree%LocalBlocks(ib)
            call falcon_check(falcon_gc_iblocks(falcon_h, int(B%iBlock - 1, c_int), gcib))
            B%GC_iBlocks_III = reshape(gcib, shape(B%GC_iBlocks_III))
            call falcon_check(falcon_gc_ntargets(falcon_h, int(B%iBlock - 1, c_int), ntg))
            if (allocated(B%GC_targets)) deallocate(B%GC_targets)
            allocate(B%GC_targets(ntg))
            B%nGC_targets = ntg
            B%nGC_sources = 0
            do it = 1, ntg
                T => B%GC_targets(it)
                call falcon_check(falcon_gc_target(falcon_h, int(B%iBlock - 1, c_int), int(it - 1, c_int), donor, &
                                                   ngc, c_null_ptr))
                allocate(ijk(3 * ngc))
                call falcon_check(falcon_gc_target(falcon_h, int(B%iBlock - 1, c_int), int(it - 1, c_int), donor, &
                                                   ngc, c_loc(ijk)))
                T%iBlock = donor
                T%if_yin = donor <= Tree%NumLeafNodes_YinYang(1)
                T%iRank = ModAllocation_GetRank(donor)
                T%nGC = ngc
                allocate(T%ijk_list(ngc, 3), T%xijk_list(ngc, 3), T%primitive_list(ngc, nvar))
                T%ijk_list = reshape(ijk, [int(ngc), 3])
                do g = 1, ngc
                    T%xijk_list(g, :) = [B%xi_I(T%ijk_list(g, 1)), B%xj_I(T%ijk_list(g, 2)), B%xk_I(T%ijk_list(g, 3))]
                end do
                T%primitive_list = 0.0d0
                deallocate(ijk)
            end do
        end do
    end subroutine falcon_host_tables

    ! Copy the host state up if it changed since falcon_attach (the driver's
    ! ModInitiation_DoAll, 3_6/test1.f90:140); called by the drop-in entry points.
    subroutine falcon_push_pending(Tree)
        type(YYTree), intent(inout), target :: Tree
        if (.not. state_pending) return
        call falcon_push_state(Tree)
        state_pending = .false.
    end subroutine falcon_push_pending

    ! host -> HBM: rho1, vr, vt, vp, br, bt, bp, s1 of every block, and psi with
    ! GLM (DivB_option=1; otherwise psi stays on the host, where no term writes
    ! it); a driver that changes primitive_IV on the host later calls it
    subroutine falcon_push_state(Tree)
        type(YYTree), intent(inout), target :: Tree
        integer                         :: ib
        type(BlockType), pointer        :: B
        do ib = 1, Tree%nLocalBlocks
            B => Tree%LocalBlocks(ib)
            call falcon_check(falcon_set_state(falcon_h, int(B%iBlock - 1, c_int), B%primitive_IV(:, :, :, 1:8)))
            if (DivB_option == 1) &
                call falcon_check(falcon_set_psi(falcon_h, int(B%iBlock - 1, c_int), B%primitive_IV(:, :, :, B%psi_)))
        end do
    end subroutine falcon_push_state

    ! HBM -> host, before ModCheck_primitive / ModSave_DoAll (3_6/test1.f90:61,69)
    subroutine falcon_sync_to_host(Tree)
        type(YYTree), intent(inout), target :: Tree
        real(c_double), allocatable     :: prim(:, :, :, :)
        integer                         :: ib
        allocate(prim(-ng+1:ni+ng, -ng+1:nj+ng, -ng+1:nk+ng, 8))
        do ib = 1, Tree%nLocalBlocks
            call falcon_check(falcon_get_state(falcon_h, int(Tree%LocalBlocks(ib)%iBlock - 1, c_int), prim))
            Tree%LocalBlocks(ib)%primitive_IV(:, :, :, 1:8) = prim
            if (DivB_option == 1) then
                call falcon_check(falcon_get_psi(falcon_h, int(Tree%LocalBlocks(ib)%iBlock - 1, c_int), prim(:, :, :, 1)))
                Tree%LocalBlocks(ib)%primitive_IV(:, :, :, Tree%LocalBlocks(ib)%psi_) = prim(:, :, :, 1)
            end if
        end do
    end subroutine falcon_sync_to_host
end module ModFalconShim
