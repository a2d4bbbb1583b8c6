! Drop-in ModCheck: compiled in place of 3_6/ModCheck.f90, same module and entry
! point, so the driver's `call ModCheck_primitive(Tree, .false.)` (3_6/test1.f90:61)
! stays unchanged.  The NaN scan runs on the blocks resident in HBM
! (falcon_check_state); the first NaN in the reference's scan order (blocks,
! then phi, theta, r, variable innermost; 3_6/ModCheck.f90:24-40) is reported
! with the reference's message, and the run stops as the reference's
! MPI_Abort does.
module ModCheck
    use iso_c_binding
    use falcon_c,       only: falcon_check_state
    use ModFalconShim,  only: falcon_h, falcon_check, falcon_push_pending
    use ModParameters,  only: ng, MpiRank
    use ModYinYang,     only: ModYinYang_CoordConv_0D
    use ModYinYangTree, only: YYTree
    use ModBlock,       only: BlockType
    implicit none
contains
    subroutine ModCheck_primitive(Tree, if_rk)
        type(YYTree), target                :: Tree
        logical, intent(in)                 :: if_rk
        integer(c_int64_t)                  :: n_nan
        integer(c_int32_t)                  :: w(5)
        integer                             :: ir, it, ip, ivar
        real(8)                             :: coord(3)
        type(BlockType), pointer            :: Block1
        call falcon_push_pending(Tree)
        call falcon_check(falcon_check_state(falcon_h, merge(1_c_int, 0_c_int, if_rk), n_nan, w))
        if (n_nan == 0) return
        Block1 => Tree%LocalBlocks(w(1) + 1)
        ir = w(2) - ng + 1; it = w(3) - ng + 1; ip = w(4) - ng + 1; ivar = w(5) + 1
        coord = [Block1%xi_I(ir), Block1%xj_I(it), Block1%xk_I(ip)]
        if (.not. Block1%if_yin) coord = ModYinYang_CoordConv_0D(coord)
        write(*,*) 'Error: Detecting NAN at MpiRank=', MpiRank, &
            'iBlock=', Block1%iBlock, 'ir,it,ip,ivar=', ir, it, ip, ivar, &
            'r,t,p=', coord
        stop 1
    end subroutine ModCheck_primitive
end module ModCheck
