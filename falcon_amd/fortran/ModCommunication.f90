! Drop-in ModCommunication for the stepping driver: compiled in place of
! 3_6/ModCommunication.f90; keeps the two entry points the driver calls
! (3_6/test1.f90:133,143-144).  The ghost fill becomes on-device block-to-block
! copies and Yin-Yang interpolation driven by the library's tables.
module ModCommunication
    use iso_c_binding
    use falcon_c,       only: falcon_sendrecv_gc
    use ModFalconShim,  only: falcon_h, falcon_check, falcon_attach, falcon_push_pending
    use ModYinYangTree, only: YYTree
    implicit none
contains
    ! 3_6/ModCommunication.f90:22-64: donors, targets, sources; here also the
    ! upload of tree and background (the state follows at the first fill)
    subroutine ModCommunication_SetGCAll(Tree)
        type(YYTree), intent(inout), target :: Tree
        call falcon_attach(Tree)
    end subroutine ModCommunication_SetGCAll

    ! 3_6/ModCommunication.f90:655-795: if_rk fills the stage buffer's ghosts
    subroutine ModCommunication_SendRecvGC_new(Tree, if_rk)
        type(YYTree), intent(inout), target :: Tree
        logical, intent(in)                 :: if_rk
        call falcon_push_pending(Tree)
        call falcon_check(falcon_sendrecv_gc(falcon_h, merge(1_c_int, 0_c_int, if_rk)))
    end subroutine ModCommunication_SendRecvGC_new
end module ModCommunication
