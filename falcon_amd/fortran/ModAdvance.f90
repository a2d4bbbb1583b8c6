! Drop-in ModAdvance: compiled in place of 3_6/ModAdvance.f90, same module and
! entry point, so the driver's `use ModAdvance` and its call
! `call ModAdvance_rk4(Tree, .true., dt)` (3_6/test1.f90:58) stay unchanged.
! The step (4 stages, artificial diffusion, 5 ghost fills; ModAdvance.f90:18-112)
! runs on the blocks resident in HBM.
module ModAdvance
    use iso_c_binding
    use falcon_c,       only: falcon_advance_rk4
    use ModFalconShim,  only: falcon_h, falcon_check, falcon_push_pending
    use ModParameters,  only: CFL
    use ModYinYangTree, only: YYTree
    implicit none
contains
    subroutine ModAdvance_rk4(Tree, if_all_same_sizes, dt_global)
        type(YYTree), intent(inout), target :: Tree
        logical, intent(in)                 :: if_all_same_sizes    ! every block is 16^3-alike on the GPU
        real(8)                             :: dt_global
        real(c_double)                      :: dt
        call falcon_push_pending(Tree)
        call falcon_check(falcon_advance_rk4(falcon_h, real(CFL, c_double), dt))
        dt_global = dt
    end subroutine ModAdvance_rk4
end module ModAdvance
