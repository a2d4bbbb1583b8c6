! ModHeatConduction (3_6/ModHeatConduction.f90:1-16 is an empty module): the
! implicit heat solve of the library, shaped as ModPFSS_setup / ModPFSS_solve
! (3_5/ModPFSS.f90:21,102).  Backward Euler for T' with kappa(r) = kappa0
! (r / r_ref)^-3, MG-preconditioned flexible CG to ||b - A T||/||b|| <= tol.
module ModHeatConduction
    use iso_c_binding
    use falcon_c
    use ModFalconShim,  only: falcon_h, falcon_check, falcon_solver
    use ModParameters,  only: ni
    use ModYinYangTree, only: YYTree
    implicit none
    real(8), save   :: heat_kappa0 = 1.0d0          ! cgs
    real(8), save   :: heat_r_ref_rsun = 0.7d0
contains
    ! T(ni,ni,ni,nLocalBlocks): T' at the old time level in, the new level out
    ! (interior cells, cubic blocks); iters = FCG iterations.
    subroutine ModHeatConduction_solve(Tree, T, dt, tol, max_iter, iters)
        type(YYTree), intent(inout), target :: Tree
        real(8), intent(inout)              :: T(:, :, :, :)
        real(8), intent(in)                 :: dt, tol
        integer, intent(in)                 :: max_iter
        integer, intent(out)                :: iters
        real(c_double)                      :: dts(1), relres(1), seconds
        integer(c_int)                      :: it(1)
        if (falcon_solver /= 1) then
            call falcon_check(falcon_heat_setup(falcon_h, heat_kappa0, heat_r_ref_rsun, 1_c_int))
            falcon_solver = 1
        end if
        dts(1) = dt
        call falcon_check(falcon_heat_set_dt(falcon_h, dts))
        call falcon_check(falcon_heat_set_T(falcon_h, T))
        call falcon_check(falcon_heat_solve(falcon_h, real(tol, c_double), int(max_iter, c_int), 0_c_int, it, relres, &
                                            seconds))
        call falcon_check(falcon_heat_get_T(falcon_h, 0_c_int, T))
        iters = it(1)
    end subroutine ModHeatConduction_solve
end module ModHeatConduction
