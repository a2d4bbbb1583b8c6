! ModPFSS (3_5/ModPFSS.f90:21-148): the potential-field source-surface solve of
! the library on the MHD blocks (cubic, power-of-two edge).  The Laplacian
! with the flux Br S of the magnetogram through the bottom faces and Phi = 0
! on the top face is solved by the MG-preconditioned flexible CG; the field is
! B0 = -grad(Phi), second order.
!
! 3_6's BlockType has no magnetogram_LL, Phi_LLL or B0_IV (3_5/ModBlock.f90
! holds them), so the magnetogram comes in and Phi and B0 go out as arrays:
!   magnetogram(ni, ni, nBottom): Br on the bottom faces (theta, phi) of the
!     bottom blocks (if_bottom), in block order: the magnetogram_LL of each;
!   Phi(ni, ni, ni, nLocalBlocks), B0(ni, ni, ni, 3, nLocalBlocks).
! The reference stores +grad(Phi) in B0_IV (3_5/ModPFSS.f90:136-141) against
! its own comment and its right-hand side b = +Br S; B0 here is -grad(Phi),
! whose radial component on the bottom face is the magnetogram.
module ModPFSS
    use iso_c_binding
    use falcon_c
    use ModFalconShim,  only: falcon_h, falcon_check, falcon_solver
    use ModYinYangTree, only: YYTree
    implicit none
contains
    subroutine ModPFSS_setup(Tree, magnetogram)
        type(YYTree), intent(inout), target :: Tree
        real(8), intent(in)                 :: magnetogram(:, :, :)
        if (falcon_solver /= 2) then
            call falcon_check(falcon_pfss_setup(falcon_h))
            falcon_solver = 2
        end if
        call falcon_check(falcon_pfss_set_magnetogram(falcon_h, magnetogram))
    end subroutine ModPFSS_setup

    ! i_option 1 (ModLinearSolver_Multigrid_CG) only: the Jacobi-CG options 2
    ! and 3 of the reference are not provided.  tol: ||b - A Phi|| / ||b||.
    subroutine ModPFSS_solve(Tree, max_iter, tol, i_option, Phi, B0, iters)
        type(YYTree), intent(inout), target :: Tree
        integer, intent(in)                 :: max_iter
        real, intent(in)                    :: tol
        integer, intent(in)                 :: i_option
        real(8), intent(out)                :: Phi(:, :, :, :), B0(:, :, :, :, :)
        integer, intent(out), optional      :: iters
        integer(c_int32_t)                  :: it
        real(c_double)                      :: relres, seconds
        if (i_option /= 1) then
            write(*, '(a,i0)') 'ModPFSS_solve: only i_option 1 (multigrid CG) is provided, got ', i_option
            stop 1
        end if
        if (falcon_solver /= 2) then
            write(*, '(a)') 'ModPFSS_solve: ModPFSS_setup has not been called'
            stop 1
        end if
        call falcon_check(falcon_pfss_solve(falcon_h, real(tol, c_double), int(max_iter, c_int), it, relres, &
                                            seconds))
        call falcon_check(falcon_pfss_get_phi(falcon_h, Phi))
        call falcon_check(falcon_pfss_field(falcon_h, B0))
        if (present(iters)) iters = it
    end subroutine ModPFSS_solve
end module ModPFSS
