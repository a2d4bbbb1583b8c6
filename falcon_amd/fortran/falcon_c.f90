! ISO_C_BINDING interfaces of libfalcon_hip.so (declarations: include/falcon_hip.h).
! Every function returns 0 on success; falcon_last_error() then holds the message.
module falcon_c
    use iso_c_binding
    implicit none

    interface
        integer(c_int) function falcon_create(h, ni, nj, nk, ng, r_lo, r_hi) bind(c, name='falcon_create')
            import :: c_int, c_double, c_ptr
            type(c_ptr), intent(out)        :: h
            integer(c_int), value           :: ni, nj, nk, ng
            real(c_double), value           :: r_lo, r_hi
        end function falcon_create

        integer(c_int) function falcon_destroy(h) bind(c, name='falcon_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
        end function falcon_destroy

        integer(c_int) function falcon_tree_set_nodes(h, n, parent, is_yin, rng6) bind(c, name='falcon_tree_set_nodes')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: n
            integer(c_int), intent(in)      :: parent(*), is_yin(*)
            real(c_double), intent(in)      :: rng6(*)
        end function falcon_tree_set_nodes

        integer(c_int) function falcon_nblocks(h, nblocks, nleaf_yin) bind(c, name='falcon_nblocks')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), intent(out)     :: nblocks, nleaf_yin
        end function falcon_nblocks

        integer(c_int) function falcon_set_background(h, ib, bg) bind(c, name='falcon_set_background')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: ib
            real(c_double), intent(in)      :: bg(*)
        end function falcon_set_background

        integer(c_int) function falcon_set_gc_all(h) bind(c, name='falcon_set_gc_all')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
        end function falcon_set_gc_all

        integer(c_int) function falcon_gc_iblocks(h, ib, gcib) bind(c, name='falcon_gc_iblocks')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: ib
            integer(c_int32_t), intent(out) :: gcib(*)
        end function falcon_gc_iblocks

        integer(c_int) function falcon_gc_ntargets(h, ib, n) bind(c, name='falcon_gc_ntargets')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: ib
            integer(c_int), intent(out)     :: n
        end function falcon_gc_ntargets

        ! ijk: c_null_ptr to query ngc, else c_loc of 3*ngc integers
        integer(c_int) function falcon_gc_target(h, ib, it, donor, ngc, ijk) bind(c, name='falcon_gc_target')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: ib, it
            integer(c_int), intent(out)     :: donor, ngc
            type(c_ptr), value              :: ijk
        end function falcon_gc_target

        integer(c_int) function falcon_set_state(h, ib, prim) bind(c, name='falcon_set_state')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: ib
            real(c_double), intent(in)      :: prim(*)
        end function falcon_set_state

        integer(c_int) function falcon_get_state(h, ib, prim) bind(c, name='falcon_get_state')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: ib
            real(c_double), intent(out)     :: prim(*)
        end function falcon_get_state

        integer(c_int) function falcon_check_state(h, which, n_nan, where) bind(c, name='falcon_check_state')
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: which
            integer(c_int64_t), intent(out) :: n_nan
            integer(c_int32_t), intent(out) :: where(5)
        end function falcon_check_state

        integer(c_int) function falcon_set_precision(h, mode) bind(c, name='falcon_set_precision')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: mode
        end function falcon_set_precision

        ! launch-shape autotuner of the current precision; shape_out (4 ints) and
        ! ms_out (8 doubles) may be c_null_ptr
        integer(c_int) function falcon_autotune(h, reps, shape_out, ms_out) bind(c, name='falcon_autotune')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: reps
            type(c_ptr), value              :: shape_out, ms_out
        end function falcon_autotune

        integer(c_int) function falcon_sendrecv_gc(h, which) bind(c, name='falcon_sendrecv_gc')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: which
        end function falcon_sendrecv_gc

        integer(c_int) function falcon_advance_rk4(h, cfl, dt) bind(c, name='falcon_advance_rk4')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            real(c_double), value           :: cfl
            real(c_double), intent(out)     :: dt
        end function falcon_advance_rk4

        integer(c_int) function falcon_heat_setup(h, kappa0, r_ref_rsun, nsys) bind(c, name='falcon_heat_setup')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            real(c_double), value           :: kappa0, r_ref_rsun
            integer(c_int), value           :: nsys
        end function falcon_heat_setup

        integer(c_int) function falcon_heat_set_dt(h, dt) bind(c, name='falcon_heat_set_dt')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            real(c_double), intent(in)      :: dt(*)
        end function falcon_heat_set_dt

        integer(c_int) function falcon_heat_set_T(h, T) bind(c, name='falcon_heat_set_T')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            real(c_double), intent(in)      :: T(*)
        end function falcon_heat_set_T

        integer(c_int) function falcon_heat_solve(h, tol, max_iter, use_corrector, iters, relres, seconds) &
                bind(c, name='falcon_heat_solve')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            real(c_double), value           :: tol
            integer(c_int), value           :: max_iter, use_corrector
            integer(c_int), intent(out)     :: iters(*)
            real(c_double), intent(out)     :: relres(*), seconds
        end function falcon_heat_solve

        integer(c_int) function falcon_heat_get_T(h, sys, T) bind(c, name='falcon_heat_get_T')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: sys
            real(c_double), intent(out)     :: T(*)
        end function falcon_heat_get_T

        integer(c_int) function falcon_set_divb(h, option) bind(c, name='falcon_set_divb')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: option
        end function falcon_set_divb

        integer(c_int) function falcon_set_psi(h, ib, psi) bind(c, name='falcon_set_psi')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: ib
            real(c_double), intent(in)      :: psi(*)
        end function falcon_set_psi

        integer(c_int) function falcon_get_psi(h, ib, psi) bind(c, name='falcon_get_psi')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            integer(c_int), value           :: ib
            real(c_double), intent(out)     :: psi(*)
        end function falcon_get_psi

        integer(c_int) function falcon_pfss_setup(h) bind(c, name='falcon_pfss_setup')
            import :: c_int, c_ptr
            type(c_ptr), value              :: h
        end function falcon_pfss_setup

        integer(c_int) function falcon_pfss_bottom_blocks(h, blocks, n) bind(c, name='falcon_pfss_bottom_blocks')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value              :: h
            integer(c_int32_t), intent(out) :: blocks(*)
            integer(c_int), intent(out)     :: n
        end function falcon_pfss_bottom_blocks

        integer(c_int) function falcon_pfss_set_magnetogram(h, br) bind(c, name='falcon_pfss_set_magnetogram')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            real(c_double), intent(in)      :: br(*)
        end function falcon_pfss_set_magnetogram

        integer(c_int) function falcon_pfss_solve(h, tol, max_iter, iters, relres, seconds) &
                bind(c, name='falcon_pfss_solve')
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value              :: h
            real(c_double), value           :: tol
            integer(c_int), value           :: max_iter
            integer(c_int32_t), intent(out) :: iters
            real(c_double), intent(out)     :: relres, seconds
        end function falcon_pfss_solve

        integer(c_int) function falcon_pfss_get_phi(h, phi) bind(c, name='falcon_pfss_get_phi')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            real(c_double), intent(out)     :: phi(*)
        end function falcon_pfss_get_phi

        integer(c_int) function falcon_pfss_field(h, B) bind(c, name='falcon_pfss_field')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value              :: h
            real(c_double), intent(out)     :: B(*)
        end function falcon_pfss_field

        integer(c_int) function falcon_last_error(buf, len) bind(c, name='falcon_last_error')
            import :: c_int, c_char
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_int), value           :: len
        end function falcon_last_error
    end interface
end module falcon_c
