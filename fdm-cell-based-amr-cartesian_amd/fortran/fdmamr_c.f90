! fdmamr_c.f90 -- ISO_C_BINDING interface to libfdmamr (include/fdmamr.h).
!
! The Fortran driver calls the MI355X hot path through these interfaces. Every entry point
! returns 0 on success and a negative code on failure (fdmamr_last_error gives the message).
module fdmamr_c
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: FDMAMR_BC_TRANSMISSIVE = 1, FDMAMR_BC_INFLOW = 2, &
                                     FDMAMR_BC_PERIODIC = 6
    integer(c_int32_t), parameter :: FDMAMR_FLUX_FP64 = 0, FDMAMR_FLUX_FP32 = 1

    type, bind(C) :: fdmamr_params
        real(c_double)    :: xmin, xmax, ymin, ymax
        real(c_double)    :: h, gamma, Ma
        integer(c_int32_t):: bc_xm, bc_xp, bc_ym, bc_yp
        real(c_double)    :: inflow(4)
        integer(c_int32_t):: flux_precision, selector, device, wls_local
    end type fdmamr_params

    interface
        integer(c_int32_t) function fdmamr_create(p, ctx) bind(C, name="fdmamr_create")
            import :: c_int32_t, c_ptr, fdmamr_params
            type(fdmamr_params), intent(in) :: p
            type(c_ptr), intent(out) :: ctx
        end function
        subroutine fdmamr_destroy(ctx) bind(C, name="fdmamr_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        integer(c_int32_t) function fdmamr_last_error(ctx, buf, n) bind(C, name="fdmamr_last_error")
            import :: c_int32_t, c_ptr, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_int32_t), value :: n
        end function
        integer(c_int32_t) function fdmamr_creat_mesh(ctx) bind(C, name="fdmamr_creat_mesh")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function fdmamr_initial(ctx, ic_kind, seed, ic_param) &
                bind(C, name="fdmamr_initial")
            import :: c_int32_t, c_int64_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: ic_kind
            integer(c_int64_t), value :: seed
            real(c_double), intent(in) :: ic_param(8)
        end function
        integer(c_int32_t) function fdmamr_build_tables(ctx) bind(C, name="fdmamr_build_tables")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function fdmamr_counts(ctx, n_leaf, n_ghost, n_wls, n_rst, l_max) &
                bind(C, name="fdmamr_counts")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: n_leaf, n_ghost, n_wls, n_rst, l_max
        end function
        integer(c_int32_t) function fdmamr_get_state(ctx, prim5) bind(C, name="fdmamr_get_state")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: prim5(*)
        end function
        integer(c_int32_t) function fdmamr_set_state(ctx, prim5) bind(C, name="fdmamr_set_state")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: prim5(*)
        end function
        integer(c_int32_t) function fdmamr_leaf_geometry(ctx, x, y, lvl, node) &
                bind(C, name="fdmamr_leaf_geometry")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: x(*), y(*)
            integer(c_int32_t), intent(out) :: lvl(*)
            type(c_ptr), value :: node
        end function
        integer(c_int32_t) function fdmamr_max_wavespeed(ctx, lam_x, lam_y, l_max) &
                bind(C, name="fdmamr_max_wavespeed")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: lam_x, lam_y
            integer(c_int32_t), intent(out) :: l_max
        end function
        integer(c_int32_t) function fdmamr_advance_rk3(ctx, dt, lam_x, lam_y, res_max, res_rms) &
                bind(C, name="fdmamr_advance_rk3")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt, lam_x, lam_y
            real(c_double), intent(out) :: res_max(4), res_rms(4)
        end function
        integer(c_int32_t) function fdmamr_regrid(ctx, lvl_amr, rg_rot, rg_div, cg_rot, cg_div, mode) &
                bind(C, name="fdmamr_regrid")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: lvl_amr, mode
            real(c_double), value :: rg_rot, rg_div, cg_rot, cg_div
        end function
    end interface

contains

    subroutine fdmamr_check(ctx, rc, what)
        type(c_ptr), intent(in) :: ctx
        integer(c_int32_t), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char) :: buf(512)
        integer :: i, dummy
        character(len=512) :: msg
        if (rc == 0) return
        dummy = fdmamr_last_error(ctx, buf, 512)
        msg = ' '
        do i = 1, 512
            if (buf(i) == c_null_char) exit
            msg(i:i) = buf(i)
        end do
        write(*,*) 'libfdmamr: ', what, ' failed (', rc, '): ', trim(msg)
        stop 'libfdmamr error'
    end subroutine fdmamr_check

end module fdmamr_c
