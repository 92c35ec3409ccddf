! fdmamr_driver.f90 -- Fortran driver with the reference's subroutine interfaces
! (MAIN.F90, solution.F90, mesh.F90, output_ADT.F90), running the per-step hot path on the
! MI355X through libfdmamr.
!
! The reference keeps its state in module globals and a pointer quadtree; here the quadtree
! and the leaf state live in the library (host tree + device tables). The driver keeps the
! reference's control flow: set_parameter, creat_mesh, idneighbor, initial, solute (lambda,
! dt, Time_advance_RK, residual, output_Res, AMR, output_error). Parameters come from an
! optional namelist file (first command-line argument), defaulting to the reference's
! hard-coded values (MAIN.F90:29-71).

module Global_Var
    use fdmamr_c
    implicit none
    integer, parameter      :: prec = kind(1.0d0)
    real(prec)              :: Xmin = 0.0, Xmax = 10.0, Ymin = 0.0, Ymax = 10.0
    real(prec)              :: h = 0.5_prec, x_factor = 1.0
    real(prec)              :: gama = 1.4          ! REAL(4) literal widened, as MAIN.F90:38
    real(prec)              :: Ma = 1.0, cv
    logical                 :: iflag_AMR = .false., iflag_localdt = .false.
    integer                 :: lvlAMR = 0, amr_sensor = 0
    real(prec)              :: Refine_goal_rot = 1.4, Refine_goal_div = 1.4
    real(prec)              :: Coarse_goal_rot = 0.4, Coarse_goal_div = 0.4
    integer                 :: kind_xpbc = 6, kind_xmbc = 6, kind_ypbc = 6, kind_ymbc = 6
    real(prec)              :: inflow(4) = 0.0_prec
    real(prec)              :: t_end = 10.0, tstep = 1e-4, CFL = 0.45
    integer                 :: it = 0, Nstep_Adapt = 0, Pstep_Adapt = 30, Pstep_Res = 100
    integer                 :: max_steps = huge(1)
    real(prec)              :: lambda_x = 0.0_prec, lambda_y = 0.0_prec, tt = 0.0_prec
    integer                 :: l_max = 0
    real(prec)              :: Res_max(4), Res_rms(4)
    integer                 :: Neq = 4, nResgrid = 0, nBCgrid = 0
    integer                 :: flux_precision = 0, ic_kind = 0, wls_local = 0
    integer(8)              :: seed = 0
    real(prec)              :: ic_param(8) = 0.0_prec
    type(c_ptr)             :: ctx = c_null_ptr
end module Global_Var


module mesh
    use Global_Var
    implicit none
contains
    subroutine creat_mesh                            ! mesh.F90:9-47
        type(fdmamr_params) :: p
        p%xmin = Xmin; p%xmax = Xmax; p%ymin = Ymin; p%ymax = Ymax
        p%h = h; p%gamma = gama; p%Ma = Ma
        p%bc_xm = kind_xmbc; p%bc_xp = kind_xpbc; p%bc_ym = kind_ymbc; p%bc_yp = kind_ypbc
        p%inflow = inflow
        p%flux_precision = flux_precision; p%selector = 0; p%device = 0; p%wls_local = wls_local
        call fdmamr_check(ctx, fdmamr_create(p, ctx), 'fdmamr_create')
        call fdmamr_check(ctx, fdmamr_creat_mesh(ctx), 'creat_mesh')
    end subroutine creat_mesh

    subroutine idneighbor                            ! neighbor.F90:399-409 (cached in the library)
    end subroutine idneighbor
end module mesh


module outfile
    use Global_Var
    implicit none
contains
    subroutine output_Res                            ! output_ADT.F90:7-15
        write(*, "(I8,I8,6E16.6)") it, nResgrid, tstep, tt, Res_rms(1:Neq)
        open(99, file='./Output/'//"Residual.dat", position="append")
        write(99, "(I8,I8,12E24.10)") it, nResgrid, tstep, tt, Res_max(1:Neq), Res_rms(1:Neq)
        close(99)
    end subroutine output_Res

    subroutine output_error                          ! output_ADT.F90:18-75 (isentropic vortex)
        ! Density errors against the exact vortex, area weighted over all interior leaves (L)
        ! and over the refined ones only (P), in the reference's file layout: a header on the
        ! first call (output_ADT.F90:35), then it, nResgrid, L1, P1, L2, P2, Loo, Poo
        ! (output_ADT.F90:65) and the second line of output_ADT.F90:67-70, whose base-grid and
        ! patch columns the reference never computes (written as zeros here).
        real(prec), allocatable :: prim(:), x(:), y(:)
        integer(c_int32_t), allocatable :: lvl(:)
        real(prec) :: dx, dy, rr, dT, rhorho, Unisum, Patchsum, L1, L2, Loo, P1, P2, Poo
        real(prec), parameter :: Pi = 3.1415926535897932d0, cc1 = 5.d0
        integer, save :: count = 1
        logical :: refined
        integer :: i
        allocate(prim(5*nResgrid), x(nResgrid), y(nResgrid), lvl(nResgrid))
        call fdmamr_check(ctx, fdmamr_get_state(ctx, prim), 'get_state')
        call fdmamr_check(ctx, fdmamr_leaf_geometry(ctx, x, y, lvl, c_null_ptr), 'leaf_geometry')
        Unisum = 0.d0; Patchsum = 0.d0; refined = .false.
        L1 = 0.d0; L2 = 0.d0; Loo = 0.d0; P1 = 0.d0; P2 = 0.d0; Poo = 0.d0
        do i = 1, nResgrid
            dy = h/2**lvl(i); dx = x_factor*dy
            Unisum = Unisum + dx*dy
            rr = (x(i)-5.d0)**2 + (y(i)-5.d0)**2
            dT = -(gama-1.d0)*(cc1*cc1)*exp(1.d0-rr)/(8.d0*gama*Pi*Pi)
            rhorho = prim(5*(i-1)+1) - (1.d0 + dT)**(1.d0/(gama - 1.d0))
            L1 = L1 + abs(rhorho)*dx*dy; L2 = L2 + rhorho*rhorho*dx*dy; Loo = max(Loo, abs(rhorho))
            if (lvl(i) /= 0) then
                refined = .true.
                Patchsum = Patchsum + dx*dy
                P1 = P1 + abs(rhorho)*dx*dy; P2 = P2 + rhorho*rhorho*dx*dy; Poo = max(Poo, abs(rhorho))
            end if
        end do
        L1 = L1/Unisum; L2 = sqrt(L2/Unisum)
        if (refined) then
            P1 = P1/Patchsum; P2 = sqrt(P2/Patchsum)
        end if
        open(83, file='./Output/accuracy_error.dat', position="append")
        if (count == 1) write(83, *) 'nstep Ngrids L1_rho P1_rho L2_rho P2_rho Loo_rho Poo_rho '
        write(83, "(2I8,6E14.6)") it, nResgrid, L1, P1, L2, P2, Loo, Poo
        write(83, "(I8, I8, 3E14.6, I8, 3E14.6, I8, 3E14.6)") it, nResgrid, L1, L2, Loo, &
            0, 0.d0, 0.d0, 0.d0, 0, 0.d0, 0.d0, 0.d0
        close(83)
        count = count + 1
        deallocate(prim, x, y, lvl)
    end subroutine output_error
end module outfile


module solution
    use Global_Var
    use mesh
    use outfile
    implicit none
contains
    subroutine initial                               ! solution.F90:9-29
        call fdmamr_check(ctx, fdmamr_initial(ctx, int(ic_kind, c_int32_t), int(seed, c_int64_t), &
                                              ic_param), 'initial')
    end subroutine initial

    subroutine IDgrid                                ! solution.F90:593-615 + table flattening
        integer(c_int32_t) :: n_leaf, n_ghost, n_wls, n_rst, lm
        call fdmamr_check(ctx, fdmamr_build_tables(ctx), 'build_tables')
        call fdmamr_check(ctx, fdmamr_counts(ctx, n_leaf, n_ghost, n_wls, n_rst, lm), 'counts')
        nResgrid = n_leaf; nBCgrid = n_ghost
    end subroutine IDgrid

    subroutine Time_advance_RK                       ! solution.F90:267-298
        call fdmamr_check(ctx, fdmamr_advance_rk3(ctx, tstep, lambda_x, lambda_y, Res_max, Res_rms), &
                          'advance_rk3')
    end subroutine Time_advance_RK

    subroutine AMR                                   ! solution.F90:408-421
        ! fdmamr_regrid runs rot_div, solute_ref, solute_gridmodify, idneighbor and IDgrid with
        ! the table flattening and the ghost fill; only the leaf counts are read back here
        integer(c_int32_t) :: n_leaf, n_ghost, n_wls, n_rst, lm
        if (it >= Nstep_Adapt .and. mod(it, Pstep_Adapt) == 1) then
            call fdmamr_check(ctx, fdmamr_regrid(ctx, int(lvlAMR, c_int32_t), Refine_goal_rot, &
                              Refine_goal_div, Coarse_goal_rot, Coarse_goal_div, &
                              int(amr_sensor, c_int32_t)), 'regrid')
            call fdmamr_check(ctx, fdmamr_counts(ctx, n_leaf, n_ghost, n_wls, n_rst, lm), 'counts')
            nResgrid = n_leaf; nBCgrid = n_ghost
        end if
    end subroutine AMR

    subroutine solute                                ! solution.F90:171-238
        use, intrinsic :: ieee_arithmetic, only: ieee_is_nan
        real(prec) :: lx, ly
        integer(c_int32_t) :: lm
        integer :: i
        call IDgrid
        write(*,*) "Start the Calculation ..."
        do while (tt < t_end .and. it < max_steps)
            it = it + 1
            call fdmamr_check(ctx, fdmamr_max_wavespeed(ctx, lx, ly, lm), 'max_wavespeed')
            lambda_x = max(lambda_x, lx); lambda_y = max(lambda_y, ly); l_max = lm
            if (iflag_localdt) then
                tstep = cfl/(lambda_x/(x_factor*h/2**l_max) + lambda_y/(h/2**l_max))
            end if
            if (tt + tstep > t_end) tstep = t_end - tt
            call Time_advance_RK
            tt = tt + tstep
            do i = 1, Neq
                Res_rms(i) = sqrt(Res_rms(i)/(1.d0*nResgrid))
                if (ieee_is_nan(Res_rms(i))) then
                    write(*,*) 'ERROR: NaN after sqrt in equation ', i, ' at it =', it
                    stop 'NaN detected in Res_rms after sqrt'
                end if
            end do
            if (mod(it, Pstep_Res) == 0 .or. tt == t_end) call output_Res
            if (iflag_AMR) call AMR
            if (tt == t_end) call output_error
        end do
    end subroutine solute
end module solution


program main                                         ! MAIN.F90:1-25
    use Global_Var
    use mesh
    use solution
    implicit none
    call set_parameter
    call creat_mesh
    call idneighbor
    call initial
    call solute
    ! a run cut short by max_steps (the reference always runs to t_end) still reports its last
    ! residual and, for the vortex, its error
    if (mod(it, Pstep_Res) /= 0 .and. tt /= t_end) call output_Res
    if (ic_kind == 0 .and. tt /= t_end) call output_error
    call fdmamr_destroy(ctx)
contains
    subroutine set_parameter                         ! MAIN.F90:29-71, overridable by namelist
        character(len=512) :: fname
        integer :: u, ios
        namelist /fdmamr/ Xmin, Xmax, Ymin, Ymax, h, gama, Ma, iflag_AMR, lvlAMR, amr_sensor, &
            Refine_goal_rot, Refine_goal_div, Coarse_goal_rot, Coarse_goal_div, kind_xpbc, &
            kind_xmbc, kind_ypbc, kind_ymbc, inflow, t_end, iflag_localdt, tstep, CFL, &
            Nstep_Adapt, Pstep_Adapt, Pstep_Res, max_steps, flux_precision, ic_kind, seed, &
            ic_param, wls_local
        if (command_argument_count() >= 1) then
            call get_command_argument(1, fname)
            open(newunit=u, file=trim(fname), status='old', action='read', iostat=ios)
            if (ios /= 0) stop 'cannot open the namelist file'
            read(u, nml=fdmamr, iostat=ios)
            if (ios /= 0) stop 'bad namelist /fdmamr/'
            close(u)
        end if
        cv = 1.d0/(gama*(gama-1.d0)*Ma*Ma)
        call execute_command_line('mkdir -p Output')
    end subroutine set_parameter
end program main
