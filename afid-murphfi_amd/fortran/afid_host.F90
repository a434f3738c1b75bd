!> Fortran host of the MI355X RK3 step: the reference's driver loop
!! (src/main.F90:281-422) and TimeMarcher call tree
!! (src/flow_solver/TimeMarcher.F90:13-165), with the arithmetic in
!! libafid_gpu through the iso_c_binding module afid_gpu.
!!
!! usage: afid_host [fused] [mixed] [seed=N]   (reads bou.in in the working directory)
!!   fused: one afid_time_marcher call per step instead of the call tree
!!   mixed: the mixed FP32/FP64 form (runs through afid_time_marcher)
module afid_host_state
    use iso_c_binding
    use afid_gpu
    implicit none
    type(afid_config) :: cfg
    type(c_ptr) :: ctx = c_null_ptr
    integer :: nsst, ntst, idtv, nread, ireset, flagMR, flagsal, flagPF, flagmelt, solidtype, pf_IC
    real(c_double) :: walltimemax, tmax, tout, tframe, save_3D, dt, resid, limitCFL, dtmin, dtmax
    real(c_double) :: pf_D, pf_A, pf_S, pf_Tm
    logical :: salinity = .false., fused = .false., mixed = .false.
end module afid_host_state

!> ReadInputFile.F90:20-76 (same line layout)
subroutine ReadInputFile
    use afid_host_state
    implicit none
    character(len=4) :: dummy
    integer :: io, nxm, nym, nzm, nxmr, nymr, nzmr, istr3, istr3r, FFscaleS
    integer :: inslws, inslwn, TfixS, TfixN, SfixS, SfixN, active_T, active_S, gAxis
    real(c_double) :: alx3, ylen, zlen, str3, rayt, prat, rays, pras, xplusU, xminusU, dPdy, dPdz
    open(newunit=io, file='bou.in', status='old')
    read(io,301) dummy; read(io,301) dummy
    read(io,*) nxm, nym, nzm, nsst
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) flagMR, nxmr, nymr, nzmr
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) flagsal, flagPF
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) nread, ireset
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) ntst, walltimemax, tmax
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) tout, tframe, save_3D
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) alx3, ylen, zlen
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) istr3, str3, istr3r
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) rayt, prat, rays, pras, FFscaleS
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) idtv, dt, resid, limitCFL, dtmin, dtmax
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) inslws, inslwn, TfixS, TfixN, SfixS, SfixN
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) active_T, active_S, gAxis
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) xplusU, xminusU, dPdy, dPdz, flagmelt
    read(io,301) dummy; read(io,301) dummy; read(io,301) dummy
    read(io,*) pf_D, pf_A, pf_S, pf_Tm, solidtype, pf_IC
301 format(a4)
    close(io)
    salinity = flagsal /= 0
    cfg = afid_config(nxm, nym, nzm, nsst, flagMR, nxmr, nymr, nzmr, flagsal, flagPF, &
                      alx3, ylen, zlen, istr3, str3, istr3r, rayt, prat, rays, pras, FFscaleS, &
                      inslws, inslwn, TfixS, TfixN, SfixS, SfixN, active_T, active_S, gAxis, &
                      xplusU, xminusU, dPdy, dPdz, AFID_FP64, 0.d0, 0, &
                      pf_A, pf_S, pf_Tm, 0.d0, 0.d0)
    if (mixed) cfg%precision = AFID_MIXED
end subroutine ReadInputFile

!> SetTempBCs.F90:18-40 and SetSalBCs (salinity.F90:105-122) for the RBC case
subroutine SetBCs
    use afid_host_state
    implicit none
    real(c_double), allocatable :: plane(:,:,:)
    integer(c_int64_t) :: lb(3), ext(3)
    real(c_double) :: tb, tt, sb, st
    if (cfg%rayt >= 0) then
        if (cfg%inslwn == 0) then
            tt = 0.d0; tb = 1.d0
        else
            tt = -0.5d0; tb = 0.5d0
        end if
    else
        tt = 0.5d0; tb = -0.5d0
    end if
    lb = [1_c_int64_t, 1_c_int64_t, 1_c_int64_t]
    ext = [1_c_int64_t, int(cfg%nym, c_int64_t), int(cfg%nzm, c_int64_t)]
    allocate(plane(1, cfg%nym, cfg%nzm))
    plane = tb
    call afid_check(ctx, afid_upload(ctx, AFID_TEMPBP, plane, lb, ext), 'upload tempbp')
    plane = tt
    call afid_check(ctx, afid_upload(ctx, AFID_TEMPTP, plane, lb, ext), 'upload temptp')
    deallocate(plane)
    if (salinity) then
        if (cfg%rays >= 0) then
            st = 0.5d0; sb = -0.5d0
        else
            st = -0.5d0; sb = 0.5d0
        end if
        ext = [1_c_int64_t, int(cfg%nymr, c_int64_t), int(cfg%nzmr, c_int64_t)]
        allocate(plane(1, cfg%nymr, cfg%nzmr))
        plane = sb
        call afid_check(ctx, afid_upload(ctx, AFID_SALBP, plane, lb, ext), 'upload salbp')
        plane = st
        call afid_check(ctx, afid_upload(ctx, AFID_SALTP, plane, lb, ext), 'upload saltp')
    end if
end subroutine SetBCs

!> TimeMarcher.F90:13-165: the call tree of one RK3 step, one C-ABI call per
!! reference routine (halo updates and transposes are internal to the library)
subroutine TimeMarcher
    use afid_host_state
    implicit none
    integer(c_int) :: ns
    if (fused) then
        call afid_check(ctx, afid_time_marcher(ctx, dt), 'TimeMarcher')
        return
    end if
    do ns = 1, nsst
        call afid_check(ctx, afid_set_substep(ctx, dt, ns), 'set_substep')
        call afid_check(ctx, afid_explicit_terms_vx(ctx), 'ExplicitTermsVX')
        call afid_check(ctx, afid_explicit_terms_vy(ctx), 'ExplicitTermsVY')
        call afid_check(ctx, afid_explicit_terms_vz(ctx), 'ExplicitTermsVZ')
        call afid_check(ctx, afid_explicit_terms_temp(ctx), 'ExplicitTermsTemp')
        if (salinity) then
            call afid_check(ctx, afid_explicit_salinity(ctx), 'ExplicitSalinity')
            if (cfg%active_s == 1) call afid_check(ctx, afid_add_sal_buoyancy(ctx), 'AddSalBuoyancy')
        end if
        call afid_check(ctx, afid_implicit_and_update_vx(ctx), 'ImplicitAndUpdateVX')
        call afid_check(ctx, afid_implicit_and_update_vy(ctx), 'ImplicitAndUpdateVY')
        call afid_check(ctx, afid_implicit_and_update_vz(ctx), 'ImplicitAndUpdateVZ')
        call afid_check(ctx, afid_implicit_and_update_temp(ctx), 'ImplicitAndUpdateTemp')
        if (salinity) call afid_check(ctx, afid_implicit_salinity(ctx), 'ImplicitSalinity')
        call afid_check(ctx, afid_calc_local_divergence(ctx), 'CalcLocalDivergence')
        call afid_check(ctx, afid_solve_pressure_correction(ctx), 'SolvePressureCorrection')
        call afid_check(ctx, afid_correct_velocity(ctx), 'CorrectVelocity')
        call afid_check(ctx, afid_correct_pressure(ctx), 'CorrectPressure')
        if (salinity) then
            call afid_check(ctx, afid_interp_vel_mgrd(ctx), 'InterpVelMgrd')
            call afid_check(ctx, afid_interp_sal_multigrid(ctx), 'InterpSalMultigrid')
        end if
    end do
end subroutine TimeMarcher

program afid_host
    use afid_host_state
    implicit none
    integer :: ntime, errorcode, narg, a
    integer(c_int64_t) :: seed
    real(c_double) :: time, instCFL, CFLmr, dmax, dmaxr, q(4)
    real(c_double) :: t0, t1, tin, minwtdt
    character(len=64) :: arg
    integer(c_int) :: rc

    seed = 1
    narg = command_argument_count()
    do a = 1, narg
        call get_command_argument(a, arg)
        if (trim(arg) == 'fused') fused = .true.
        if (trim(arg) == 'mixed') then
            mixed = .true.; fused = .true.
        end if
        if (arg(1:5) == 'seed=') read(arg(6:), *) seed
    end do

    call ReadInputFile
    rc = afid_gpu_init(cfg, ctx)
    call afid_check(ctx, rc, 'afid_gpu_init')
    call SetBCs
    call afid_check(ctx, afid_create_initial_conditions(ctx, seed, 1.d-3), 'CreateInitialConditions')

    call cpu_time(tin)
    time = 0.d0
    minwtdt = huge(0.d0)
    errorcode = 0
    dmax = 0.d0; dmaxr = 0.d0
    ! main.F90:281-422
    do ntime = 0, ntst
        call cpu_time(t0)
        call afid_check(ctx, afid_calc_max_cfl(ctx, instCFL, CFLmr), 'CalcMaxCFL')
        if (idtv /= 0) then
            if (ntime > 1) then
                if (CFLmr < 1.0d-8) then
                    dt = dtmax
                else
                    dt = limitCFL / CFLmr
                end if
                if (dt > dtmax) dt = dtmax
            end if
            if (dt < dtmin) errorcode = 166
        else
            CFLmr = CFLmr * dt
            if (CFLmr > limitCFL) errorcode = 165
        end if

        call TimeMarcher
        time = time + dt

        if (ntime == 1 .or. mod(time, tout) < dt) then
            call afid_check(ctx, afid_check_divergence(ctx, dmax, dmaxr), 'CheckDivergence')
            if (abs(dmax) > resid) errorcode = 169
        end if
        if (time > tmax) errorcode = 333
        call afid_check(ctx, afid_synchronize(ctx), 'synchronize')
        call cpu_time(t1)
        minwtdt = min(minwtdt, t1 - t0)
        if (t1 - tin > walltimemax) errorcode = 334
        if (ntime == ntst) errorcode = 555
        if (errorcode /= 0) exit
    end do

    call afid_check(ctx, afid_check_divergence(ctx, dmax, dmaxr), 'CheckDivergence')
    call afid_check(ctx, afid_global_quantities(ctx, q), 'GlobalQuantities')
    write(*, '(a,i0,a,i0)') 'steps=', ntime + 1, ' errorcode=', errorcode
    write(*, '(a,es24.16)') 'time=', time
    write(*, '(a,es24.16,a,es24.16)') 'maxdiv=', dmax, ' maxdivr=', dmaxr
    write(*, '(a,es24.16,a,es24.16)') 'nu=', q(1), ' ke=', q(2)
    write(*, '(a,f10.4,a)') 'Minimum Iteration Time = ', minwtdt, ' sec.'
    rc = afid_gpu_finalize(ctx)
end program afid_host
