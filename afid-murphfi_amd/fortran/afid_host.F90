!> Fortran host of the MI355X RK3 step: the reference's driver loop
!! (src/main.F90:281-422) and TimeMarcher call tree
!! (src/flow_solver/TimeMarcher.F90:13-165), with the arithmetic in
!! libafid_gpu through the iso_c_binding module afid_gpu.
!!
!! usage: afid_host [fused] [mixed] [seed=N]   (reads bou.in in the working directory)
!!   files: outputdir/continua_*.h5 at a normal end, outputdir/fields/ every SAVE_3D,
!!   outputdir/means.h5 every TOUT; NREAD = 1 restarts from outputdir/continua_*.h5
!!   fused: one afid_time_marcher call per step instead of the call tree
!!   mixed: the mixed FP32/FP64 form (runs through afid_time_marcher)
!!   moisture (src/moisture.F90) is on when humid.in exists, as in main.F90:38-39
module afid_host_state
    use iso_c_binding
    use afid_gpu
    implicit none
    type(afid_config) :: cfg
    type(c_ptr) :: ctx = c_null_ptr
    integer :: nsst, ntst, idtv, nread, ireset, flagMR, flagsal, flagPF, flagmelt, solidtype, pf_IC
    real(c_double) :: walltimemax, tmax, tout, tframe, save_3D, dt, resid, limitCFL, dtmin, dtmax
    real(c_double) :: pf_D, pf_A, pf_S, pf_Tm
    logical :: salinity = .false., fused = .false., mixed = .false., moist = .false.
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
                      pf_A, pf_S, pf_Tm, 0.d0, 0.d0, &
                      0, 0.d0, 0.d0, 0.d0, 0.d0, 0, 1, 1, 1.d0)
    if (mixed) cfg%precision = AFID_MIXED
    ! main.F90:38-39 and InitMoistVariables (moisture.F90:31-61)
    inquire(file='humid.in', exist=moist)
    if (moist) call ReadMoistParameters
end subroutine ReadInputFile

!> ReadMoistParameters, moisture.F90:64-117: the humid.in lines; dtmax is
!! capped at a tenth of the condensation time (InitMoistVariables, :57)
subroutine ReadMoistParameters
    use afid_host_state
    implicit none
    integer :: io, i
    logical :: ifdiff
    real(c_double) :: alpha_q, beta_q, gamma_q, tau_q, sm, pect
    integer :: qfixn, qfixs
    open(newunit=io, file='humid.in', status='old', action='read')
    do i = 1, 4
        read(io,*)
    end do
    read(io,*) alpha_q
    read(io,*); read(io,*)
    read(io,*) beta_q
    read(io,*); read(io,*)
    read(io,*) gamma_q
    read(io,*); read(io,*)
    read(io,*) tau_q
    read(io,*); read(io,*)
    read(io,*) ifdiff
    read(io,*); read(io,*); read(io,*)
    read(io,*) qfixn, qfixs
    read(io,*); read(io,*)
    read(io,*) sm
    close(io)
    cfg%moist = 1
    cfg%q_alpha = alpha_q; cfg%q_beta = beta_q; cfg%q_gamma = gamma_q; cfg%q_tau = tau_q
    cfg%q_ifdiff = merge(1, 0, ifdiff); cfg%q_fixn = qfixn; cfg%q_fixs = qfixs; cfg%q_sm = sm
    ! ReadInputFile.F90:101-122 (pect), moisture.F90:99 (tau in diffusive units)
    pect = sqrt(abs(cfg%rayt) / cfg%prat) * cfg%prat
    if (salinity .and. cfg%ffscales == 1) pect = sqrt(abs(cfg%rays) / cfg%pras) * cfg%prat
    if (ifdiff) tau_q = pect * tau_q
    dtmax = min(dtmax, 0.1d0 * tau_q)
end subroutine ReadMoistParameters

!> CreateInitialHumidity, moisture.F90:150-215 (no drizzle.h5): T = x (1-x)
!! times 1e-3 noise and a humid blob q = 5 exp(-r^2/0.005) around (0.1, Ly/2,
!! Lz/2) plus 1e-3 noise, then qsat (main.F90:202).  The reference draws the
!! noise from the system-seeded random_number; here it is seeded with seed.
subroutine CreateInitialHumidity(seed)
    use afid_host_state
    implicit none
    integer(c_int64_t), intent(in) :: seed
    real(c_double), allocatable :: t(:,:,:), q(:,:,:), xm(:)
    integer(c_int64_t) :: lb(3), ext(3), n
    integer :: ic, jc, kc, ns, i_
    integer, allocatable :: put(:)
    real(c_double) :: rnum, r2, ym, zm
    allocate(t(cfg%nxm + 1, cfg%nym, cfg%nzm), q(cfg%nxm + 1, cfg%nym, cfg%nzm), xm(cfg%nxm))
    n = int(cfg%nxm, c_int64_t)
    call afid_check(ctx, afid_grid_array(ctx, 'xm'//c_null_char, xm, n, n), 'grid xm')
    call random_seed(size=ns)
    allocate(put(ns))
    put = int(seed) + 7919 * [(i_, i_ = 1, ns)]
    call random_seed(put=put)
    t = 0.d0; q = 0.d0
    do ic = 1, cfg%nzm
        do jc = 1, cfg%nym
            do kc = 1, cfg%nxm
                call random_number(rnum)
                t(kc, jc, ic) = (1.d0 - xm(kc)) * xm(kc) * rnum * 1.d-3
            end do
        end do
    end do
    do ic = 1, cfg%nzm
        zm = (ic - 0.5d0) * cfg%zlen / cfg%nzm
        do jc = 1, cfg%nym
            ym = (jc - 0.5d0) * cfg%ylen / cfg%nym
            do kc = 1, cfg%nxm
                call random_number(rnum)
                r2 = (ym - 0.5d0 * cfg%ylen)**2 + (xm(kc) - 0.1d0)**2 + (zm - 0.5d0 * cfg%zlen)**2
                q(kc, jc, ic) = 5.d0 * exp(-r2 / 0.005d0) + 1.d-3 * rnum
            end do
        end do
    end do
    lb = [1_c_int64_t, 1_c_int64_t, 1_c_int64_t]
    ext = [int(cfg%nxm + 1, c_int64_t), int(cfg%nym, c_int64_t), int(cfg%nzm, c_int64_t)]
    call afid_check(ctx, afid_upload(ctx, AFID_TEMP, t, lb, ext), 'upload temp')
    call afid_check(ctx, afid_upload(ctx, AFID_HUMID, q, lb, ext), 'upload humid')
    call afid_check(ctx, afid_update_saturation(ctx), 'UpdateSaturation')
end subroutine CreateInitialHumidity

!> SetTempBCs.F90:18-59 (RBC and phase-field wall values) and SetSalBCs
!! (salinity.F90:105-122)
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
    ! SetTempBCs.F90:43-59: the phase-field case holds T at 1 and 0
    if (cfg%phasefield /= 0) then
        if (cfg%rayt >= 0) then
            tt = 0.d0; tb = 1.d0
        else
            tt = 1.d0; tb = 0.d0
        end if
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
    ! main.F90:155
    if (moist) call afid_check(ctx, afid_set_humidity_bcs(ctx), 'SetHumidityBCs')
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
        ! TimeMarcher.F90:66-77 (the salt terms of the interface belong to the
        ! salinity + phase-field case, which the library does not support)
        if (cfg%phasefield /= 0) then
            call afid_check(ctx, afid_explicit_phase(ctx), 'ExplicitPhase')
            call afid_check(ctx, afid_add_volume_penalty(ctx), 'AddVolumePenalty')
            call afid_check(ctx, afid_implicit_phase(ctx), 'ImplicitPhase')
            call afid_check(ctx, afid_add_latent_heat(ctx), 'AddLatentHeat')
        end if
        ! TimeMarcher.F90:79-81
        if (moist) then
            call afid_check(ctx, afid_explicit_humidity(ctx), 'ExplicitHumidity')
            call afid_check(ctx, afid_add_condensation(ctx), 'AddCondensation')
        end if
        ! with the phase field on, the velocity solves are the masked
        ! SolveImpEqnUpdate_X_pf / _YZ_pf (phasefield.F90:821-930)
        call afid_check(ctx, afid_implicit_and_update_vx(ctx), 'ImplicitAndUpdateVX')
        call afid_check(ctx, afid_implicit_and_update_vy(ctx), 'ImplicitAndUpdateVY')
        call afid_check(ctx, afid_implicit_and_update_vz(ctx), 'ImplicitAndUpdateVZ')
        call afid_check(ctx, afid_implicit_and_update_temp(ctx), 'ImplicitAndUpdateTemp')
        if (salinity) call afid_check(ctx, afid_implicit_salinity(ctx), 'ImplicitSalinity')
        if (moist) call afid_check(ctx, afid_implicit_humidity(ctx), 'ImplicitHumidity')     ! :97
        call afid_check(ctx, afid_calc_local_divergence(ctx), 'CalcLocalDivergence')
        call afid_check(ctx, afid_solve_pressure_correction(ctx), 'SolvePressureCorrection')
        call afid_check(ctx, afid_correct_velocity(ctx), 'CorrectVelocity')
        call afid_check(ctx, afid_correct_pressure(ctx), 'CorrectPressure')
        if (salinity) then
            call afid_check(ctx, afid_interp_vel_mgrd(ctx), 'InterpVelMgrd')
            call afid_check(ctx, afid_interp_sal_multigrid(ctx), 'InterpSalMultigrid')
        end if
        ! TimeMarcher.F90:147-152
        if (cfg%phasefield /= 0) then
            call afid_check(ctx, afid_interp_temp_multigrid(ctx), 'InterpTempMultigrid')
            call afid_check(ctx, afid_interp_phi_multigrid(ctx), 'InterpPhiMultigrid')
        end if
        if (moist) call afid_check(ctx, afid_update_saturation(ctx), 'UpdateSaturation')       ! :154
    end do
end subroutine TimeMarcher

!> wall-clock seconds (the reference times its loop with MPI_WTIME, main.F90:282,383)
function wtime() result(t)
    use iso_c_binding, only: c_double, c_int64_t
    implicit none
    real(c_double) :: t
    integer(c_int64_t) :: cnt, rate
    call system_clock(cnt, rate)
    t = real(cnt, c_double) / real(rate, c_double)
end function wtime

!> extent (x incl. the top point, y, z) of a base or refined module array
function field_extent(refined) result(ext)
    use afid_host_state
    implicit none
    logical, intent(in) :: refined
    integer(c_int64_t) :: ext(3)
    if (refined) then
        ext = [int(cfg%nxmr + 1, c_int64_t), int(cfg%nymr, c_int64_t), int(cfg%nzmr, c_int64_t)]
    else
        ext = [int(cfg%nxm + 1, c_int64_t), int(cfg%nym, c_int64_t), int(cfg%nzm, c_int64_t)]
    end if
end function field_extent

!> one module array to / from its HDF5 file (HdfWriteRealHalo3D,
!! HdfRoutines.F90:428-520 / HdfReadContinua, HdfReadContinua.F90:11-107)
subroutine field_file(field, refined, path, write_it)
    use afid_host_state
    use afid_io
    implicit none
    integer(c_int), intent(in) :: field
    logical, intent(in) :: refined, write_it
    character(len=*), intent(in) :: path
    interface
        function field_extent(refined) result(ext)
            import :: c_int64_t
            logical, intent(in) :: refined
            integer(c_int64_t) :: ext(3)
        end function
    end interface
    real(c_double), allocatable :: buf(:)
    integer(c_int64_t) :: ext(3), lb(3)
    lb = 1
    ext = field_extent(refined)
    allocate(buf(ext(1) * ext(2) * ext(3)))
    if (write_it) then
        call afid_check(ctx, afid_download(ctx, field, buf, lb, ext), 'download '//path)
        call h5_check(afid_h5_write_field3d(cstr(path), buf, ext(1), ext(2), ext(3)), 'write '//path)
    else
        call h5_check(afid_h5_read_field3d(cstr(path), buf, ext(1), ext(2), ext(3)), 'read '//path)
        call afid_check(ctx, afid_upload(ctx, field, buf, lb, ext), 'upload '//path)
    end if
end subroutine field_file

!> WriteFlowField.F90:11-86
subroutine WriteFlowField(final, time)
    use afid_host_state
    use afid_io
    implicit none
    logical, intent(in) :: final
    real(c_double), intent(in) :: time
    character(len=64) :: base, frame, master
    base = 'outputdir/continua'
    if (.not. final) then
        write(frame, '(i5.5)') nint(time / save_3D)
        base = 'outputdir/fields/' // trim(frame)
    end if
    call field_file(AFID_TEMP, .false., trim(base)//'_temp.h5', .true.)
    call field_file(AFID_VX, .false., trim(base)//'_vx.h5', .true.)
    call field_file(AFID_VY, .false., trim(base)//'_vy.h5', .true.)
    call field_file(AFID_VZ, .false., trim(base)//'_vz.h5', .true.)
    if (moist) call field_file(AFID_HUMID, .false., trim(base)//'_humid.h5', .true.)
    if (salinity) call field_file(AFID_SAL, .true., trim(base)//'_sal.h5', .true.)
    if (cfg%phasefield /= 0) call field_file(AFID_PHI, .true., trim(base)//'_phi.h5', .true.)
    if (.not. final) return
    call field_file(AFID_PR, .false., trim(base)//'_pr.h5', .true.)
    master = 'outputdir/continua_master.h5'
    call h5_check(afid_h5_create_file(cstr(master)), 'create '//master)
    call h5_check(afid_h5_write_int_scalar(cstr(master), cstr('nx'), cfg%nxm + 1), 'nx')
    call h5_check(afid_h5_write_int_scalar(cstr(master), cstr('ny'), cfg%nym + 1), 'ny')
    call h5_check(afid_h5_write_int_scalar(cstr(master), cstr('nz'), cfg%nzm + 1), 'nz')
    call h5_check(afid_h5_write_real_scalar(cstr(master), cstr('ylen'), cfg%ylen), 'ylen')
    call h5_check(afid_h5_write_real_scalar(cstr(master), cstr('zlen'), cfg%zlen), 'zlen')
    call h5_check(afid_h5_write_real_scalar(cstr(master), cstr('time'), time), 'time')
    call h5_check(afid_h5_write_int_scalar(cstr(master), cstr('istr3'), cfg%istr3), 'istr3')
    call h5_check(afid_h5_write_real_scalar(cstr(master), cstr('str3'), cfg%str3), 'str3')
    if (cfg%multires /= 0) then
        call h5_check(afid_h5_write_int_scalar(cstr(master), cstr('nxr'), cfg%nxmr + 1), 'nxr')
        call h5_check(afid_h5_write_int_scalar(cstr(master), cstr('nyr'), cfg%nymr + 1), 'nyr')
        call h5_check(afid_h5_write_int_scalar(cstr(master), cstr('nzr'), cfg%nzmr + 1), 'nzr')
        call h5_check(afid_h5_write_int_scalar(cstr(master), cstr('istr3r'), cfg%istr3r), 'istr3r')
    end if
end subroutine WriteFlowField

!> ReadFlowInterp.F90:1-179: the same grid reads the fields (:143-171);
!! another grid reads them with the old dims and interpolates them onto the
!! configured grids (:90-141: InterpInputVel, InterpInputSal, InterpInputPhi)
subroutine ReadFlowInterp(time)
    use afid_host_state
    use afid_io
    implicit none
    real(c_double), intent(out) :: time
    character(len=*), parameter :: master = 'outputdir/continua_master.h5'
    integer(c_int32_t) :: nxo, nyo, nzo, istr3o, nxro, nyro, nzro, istr3ro
    real(c_double) :: str3o, yleno, zleno
    logical :: newgrid, ex
    type(afid_old_grid) :: og
    real(c_double), allocatable :: vxo(:), vyo(:), vzo(:), tempo(:), fo(:)
    integer(c_int64_t) :: n3, m3
    call h5_check(afid_h5_read_int_scalar(cstr(master), cstr('nx'), nxo), 'read nx')
    call h5_check(afid_h5_read_int_scalar(cstr(master), cstr('ny'), nyo), 'read ny')
    call h5_check(afid_h5_read_int_scalar(cstr(master), cstr('nz'), nzo), 'read nz')
    call h5_check(afid_h5_read_real_scalar(cstr(master), cstr('time'), time), 'read time')
    call h5_check(afid_h5_read_int_scalar(cstr(master), cstr('istr3'), istr3o), 'read istr3')
    call h5_check(afid_h5_read_real_scalar(cstr(master), cstr('str3'), str3o), 'read str3')
    call h5_check(afid_h5_read_real_scalar(cstr(master), cstr('ylen'), yleno), 'read ylen')
    call h5_check(afid_h5_read_real_scalar(cstr(master), cstr('zlen'), zleno), 'read zlen')
    ! ReadFlowInterp.F90:76-81: the domain lengths are checked first, for
    ! the same grid as for a new one
    if (abs(cfg%ylen - yleno) > 1d-8 .or. abs(cfg%zlen - zleno) > 1d-8) then
        write(0, '(a)') 'Continua domain size does not match bou.in'
        write(0, '(a,2es16.8)') 'old Ly, Lz: ', yleno, zleno
        write(0, '(a,2es16.8)') 'current Ly, Lz: ', cfg%ylen, cfg%zlen
        error stop 1
    end if
    nxro = 0; nyro = 0; nzro = 0; istr3ro = 0
    newgrid = (cfg%nxm + 1 /= nxo) .or. (cfg%nym + 1 /= nyo) .or. (cfg%nzm + 1 /= nzo) .or. &
              (cfg%istr3 /= istr3o) .or. (abs(cfg%str3 - str3o) > 1d-8)
    if (cfg%multires /= 0) then
        call h5_check(afid_h5_read_int_scalar(cstr(master), cstr('nxr'), nxro), 'read nxr')
        call h5_check(afid_h5_read_int_scalar(cstr(master), cstr('nyr'), nyro), 'read nyr')
        call h5_check(afid_h5_read_int_scalar(cstr(master), cstr('nzr'), nzro), 'read nzr')
        call h5_check(afid_h5_read_int_scalar(cstr(master), cstr('istr3r'), istr3ro), 'read istr3r')
        newgrid = newgrid .or. (cfg%nxmr + 1 /= nxro) .or. (cfg%nymr + 1 /= nyro) .or. &
                  (cfg%nzmr + 1 /= nzro) .or. (cfg%istr3r /= istr3ro)
    end if
    if (newgrid) then
        ! :90-141
        og = afid_old_grid(nxo, nyo, nzo, istr3o, str3o, yleno, zleno, nxro, nyro, nzro, istr3ro)
        n3 = int(nxo, c_int64_t) * (nyo - 1) * (nzo - 1)
        allocate(vxo(n3), vyo(n3), vzo(n3), tempo(n3))
        call h5_check(afid_h5_read_field3d(cstr('outputdir/continua_vx.h5'), vxo, int(nxo, c_int64_t), &
                      int(nyo - 1, c_int64_t), int(nzo - 1, c_int64_t)), 'read vx')
        call h5_check(afid_h5_read_field3d(cstr('outputdir/continua_vy.h5'), vyo, int(nxo, c_int64_t), &
                      int(nyo - 1, c_int64_t), int(nzo - 1, c_int64_t)), 'read vy')
        call h5_check(afid_h5_read_field3d(cstr('outputdir/continua_vz.h5'), vzo, int(nxo, c_int64_t), &
                      int(nyo - 1, c_int64_t), int(nzo - 1, c_int64_t)), 'read vz')
        call h5_check(afid_h5_read_field3d(cstr('outputdir/continua_temp.h5'), tempo, int(nxo, c_int64_t), &
                      int(nyo - 1, c_int64_t), int(nzo - 1, c_int64_t)), 'read temp')
        call afid_check(ctx, afid_interp_input_vel(ctx, og, vxo, vyo, vzo, tempo), 'InterpInputVel')
        if (moist) then                 ! ReadFlowInterp.F90:133-135, InterpInputHum
            call h5_check(afid_h5_read_field3d(cstr('outputdir/continua_humid.h5'), tempo, int(nxo, c_int64_t), &
                          int(nyo - 1, c_int64_t), int(nzo - 1, c_int64_t)), 'read humid')
            call afid_check(ctx, afid_interp_input_humid(ctx, og, tempo), 'InterpInputHum')
        end if
        deallocate(vxo, vyo, vzo, tempo)
        if (salinity .or. cfg%phasefield /= 0) then
            m3 = int(nxro, c_int64_t) * (nyro - 1) * (nzro - 1)
            allocate(fo(m3))
            if (salinity) then
                call h5_check(afid_h5_read_field3d(cstr('outputdir/continua_sal.h5'), fo, int(nxro, c_int64_t), &
                              int(nyro - 1, c_int64_t), int(nzro - 1, c_int64_t)), 'read sal')
                call afid_check(ctx, afid_interp_input_sal(ctx, og, fo), 'InterpInputSal')
            end if
            if (cfg%phasefield /= 0) then
                call h5_check(afid_h5_read_field3d(cstr('outputdir/continua_phi.h5'), fo, int(nxro, c_int64_t), &
                              int(nyro - 1, c_int64_t), int(nzro - 1, c_int64_t)), 'read phi')
                call afid_check(ctx, afid_interp_input_phi(ctx, og, fo), 'InterpInputPhi')
            end if
            deallocate(fo)
        end if
        if (ireset /= 0) time = 0.d0
        return
    end if
    call field_file(AFID_VX, .false., 'outputdir/continua_vx.h5', .false.)
    call field_file(AFID_VY, .false., 'outputdir/continua_vy.h5', .false.)
    call field_file(AFID_VZ, .false., 'outputdir/continua_vz.h5', .false.)
    call field_file(AFID_TEMP, .false., 'outputdir/continua_temp.h5', .false.)
    if (moist) call field_file(AFID_HUMID, .false., 'outputdir/continua_humid.h5', .false.)
    if (salinity) call field_file(AFID_SAL, .true., 'outputdir/continua_sal.h5', .false.)
    if (cfg%phasefield /= 0) call field_file(AFID_PHI, .true., 'outputdir/continua_phi.h5', .false.)
    inquire(file='outputdir/continua_pr.h5', exist=ex)
    if (ex) call field_file(AFID_PR, .false., 'outputdir/continua_pr.h5', .false.)
    if (ireset /= 0) time = 0.d0
end subroutine ReadFlowInterp

!> the file part of CalcMeanProfiles (CalcMeanProfiles.F90:209-283), CalcSalStats
!! (salinity.F90:588-603) and CalcPhiStats (phasefield.F90:786-793)
subroutine WriteMeans(time)
    use afid_host_state
    use afid_io
    implicit none
    real(c_double), intent(in) :: time
    character(len=*), parameter :: path = 'outputdir/means.h5'
    character(len=7), parameter :: mnames(15) = [character(len=7) :: 'Tbar', 'vybar', 'vzbar', 'vxT', 'vyT', &
        'vzT', 'Trms', 'vxrms', 'vyrms', 'vzrms', 'vxvy', 'vxvz', 'vyvz', 'chiT', 'epsilon']
    character(len=7), parameter :: snames(6) = [character(len=7) :: 'Sbar', 'Srms', 'vxS', 'vyS', 'vzS', 'chiS']
    character(len=7), parameter :: pnames(2) = [character(len=7) :: 'phibar', 'phirms']
    character(len=7), parameter :: qnames(7) = [character(len=7) :: 'qbar', 'qsbar', 'qrms', 'qrel', 'vxq', 'vyq', 'vzq']
    real(c_double), allocatable :: buf(:)
    character(len=5) :: nstat
    logical :: ex
    integer :: q
    integer(c_int64_t) :: n
    inquire(file=path, exist=ex)
    if (.not. ex) then
        call h5_check(afid_h5_create_file(cstr(path)), 'create '//path)
        do q = 1, 15
            call h5_check(afid_h5_create_group(cstr(path), cstr(mnames(q))), 'group')
        end do
        if (salinity) then
            do q = 1, 6
                call h5_check(afid_h5_create_group(cstr(path), cstr(snames(q))), 'group')
            end do
        end if
        if (cfg%phasefield /= 0) then
            do q = 1, 2
                call h5_check(afid_h5_create_group(cstr(path), cstr(pnames(q))), 'group')
            end do
        end if
        if (moist) then                   ! CreateMoistH5Groups, moisture.F90:509-535
            do q = 1, 7
                call h5_check(afid_h5_create_group(cstr(path), cstr(qnames(q))), 'group')
            end do
        end if
    end if
    write(nstat, '(i5.5)') nint(time / tout)
    n = cfg%nxm
    allocate(buf(15 * n))
    call afid_check(ctx, afid_calc_mean_profiles(ctx, buf), 'CalcMeanProfiles')
    do q = 1, 15
        call h5_check(afid_h5_write_real1d(cstr(path), cstr(trim(mnames(q))//'/'//nstat), buf((q - 1) * n + 1), n), &
                      'write '//mnames(q))
    end do
    deallocate(buf)
    if (moist) then                       ! CalcMoistStats, moisture.F90:485-500
        allocate(buf(7 * n))
        call afid_check(ctx, afid_calc_moist_stats(ctx, buf), 'CalcMoistStats')
        do q = 1, 7
            call h5_check(afid_h5_write_real1d(cstr(path), cstr(trim(qnames(q))//'/'//nstat), buf((q - 1) * n + 1), n), &
                          'write '//qnames(q))
        end do
        deallocate(buf)
    end if
    n = cfg%nxmr
    if (salinity) then
        allocate(buf(6 * n))
        call afid_check(ctx, afid_calc_sal_stats(ctx, buf), 'CalcSalStats')
        do q = 1, 6
            call h5_check(afid_h5_write_real1d(cstr(path), cstr(trim(snames(q))//'/'//nstat), buf((q - 1) * n + 1), n), &
                          'write '//snames(q))
        end do
        deallocate(buf)
    end if
    if (cfg%phasefield /= 0) then
        allocate(buf(2 * n))
        call afid_check(ctx, afid_calc_phi_stats(ctx, buf), 'CalcPhiStats')
        do q = 1, 2
            call h5_check(afid_h5_write_real1d(cstr(path), cstr(trim(pnames(q))//'/'//nstat), buf((q - 1) * n + 1), n), &
                          'write '//pnames(q))
        end do
    end if
end subroutine WriteMeans

program afid_host
    use afid_host_state
    implicit none
    integer :: ntime, errorcode, narg, a
    integer(c_int64_t) :: seed
    real(c_double) :: time, instCFL, CFLmr, dmax, dmaxr, q(4)
    real(c_double) :: t0, t1, tin, minwtdt
    character(len=64) :: arg
    integer(c_int) :: rc
    real(c_double), external :: wtime

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
    ! main.F90:77, before anything touches the GPU
    call execute_command_line('mkdir -p outputdir/fields')
    rc = afid_gpu_init(cfg, ctx)
    call afid_check(ctx, rc, 'afid_gpu_init')
    call SetBCs
    if (nread /= 0) then
        ! main.F90:156-161, then the start-up transfers of :205-213
        call ReadFlowInterp(time)
        if (salinity) then
            call afid_check(ctx, afid_interp_vel_mgrd(ctx), 'InterpVelMgrd')
            call afid_check(ctx, afid_interp_sal_multigrid(ctx), 'InterpSalMultigrid')
        end if
        if (cfg%phasefield /= 0) then
            call afid_check(ctx, afid_interp_temp_multigrid(ctx), 'InterpTempMultigrid')
            call afid_check(ctx, afid_interp_phi_multigrid(ctx), 'InterpPhiMultigrid')
        end if
        if (moist) call afid_check(ctx, afid_update_saturation(ctx), 'UpdateSaturation')     ! main.F90:202
    else
        time = 0.d0
        call afid_check(ctx, afid_create_initial_conditions(ctx, seed, 1.d-3), 'CreateInitialConditions')
        if (moist) call CreateInitialHumidity(seed)        ! main.F90:174
    end if
    ! main.F90:227 and :241
    call WriteMeans(time)
    call WriteFlowField(.false., time)

    tin = wtime()
    minwtdt = huge(0.d0)
    errorcode = 0
    dmax = 0.d0; dmaxr = 0.d0
    ! main.F90:281-422
    do ntime = 0, ntst
        t0 = wtime()
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
        ! main.F90:315-320
        if (mod(time, tout) < dt) call WriteMeans(time)

        if (ntime == 1 .or. mod(time, tout) < dt) then
            call afid_check(ctx, afid_check_divergence(ctx, dmax, dmaxr), 'CheckDivergence')
            if (abs(dmax) > resid) errorcode = 169
        end if
        if (time > tmax) errorcode = 333
        ! main.F90:376-380
        if (mod(time, save_3D) < dt .and. floor(time / tframe) /= 0) call WriteFlowField(.false., time)
        call afid_check(ctx, afid_synchronize(ctx), 'synchronize')
        t1 = wtime()
        minwtdt = min(minwtdt, t1 - t0)
        if (t1 - tin > walltimemax) errorcode = 334
        if (ntime == ntst) errorcode = 555
        if (errorcode /= 0) exit
    end do

    ! QuitRoutine.F90:34-38: a normal end (333, 334, 555) writes the checkpoint
    if (errorcode == 333 .or. errorcode == 334 .or. errorcode == 555) call WriteFlowField(.true., time)
    call afid_check(ctx, afid_check_divergence(ctx, dmax, dmaxr), 'CheckDivergence')
    call afid_check(ctx, afid_global_quantities(ctx, q), 'GlobalQuantities')
    write(*, '(a,i0,a,i0)') 'steps=', ntime + 1, ' errorcode=', errorcode
    write(*, '(a,es24.16)') 'time=', time
    write(*, '(a,es24.16,a,es24.16)') 'maxdiv=', dmax, ' maxdivr=', dmaxr
    write(*, '(a,es24.16,a,es24.16)') 'nu=', q(1), ' ke=', q(2)
    write(*, '(a,f10.4,a)') 'Minimum Iteration Time = ', minwtdt, ' sec.'
    rc = afid_gpu_finalize(ctx)
end program afid_host
