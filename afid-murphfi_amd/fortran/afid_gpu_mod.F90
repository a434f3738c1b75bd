!> iso_c_binding interface of libafid_gpu (include/afid_gpu.h).
!! Each binding carries the name of the reference routine it stands for.
module afid_gpu
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: AFID_FP64 = 0, AFID_MIXED = 1
    integer(c_int), parameter :: AFID_VX = 0, AFID_VY = 1, AFID_VZ = 2, AFID_TEMP = 3, AFID_PR = 4
    integer(c_int), parameter :: AFID_SAL = 14, AFID_VXR = 18, AFID_VYR = 19, AFID_VZR = 20, AFID_SALC = 21
    integer(c_int), parameter :: AFID_TEMPBP = 22, AFID_TEMPTP = 23, AFID_SALBP = 24, AFID_SALTP = 25
    integer(c_int), parameter :: AFID_PHI = 26
    integer(c_int), parameter :: AFID_HUMID = 31, AFID_RUHUMID = 32, AFID_HHUMID = 33, AFID_QSAT = 34
    integer(c_int), parameter :: AFID_HUMBP = 35, AFID_HUMTP = 36

    !> the bou.in value lines (ReadInputFile.F90:20-76) + arithmetic form
    type, bind(C) :: afid_config
        integer(c_int32_t) :: nxm, nym, nzm, nsst
        integer(c_int32_t) :: multires, nxmr, nymr, nzmr
        integer(c_int32_t) :: salinity, phasefield
        real(c_double)     :: alx3, ylen, zlen
        integer(c_int32_t) :: istr3
        real(c_double)     :: str3
        integer(c_int32_t) :: istr3r
        real(c_double)     :: rayt, prat, rays, pras
        integer(c_int32_t) :: ffscales
        integer(c_int32_t) :: inslws, inslwn, tfixs, tfixn, sfixs, sfixn
        integer(c_int32_t) :: active_t, active_s, gaxis
        real(c_double)     :: xplusu, xminusu, dpdy, dpdz
        integer(c_int32_t) :: precision
        real(c_double)     :: div_tol
        integer(c_int32_t) :: max_sweeps
        real(c_double)     :: pf_a, pf_s, pf_tm, pf_eps, pf_r0
        !> moisture (moisture.F90, humid.in): on/off, alpha_q, beta_q,
        !! gamma_q, tau_q, ifdiff, qfixN, qfixS, Sm
        integer(c_int32_t) :: moist
        real(c_double)     :: q_alpha, q_beta, q_gamma, q_tau
        integer(c_int32_t) :: q_ifdiff, q_fixn, q_fixs
        real(c_double)     :: q_sm
    end type afid_config

    !> batched 2-D ensemble (include/afid_ensemble.h): the time-stepping line
    !! of each run's bou.in and its NTST / TMAX (main.F90:287-302, 355-385)
    type, bind(C) :: afid_run_control
        integer(c_int32_t) :: variable_dt
        real(c_double)     :: dt, resid, limit_cfl, dtmin, dtmax
        integer(c_int32_t) :: ntst
        real(c_double)     :: tmax
    end type afid_run_control

    type, bind(C) :: afid_ens_tuning
        integer(c_int32_t) :: threads, workgroups, lines, phased
        real(c_double)     :: ms_per_step
    end type afid_ens_tuning

    !> the old run's grid of a checkpoint (continua_master.h5; sizes count the faces)
    type, bind(C) :: afid_old_grid
        integer(c_int32_t) :: nxo, nyo, nzo, istr3o
        real(c_double)     :: str3o, ylen, zlen
        integer(c_int32_t) :: nxro, nyro, nzro, istr3ro
    end type afid_old_grid

    type, bind(C) :: afid_run_state
        real(c_double)     :: time, dt, cfl, max_div
        integer(c_int32_t) :: ntime, status, steps
    end type afid_run_state

    interface
        integer(c_int) function afid_gpu_init(cfg, ctx) bind(C, name='afid_gpu_init')
            import; type(afid_config), intent(in) :: cfg; type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function afid_gpu_finalize(ctx) bind(C, name='afid_gpu_finalize')
            import; type(c_ptr), value :: ctx
        end function
        type(c_ptr) function afid_last_error(ctx) bind(C, name='afid_last_error')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_upload(ctx, field, host, lb, ext) bind(C, name='afid_upload')
            import; type(c_ptr), value :: ctx; integer(c_int), value :: field
            real(c_double), intent(in) :: host(*); integer(c_int64_t), intent(in) :: lb(3), ext(3)
        end function
        integer(c_int) function afid_download(ctx, field, host, lb, ext) bind(C, name='afid_download')
            import; type(c_ptr), value :: ctx; integer(c_int), value :: field
            real(c_double), intent(out) :: host(*); integer(c_int64_t), intent(in) :: lb(3), ext(3)
        end function
        !> TimeMarcher.F90:28,38-40
        integer(c_int) function afid_set_substep(ctx, dt, ns) bind(C, name='afid_set_substep')
            import; type(c_ptr), value :: ctx; real(c_double), value :: dt; integer(c_int), value :: ns
        end function
        integer(c_int) function afid_explicit_terms_vx(ctx) bind(C, name='afid_explicit_terms_vx')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_explicit_terms_vy(ctx) bind(C, name='afid_explicit_terms_vy')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_explicit_terms_vz(ctx) bind(C, name='afid_explicit_terms_vz')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_explicit_terms_temp(ctx) bind(C, name='afid_explicit_terms_temp')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_explicit_salinity(ctx) bind(C, name='afid_explicit_salinity')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_add_sal_buoyancy(ctx) bind(C, name='afid_add_sal_buoyancy')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_and_update_vx(ctx) bind(C, name='afid_implicit_and_update_vx')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_and_update_vy(ctx) bind(C, name='afid_implicit_and_update_vy')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_and_update_vz(ctx) bind(C, name='afid_implicit_and_update_vz')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_and_update_temp(ctx) bind(C, name='afid_implicit_and_update_temp')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_salinity(ctx) bind(C, name='afid_implicit_salinity')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_calc_local_divergence(ctx) bind(C, name='afid_calc_local_divergence')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_solve_pressure_correction(ctx) bind(C, name='afid_solve_pressure_correction')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_correct_velocity(ctx) bind(C, name='afid_correct_velocity')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_correct_pressure(ctx) bind(C, name='afid_correct_pressure')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_interp_vel_mgrd(ctx) bind(C, name='afid_interp_vel_mgrd')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_interp_sal_multigrid(ctx) bind(C, name='afid_interp_sal_multigrid')
            import; type(c_ptr), value :: ctx
        end function
        !> InterpTempMultigrid / InterpPhiMultigrid (phasefield.F90)
        ! restart onto a new grid (ReadFlowInterp.F90:90-141): InterpInputVel,
        ! InterpInputSal, InterpInputPhi; old arrays of the checkpoint's dims
        integer(c_int) function afid_interp_input_vel(ctx, og, vx, vy, vz, temp) bind(C, name='afid_interp_input_vel')
            import; type(c_ptr), value :: ctx; type(afid_old_grid), intent(in) :: og
            real(c_double), intent(in) :: vx(*), vy(*), vz(*), temp(*)
        end function
        integer(c_int) function afid_interp_input_sal(ctx, og, sal) bind(C, name='afid_interp_input_sal')
            import; type(c_ptr), value :: ctx; type(afid_old_grid), intent(in) :: og; real(c_double), intent(in) :: sal(*)
        end function
        integer(c_int) function afid_interp_input_phi(ctx, og, phi) bind(C, name='afid_interp_input_phi')
            import; type(c_ptr), value :: ctx; type(afid_old_grid), intent(in) :: og; real(c_double), intent(in) :: phi(*)
        end function
        ! phase field (phasefield.F90:345-725)
        integer(c_int) function afid_explicit_phase(ctx) bind(C, name='afid_explicit_phase')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_add_volume_penalty(ctx) bind(C, name='afid_add_volume_penalty')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_phase(ctx) bind(C, name='afid_implicit_phase')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_add_latent_heat(ctx) bind(C, name='afid_add_latent_heat')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_interp_temp_multigrid(ctx) bind(C, name='afid_interp_temp_multigrid')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_interp_phi_multigrid(ctx) bind(C, name='afid_interp_phi_multigrid')
            import; type(c_ptr), value :: ctx
        end function
        ! moisture.F90
        integer(c_int) function afid_set_humidity_bcs(ctx) bind(C, name='afid_set_humidity_bcs')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_explicit_humidity(ctx) bind(C, name='afid_explicit_humidity')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_add_condensation(ctx) bind(C, name='afid_add_condensation')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_humidity(ctx) bind(C, name='afid_implicit_humidity')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_update_saturation(ctx) bind(C, name='afid_update_saturation')
            import; type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_calc_moist_stats(ctx, out) bind(C, name='afid_calc_moist_stats')
            import; type(c_ptr), value :: ctx; real(c_double) :: out(*)
        end function
        integer(c_int) function afid_interp_input_humid(ctx, og, humid) bind(C, name='afid_interp_input_humid')
            import; type(c_ptr), value :: ctx; type(afid_old_grid), intent(in) :: og; real(c_double), intent(in) :: humid(*)
        end function
        integer(c_int) function afid_grid_array(ctx, name, out, n, count) bind(C, name='afid_grid_array')
            import; type(c_ptr), value :: ctx; character(kind=c_char), intent(in) :: name(*)
            real(c_double) :: out(*); integer(c_int64_t), value :: n; integer(c_int64_t) :: count
        end function
        integer(c_int) function afid_time_marcher(ctx, dt) bind(C, name='afid_time_marcher')
            import; type(c_ptr), value :: ctx; real(c_double), value :: dt
        end function
        integer(c_int) function afid_calc_max_cfl(ctx, cflm, cflmr) bind(C, name='afid_calc_max_cfl')
            import; type(c_ptr), value :: ctx; real(c_double), intent(out) :: cflm, cflmr
        end function
        integer(c_int) function afid_check_divergence(ctx, qmax, qmaxr) bind(C, name='afid_check_divergence')
            import; type(c_ptr), value :: ctx; real(c_double), intent(out) :: qmax, qmaxr
        end function
        integer(c_int) function afid_global_quantities(ctx, out) bind(C, name='afid_global_quantities')
            import; type(c_ptr), value :: ctx; real(c_double), intent(out) :: out(4)
        end function
        ! CalcMeanProfiles: out(nxm, 15) in the order of means.h5 (Tbar, vybar, ..., chiT, epsilon)
        integer(c_int) function afid_calc_mean_profiles(ctx, out) bind(C, name='afid_calc_mean_profiles')
            import; type(c_ptr), value :: ctx; real(c_double), intent(out) :: out(*)
        end function
        ! CalcSalStats: out(nxmr, 6) in the order of means.h5 (Sbar, Srms, vxS, vyS, vzS, chiS)
        integer(c_int) function afid_calc_sal_stats(ctx, out) bind(C, name='afid_calc_sal_stats')
            import; type(c_ptr), value :: ctx; real(c_double), intent(out) :: out(*)
        end function
        ! CalcPhiStats: out(nxmr, 2) = phibar, phirms
        integer(c_int) function afid_calc_phi_stats(ctx, out) bind(C, name='afid_calc_phi_stats')
            import; type(c_ptr), value :: ctx; real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function afid_create_initial_conditions(ctx, seed, eps) bind(C, name='afid_create_initial_conditions')
            import; type(c_ptr), value :: ctx; integer(c_int64_t), value :: seed; real(c_double), value :: eps
        end function
        integer(c_int) function afid_synchronize(ctx) bind(C, name='afid_synchronize')
            import; type(c_ptr), value :: ctx
        end function
        !> ensemble: many 2-D runs advanced together (main.F90:281-422 per run)
        integer(c_int) function afid_ens_init(cfgs, ctl, nruns, ens) bind(C, name='afid_ens_init')
            import; type(afid_config), intent(in) :: cfgs(*); type(afid_run_control), intent(in) :: ctl(*)
            integer(c_int32_t), value :: nruns; type(c_ptr), intent(out) :: ens
        end function
        integer(c_int) function afid_ens_finalize(ens) bind(C, name='afid_ens_finalize')
            import; type(c_ptr), value :: ens
        end function
        type(c_ptr) function afid_ens_last_error(ens) bind(C, name='afid_ens_last_error')
            import; type(c_ptr), value :: ens
        end function
        integer(c_int) function afid_ens_upload(ens, run, field, host, lb, ext) bind(C, name='afid_ens_upload')
            import; type(c_ptr), value :: ens; integer(c_int32_t), value :: run, field
            real(c_double), intent(in) :: host(*); integer(c_int64_t), intent(in) :: lb(3), ext(3)
        end function
        integer(c_int) function afid_ens_download(ens, run, field, host, lb, ext) bind(C, name='afid_ens_download')
            import; type(c_ptr), value :: ens; integer(c_int32_t), value :: run, field
            real(c_double), intent(out) :: host(*); integer(c_int64_t), intent(in) :: lb(3), ext(3)
        end function
        integer(c_int) function afid_ens_create_initial_conditions(ens, seed0, eps) &
                bind(C, name='afid_ens_create_initial_conditions')
            import; type(c_ptr), value :: ens; integer(c_int64_t), value :: seed0; real(c_double), value :: eps
        end function
        integer(c_int) function afid_ens_reset(ens) bind(C, name='afid_ens_reset')
            import; type(c_ptr), value :: ens
        end function
        integer(c_int) function afid_ens_step(ens, nsteps, graph) bind(C, name='afid_ens_step')
            import; type(c_ptr), value :: ens; integer(c_int32_t), value :: nsteps, graph
        end function
        integer(c_int) function afid_ens_run_state(ens, run, st) bind(C, name='afid_ens_run_state')
            import; type(c_ptr), value :: ens; integer(c_int32_t), value :: run; type(afid_run_state), intent(out) :: st
        end function
        integer(c_int) function afid_ens_autotune(ens, steps, best, log, maxlog, nlog) bind(C, name='afid_ens_autotune')
            import; type(c_ptr), value :: ens; integer(c_int32_t), value :: steps, maxlog
            type(afid_ens_tuning), intent(out) :: best; type(afid_ens_tuning), intent(out) :: log(*)
            integer(c_int32_t), intent(out) :: nlog
        end function
    end interface

contains

    !> stop with the library's message if a call failed
    subroutine afid_check(ctx, rc, what)
        type(c_ptr), intent(in) :: ctx
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (rc == 0) return
        call c_f_pointer(afid_last_error(ctx), msg, [1024])
        n = 0
        do while (n < 1024)
            if (msg(n + 1) == c_null_char) exit
            n = n + 1
        end do
        write(*, '(a,a,i0,a,1024a)') what, ' failed (', rc, '): ', msg(1:n)
        error stop 1
    end subroutine afid_check

end module afid_gpu
