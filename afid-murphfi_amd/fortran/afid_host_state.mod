﻿!mod$ v1 sum:d33297127095e113
!need$ 55e281dce120dcec n afid_gpu
!need$ 0bde2ac47243ead2 i iso_c_binding
module afid_host_state
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
use afid_gpu,only:afid_fp64
use afid_gpu,only:afid_mixed
use afid_gpu,only:afid_vx
use afid_gpu,only:afid_vy
use afid_gpu,only:afid_vz
use afid_gpu,only:afid_temp
use afid_gpu,only:afid_pr
use afid_gpu,only:afid_sal
use afid_gpu,only:afid_vxr
use afid_gpu,only:afid_vyr
use afid_gpu,only:afid_vzr
use afid_gpu,only:afid_salc
use afid_gpu,only:afid_tempbp
use afid_gpu,only:afid_temptp
use afid_gpu,only:afid_salbp
use afid_gpu,only:afid_saltp
use afid_gpu,only:afid_config
use afid_gpu,only:afid_gpu_init
use afid_gpu,only:afid_gpu_finalize
use afid_gpu,only:afid_last_error
use afid_gpu,only:afid_upload
use afid_gpu,only:afid_download
use afid_gpu,only:afid_set_substep
use afid_gpu,only:afid_explicit_terms_vx
use afid_gpu,only:afid_explicit_terms_vy
use afid_gpu,only:afid_explicit_terms_vz
use afid_gpu,only:afid_explicit_terms_temp
use afid_gpu,only:afid_explicit_salinity
use afid_gpu,only:afid_add_sal_buoyancy
use afid_gpu,only:afid_implicit_and_update_vx
use afid_gpu,only:afid_implicit_and_update_vy
use afid_gpu,only:afid_implicit_and_update_vz
use afid_gpu,only:afid_implicit_and_update_temp
use afid_gpu,only:afid_implicit_salinity
use afid_gpu,only:afid_calc_local_divergence
use afid_gpu,only:afid_solve_pressure_correction
use afid_gpu,only:afid_correct_velocity
use afid_gpu,only:afid_correct_pressure
use afid_gpu,only:afid_interp_vel_mgrd
use afid_gpu,only:afid_interp_sal_multigrid
use afid_gpu,only:afid_time_marcher
use afid_gpu,only:afid_calc_max_cfl
use afid_gpu,only:afid_check_divergence
use afid_gpu,only:afid_global_quantities
use afid_gpu,only:afid_create_initial_conditions
use afid_gpu,only:afid_synchronize
use afid_gpu,only:afid_check
type(afid_config)::cfg
type(c_ptr)::ctx
integer(4)::nsst
integer(4)::ntst
integer(4)::idtv
integer(4)::nread
integer(4)::ireset
integer(4)::flagmr
integer(4)::flagsal
integer(4)::flagpf
integer(4)::flagmelt
integer(4)::solidtype
integer(4)::pf_ic
real(8)::walltimemax
real(8)::tmax
real(8)::tout
real(8)::tframe
real(8)::save_3d
real(8)::dt
real(8)::resid
real(8)::limitcfl
real(8)::dtmin
real(8)::dtmax
real(8)::pf_d
real(8)::pf_a
real(8)::pf_s
real(8)::pf_tm
logical(4)::salinity
logical(4)::fused
end
