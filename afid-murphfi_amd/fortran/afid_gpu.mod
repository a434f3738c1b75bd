﻿!mod$ v1 sum:55e281dce120dcec
!need$ 0bde2ac47243ead2 i iso_c_binding
module afid_gpu
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
integer(4),parameter::afid_fp64=0_4
integer(4),parameter::afid_mixed=1_4
integer(4),parameter::afid_vx=0_4
integer(4),parameter::afid_vy=1_4
integer(4),parameter::afid_vz=2_4
integer(4),parameter::afid_temp=3_4
integer(4),parameter::afid_pr=4_4
integer(4),parameter::afid_sal=14_4
integer(4),parameter::afid_vxr=18_4
integer(4),parameter::afid_vyr=19_4
integer(4),parameter::afid_vzr=20_4
integer(4),parameter::afid_salc=21_4
integer(4),parameter::afid_tempbp=22_4
integer(4),parameter::afid_temptp=23_4
integer(4),parameter::afid_salbp=24_4
integer(4),parameter::afid_saltp=25_4
type,bind(c)::afid_config
integer(4)::nxm
integer(4)::nym
integer(4)::nzm
integer(4)::nsst
integer(4)::multires
integer(4)::nxmr
integer(4)::nymr
integer(4)::nzmr
integer(4)::salinity
integer(4)::phasefield
real(8)::alx3
real(8)::ylen
real(8)::zlen
integer(4)::istr3
real(8)::str3
integer(4)::istr3r
real(8)::rayt
real(8)::prat
real(8)::rays
real(8)::pras
integer(4)::ffscales
integer(4)::inslws
integer(4)::inslwn
integer(4)::tfixs
integer(4)::tfixn
integer(4)::sfixs
integer(4)::sfixn
integer(4)::active_t
integer(4)::active_s
integer(4)::gaxis
real(8)::xplusu
real(8)::xminusu
real(8)::dpdy
real(8)::dpdz
integer(4)::precision
end type
interface
function afid_gpu_init(cfg,ctx) bind(c,name="afid_gpu_init")
import::afid_config
import::c_ptr
type(afid_config),intent(in)::cfg
type(c_ptr),intent(out)::ctx
integer(4)::afid_gpu_init
end
end interface
interface
function afid_gpu_finalize(ctx) bind(c,name="afid_gpu_finalize")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_gpu_finalize
end
end interface
interface
function afid_last_error(ctx) bind(c,name="afid_last_error")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr)::afid_last_error
end
end interface
interface
function afid_upload(ctx,field,host,lb,ext) bind(c,name="afid_upload")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::field
real(8),intent(in)::host(1_8:*)
integer(8),intent(in)::lb(1_8:3_8)
integer(8),intent(in)::ext(1_8:3_8)
integer(4)::afid_upload
end
end interface
interface
function afid_download(ctx,field,host,lb,ext) bind(c,name="afid_download")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::field
real(8),intent(out)::host(1_8:*)
integer(8),intent(in)::lb(1_8:3_8)
integer(8),intent(in)::ext(1_8:3_8)
integer(4)::afid_download
end
end interface
interface
function afid_set_substep(ctx,dt,ns) bind(c,name="afid_set_substep")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::dt
integer(4),value::ns
integer(4)::afid_set_substep
end
end interface
interface
function afid_explicit_terms_vx(ctx) bind(c,name="afid_explicit_terms_vx")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_explicit_terms_vx
end
end interface
interface
function afid_explicit_terms_vy(ctx) bind(c,name="afid_explicit_terms_vy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_explicit_terms_vy
end
end interface
interface
function afid_explicit_terms_vz(ctx) bind(c,name="afid_explicit_terms_vz")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_explicit_terms_vz
end
end interface
interface
function afid_explicit_terms_temp(ctx) bind(c,name="afid_explicit_terms_temp")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_explicit_terms_temp
end
end interface
interface
function afid_explicit_salinity(ctx) bind(c,name="afid_explicit_salinity")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_explicit_salinity
end
end interface
interface
function afid_add_sal_buoyancy(ctx) bind(c,name="afid_add_sal_buoyancy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_add_sal_buoyancy
end
end interface
interface
function afid_implicit_and_update_vx(ctx) bind(c,name="afid_implicit_and_update_vx")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_implicit_and_update_vx
end
end interface
interface
function afid_implicit_and_update_vy(ctx) bind(c,name="afid_implicit_and_update_vy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_implicit_and_update_vy
end
end interface
interface
function afid_implicit_and_update_vz(ctx) bind(c,name="afid_implicit_and_update_vz")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_implicit_and_update_vz
end
end interface
interface
function afid_implicit_and_update_temp(ctx) bind(c,name="afid_implicit_and_update_temp")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_implicit_and_update_temp
end
end interface
interface
function afid_implicit_salinity(ctx) bind(c,name="afid_implicit_salinity")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_implicit_salinity
end
end interface
interface
function afid_calc_local_divergence(ctx) bind(c,name="afid_calc_local_divergence")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_calc_local_divergence
end
end interface
interface
function afid_solve_pressure_correction(ctx) bind(c,name="afid_solve_pressure_correction")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_solve_pressure_correction
end
end interface
interface
function afid_correct_velocity(ctx) bind(c,name="afid_correct_velocity")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_correct_velocity
end
end interface
interface
function afid_correct_pressure(ctx) bind(c,name="afid_correct_pressure")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_correct_pressure
end
end interface
interface
function afid_interp_vel_mgrd(ctx) bind(c,name="afid_interp_vel_mgrd")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_interp_vel_mgrd
end
end interface
interface
function afid_interp_sal_multigrid(ctx) bind(c,name="afid_interp_sal_multigrid")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_interp_sal_multigrid
end
end interface
interface
function afid_time_marcher(ctx,dt) bind(c,name="afid_time_marcher")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::dt
integer(4)::afid_time_marcher
end
end interface
interface
function afid_calc_max_cfl(ctx,cflm,cflmr) bind(c,name="afid_calc_max_cfl")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::cflm
real(8),intent(out)::cflmr
integer(4)::afid_calc_max_cfl
end
end interface
interface
function afid_check_divergence(ctx,qmax,qmaxr) bind(c,name="afid_check_divergence")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::qmax
real(8),intent(out)::qmaxr
integer(4)::afid_check_divergence
end
end interface
interface
function afid_global_quantities(ctx,out) bind(c,name="afid_global_quantities")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::out(1_8:4_8)
integer(4)::afid_global_quantities
end
end interface
interface
function afid_create_initial_conditions(ctx,seed,eps) bind(c,name="afid_create_initial_conditions")
import::c_ptr
type(c_ptr),value::ctx
integer(8),value::seed
real(8),value::eps
integer(4)::afid_create_initial_conditions
end
end interface
interface
function afid_synchronize(ctx) bind(c,name="afid_synchronize")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::afid_synchronize
end
end interface
contains
subroutine afid_check(ctx,rc,what)
type(c_ptr),intent(in)::ctx
integer(4),intent(in)::rc
character(*,1),intent(in)::what
end
end
