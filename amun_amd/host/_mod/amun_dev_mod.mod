﻿!mod$ v1 sum:49faf7fc53b1d3c7
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module amun_dev_mod
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
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
integer(4),parameter::amun_eqsys_hd=1_4
integer(4),parameter::amun_eqsys_mhd=2_4
integer(4),parameter::amun_eqsys_srmhd=4_4
integer(4),parameter::amun_eos_adi=1_4
integer(4),parameter::amun_rec_tvd=1_4
integer(4),parameter::amun_rec_weno5z=5_4
integer(4),parameter::amun_lim_zero=0_4
integer(4),parameter::amun_lim_minmod=1_4
integer(4),parameter::amun_lim_mc=2_4
integer(4),parameter::amun_lim_superbee=3_4
integer(4),parameter::amun_lim_vanleer=4_4
integer(4),parameter::amun_lim_vanalbada=5_4
integer(4),parameter::amun_rs_hll=1_4
integer(4),parameter::amun_rs_hllc=2_4
integer(4),parameter::amun_rs_hlld=3_4
integer(4),parameter::amun_ta_euler=0_4
integer(4),parameter::amun_ta_rk2=1_4
integer(4),parameter::amun_ta_rk3=2_4
integer(4),parameter::amun_prec_fp64=0_4
integer(4),parameter::amun_prec_mixed=1_4
type,bind(c)::amun_cfg
integer(4)::ndims
integer(4)::eqsys
integer(4)::eos
integer(4)::nv
integer(4)::ncells
integer(4)::nghosts
integer(4)::xblocks
integer(4)::yblocks
integer(4)::zblocks
integer(4)::reconstruction
integer(4)::limiter
integer(4)::riemann_solver
integer(4)::time_advance
integer(4)::precision
integer(4)::fix_positivity
integer(4)::clip_extrema
real(8)::gamma
real(8)::cfl
real(8)::eps
real(8)::alpha
real(8)::xmin
real(8)::xmax
real(8)::ymin
real(8)::ymax
real(8)::zmin
real(8)::zmax
real(8)::tolerance
real(8)::pmin
integer(4)::nr_maxit
integer(4)::nr_extra
end type
interface
subroutine amun_dev_default_cfg(cfg) bind(c,name="amun_dev_default_cfg")
import::amun_cfg
type(amun_cfg),intent(out)::cfg
end
end interface
interface
function amun_dev_init(cfg,ctx) bind(c,name="amun_dev_init")
import::amun_cfg
import::c_ptr
type(amun_cfg),intent(in)::cfg
type(c_ptr),intent(out)::ctx
integer(4)::amun_dev_init
end
end interface
interface
subroutine amun_dev_finalize(ctx) bind(c,name="amun_dev_finalize")
import::c_ptr
type(c_ptr),value::ctx
end
end interface
interface
function amun_dev_upload_blocks(ctx,u) bind(c,name="amun_dev_upload_blocks")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::u(1_8:*)
integer(4)::amun_dev_upload_blocks
end
end interface
interface
function amun_dev_download_blocks(ctx,u,q) bind(c,name="amun_dev_download_blocks")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::u(1_8:*)
type(c_ptr),value::q
integer(4)::amun_dev_download_blocks
end
end interface
interface
function amun_dev_new_time_step(ctx,dtnext,dt,cmax) bind(c,name="amun_dev_new_time_step")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::dtnext
real(8),intent(out)::dt
real(8),intent(out)::cmax
integer(4)::amun_dev_new_time_step
end
end interface
interface
function amun_dev_step(ctx,dtnext,dt,cmax) bind(c,name="amun_dev_step")
import::c_ptr
type(c_ptr),value::ctx
real(8),value::dtnext
real(8),intent(out)::dt
real(8),intent(out)::cmax
integer(4)::amun_dev_step
end
end interface
interface
function amun_dev_run(ctx,nsteps,dtnext,dts) bind(c,name="amun_dev_run")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nsteps
real(8),value::dtnext
real(8),intent(out)::dts(1_8:*)
integer(4)::amun_dev_run
end
end interface
interface
function amun_dev_integrals(ctx,out) bind(c,name="amun_dev_integrals")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::out(1_8:8_8)
integer(4)::amun_dev_integrals
end
end interface
interface
function amun_dev_get_time(ctx,time,step) bind(c,name="amun_dev_get_time")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::time
integer(8),intent(out)::step
integer(4)::amun_dev_get_time
end
end interface
interface
function amun_dev_block_coords(ctx,coords) bind(c,name="amun_dev_block_coords")
import::c_ptr
type(c_ptr),value::ctx
integer(4),intent(out)::coords(1_8:*)
integer(4)::amun_dev_block_coords
end
end interface
interface
function amun_dev_c2p_events(ctx,n) bind(c,name="amun_dev_c2p_events")
import::c_ptr
type(c_ptr),value::ctx
integer(8),intent(out)::n(1_8:3_8)
integer(4)::amun_dev_c2p_events
end
end interface
interface
function amun_dev_last_error_c() bind(c,name="amun_dev_last_error")
import::c_ptr
type(c_ptr)::amun_dev_last_error_c
end
end interface
contains
function amun_dev_last_error() result(msg)
character(256_4,1)::msg
end
end
