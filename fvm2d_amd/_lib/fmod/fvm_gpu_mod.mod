﻿!mod$ v1 sum:1cbc09c3f202c13e
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module fvm_gpu_mod
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
integer(4),parameter::fvm_ok=0_4
integer(4),parameter::fvm_converged=1_4
integer(4),parameter::fvm_iter_limit=2_4
integer(4),parameter::fvm_e_degenerate=-1_4
integer(4),parameter::fvm_e_area=-2_4
integer(4),parameter::fvm_e_orphan_edge=-3_4
integer(4),parameter::fvm_e_neg_sound=-4_4
integer(4),parameter::fvm_e_nonpositive=-5_4
integer(4),parameter::fvm_e_zero_res1=-6_4
integer(4),parameter::fvm_e_input=-7_4
type,bind(c)::fvm_params
integer(4)::iflow
integer(4)::iflux
integer(4)::ilimit
integer(4)::vortex
integer(4)::cell_type
integer(4)::iterlast
integer(4)::maxiter
integer(4)::saveinterval
real(8)::mach_inf
real(8)::aoa_deg
real(8)::rey
real(8)::cfl
real(8)::minres
real(8)::xref
real(8)::yref
integer(4)::sa_compat
integer(4)::hllc_rotate
integer(4)::precision
integer(4)::patch_size
real(8)::nut_factor
end type
type,bind(c)::fvm_mesh
integer(4)::np
integer(4)::nt
integer(4)::ne
integer(4)::netot
integer(4)::nsp
integer(4)::nswe1
integer(4)::nswe2
integer(4)::nffe1
integer(4)::nffe2
integer(4)::reserved
type(c_ptr)::coord
type(c_ptr)::elem
type(c_ptr)::edge
type(c_ptr)::iedge
type(c_ptr)::beindx
type(c_ptr)::ptype
type(c_ptr)::spts
type(c_ptr)::ds
type(c_ptr)::elarea
type(c_ptr)::cvarea
type(c_ptr)::mcarea
type(c_ptr)::drmin
type(c_ptr)::wd
end type
interface
function fvm_geometric(np,nt,coord,elem,ptype,cell_type,iflow,sa_compat,max_edges,edge,edneigh,iedge,ds,dsb,elarea,cvarea,mcarea,drmin,wd,wd1,beindx,spts,dims) bind(c,name="fvm_geometric")
integer(4),value::np
integer(4),value::nt
real(8),intent(in)::coord(1_8:*)
integer(4),intent(inout)::elem(1_8:*)
integer(4),intent(in)::ptype(1_8:*)
integer(4),value::cell_type
integer(4),value::iflow
integer(4),value::sa_compat
integer(4),value::max_edges
integer(4),intent(out)::edge(1_8:*)
integer(4),intent(out)::edneigh(1_8:*)
integer(4),intent(out)::iedge(1_8:*)
real(8),intent(out)::ds(1_8:*)
real(8),intent(out)::dsb(1_8:*)
real(8),intent(out)::elarea(1_8:*)
real(8),intent(out)::cvarea(1_8:*)
real(8),intent(out)::mcarea(1_8:*)
real(8),intent(out)::drmin(1_8:*)
real(8),intent(out)::wd(1_8:*)
real(8),intent(out)::wd1(1_8:*)
integer(4),intent(out)::beindx(1_8:*)
integer(4),intent(out)::spts(1_8:*)
integer(4),intent(out)::dims(1_8:*)
integer(4)::fvm_geometric
end
end interface
interface
function fvm_gpu_wall_dist(np,coord,ptype,netot,edge,sa_compat,wd,wd1) bind(c,name="fvm_gpu_wall_dist")
integer(4),value::np
real(8),intent(in)::coord(1_8:*)
integer(4),intent(in)::ptype(1_8:*)
integer(4),value::netot
integer(4),intent(in)::edge(1_8:*)
integer(4),value::sa_compat
real(8),intent(out)::wd(1_8:*)
real(8),intent(out)::wd1(1_8:*)
integer(4)::fvm_gpu_wall_dist
end
end interface
interface
function fvm_gpu_create(par,mesh,ctx) bind(c,name="fvm_gpu_create")
import::c_ptr
import::fvm_mesh
import::fvm_params
type(fvm_params),intent(in)::par
type(fvm_mesh),intent(in)::mesh
type(c_ptr),intent(out)::ctx
integer(4)::fvm_gpu_create
end
end interface
interface
function fvm_gpu_destroy(ctx) bind(c,name="fvm_gpu_destroy")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::fvm_gpu_destroy
end
end interface
interface
function fvm_gpu_set_state(ctx,prim,nut) bind(c,name="fvm_gpu_set_state")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::prim(1_8:*)
real(8),intent(in)::nut(1_8:*)
integer(4)::fvm_gpu_set_state
end
end interface
interface
function fvm_gpu_get_state(ctx,prim,nut,qx,qy,mu) bind(c,name="fvm_gpu_get_state")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::prim(1_8:*)
real(8),intent(out)::nut(1_8:*)
real(8),intent(out)::qx(1_8:*)
real(8),intent(out)::qy(1_8:*)
real(8),intent(out)::mu(1_8:*)
integer(4)::fvm_gpu_get_state
end
end interface
interface
function fvm_gpu_iterate(ctx,n,hist,ndone,err_index) bind(c,name="fvm_gpu_iterate")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::n
real(8),intent(out)::hist(1_8:*)
integer(4),intent(out)::ndone
integer(4),intent(out)::err_index
integer(4)::fvm_gpu_iterate
end
end interface
interface
function fvm_gpu_monitor(ctx,mon) bind(c,name="fvm_gpu_monitor")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::mon(1_8:*)
integer(4)::fvm_gpu_monitor
end
end interface
interface
function fvm_status_string(code) bind(c,name="fvm_status_string")
import::c_ptr
integer(4),value::code
type(c_ptr)::fvm_status_string
end
end interface
contains
function status_text(code) result(s)
integer(4),intent(in)::code
character(128_4,1)::s
end
end
