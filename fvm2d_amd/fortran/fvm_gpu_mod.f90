! fvm_gpu_mod.f90 -- iso_c_binding interface to libfvm2d_gpu (the C-ABI
! declared in include/fvm2d_gpu.h).  A Fortran host keeps its own arrays in
! the reference's layout (prim(4,np), coord(2,np), elem(3,nt), 1-based) and
! passes them by address; every pointer is borrowed for one call.
module fvm_gpu_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: FVM_OK = 0, FVM_CONVERGED = 1, FVM_ITER_LIMIT = 2
  integer(c_int), parameter :: FVM_E_DEGENERATE = -1, FVM_E_AREA = -2, FVM_E_ORPHAN_EDGE = -3
  integer(c_int), parameter :: FVM_E_NEG_SOUND = -4, FVM_E_NONPOSITIVE = -5
  integer(c_int), parameter :: FVM_E_ZERO_RES1 = -6, FVM_E_INPUT = -7, FVM_E_SIZE = -8
  integer(c_int), parameter :: FVM_E_HIP = -10, FVM_E_NOMEM = -11

  ! fvm_params: the COMMON values of the deck (src/read_input.f:20-66)
  type, bind(C) :: fvm_params
    integer(c_int) :: iflow, iflux, ilimit, vortex, cell_type, iterlast, maxiter, saveinterval
    real(c_double) :: mach_inf, aoa_deg, rey, cfl, minres, xref, yref
    integer(c_int) :: sa_compat, hllc_rotate, precision, patch_size
    real(c_double) :: nut_factor
    integer(c_int) :: block_threads = 0, reserved_params = 0
  end type fvm_params

  ! fvm_mesh: what geometric() leaves in COMMON (src/geometric.f:15-60)
  type, bind(C) :: fvm_mesh
    integer(c_int) :: np, nt, ne, netot, nsp, nswe1, nswe2, nffe1, nffe2, reserved
    type(c_ptr) :: coord, elem, edge, iedge, beindx, ptype, spts, ds, elarea, cvarea, mcarea, &
                   drmin, wd
  end type fvm_mesh

  interface
    integer(c_int) function fvm_geometric(np, nt, coord, elem, ptype, cell_type, iflow, &
        sa_compat, max_edges, edge, edneigh, iedge, ds, dsb, elarea, cvarea, mcarea, drmin, &
        wd, wd1, beindx, spts, dims) bind(C, name="fvm_geometric")
      import :: c_int, c_double
      integer(c_int), value :: np, nt, cell_type, iflow, sa_compat, max_edges
      real(c_double), intent(in) :: coord(*)
      integer(c_int), intent(inout) :: elem(*)
      integer(c_int), intent(in) :: ptype(*)
      integer(c_int), intent(out) :: edge(*), edneigh(*), iedge(*), beindx(*), spts(*), dims(*)
      real(c_double), intent(out) :: ds(*), dsb(*), elarea(*), cvarea(*), mcarea(*), drmin(*), &
                                     wd(*), wd1(*)
    end function

    integer(c_int) function fvm_gpu_wall_dist(np, coord, ptype, netot, edge, sa_compat, wd, &
        wd1) bind(C, name="fvm_gpu_wall_dist")
      import :: c_int, c_double
      integer(c_int), value :: np, netot, sa_compat
      real(c_double), intent(in) :: coord(*)
      integer(c_int), intent(in) :: ptype(*), edge(*)
      real(c_double), intent(out) :: wd(*), wd1(*)
    end function

    integer(c_int) function fvm_gpu_create(par, mesh, ctx) bind(C, name="fvm_gpu_create")
      import :: c_int, c_ptr, fvm_params, fvm_mesh
      type(fvm_params), intent(in) :: par
      type(fvm_mesh), intent(in) :: mesh
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function fvm_gpu_destroy(ctx) bind(C, name="fvm_gpu_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function fvm_gpu_set_state(ctx, prim, nut) bind(C, name="fvm_gpu_set_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: prim(*), nut(*)
    end function

    integer(c_int) function fvm_gpu_get_state(ctx, prim, nut, qx, qy, mu) &
        bind(C, name="fvm_gpu_get_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: prim(*), nut(*), qx(*), qy(*), mu(*)
    end function

    integer(c_int) function fvm_gpu_iterate(ctx, n, hist, ndone, err_index) &
        bind(C, name="fvm_gpu_iterate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      real(c_double), intent(out) :: hist(*)
      integer(c_int), intent(out) :: ndone, err_index
    end function

    ! residu(...) of src/residu.f:2-3 on the given state: res(4,np) in the
    ! caller's order (the parity hook; the iteration state is left as it was)
    integer(c_int) function fvm_gpu_residu(ctx, prim, nut, res) bind(C, name="fvm_gpu_residu")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: prim(*), nut(*)
      real(c_double), intent(out) :: res(*)
    end function

    ! COMMON /liftdrag/ cl, read by the farfield vortex correction
    ! (src/farfield_flux.f:28-44)
    integer(c_int) function fvm_gpu_set_cl(ctx, cl) bind(C, name="fvm_gpu_set_cl")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cl
    end function

    integer(c_int) function fvm_gpu_tune(ctx, variant, vertices_per_block, xcd_remap) &
        bind(C, name="fvm_gpu_tune")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: variant, vertices_per_block, xcd_remap
    end function

    integer(c_int) function fvm_gpu_layout_info(ctx, info) bind(C, name="fvm_gpu_layout_info")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), intent(out) :: info(*)
    end function

    integer(c_int) function fvm_gpu_workgroup_threads(ctx) bind(C, name="fvm_gpu_workgroup_threads")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_long_long) function fvm_gpu_footprint(ctx) bind(C, name="fvm_gpu_footprint")
      import :: c_long_long, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_long_long) function fvm_gpu_footprint_estimate(np, nt, netot, iflow, precision) &
        bind(C, name="fvm_gpu_footprint_estimate")
      import :: c_long_long, c_int
      integer(c_int), value :: np, nt, netot, iflow, precision
    end function

    ! batches: many solves of one kind advanced by one launch per RK stage
    integer(c_int) function fvm_batch_create(batch) bind(C, name="fvm_batch_create")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: batch
    end function

    integer(c_int) function fvm_batch_add(batch, ctx) bind(C, name="fvm_batch_add")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch, ctx
    end function

    integer(c_int) function fvm_batch_remove(batch, ctx) bind(C, name="fvm_batch_remove")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch, ctx
    end function

    integer(c_int) function fvm_batch_size(batch) bind(C, name="fvm_batch_size")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function fvm_batch_iterate(batch, n, status) bind(C, name="fvm_batch_iterate")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: n
      integer(c_int), intent(out) :: status(*)
    end function

    integer(c_int) function fvm_batch_launch(batch, n) bind(C, name="fvm_batch_launch")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: n
    end function

    integer(c_int) function fvm_batch_query(batch) bind(C, name="fvm_batch_query")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function fvm_batch_wait(batch, status) bind(C, name="fvm_batch_wait")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), intent(out) :: status(*)
    end function

    integer(c_int) function fvm_batch_destroy(batch) bind(C, name="fvm_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function fvm_gpu_monitor(ctx, mon) bind(C, name="fvm_gpu_monitor")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: mon(*)
    end function

    type(c_ptr) function fvm_status_string(code) bind(C, name="fvm_status_string")
      import :: c_int, c_ptr
      integer(c_int), value :: code
    end function
  end interface

contains

  ! the status message as a Fortran string
  function status_text(code) result(s)
    integer(c_int), intent(in) :: code
    character(len=128) :: s
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: i
    s = ' '
    cp = fvm_status_string(code)
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [128])
    do i = 1, 128
      if (p(i) == c_null_char) exit
      s(i:i) = p(i)
    end do
  end function status_text

end module fvm_gpu_mod
