! fvm2d_gpu.f90 -- host program of the GPU solver, in the reference's own
! terms: it reads the same input deck (src/read_input.f:20-66) and amdba grid
! (src/read_grid.f:19-39), restarts from the same SOL file
! (src/initialize.f:91-103), writes RES.DAT rows in the same format
! (src/main.f:63), WALL.DAT at every SaveInt (src/result.f:92-118), and SOL
! and PARAM.DAT at the end (src/finalize.f:14-59).  The preprocessing and the
! iterations go through the C-ABI of libfvm2d_gpu (fvm_gpu_mod); the loop
! condition of src/main.f:42-76 is evaluated on the device.
!
!     fvm2d_gpu <deck> [<precision 0|1> [residu | batch=K]]
!
! residu: writes RESIDU.DAT, the residual of the start state through the
! parity hook fvm_gpu_residu (src/residu.f:2-3), res(4,np) at full
! precision, and stops.  batch=K: K solves of the deck at AOA, AOA+2, ...,
! AOA+2(K-1) deg advanced together (fvm_batch_*), each member's last RES.DAT
! row in BATCH.DAT at full precision.
!
! Plots (gnuplot, mayavi, vigie files) are not written.
program fvm2d_gpu
  use iso_c_binding
  use fvm_gpu_mod
  implicit none

  real(c_double), parameter :: GAMMA = 1.4d0, GAMMA1 = GAMMA - 1.0d0
  integer, parameter :: SCRATCH = 1, RESTART = 2, INVISCID = 1, TURBULENT = 3

  ! deck
  integer :: istart, iflow, iterlast, maxiter, saveinterval, niso, iflux, ilimit, vortex, &
             cell_type
  real(c_double) :: mach_inf, aoa_deg, rey, cfl, minres, xref, yref
  character(len=256) :: deckfile, gridfile, arg, mode
  integer :: precision, nbatch
  ! range of the solution at the last write_result (src/result.f:16-48), for PARAM.DAT
  real(c_double) :: range_lo(7), range_hi(7)

  ! grid and geometry (the reference's layout)
  integer(c_int) :: np, nt, me
  real(c_double), allocatable, target :: coord(:,:), ds(:,:), dsb(:,:), elarea(:), cvarea(:), &
                                         mcarea(:), drmin(:), wd(:), wd1(:)
  integer(c_int), allocatable, target :: elem(:,:), ptype(:), edge(:,:), edneigh(:,:), &
                                         iedge(:), beindx(:), spts(:)
  integer(c_int) :: dims(12)

  ! state
  real(c_double), allocatable :: prim(:,:), nut(:), qx(:,:), qy(:,:), mu(:), hist(:,:)
  real(c_double) :: mon(6), p_inf, r_inf, q_inf, u1, u2, u3, u4, u5
  type(fvm_params) :: par
  type(fvm_mesh) :: mesh
  type(c_ptr) :: ctx
  integer(c_int) :: rc, ndone, erridx, n, iter, k
  integer :: i, j, ngrid

  call get_command_argument(1, deckfile)
  if (len_trim(deckfile) == 0) then
    print *, 'You must specify an input file.'
    stop
  end if
  precision = 0
  call get_command_argument(2, arg)
  if (len_trim(arg) > 0) read (arg, *) precision
  call get_command_argument(3, mode)
  nbatch = 0
  if (mode(1:6) == 'batch=') read (mode(7:), *) nbatch
  call read_deck()

  ! ---- grid (src/read_grid.f:19-39) ----
  print *, 'Reading grid from file ', trim(gridfile)
  ngrid = 10
  open (ngrid, file=gridfile, status='old')
  read (ngrid, *) np, nt
  print *, 'Number of points    : ', np
  print *, 'Number of triangles : ', nt
  allocate (coord(2, np), ptype(np), elem(3, nt))
  do i = 1, np
    read (ngrid, *) k, coord(1, i), coord(2, i), ptype(i)
  end do
  do i = 1, nt
    read (ngrid, *) k, elem(1, i), elem(2, i), elem(3, i), j
  end do
  close (ngrid)

  ! ---- geometric (src/geometric.f:15-60) ----
  me = 3 * nt + 3
  allocate (edge(2, me), edneigh(2, me), iedge(me), ds(2, me), dsb(2, np), elarea(nt), &
            cvarea(np), mcarea(np), drmin(np), wd(np), wd1(np), beindx(me), spts(np))
  wd = 0.0d0
  wd1 = 0.0d0
  rc = fvm_geometric(np, nt, coord, elem, ptype, cell_type, INVISCID, 1, me, edge, edneigh, &
                     iedge, ds, dsb, elarea, cvarea, mcarea, drmin, wd, wd1, beindx, spts, dims)
  if (rc /= FVM_OK) then
    print *, trim(status_text(rc)), ' at ', dims(12)
    stop
  end if
  if (iflow == TURBULENT) then     ! src/wall_dist.f on the GPU
    rc = fvm_gpu_wall_dist(np, coord, ptype, dims(4), edge, 1, wd, wd1)
    if (rc /= FVM_OK) then
      print *, trim(status_text(rc))
      stop
    end if
  end if

  ! ---- context: mesh resident on the device, free-stream state ----
  par%iflow = iflow; par%iflux = iflux; par%ilimit = ilimit; par%vortex = vortex
  par%cell_type = cell_type; par%iterlast = iterlast; par%maxiter = maxiter
  par%saveinterval = saveinterval
  par%mach_inf = mach_inf; par%aoa_deg = aoa_deg; par%rey = rey; par%cfl = cfl
  par%minres = minres; par%xref = xref; par%yref = yref
  par%sa_compat = 1; par%hllc_rotate = 0; par%precision = precision; par%patch_size = 0
  par%nut_factor = 0.1d0
  mesh%np = dims(1); mesh%nt = dims(2); mesh%ne = dims(3); mesh%netot = dims(4)
  mesh%nsp = dims(5); mesh%nswe1 = dims(8); mesh%nswe2 = dims(9); mesh%nffe1 = dims(10)
  mesh%nffe2 = dims(11); mesh%reserved = 0
  mesh%coord = c_loc(coord); mesh%elem = c_loc(elem); mesh%edge = c_loc(edge)
  mesh%iedge = c_loc(iedge); mesh%beindx = c_loc(beindx); mesh%ptype = c_loc(ptype)
  mesh%spts = c_loc(spts); mesh%ds = c_loc(ds); mesh%elarea = c_loc(elarea)
  mesh%cvarea = c_loc(cvarea); mesh%mcarea = c_loc(mcarea); mesh%drmin = c_loc(drmin)
  mesh%wd = c_loc(wd)
  rc = fvm_gpu_create(par, mesh, ctx)
  if (rc /= FVM_OK) then
    print *, 'fvm_gpu_create: ', trim(status_text(rc))
    stop
  end if

  q_inf = 1.0d0
  r_inf = 1.0d0
  p_inf = 1.0d0 / (GAMMA * mach_inf**2)
  allocate (prim(4, np), nut(np), qx(4, np), qy(4, np), mu(np))
  nut = 0.0d0

  ! ---- restart from SOL (src/initialize.f:91-103) ----
  if (istart == RESTART) then
    print *, 'Initializing solution to old values from SOL'
    open (unit=20, file='SOL', status='old')
    do j = 1, np
      read (20, *) u1, u2, u3, u4, u5
      prim(1, j) = u1
      prim(2, j) = u2 / u1
      prim(3, j) = u3 / u1
      prim(4, j) = GAMMA1 * (u4 - 0.5d0 * (u2**2 + u3**2) / u1)
      nut(j) = u5
    end do
    close (20)
    rc = fvm_gpu_set_state(ctx, prim, nut)
    if (rc /= FVM_OK) then
      print *, 'fvm_gpu_set_state: ', trim(status_text(rc))
      stop
    end if
  else
    print *, 'Initializing solution to free stream values'
  end if

  if (trim(mode) == 'residu') then
    call write_residu()
    rc = fvm_gpu_destroy(ctx)
    stop
  end if
  if (nbatch > 0) then
    call run_batch()
    rc = fvm_gpu_destroy(ctx)
    stop
  end if

  ! ---- iterations (src/main.f:36-76) ----
  print *, 'Beginning of iterations ...'
  open (unit=11, file='RES.DAT', status='replace')
  iter = iterlast
  allocate (hist(6, 4096))
  do
    ! up to the next multiple of SaveInt, where the reference writes results
    n = min(saveinterval - mod(iter, saveinterval), 4096)
    rc = fvm_gpu_iterate(ctx, n, hist, ndone, erridx)
    do k = 1, ndone
      write (11, '(I8,E14.6,I8,3E14.6)') nint(hist(1, k)), hist(2, k), nint(hist(3, k)), &
                                          hist(4, k), hist(5, k), hist(6, k)
    end do
    iter = iter + ndone
    if (rc < 0) then
      call stop_message(rc, erridx)
      stop
    end if
    if (ndone > 0 .and. mod(iter, saveinterval) == 0) then
      call flush (11)
      call write_result()
      if (iter >= maxiter) then                    ! src/main.f:69-73
        print *, '*** Stopping execution ***'
        print *, 'Maximum number of iterations reached'
        exit
      end if
    end if
    if (rc == FVM_CONVERGED .or. rc == FVM_ITER_LIMIT) then   ! loop condition, main.f:42
      close (11)
      call write_result()
      print *, '*** Stopping execution ***'
      print *, 'Residue has been reduced below MINRES =', minres
      exit
    end if
  end do
  close (11)
  call finalize()
  rc = fvm_gpu_destroy(ctx)

contains

  subroutine read_deck()
    character(len=32) :: s
    integer :: inp
    inp = 11
    open (unit=inp, file=deckfile, status='old')
    print *, 'Reading parameters from ', trim(deckfile)
    read (inp, *) s, istart;       call check_string('Start', s)
    read (inp, *) s, iflow;        call check_string('FlowType', s)
    read (inp, *) s, mach_inf;     call check_string('Mach', s)
    read (inp, *) s, aoa_deg;      call check_string('AOA', s)
    read (inp, *) s, rey;          call check_string('Reynolds', s)
    read (inp, *) s, cfl;          call check_string('CFL', s)
    read (inp, *) s, iterlast;     call check_string('LastIter', s)
    read (inp, *) s, maxiter;      call check_string('MaxIter', s)
    read (inp, *) s, minres;       call check_string('MinRes', s)
    read (inp, *) s, saveinterval; call check_string('SaveInt', s)
    read (inp, *) s, niso;         call check_string('NoIso', s)
    read (inp, *) s, iflux;        call check_string('Flux', s)
    read (inp, *) s, ilimit;       call check_string('Limiter', s)
    read (inp, *) s, vortex, xref, yref; call check_string('FarBC', s)
    read (inp, *) s, cell_type;    call check_string('CellType', s)
    read (inp, *) s, gridfile;     call check_string('GridFile', s)
    close (inp)
    if (istart /= SCRATCH .and. istart /= RESTART) then
      print *, 'Unknown start option', istart
      stop
    end if
    if (saveinterval <= 0) saveinterval = maxiter
  end subroutine read_deck

  subroutine check_string(expect, got)
    character(len=*), intent(in) :: expect, got
    if (trim(expect) /= trim(got)) then
      print *, 'Expecting ', expect, ' but found ', trim(got)
      stop
    end if
  end subroutine check_string

  ! the reference's message for each STOP condition
  subroutine stop_message(code, idx)
    integer(c_int), intent(in) :: code, idx
    integer :: ierr
    select case (code)
    case (FVM_E_NEG_SOUND)
      print *, 'Sonic speed is negative'                            ! src/roe_flux.f:63
    case (FVM_E_NONPOSITIVE)
      ierr = fvm_gpu_get_state(ctx, prim, nut, qx, qy, mu)
      print *, 'Density/pressure is negative at point ', idx        ! src/check_positivity.f:15
      if (idx >= 1 .and. idx <= np) then
        print *, 'Type    = ', ptype(idx)
        print *, 'Coord   = ', coord(1, idx), coord(2, idx)
        print *, 'Density = ', prim(1, idx)
        print *, 'u vel   = ', prim(2, idx)
        print *, 'v vel   = ', prim(3, idx)
        print *, 'Pressure= ', prim(4, idx)
        print *, 'Visc    = ', nut(idx)
      end if
    case (FVM_E_ZERO_RES1)
      print *, 'Residue in first iteration = ', 0.0d0               ! src/conres.f:27
    case default
      print *, trim(status_text(code)), ' at ', idx
    end select
  end subroutine stop_message

  ! RESIDU.DAT: the residual of the current state through the hook
  subroutine write_residu()
    real(c_double), allocatable :: res(:,:)
    integer :: ierr, ii
    allocate (res(4, np))
    ierr = fvm_gpu_get_state(ctx, prim, nut, qx, qy, mu)
    ierr = fvm_gpu_residu(ctx, prim, nut, res)
    if (ierr /= FVM_OK) then
      call stop_message(ierr, 0)
      stop
    end if
    open (unit=20, file='RESIDU.DAT', status='replace')
    do ii = 1, np
      write (20, '(4es26.17e3)') res(:, ii)
    end do
    close (20)
  end subroutine write_residu

  ! K solves of this deck, alpha stepped by 2 deg, as one batch
  subroutine run_batch()
    type(c_ptr), allocatable :: member(:)
    type(c_ptr) :: batch
    type(fvm_params) :: pk
    integer(c_int), allocatable :: st(:)
    integer(c_int) :: ierr, left
    integer :: kk
    allocate (member(nbatch), st(nbatch))
    ierr = fvm_batch_create(batch)
    do kk = 1, nbatch
      pk = par
      pk%aoa_deg = aoa_deg + 2.0d0 * (kk - 1)
      ierr = fvm_gpu_create(pk, mesh, member(kk))
      if (ierr == FVM_OK) ierr = fvm_batch_add(batch, member(kk))
      if (ierr /= FVM_OK) then
        print *, 'batch member ', kk, ': ', trim(status_text(ierr))
        stop
      end if
    end do
    left = maxiter
    do while (left > 0)
      ierr = fvm_batch_iterate(batch, min(left, 4096), st)
      if (ierr /= FVM_OK) then
        print *, 'fvm_batch_iterate: ', trim(status_text(ierr))
        stop
      end if
      left = left - min(left, 4096)
      if (all(st /= FVM_OK)) exit
    end do
    open (unit=20, file='BATCH.DAT', status='replace')
    do kk = 1, nbatch
      ierr = fvm_gpu_monitor(member(kk), mon)
      write (20, '(i4,i4,es26.17e3,i8,es26.17e3,i8,3es26.17e3)') kk, ierr, aoa_deg + 2.0d0 * (kk - 1), &
          nint(mon(1)), mon(2), nint(mon(3)), mon(4), mon(5), mon(6)
    end do
    close (20)
    ierr = fvm_batch_destroy(batch)
    do kk = 1, nbatch
      ierr = fvm_gpu_destroy(member(kk))
    end do
  end subroutine run_batch

  ! Wall quantities at solid point is (WALL.DAT columns, src/result.f:92-118):
  ! the pressure coefficient and, with viscosity, the skin friction from the
  ! wall shear mu d(u.t)/dn (t the wall tangent, n its normal from dsb), the
  ! off-wall distance wd1 and y+ = rho u_tau wd1 / mu.
  subroutine wall_row(ii, is, row, ncol)
    integer, intent(in) :: ii, is
    real(c_double), intent(out) :: row(5)
    integer, intent(out) :: ncol
    real(c_double) :: qdyn, n(2), dudn(2), shear, utau
    qdyn = 0.5d0 * r_inf * q_inf**2
    row = 0.0d0
    row(1) = coord(1, is)
    row(2) = (p_inf - prim(4, is)) / qdyn
    ncol = 2
    if (iflow == INVISCID) return
    n = dsb(:, is) / sqrt(dsb(1, is)**2 + dsb(2, is)**2)
    ! normal derivative of the velocity, then its tangential part (t = (n2, -n1))
    dudn(1) = qx(2, is) * n(1) + qy(2, is) * n(2)
    dudn(2) = qx(3, is) * n(1) + qy(3, is) * n(2)
    shear = mu(is) * (n(2) * dudn(1) - n(1) * dudn(2))
    utau = sqrt(abs(shear) / prim(1, is))
    row(3) = shear / qdyn
    row(4) = wd1(ii)
    row(5) = prim(1, is) * utau * wd1(ii) / mu(is)
    ncol = 5
  end subroutine wall_row

  ! WALL.DAT at every SaveInt and at the end, and the range of the solution
  ! that PARAM.DAT reports
  subroutine write_result()
    integer :: ii, ncol, ierr
    real(c_double) :: row(5), var(7), speed2
    ierr = fvm_gpu_get_state(ctx, prim, nut, qx, qy, mu)
    ierr = fvm_gpu_monitor(ctx, mon)
    print *, 'Iterations =', nint(mon(1)), ' Cl, Cd =', mon(5), mon(6), ' Residue =', mon(2)
    ! density, pressure, Mach number, x and y velocity, log10 of the entropy
    ! ratio s/s_inf, nu_tilde
    range_lo = 1.0d10
    range_hi = -1.0d10
    do ii = 1, np
      speed2 = prim(2, ii)**2 + prim(3, ii)**2
      var = [prim(1, ii), prim(4, ii), sqrt(speed2 * prim(1, ii) / (GAMMA * prim(4, ii))), prim(2, ii), &
             prim(3, ii), log10(prim(4, ii) / prim(1, ii)**GAMMA / (p_inf / r_inf**GAMMA)), nut(ii)]
      range_lo = min(range_lo, var)
      range_hi = max(range_hi, var)
    end do
    open (unit=10, file='WALL.DAT', status='replace')
    do ii = 1, dims(5)
      call wall_row(ii, spts(ii), row, ncol)
      write (10, *) row(1:ncol)
    end do
    close (10)
  end subroutine write_result

  ! SOL in conserved variables and PARAM.DAT (src/finalize.f:14-61)
  subroutine finalize()
    integer :: ierr, ii, ip
    real(c_double) :: v(5)
    real :: cpu(2), total
    character(len=20), parameter :: what(7) = [character(len=20) :: 'density   =', 'pressure  =', &
        'mach no   =', 'x vel     =', 'y vel     =', 'entropy   =', 'viscosity =']
    ierr = fvm_gpu_get_state(ctx, prim, nut, qx, qy, mu)
    ierr = fvm_gpu_monitor(ctx, mon)
    open (unit=20, file='SOL', status='replace')
    do ii = 1, np
      v(1) = prim(1, ii)
      v(2) = v(1) * prim(2, ii)
      v(3) = v(1) * prim(3, ii)
      v(4) = prim(4, ii) / GAMMA1 + 0.5d0 * prim(1, ii) * (prim(2, ii)**2 + prim(3, ii)**2)
      v(5) = nut(ii)
      write (20, *) v(1), v(2), v(3), v(4), v(5)
    end do
    close (20)
    open (unit=20, file='PARAM.DAT', status='replace')
    write (20, '(70a)') ('-', ii = 1, 70)
    write (20, '(" Mach no  =",f6.3,", AOA  =",f6.3, ", Rey =",e10.4)') mach_inf, aoa_deg, rey
    write (20, '(" CFL      =",f6.3,", Flux =",i6, ", Lim =",i2)') cfl, iflux, ilimit
    write (20, '(" Gridfile = ",a30)') gridfile
    write (20, '(70a)') ('-', ii = 1, 70)
    write (20, '(" Iterations        =",i12)') nint(mon(1))
    write (20, '(" Cl                =",f12.6)') mon(5)
    write (20, '(" Cd                =",f12.6)') mon(6)
    write (20, '(" L2 residue        =",f12.6)') log10(mon(2))
    write (20, '(" Linf residue      =",e16.6)') mon(4)
    ip = nint(mon(3))
    if (ip >= 1 .and. ip <= np) then
      write (20, '(" Linf point        =",f12.6,f12.6,i8,i4)') coord(1, ip), coord(2, ip), ip, ptype(ip)
    end if
    do ii = 1, 7
      if (ii < 7) then
        write (20, '(" Minimum ",a,f12.6)') what(ii)(1:11), range_lo(ii)
        write (20, '(" Maximum ",a,f12.6)') what(ii)(1:11), range_hi(ii)
      else
        write (20, '(" Minimum ",a,e16.6)') what(ii)(1:11), range_lo(ii)
        write (20, '(" Maximum ",a,e16.6)') what(ii)(1:11), range_hi(ii)
      end if
    end do
    total = etime(cpu)
    write (20, '(" Elapsed time      =",f12.6, " min")') total / 60.0
    write (20, '(" User    time      =",f12.6, " min")') cpu(1) / 60.0
    write (20, '(" System  time      =",f12.6, " min")') cpu(2) / 60.0
    close (20)
  end subroutine finalize

end program fvm2d_gpu
