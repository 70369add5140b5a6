C ref_dump.f -- golden-vector harness around the reference's own routines.
C
C TEST INFRASTRUCTURE ONLY.  Linked by oracle/Makefile against the
C reference sources compiled where they lie (src/*.f); never shipped.
C It replaces main.f's loop (main.f:42-76) so that it can write full
C precision arrays (the reference's own files carry 7 digits, main.f:63),
C and ends with the reference's write_result and finalize (WALL.DAT, SOL,
C PARAM.DAT in the output directory):
C   argv(1) deck, argv(2) number of iterations, argv(3) output directory.
C The sa_model call passes nut as the 7th argument; main.f:54-55 omits it.
      program refdump
      implicit none
      include 'param.h'
      include 'gdata.h'

      double precision coord(2, npmax), elarea(ntmax),
     &                 cvarea(npmax), ds(2,nemax), dsb(2,npmax),
     &                 dt(npmax), res(nvar,npmax), prim(nvar,npmax),
     &                 prim_old(nvar,npmax), qx(nvar,npmax),
     &                 qy(nvar,npmax), drmin(npmax), mu(npmax),
     &                 nut(npmax), mul(npmax), mcarea(npmax), wd(npmax)
      double precision hrow(6), cst(20)
      integer          irk, nit, ndims(12), nhist
      character*256    arg, od

      call math
      call read_input
      call getarg(2, arg)
      read(arg,*) nit
      call getarg(3, od)
      call read_grid(coord, elem, ptype, spts, fpts, opts, bpts)
      call geometric(coord, elarea, cvarea, ds, dsb, mcarea, wd,
     &               drmin)
      call initialize(prim, nut, mul, mu)

      ndims(1) = np
      ndims(2) = nt
      ndims(3) = ne
      ndims(4) = netot
      ndims(5) = nsp
      ndims(6) = nfp
      ndims(7) = nbe
      ndims(8) = nswe1
      ndims(9) = nswe2
      ndims(10) = nffe1
      ndims(11) = nffe2
      ndims(12) = 0
      call dumpi(od, 'dims', ndims, 12)
      call dumpi(od, 'elem', elem, 3*nt)
      call dumpi(od, 'edge', edge, 2*netot)
      call dumpi(od, 'edneigh', edneigh, 2*netot)
      call dumpi(od, 'iedge', iedge, ne)
      call dumpi(od, 'beindx', beindx, nbe)
      call dumpd(od, 'ds', ds, 2*netot)
      call dumpd(od, 'dsb', dsb, 2*np)
      call dumpd(od, 'elarea', elarea, nt)
      call dumpd(od, 'cvarea', cvarea, np)
      call dumpd(od, 'mcarea', mcarea, np)
      call dumpd(od, 'drmin', drmin, np)
      call dumpd(od, 'wd', wd, np)

      cst(1) = GAMMA1
      cst(2) = M_PI
      cst(3) = p_inf
      cst(4) = u_inf
      cst(5) = v_inf
      cst(6) = a_inf
      cst(7) = SCONST
      cst(8) = Cw1
      cst(9) = Cv11
      cst(10) = Cw31
      cst(11) = Cw32
      cst(12) = kolm2
      cst(13) = Cb2Sig1
      cst(14) = Cb2Sig2
      cst(15) = nut_inf
      cst(16) = ALBADA12
      cst(17) = ALBADA22
      cst(18) = T_inf
      cst(19) = aoa
      cst(20) = wdmin
      call dumpd(od, 'consts', cst, 20)

      call dumpd(od, 'prim0', prim, nvar*np)
      call dumpd(od, 'nut0', nut, np)
      call residu(coord, elarea, ds, prim, qx, qy, mu, mul,
     &            nut, mcarea, res)
      call dumpd(od, 'res0', res, nvar*np)

      iter  = iterlast
      fres  = MINRESIDUE + 1.0d0
      fres1 = 0.0d0
      nhist = 0
      open(unit=31, file=trim(od)//'/hist.bin', access='stream',
     &     form='unformatted', status='replace')
      do while(fres .gt. MINRESIDUE .and. iter .lt. iterlast+nit)
         call save_old(prim, prim_old)
         call time_step(prim, mu, drmin, dt)
         do irk=1,NIRK
            call residu(coord, elarea, ds, prim, qx, qy, mu, mul,
     &                  nut, mcarea, res)
            call update(irk, res, prim_old, prim, cvarea, dt)
         enddo
         if(iflow .eq. turbulent)then
            call sa_model(coord, elarea, cvarea, prim, prim_old, wd,
     &                    nut, mul, ds, dt, qx, qy)
         endif
         call check_positivity(prim, nut, coord)
         call clcd(coord, prim, mu, qx, qy)
         call conres(prim, prim_old, dt)
         hrow(1) = dble(iter)
         hrow(2) = fres
         hrow(3) = dble(iresi)
         hrow(4) = fresi
         hrow(5) = cl
         hrow(6) = cd
         write(31) hrow
         nhist = nhist + 1
         if(iter .eq. iterlast)then
            call dumpd(od, 'prim1', prim, nvar*np)
            call dumpd(od, 'nut1', nut, np)
            call dumpd(od, 'dt1', dt, np)
         endif
         iter = iter + 1
      enddo
      close(31)

C     the reference's own output at the end of the run (main.f:79-84):
C     WALL.DAT and PARAM.DAT, with the last stage's gradients and viscosity
      call write_result(prim, nut, mu, qx, qy, coord, dsb)
      call finalize(prim, nut, coord)

      call dumpd(od, 'primN', prim, nvar*np)
      call dumpd(od, 'nutN', nut, np)
      call residu(coord, elarea, ds, prim, qx, qy, mu, mul,
     &            nut, mcarea, res)
      call dumpd(od, 'resN', res, nvar*np)
      call dumpd(od, 'qxN', qx, nvar*np)
      call dumpd(od, 'qyN', qy, nvar*np)
      call dumpd(od, 'muN', mu, np)
      call dumpd(od, 'mulN', mul, np)

      stop
      end

      subroutine dumpd(od, name, a, n)
      implicit none
      character*(*)    od, name
      integer          n
      double precision a(n)
      open(unit=32, file=trim(od)//'/'//name//'.bin', access='stream',
     &     form='unformatted', status='replace')
      write(32) a
      close(32)
      return
      end

      subroutine dumpi(od, name, a, n)
      implicit none
      character*(*)    od, name
      integer          n, a(n)
      open(unit=32, file=trim(od)//'/'//name//'.bin', access='stream',
     &     form='unformatted', status='replace')
      write(32) a
      close(32)
      return
      end
