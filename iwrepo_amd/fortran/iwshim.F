c-----------------------------------------------------------------------
c     Fortran drop-in shims of the mode1_bbl hot-path routines.
c
c     Each subroutine keeps the name, argument list and COMMON-block
c     contract of the reference routine it replaces (file:line below),
c     so a caller such as main.f links against these objects in place
c     of the reference's and runs unchanged.  The work is done on the
c     GPU by libiwgpu.so through its C-ABI (include/iwgpu.h).
c
c     The shims are built with the caller's own include path, because
c     the reference sizes its COMMON blocks at compile time (dim.h), and
c     with the caller's REAL size:
c        flang -cpp -fdefault-real-8 -fdefault-double-8 -I<src> -c iwshim.F
c     Define IW_NOSLIP when the caller is built with calcuvwt_noslip.f
c     (no-slip bottom) rather than calcuvwt.f (free slip).
c
c     Data flow: each call copies the host arrays the reference routine
c     reads to the device instance and copies back those it writes, so
c     on return the caller's arrays hold what the reference routine
c     would have left there.  One device instance is created on the
c     first call from the parameters then held in COMMON.
c-----------------------------------------------------------------------
      module iwshim
      use iso_c_binding
      implicit none
      integer, parameter :: iw_maxsub = 64
c     layout of iw_params (include/iwgpu.h)
      type, bind(c) :: iw_params_t
        integer(c_int) :: nx, ny, nzloc, nsubd, bc_uv_bottom
        integer(c_int) :: ispecfilternl, ispecfilterexp
        integer(c_int) :: nspecfilterexp, navgintrfc
        integer(c_int) :: ifilter, nstepfilter, npostp
        integer(c_int) :: imaxcalc, dtadapt, n3doutput, ncalclegspec
        real(c_double) :: xlen, ylen, zlen
        real(c_double) :: xnu, xkappa, brunt, grav, rho0, omz
        real(c_double) :: dt
        real(c_double) :: amp, p, pf, fackc
        real(c_double) :: facdir, facneumn, facrobin
        real(c_double) :: zcen, xbcrit
        real(c_double) :: xacrit, umax, fwave
        real(c_double) :: fackcbv
        real(c_double) :: fcshear, uo
        real(c_double) :: z0(iw_maxsub), zh(iw_maxsub)
      end type
      type(c_ptr), save :: iwh = c_null_ptr
      interface
        integer(c_int) function iw_params_size() bind(c)
          import
        end function
        type(c_ptr) function iw_last_error() bind(c)
          import
        end function
        integer(c_int) function iw_create(prm, h) bind(c)
          import
          type(iw_params_t), intent(in) :: prm
          type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function iw_set_state(h, idx, a) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: idx
          real(c_double) :: a(*)
        end function
        integer(c_int) function iw_get_state(h, idx, a) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: idx
          real(c_double) :: a(*)
        end function
        integer(c_int) function iw_set_pressbc(h, a, b) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: a(*), b(*)
        end function
        integer(c_int) function iw_get_pressbc(h, a, b) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: a(*), b(*)
        end function
        integer(c_int) function iw_set_coeffs(h, t, fac, bdf, bdfv)
     &       bind(c)
          import
          type(c_ptr), value :: h
          real(c_double), value :: t, bdfv
          real(c_double) :: fac(3), bdf(3)
        end function
        integer(c_int) function iw_set_dt(h, dt) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double), value :: dt
        end function
        integer(c_int) function iw_horfft(h, b, is, nk) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: b(*)
          integer(c_int), value :: is, nk
        end function
        integer(c_int) function iw_dpdz(h, a, b) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: a(*), b(*)
        end function
        integer(c_int) function iw_penalty_params(h, eps, ibot, itop,
     &       al, be, ga, de, t1, t2) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double), value :: eps
          integer(c_int), value :: ibot, itop
          real(c_double) :: al(*), be(*), ga(*), de(*), t1(*), t2(*)
        end function
        integer(c_int) function iw_param_set(h, ifbc, eps, ybar,
     &       a0, b0, t0) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: ifbc
          real(c_double), value :: eps, ybar
          real(c_double) :: a0, b0, t0
        end function
        integer(c_int) function iw_penalty_solve(h, eps, ibot, itop,
     &       brb, bib, brt, bit, rr, ri, fr, fi) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double), value :: eps, brb, bib, brt, bit
          integer(c_int), value :: ibot, itop
          real(c_double) :: rr(*), ri(*), fr(*), fi(*)
        end function
        integer(c_int) function iw_filter_3dlegenfourier(h, f, ifo, ile)
     &       bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: f(*)
          integer(c_int), value :: ifo, ile
        end function
        integer(c_int) function iw_interface_avg3d(h, f, ic) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: f(*)
          integer(c_int) :: ic
        end function
        integer(c_int) function iw_convec(h, istep, inl) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: istep, inl
        end function
        integer(c_int) function iw_calcuvwt(h, icpen) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: icpen
        end function
        integer(c_int) function iw_filtering(h, isgs) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: isgs
        end function
        integer(c_int) function iw_get_sgs(h, idx, a) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: idx
          real(c_double) :: a(*)
        end function
        integer(c_int) function iw_set_sgs(h, idx, a) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: idx
          real(c_double) :: a(*)
        end function
        integer(c_int) function iw_postp(h, e) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: e(*)
        end function
        integer(c_int) function iw_speclegen(h, idx, sp) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: idx
          real(c_double) :: sp(*)
        end function
        integer(c_int) function iw_get_coeffs(h, t, fac, bdf, bdfv)
     &       bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: t, bdfv
          real(c_double) :: fac(3), bdf(3)
        end function
        integer(c_int) function iw_set_meanwave(h, mw) bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: mw(*)
        end function
        integer(c_int) function iw_maxcalc(h, istep, rec) bind(c)
          import
          type(c_ptr), value :: h
          integer(c_int), value :: istep
          real(c_double) :: rec(*)
        end function
        integer(c_int) function iw_get_dtstate(h, dt, cdt, iv, idt)
     &       bind(c)
          import
          type(c_ptr), value :: h
          real(c_double) :: dt, cdt
          integer(c_int) :: iv, idt
        end function
        integer(c_int) function iw_set_dtstate(h, cdt, iv, idt)
     &       bind(c)
          import
          type(c_ptr), value :: h
          real(c_double), value :: cdt
          integer(c_int), value :: iv, idt
        end function
        real(c_double) function iw_abvar1(a, b, c) bind(c)
          import
          real(c_double), value :: a, b, c
        end function
        real(c_double) function iw_abvar2(a, b, c) bind(c)
          import
          real(c_double), value :: a, b, c
        end function
        real(c_double) function iw_abvar3(a, b, c) bind(c)
          import
          real(c_double), value :: a, b, c
        end function
        real(c_double) function iw_bdfvarv(a, b, c) bind(c)
          import
          real(c_double), value :: a, b, c
        end function
        real(c_double) function iw_bdfvar1(a, b, c) bind(c)
          import
          real(c_double), value :: a, b, c
        end function
        real(c_double) function iw_bdfvar2(a, b, c) bind(c)
          import
          real(c_double), value :: a, b, c
        end function
        real(c_double) function iw_bdfvar3(a, b, c) bind(c)
          import
          real(c_double), value :: a, b, c
        end function
      end interface
      contains
c     a non-zero return code of the C-ABI stops the program with the
c     library's message (the reference has no error returns)
        subroutine iwshim_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (rc == 0) return
        call c_f_pointer(iw_last_error(), msg, [1024])
        n = 0
        do while (n < 1024)
          if (msg(n + 1) == c_null_char) exit
          n = n + 1
        end do
        write(*, '(3a,1024a1)') 'iwshim: ', what, ': ', msg(1:n)
        error stop 1
        end subroutine
      end module iwshim

c-----------------------------------------------------------------------
c     Device instance from the caller's dim.h and COMMON parameters
c     (created once; the same values setup and main.f hand to the
c     reference routines).
c-----------------------------------------------------------------------
      subroutine iwshim_init
      use iwshim
      include 'dim.h'
      include 'comparamphys.h'
      include 'comsubd.h'
      include 'comtime.h'
      include 'comspecfilter.h'
      include 'compencoeff.h'
      include 'comcritlayer.h'
      include 'comsoliton.h'
      common /dtanimation/ dtanim, fwave
      type(iw_params_t) :: iwp
      integer ks
      if (c_associated(iwh)) return
      if (iw_params_size() .ne. c_sizeof(iwp)) then
        write(*,*) 'iwshim: iw_params layout differs from libiwgpu'
        error stop 1
      endif
      if (nsubd .gt. iw_maxsub) then
        write(*,*) 'iwshim: nsubd exceeds ', iw_maxsub
        error stop 1
      endif
c     mean-u damping width: calcuvwt.f:343-361 uses zdim, set to xlen
c     by main.f:339
      if (zdim .ne. xlen) then
        write(*,*) 'iwshim: zdim must equal xlen (main.f:339)'
        error stop 1
      endif
      iwp%nx = nx
      iwp%ny = ny
      iwp%nzloc = nzloc
      iwp%nsubd = nsubd
#ifdef IW_NOSLIP
      iwp%bc_uv_bottom = 1
#else
      iwp%bc_uv_bottom = 0
#endif
c     the step-loop flags are arguments of convec/calcuvwt here
      iwp%ispecfilternl = 1
      iwp%ispecfilterexp = 1
      iwp%nspecfilterexp = 1
      iwp%navgintrfc = 1
c     the estimation filter is an argument of filtering here
      iwp%ifilter = 0
      iwp%nstepfilter = 1
      iwp%npostp = 1
c     maxcalc is its own shim; adaptive dt as in various.f:911-985 with
c     Tso kept between calls
      iwp%imaxcalc = 1
      iwp%dtadapt = 1
c     the step loop (and its plotf calls) is main.f's own here
      iwp%n3doutput = 0
      iwp%ncalclegspec = 0
      iwp%xlen = xlen
      iwp%ylen = ylen
      iwp%zlen = zlen
      iwp%xnu = xnu
      iwp%xkappa = xkappa
      iwp%brunt = brunt
      iwp%grav = grav
      iwp%rho0 = rho0
      iwp%omz = omz
      iwp%dt = dt
      iwp%amp = amp
      iwp%p = p
      iwp%pf = pf
      iwp%fackc = fackc
      iwp%facdir = facdir
      iwp%facneumn = facneumn
      iwp%facrobin = facrobin
      iwp%zcen = zcen
      iwp%xbcrit = xbcrit
      iwp%xacrit = xacrit
      iwp%umax = umax
      iwp%fwave = fwave
      iwp%fackcbv = fackcbv
      iwp%fcshear = fcshear
      iwp%uo = uo
      iwp%z0 = 0.d0
      iwp%zh = 0.d0
      do ks = 1, nsubd
        iwp%z0(ks) = z0(ks)
        iwp%zh(ks) = zh(ks)
      enddo
      call iwshim_check(iw_create(iwp, iwh), 'iw_create')
c     the nine /sol2dvar/ arrays follow umwv in the COMMON block
      call iwshim_check(iw_set_meanwave(iwh, umwv), 'meanwave')
      end

c     /flow/ and /matrix/ (comflow.h, commatrix.h) to the device
      subroutine iwshim_put_state
      use iwshim
      include 'dim.h'
      include 'comflow.h'
      include 'commatrix.h'
      call iwshim_check(iw_set_state(iwh, 0, u), 'state')
      call iwshim_check(iw_set_state(iwh, 1, v), 'state')
      call iwshim_check(iw_set_state(iwh, 2, w), 'state')
      call iwshim_check(iw_set_state(iwh, 3, temp), 'state')
      call iwshim_check(iw_set_state(iwh, 4, un), 'state')
      call iwshim_check(iw_set_state(iwh, 5, vn), 'state')
      call iwshim_check(iw_set_state(iwh, 6, wn), 'state')
      call iwshim_check(iw_set_state(iwh, 7, tn), 'state')
      call iwshim_check(iw_set_state(iwh, 8, unm), 'state')
      call iwshim_check(iw_set_state(iwh, 9, vnm), 'state')
      call iwshim_check(iw_set_state(iwh, 10, wnm), 'state')
      call iwshim_check(iw_set_state(iwh, 11, tnm), 'state')
      call iwshim_check(iw_set_state(iwh, 12, su), 'state')
      call iwshim_check(iw_set_state(iwh, 13, sv), 'state')
      call iwshim_check(iw_set_state(iwh, 14, sw), 'state')
      call iwshim_check(iw_set_state(iwh, 15, st), 'state')
      call iwshim_check(iw_set_state(iwh, 16, sun), 'state')
      call iwshim_check(iw_set_state(iwh, 17, svn), 'state')
      call iwshim_check(iw_set_state(iwh, 18, swn), 'state')
      call iwshim_check(iw_set_state(iwh, 19, stn), 'state')
      call iwshim_check(iw_set_state(iwh, 20, sunm), 'state')
      call iwshim_check(iw_set_state(iwh, 21, svnm), 'state')
      call iwshim_check(iw_set_state(iwh, 22, swnm), 'state')
      call iwshim_check(iw_set_state(iwh, 23, stnm), 'state')
      end

c     device to /flow/ and /matrix/
      subroutine iwshim_get_state
      use iwshim
      include 'dim.h'
      include 'comflow.h'
      include 'commatrix.h'
      call iwshim_check(iw_get_state(iwh, 0, u), 'state')
      call iwshim_check(iw_get_state(iwh, 1, v), 'state')
      call iwshim_check(iw_get_state(iwh, 2, w), 'state')
      call iwshim_check(iw_get_state(iwh, 3, temp), 'state')
      call iwshim_check(iw_get_state(iwh, 4, un), 'state')
      call iwshim_check(iw_get_state(iwh, 5, vn), 'state')
      call iwshim_check(iw_get_state(iwh, 6, wn), 'state')
      call iwshim_check(iw_get_state(iwh, 7, tn), 'state')
      call iwshim_check(iw_get_state(iwh, 8, unm), 'state')
      call iwshim_check(iw_get_state(iwh, 9, vnm), 'state')
      call iwshim_check(iw_get_state(iwh, 10, wnm), 'state')
      call iwshim_check(iw_get_state(iwh, 11, tnm), 'state')
      call iwshim_check(iw_get_state(iwh, 12, su), 'state')
      call iwshim_check(iw_get_state(iwh, 13, sv), 'state')
      call iwshim_check(iw_get_state(iwh, 14, sw), 'state')
      call iwshim_check(iw_get_state(iwh, 15, st), 'state')
      call iwshim_check(iw_get_state(iwh, 16, sun), 'state')
      call iwshim_check(iw_get_state(iwh, 17, svn), 'state')
      call iwshim_check(iw_get_state(iwh, 18, swn), 'state')
      call iwshim_check(iw_get_state(iwh, 19, stn), 'state')
      call iwshim_check(iw_get_state(iwh, 20, sunm), 'state')
      call iwshim_check(iw_get_state(iwh, 21, svnm), 'state')
      call iwshim_check(iw_get_state(iwh, 22, swnm), 'state')
      call iwshim_check(iw_get_state(iwh, 23, stnm), 'state')
      end

c     clock, dt and the SS/BDF coefficients of /timevars/, /ab/, /bdf/
      subroutine iwshim_put_coeffs
      use iwshim
      include 'dim.h'
      include 'comtime.h'
      include 'comtimeadvance.h'
      real(c_double) :: fac(3), bdf(3)
      fac(1) = fac1
      fac(2) = fac2
      fac(3) = fac3
      bdf(1) = bdf1
      bdf(2) = bdf2
      bdf(3) = bdf3
      call iwshim_check(iw_set_dt(iwh, dt), 'dt')
      call iwshim_check(iw_set_coeffs(iwh, t, fac, bdf, bdfv), 'coeffs')
      end

c-----------------------------------------------------------------------
c     horfft.f:2   subroutine horfft(b,is,nk)
c-----------------------------------------------------------------------
      subroutine horfft(b, is, nk)
      use iwshim
      dimension b(*)
      call iwshim_init
      call iwshim_check(iw_horfft(iwh, b, is, nk), 'horfft')
      end

c-----------------------------------------------------------------------
c     derv.f:212   subroutine dpdz(a,b)
c-----------------------------------------------------------------------
      subroutine dpdz(a, b)
      use iwshim
      dimension a(*), b(*)
      call iwshim_init
      call iwshim_check(iw_dpdz(iwh, a, b), 'dpdz')
      end

c-----------------------------------------------------------------------
c     filtering.f:811   subroutine filter_3dlegenfourier(fin3d,
c     ifourier,ilegen,iout); iout only selects printing there
c-----------------------------------------------------------------------
      subroutine filter_3dlegenfourier(fin3d, ifourier, ilegen, iout)
      use iwshim
      dimension fin3d(*)
      call iwshim_init
      call iwshim_check(iw_filter_3dlegenfourier(iwh, fin3d, ifourier,
     &     ilegen), 'filter_3dlegenfourier')
      end

c-----------------------------------------------------------------------
c     penalty.f:714   subroutine interface_avg3d(fin,ic)
c-----------------------------------------------------------------------
      subroutine interface_avg3d(fin, ic)
      use iwshim
      dimension fin(*)
      integer(c_int) :: icc
      call iwshim_init
      call iwshim_check(iw_interface_avg3d(iwh, fin, icc),
     &     'interface_avg3d')
      ic = icc
      end

c-----------------------------------------------------------------------
c     solver.f:6   subroutine convec(istep,istart,ispecfilternl)
c     reads /flow/, /matrix/, /timevars/, /ab/, /bdf/; writes /flow/,
c     /matrix/ (new NL terms, BDF-assembled fields, shifted histories)
c     and /pressbc/ (comrotpressbc.h).  istart only selects printing.
c-----------------------------------------------------------------------
      subroutine convec(istep, istart, ispecfilternl)
      use iwshim
      include 'dim.h'
      include 'comrotpressbc.h'
      call iwshim_init
      call iwshim_put_state
      call iwshim_put_coeffs
      call iwshim_check(iw_convec(iwh, istep, ispecfilternl), 'convec')
      call iwshim_get_state
      call iwshim_check(iw_get_pressbc(iwh, wnl, wviscrot), 'convec')
      end

c-----------------------------------------------------------------------
c     calcuvwt.f:2 / calcuvwt_noslip.f:2
c     subroutine calcuvwt(izmode,ispecfilter,irelax,cpenspecfilter)
c     izmode, ispecfilter and irelax are not read by the reference.
c-----------------------------------------------------------------------
      subroutine calcuvwt(izmode, ispecfilter, irelax, cpenspecfilter)
      use iwshim
      include 'dim.h'
      include 'comrotpressbc.h'
      logical cpenspecfilter
      integer(c_int) :: icpen
      call iwshim_init
      call iwshim_put_state
      call iwshim_put_coeffs
      call iwshim_check(iw_set_pressbc(iwh, wnl, wviscrot), 'calcuvwt')
      icpen = 0
      if (cpenspecfilter) icpen = 1
      call iwshim_check(iw_calcuvwt(iwh, icpen), 'calcuvwt')
      call iwshim_get_state
      end

c-----------------------------------------------------------------------
c     various.f:665   subroutine maxcalc(istep)
c     the extrema and CFL maxima of the physical u,v,w,temp, the horfft
c     round trip of /flow/ u,v,w,temp and the adaptive step of
c     various.f:887-985 (dt, cdt, /ab/, /bdf/, /fachelm/, /vardtflags/)
c     on the device; the reference's outputs (uveltest.dat, units 50, 51
c     and 65, the printed lines, various.f:744-985) in its formats
c-----------------------------------------------------------------------
      subroutine maxcalc(istep)
      use iwshim
      include 'dim.h'
      include 'comflow.h'
      include 'comtime.h'
      include 'comtimeadvance.h'
      include 'comparamphys.h'
      real(c_double) :: rec(42), fac(3), bdf(3), tt, bv, dtn, cdtn
      integer(c_int) :: iv, idt
      integer ir(42), m
      call iwshim_init
      call iwshim_check(iw_set_state(iwh, 0, u), 'maxcalc')
      call iwshim_check(iw_set_state(iwh, 1, v), 'maxcalc')
      call iwshim_check(iw_set_state(iwh, 2, w), 'maxcalc')
      call iwshim_check(iw_set_state(iwh, 3, temp), 'maxcalc')
      call iwshim_put_coeffs
      call iwshim_check(iw_set_dtstate(iwh, cdt, ivardt, idtstep),
     &     'maxcalc')
      call iwshim_check(iw_maxcalc(iwh, istep, rec), 'maxcalc')
      call iwshim_check(iw_get_state(iwh, 0, u), 'maxcalc')
      call iwshim_check(iw_get_state(iwh, 1, v), 'maxcalc')
      call iwshim_check(iw_get_state(iwh, 2, w), 'maxcalc')
      call iwshim_check(iw_get_state(iwh, 3, temp), 'maxcalc')
      do m = 1, 42
        ir(m) = 0
        if (rec(m) .eq. rec(m)) ir(m) = nint(rec(m))
      enddo
      open(700, file='uveltest.dat', position='append')
      write(700,*) t, rec(41), rec(42)
      close(700)
      write(50,'(1x,I5,2x,4(E12.6,1x,E12.6,1x,I3,1x,I3,2x))')
     &  istep, rec(3), rec(4), ir(5), ir(6), rec(7), rec(8), ir(9),
     &  ir(10), rec(11), rec(12), ir(13), ir(14), rec(15), rec(16),
     &  ir(17), ir(18)
      write(*,'(1x,I5,2x,e12.6,2x,e12.6,2x,i3,2x,i3)')
     &  istep-1, rec(15), rec(16), ir(17), ir(18)
      write(51,'(1x,e12.6,2x,3(1x,e12.6,2x,3(I5,2x)))')
     &  t, rec(23), ir(24), ir(25), ir(26), rec(27), ir(28), ir(29),
     &  ir(30), rec(31), ir(32), ir(33), ir(34)
      write(*,*) 'Temp. min at i,j,k', ir(19), ir(20), ir(17)
      write(*,*) 'Temp. max at i,j,k', ir(21), ir(22), ir(18)
      write(*,*) 'At max CFLZ dz= ', rec(35), rec(31)
      if (ir(38) .ne. 0) then
        call iwshim_check(iw_get_dtstate(iwh, dtn, cdtn, iv, idt),
     &       'maxcalc')
        call iwshim_check(iw_get_coeffs(iwh, tt, fac, bdf, bv),
     &       'maxcalc')
        dt = dtn
        cdt = cdtn
        ivardt = iv
        idtstep = idt
        fac1 = fac(1)
        fac2 = fac(2)
        fac3 = fac(3)
        bdf1 = bdf(1)
        bdf2 = bdf(2)
        bdf3 = bdf(3)
        bdfv = bv
        fact = 1./(xkappa*dt)
        facvel = 1./(pr*dt)
        write(65,*) t, dt
        write(*,*) 'At time= ', t
        if (cdt .gt. 1.) write(*,*)
     &    'TIMESTEP *REDUCED* by a factor of 4/5 to dt=', dt
        if (cdt .lt. 1.) write(*,*)
     &    'TIMESTEP *INCREASED* by a factor of 5/4 to dt=', dt
      endif
      end

c-----------------------------------------------------------------------
c     filtering.f:8   subroutine filtering(isgscalc)
c     estimation-model filter of /flow/ u,v,w; with isgscalc = 1 the SGS
c     tensors of /comsgs/ (comsgs.h, sgscalc.f:4-223) as well
c-----------------------------------------------------------------------
      subroutine filtering(isgscalc)
      use iwshim
      include 'dim.h'
      include 'comsgs.h'
      call iwshim_init
      call iwshim_put_state
      call iwshim_check(iw_filtering(iwh, isgscalc), 'filtering')
      call iwshim_get_state
      if (isgscalc .eq. 1) then
        call iwshim_check(iw_get_sgs(iwh, 0, t11), 'filtering')
        call iwshim_check(iw_get_sgs(iwh, 1, t12), 'filtering')
        call iwshim_check(iw_get_sgs(iwh, 2, t13), 'filtering')
        call iwshim_check(iw_get_sgs(iwh, 3, t22), 'filtering')
        call iwshim_check(iw_get_sgs(iwh, 4, t23), 'filtering')
        call iwshim_check(iw_get_sgs(iwh, 5, t33), 'filtering')
        call iwshim_check(iw_get_sgs(iwh, 6, th1), 'filtering')
        call iwshim_check(iw_get_sgs(iwh, 7, th2), 'filtering')
        call iwshim_check(iw_get_sgs(iwh, 8, th3), 'filtering')
      endif
      end

c-----------------------------------------------------------------------
c     postp.f:9   subroutine postp(initenergy,t,brunt,tstart)
c     the TKE budget of energycalc_postp (postp.f:415-732) of /flow/
c     u,v,w,temp with the /comsgs/ tensors, on the device: /energetics/
c     and /energeticsw/ (comenergy.h, comenergyw.h), the state after
c     postp's horfft round trips and the reference's output (units 60,
c     61, 62 and the printed lines, postp.f:21, 53-73, 724-726).  The
c     wake-interior values are 0, as in the reference, where the wake
c     sizes of /lengthscales/ are never set (wakelengthcalc is not
c     called, postp.f:39-41).  The x-means of /meanv/ (commeanv.h),
c     read only by the relaxation procedure, are not returned.
c-----------------------------------------------------------------------
      subroutine postp(initenergy, t, brunt, tstart)
      use iwshim
      include 'dim.h'
      include 'comflow.h'
      include 'comsgs.h'
      include 'comenergy.h'
      include 'comenergyw.h'
      include 'comlengthsc.h'
      real(c_double) :: e(16)
      call iwshim_init
      write(*,*) 'POSTPROCESSING DATA'
      call iwshim_check(iw_set_state(iwh, 0, u), 'postp')
      call iwshim_check(iw_set_state(iwh, 1, v), 'postp')
      call iwshim_check(iw_set_state(iwh, 2, w), 'postp')
      call iwshim_check(iw_set_state(iwh, 3, temp), 'postp')
      call iwshim_check(iw_set_sgs(iwh, 0, t11), 'postp')
      call iwshim_check(iw_set_sgs(iwh, 1, t12), 'postp')
      call iwshim_check(iw_set_sgs(iwh, 2, t13), 'postp')
      call iwshim_check(iw_set_sgs(iwh, 3, t22), 'postp')
      call iwshim_check(iw_set_sgs(iwh, 4, t23), 'postp')
      call iwshim_check(iw_set_sgs(iwh, 5, t33), 'postp')
      call iwshim_check(iw_set_sgs(iwh, 6, th1), 'postp')
      call iwshim_check(iw_set_sgs(iwh, 7, th2), 'postp')
      call iwshim_check(iw_set_sgs(iwh, 8, th3), 'postp')
      call iwshim_check(iw_postp(iwh, e), 'postp')
      call iwshim_check(iw_get_state(iwh, 0, u), 'postp')
      call iwshim_check(iw_get_state(iwh, 1, v), 'postp')
      call iwshim_check(iw_get_state(iwh, 2, w), 'postp')
      call iwshim_check(iw_get_state(iwh, 3, temp), 'postp')
      tketot = e(1)
      u1tke = e(2)
      u2tke = e(3)
      u3tke = e(4)
      prody = e(5)
      prodz = e(6)
      prod = e(7)
      bflux = e(8)
      epsilon = e(9)
      epssgs = e(10)
      epstot = e(11)
      chi = e(12)
      chisgs = e(13)
      chitot = e(14)
      tkemean = e(15)
      tempvar = e(16)
      tketotw = 0.
      u1tkew = 0.
      u2tkew = 0.
      u3tkew = 0.
      prodyw = 0.
      prodzw = 0.
      prodw = 0.
      bfluxw = 0.
      epsilonw = 0.
      epssgsw = 0.
      epstotw = 0.
      tempvarw = 0.
      chiw = 0.
      chisgsw = 0.
      chitotw = 0.
      write(*,*) 'DISSIPATION RATES',epssgs,epsilon
      write(*,*) 'Scalar DRs',chisgs,chi
      write(60,'(1x,F10.4,17(2x,E12.6))')
     &  (t-tstart),(t-tstart)*brunt,
     &  tketot,u1tke,u2tke,u3tke,prody,prodz,prod,
     &  bflux,epsilon,epssgs,epstot,
     &  chi,chisgs,chitot,
     &  tkemean,tempvar
      write(61,'(1x,F10.4,17(2x,E12.6))')
     &  (t-tstart),(t-tstart)*brunt,
     &  tketotw,u1tkew,u2tkew,u3tkew,prodyw,prodzw,prodw,
     &  bfluxw,epsilonw,epssgsw,epstotw,
     &  chiw,chisgsw,chitotw,
     &  tkemean,tempvarw
      write(62,'(1x,F10.4,8(2x,E12.6))')
     &  (t-tstart),(t-tstart)*brunt,
     &  (udefy+udefz)/2.,udefy,udefz,wakely,wakelz,yc,zc
      write(*,*) 'Dissipation rate estimate#2',epsilon,epssgs
      end

c-----------------------------------------------------------------------
c     filtering.f:1110   subroutine speclegencalc(filein,t,fin)
c     the x-averaged Legendre spectra of fin at j = ny/2 on the device,
c     written to speclegen<filein><t>.dat in the reference's name and
c     format (filtering.f:1131-1141, 1194-1216), and fin after its
c     horfft round trip.  The device's u slot carries fin: every shim
c     loads the state it reads, so nothing else depends on that slot.
c-----------------------------------------------------------------------
      subroutine speclegencalc(filein, t, fin)
      use iwshim
      include 'dim.h'
      character*1 filein
      character*30 fileout
      character*10 fmt1
      dimension fin(nxpp,ny,nz)
      real(c_double) :: sp(nzloc*nsubd)
      integer kloc, ks
      call iwshim_init
      call iwshim_check(iw_set_state(iwh, 0, fin), 'speclegencalc')
      call iwshim_check(iw_speclegen(iwh, 0, sp), 'speclegencalc')
      call iwshim_check(iw_get_state(iwh, 0, fin), 'speclegencalc')
      fileout = 'speclegen'
      write(unit=fileout(10:10),fmt='(a1)') filein
      fmt1 = '(f5.3,a4)'
      if (nint(t).ge.10) fmt1 = '(f6.3,a4)'
      if (nint(t).ge.100) fmt1 = '(f7.3,a4)'
      if (nint(t).ge.1000) fmt1 = '(f8.3,a4)'
      write(unit=fileout(11:23),fmt=fmt1) t,'.dat'
      write(*,*) fileout
      open(405,file=fileout)
      do kloc = 1, nzloc
        write(405,150) kloc, (sp((kloc-1)*nsubd+ks), ks=1,nsubd)
      enddo
      close(405)
 150  format (1x,I2,2x,10(e12.6,2x))
      end

c-----------------------------------------------------------------------
c     multistep.f:7-89 variable-step AB3/BDF3 coefficients
c-----------------------------------------------------------------------
c-----------------------------------------------------------------------
c     penalty.f:9   subroutine penalty_setup(epsilon): the penalty
c     parameters of the interior subdomains into /penaltyparams/ and
c     omega into /comomega/ (comparampen.h)
c-----------------------------------------------------------------------
      subroutine penalty_setup(epsilon)
      use iwshim
      include 'dim.h'
      include 'comparampen.h'
      real(c_double) :: pa(nsubd), pb(nsubd), pg(nsubd), pd(nsubd)
      real(c_double) :: pt1(nsubd), pt2(nsubd)
      integer ks
      call iwshim_init
      call iwshim_check(iw_penalty_params(iwh, epsilon, 2, 2, pa, pb,
     &     pg, pd, pt1, pt2), 'penalty_setup')
      omega = 2./((nzloc+1)*(nzloc+2))
      do ks = 2, nsubd - 1
        alpha(ks) = pa(ks)
        beta(ks) = pb(ks)
        tau1(ks) = pt1(ks)
        gamma(ks) = pg(ks)
        delta(ks) = pd(ks)
        tau2(ks) = pt2(ks)
      enddo
      end

c-----------------------------------------------------------------------
c     penalty.f:57  subroutine param_set(ifbc,alpha0,beta0,tau0,
c     epsilon,ybar,ks)
c-----------------------------------------------------------------------
      subroutine param_set(ifbc, alpha0, beta0, tau0, epsilon, ybar, ks)
      use iwshim
      call iwshim_init
      call iwshim_check(iw_param_set(iwh, ifbc, epsilon, ybar, alpha0,
     &     beta0, tau0), 'param_set')
      end

c-----------------------------------------------------------------------
c     penalty.f:162 subroutine penalty_main(epsilon,ifuncreal,bcrbot,
c     bcibot,bcrtop,bcitop,ifbcbot,ifbctop,rhsinr,rhsini,fr,fi): the
c     right-hand sides times -epsilon in place, the boundary parameters
c     into /penaltyparams/, then helmv_penalty
c-----------------------------------------------------------------------
      subroutine penalty_main(epsilon, ifuncreal, bcrbot, bcibot,
     &     bcrtop, bcitop, ifbcbot, ifbctop, rhsinr, rhsini, fr, fi)
      use iwshim
      include 'dim.h'
      include 'comparampen.h'
      dimension fr(nz), fi(nz), rhsinr(nz), rhsini(nz)
      real(c_double) :: pa(nsubd), pb(nsubd), pg(nsubd), pd(nsubd)
      real(c_double) :: pt1(nsubd), pt2(nsubd)
      call iwshim_init
      do k = 1, nz
        rhsinr(k) = -rhsinr(k)*epsilon
        rhsini(k) = -rhsini(k)*epsilon
      enddo
      call iwshim_check(iw_penalty_params(iwh, epsilon, ifbcbot,
     &     ifbctop, pa, pb, pg, pd, pt1, pt2), 'penalty_main')
      alpha(1) = pa(1)
      beta(1) = pb(1)
      tau1(1) = pt1(1)
      if (nsubd .gt. 1) then
        gamma(1) = pg(1)
        delta(1) = pd(1)
        tau2(1) = pt2(1)
      endif
      gamma(nsubd) = pg(nsubd)
      delta(nsubd) = pd(nsubd)
      tau2(nsubd) = pt2(nsubd)
      if (nsubd .gt. 1) then
        alpha(nsubd) = pa(nsubd)
        beta(nsubd) = pb(nsubd)
        tau1(nsubd) = pt1(nsubd)
      endif
      call helmv_penalty(fr, fi, epsilon, rhsinr, rhsini, bcrtop,
     &     bcitop, bcrbot, bcibot, ifbctop, ifbcbot, ifuncreal)
      end

c-----------------------------------------------------------------------
c     penalty.f:255 subroutine helmv_penalty(fvr,fvi,epsilon,rhsinr,
c     rhsini,bcrtop,bcitop,bcrbot,bcibot,ifbctop,ifbcbot,ifuncreal):
c     assembly (matrix_elmsetup), banded LU and solve (bandlu_solve) on
c     the device.  With ifuncreal = 1 only the real column is solved and
c     fvi keeps the right-hand side with its boundary terms
c     (band_lu.f:41-43, penalty.f:583-604).
c-----------------------------------------------------------------------
      subroutine helmv_penalty(fvr, fvi, epsilon, rhsinr, rhsini,
     &     bcrtop, bcitop, bcrbot, bcibot, ifbctop, ifbcbot, ifuncreal)
      use iwshim
      include 'dim.h'
      include 'comparampen.h'
      dimension fvr(nz), fvi(nz), rhsinr(nz), rhsini(nz)
      call iwshim_init
      call iwshim_check(iw_penalty_solve(iwh, epsilon, ifbcbot,
     &     ifbctop, bcrbot, bcibot, bcrtop, bcitop, rhsinr, rhsini,
     &     fvr, fvi), 'helmv_penalty')
      if (ifuncreal .eq. 1) then
        do k = 1, nz
          fvi(k) = rhsini(k)
        enddo
        if (ifbcbot .eq. 0) then
          fvi(1) = rhsini(1) + tau1(1)*bcibot
        else
          fvi(1) = rhsini(1) - tau1(1)*bcibot*epsilon
        endif
        if (ifbctop .eq. 0) then
          fvi(nz) = rhsini(nz) + tau2(nsubd)*bcitop
        else
          fvi(nz) = rhsini(nz) + tau2(nsubd)*bcitop*epsilon
        endif
      endif
      end

      real function abvar1(dt1, dt2, dt3)
      use iwshim
      abvar1 = iw_abvar1(dt1, dt2, dt3)
      end
      real function abvar2(dt1, dt2, dt3)
      use iwshim
      abvar2 = iw_abvar2(dt1, dt2, dt3)
      end
      real function abvar3(dt1, dt2, dt3)
      use iwshim
      abvar3 = iw_abvar3(dt1, dt2, dt3)
      end
      real function bdfvarv(dt1, dt2, dt3)
      use iwshim
      bdfvarv = iw_bdfvarv(dt1, dt2, dt3)
      end
      real function bdfvar1(dt1, dt2, dt3)
      use iwshim
      bdfvar1 = iw_bdfvar1(dt1, dt2, dt3)
      end
      real function bdfvar2(dt1, dt2, dt3)
      use iwshim
      bdfvar2 = iw_bdfvar2(dt1, dt2, dt3)
      end
      real function bdfvar3(dt1, dt2, dt3)
      use iwshim
      bdfvar3 = iw_bdfvar3(dt1, dt2, dt3)
      end
