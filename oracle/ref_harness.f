c=======================================================================
c  Test driver for the reference Fortran (oracle/_ref).  TEST
c  INFRASTRUCTURE ONLY: it is compiled against the reference sources
c  where they lie (symlinked by oracle/build_ref.py) and is never part
c  of the product.
c
c  It replays the step sequence of main.f:653-1066 (jmsegret/IWrepo,
c  mode1_bbl) without main.f's file output:
c    * fcshr sets the shear-current factor fcshear (main.f:120 has 0)
c      and mwfile names a stream file with the nine /sol2dvar/ arrays;
c    * with imaxc=0, maxcalc (various.f:665) is replaced by its
c      state-changing part, the horfft(+1)/horfft(-1) round trip of
c      u,v,w,temp (various.f:739-742, 854-857), so dt stays fixed; with
c      imaxc=1 (builds compiled with -DIW_MAXCALC, where various.f is
c      part of the build) the reference maxcalc itself runs, adaptive dt
c      included, and t, dt, cdt after it go to dtlog.txt;
c    * convec, calcuvwt, interface_avg3d are the reference routines;
c    * the state-changing output routines run where main.f calls them:
c      plotf (output.f:125) after the initial transform (main.f:543),
c      after a restart (main.f:607) and every n3doutput steps
c      (main.f:996-1039), its dump going to p_out.dat; in builds
c      compiled with -DIW_POSTP, postp (postp.f:9) after a restart
c      (main.f:635) and every npostp steps (main.f:935), the sixteen
c      /energetics/ values (comenergy.h) going to energy.txt at full
c      precision in the order of postp's fort.60 line; speclegencalc
c      (filtering.f:1110) of u,v,w,temp every ncalclegspec steps
c      (main.f:905-911), writing its speclegen*.dat files;
c    * the SS/BDF coefficient selection follows main.f:660-664,
c      700-777 and 813-856.
c  Without -DIW_MAXCALC, various.f is not compiled (flang rejects it for
c  ny=1, various.f:194-204), so the two one-line helpers that the hot
c  path uses from it, zero and copy (various.f:2-26), are given below.
c
c  Modes (namelist &harness, file harness.nml in the working dir):
c    mode=0  full run: read physical IC, nrun steps, dump state
c    mode=1  horfft forward of the IC fields, dump spectral fields
c    mode=2  dpdz of the IC fields (taken as given), dump
c    mode=3  filter_3dlegenfourier of the IC fields (spectral), dump
c    mode=4  interface_avg3d of the IC fields (spectral), dump + counts
c    mode=5  filtering(isgs) of the IC fields (spectral), dump + the
c            nine SGS tensors of /comsgs/
c    mode=6  postp of the IC fields (spectral; -DIW_POSTP builds),
c            after filtering(1) when isgs=1, energies to energy.txt
c    mode=7  kinetic and potential energy of the IC fields (physical),
c            the z integral by the reference's integrcol (derv.f:479),
c            to kepe.txt
c    mode=8  penalty_setup + penalty_main (testhelm.f:60-100) of IC
c            columns for four (epsilon, boundary types and values),
c            solutions and scaled right-hand sides to the diagnostics file
c  istart=1 (namelist &inputs) in mode 0 restarts from start.dat with
c  restart.f and the first-step coefficients of main.f:619-634, 703-734
c  and 818-825; iwrite=1 writes restart_out.dat with output.f at the end.
c=======================================================================
      program iwref
      include 'dim.h'
      include 'comwave.h'
      include 'comleng.h'
      include 'comtime.h'
      include 'comflow.h'
      include 'comfft.h'
      include 'commatrix.h'
      include 'comparamphys.h'
      include 'comsubd.h'
      include 'comspecfilter.h'
      include 'compencoeff.h'
      include 'comtimeadvance.h'
      include 'comsoliton.h'
      include 'comgridptxyz.h'
      include 'comcritlayer.h'
      include 'comscratch.h'
      include 'comsgs.h'
      include 'comenergy.h'

      logical cpenspecfilter, ldtadapt
      character*256 icfile, outfile, diagfile
      dimension wrk(nxpp,ny,nz)
      dimension rhr8(nz), rhi8(nz), fr8(nz), fi8(nz)

      common /Dump/ Tdump,Tx
      common /anim/ianim,ianimc
      common /Misc/tzmode
      common /domlength/ ZL,Zcur
      common /dtanimation/dtanim,fwave

      namelist/inputs/ istart,dtin,nsteps,tanim,iform,cdtin,waveperiods,
     >fwave,dtanim
      namelist/params/ xlen,ylen,zlen,brunt,xnu,xkappa,grav,
     >expan,umax,rho0,xacrit,xbcrit,zcen,xacu,zsh
      namelist/mflags/irelax,ispecfilternl,
     >ispecfilterexp,nspecfilterexp,amp,p,pf,fackc,fackcbv,
     >facdir,facneumn,facrobin,navgintrfc,ncalclegspec,
     >ifilter,nstepfilter,
     >tstart,npostp,itempreset,nrestartdump,n3doutput,i0
      character*256 mwfile
      namelist/harness/ icfile, outfile, diagfile, nrun, mode, iprobe,
     >ipx, kp1, kp2, isgs, iwrite, imaxc, fcshr, mwfile

      imaxc = 0
      fcshr = 0.
      mwfile = ' '
      mode = 0
      isgs = 0
      iwrite = 0
      iprobe = 0
      ipx = 1
      kp1 = 1
      kp2 = 1
      nrun = 1
      diagfile = 'diag.txt'
      open(7,file='input.dat',status='old')
      read(7,inputs)
      read(7,params)
      read(7,mflags)
      close(7)
      open(7,file='harness.nml',status='old')
      read(7,harness)
      close(7)

c-FFT tables (main.f:210-218)
      call fftfax(nx,ifaxx,trigsx)
      if (ndims.eq.3) then
        call cftfax(ny,ifaxy,trigsy)
        call fftfax(ny,ifaxy1d,trigsy1d)
      endif
      scaley = 1./float(ny)
      pi2 = 8.*atan(1.0)
      ZL = zlen
      Zcur = zsh
c-Derived parameters (main.f:328-352, 416-472)
      alpha = pi2/xlen
      beta  = pi2/ylen
      gamma = pi2/zlen
      zdim = xlen
      xkcrit = pi2/xlen
      xmcrit = pi2/(2.*zlen)
      xomega = (bruntz(0.0,zlen,brunt)*xkcrit)
     >           /sqrt(xkcrit**2+xmcrit**2)
      uo = 0.07
c-main.f:120 sets fcshear = 0; the harness may set the shear current
      fcshear = fcshr
      omz = 0.
      pr = xnu
      dt = dtin
      cdt = cdtin
      fact = 1./(xkappa*dt)
      facvel = 1./(pr*dt)
      ianim = 0
      Tdump = 0.
      izmode = 0
      ispecfilter = 0

      call setup

c-/sol2dvar/ mean-wave arrays (comsoliton.h:9-13), zero in the reference
c-(wavesetup is not called, setup.f:191) unless a restart file holds them
      if ((mwfile.ne.' ').and.(istart.eq.0)) then
        open(10,file=mwfile,form='unformatted',access='stream',
     >       status='old')
        read(10) umwv,wmwv,dumdx,dumdz,dwmdx,dwmdz,rhomwv,drhomdx,
     >           drhomdz
        close(10)
      endif

      if (istart.ne.0) then
c-main.f:563-634: restart file, set-up again, first-step coefficients
        open(8,file='start.dat',form='unformatted',status='old')
        call restart(8,t)
        close(8)
        call setup
        dt1 = dt
        dt2 = cdt*dt
        dt3 = cdt*dt
        fac1 = abvar1(dt1,dt2,dt3)
        fac2 = abvar2(dt1,dt2,dt3)
        fac3 = abvar3(dt1,dt2,dt3)
        bdf1 = bdfvar1(dt1,dt2,dt3)
        bdf2 = bdfvar2(dt1,dt2,dt3)
        bdf3 = bdfvar3(dt1,dt2,dt3)
        bdfv = bdfvarv(dt1,dt2,dt3)
        open(12,file=diagfile,status='unknown')
        open(3,file='p_out.dat',form='unformatted',status='unknown')
        call plotf(3,iform,t)
        close(3)
#ifdef IW_POSTP
        open(14,file='energy.txt',status='unknown')
        initenergy = 1
        call postp(initenergy,t,brunt,tstart)
        call wenergy(0,t)
#endif
        goto 500
      endif

c-Initial condition: 4 physical fields (nxpp,ny,nz), pads ignored
      open(10,file=icfile,form='unformatted',access='stream',
     >     status='old')
      read(10) u
      read(10) v
      read(10) w
      read(10) temp
      close(10)

      open(12,file=diagfile,status='unknown')

      if (mode.eq.1) then
        call horfft(u,-1,nz)
        call horfft(v,-1,nz)
        call horfft(w,-1,nz)
        call horfft(temp,-1,nz)
        goto 900
      endif
      if (mode.eq.2) then
        call dpdz(u,wrk)
        call copy(wrk,u,ntot)
        call dpdz(v,wrk)
        call copy(wrk,v,ntot)
        call dpdz(w,wrk)
        call copy(wrk,w,ntot)
        call dpdz(temp,wrk)
        call copy(wrk,temp,ntot)
        goto 900
      endif
      if (mode.eq.3) then
        call filter_3dlegenfourier(u,1,1,0)
        call filter_3dlegenfourier(v,1,1,0)
        call filter_3dlegenfourier(w,1,1,0)
        call filter_3dlegenfourier(temp,1,1,0)
        goto 900
      endif
      if (mode.eq.5) then
        call filtering(isgs)
        goto 900
      endif
      if (mode.eq.4) then
        call interface_avg3d(u,icu)
        call interface_avg3d(v,icv)
        call interface_avg3d(w,icw)
        call interface_avg3d(temp,ict)
        write(12,'(4(1x,I10))') icu,icv,icw,ict
        goto 900
      endif
      if (mode.eq.8) then
c-testhelm.f:60-100: penalty_setup, then penalty_main of the column
c-i=icase+1, j=1 of u (real) and v (imaginary), for four (epsilon,
c-boundary types, boundary values); solution and the right-hand sides
c-as penalty_main leaves them to the diagnostics file
        do icase = 1, 4
          if (icase.eq.1) then
            eps8 = 0.01
            ifbb8 = 0
            ifbt8 = 0
            bcrb8 = 0.
            bcib8 = 0.
            bcrt8 = 0.
            bcit8 = 0.
          elseif (icase.eq.2) then
            eps8 = 1./17.
            ifbb8 = 1
            ifbt8 = 1
            bcrb8 = 0.3
            bcib8 = -0.2
            bcrt8 = 0.1
            bcit8 = 0.05
          elseif (icase.eq.3) then
            eps8 = 1./(17.+11./6./(xnu*dt))
            ifbb8 = 0
            ifbt8 = 1
            bcrb8 = 1.
            bcib8 = 0.5
            bcrt8 = -0.25
            bcit8 = 0.75
          else
            eps8 = 1./(5.+11./6./(xkappa*dt))
            ifbb8 = 1
            ifbt8 = 0
            bcrb8 = 0.
            bcib8 = 0.
            bcrt8 = 0.2
            bcit8 = -0.1
          endif
          do k = 1, nz
            rhr8(k) = u(icase+1,1,k)
            rhi8(k) = v(icase+1,1,k)
          enddo
          call penalty_setup(eps8)
          call penalty_main(eps8,0,bcrb8,bcib8,bcrt8,bcit8,ifbb8,ifbt8,
     >                      rhr8,rhi8,fr8,fi8)
          do k = 1, nz
            write(12,'(4(1x,E25.17))') fr8(k),fi8(k),rhr8(k),rhi8(k)
          enddo
        enddo
        goto 900
      endif
      if (mode.eq.7) then
        open(14,file='kepe.txt',status='unknown')
        call kepe(u,v,w,temp)
        close(14)
        goto 900
      endif
#ifdef IW_POSTP
      if (mode.eq.6) then
        if (isgs.eq.1) call filtering(1)
        open(14,file='energy.txt',status='unknown')
        initenergy = 0
        call postp(initenergy,t,brunt,tstart)
        call wenergy(0,t)
        goto 900
      endif
#endif

c-mode 0: full run from a physical-space initial condition
      call horfft(u,-1,nz)
      call horfft(v,-1,nz)
      call horfft(w,-1,nz)
      call horfft(temp,-1,nz)

      t = 0.
      open(3,file='p_out.dat',form='unformatted',status='unknown')
      call plotf(3,iform,t)
      close(3)
#ifdef IW_POSTP
      open(14,file='energy.txt',status='unknown')
#endif
      fac1 = 1.0
      fac2 = 0.
      fac3 = 0.
      bdf1 = 1.
      bdf2 = 0.
      bdf3 = 0.
      bdfv = 1.

 500  continue
      if (imaxc.eq.1) open(13,file='dtlog.txt',status='unknown')
      do 1000 istep=1,nrun
        if (imaxc.eq.1) then
#ifdef IW_MAXCALC
          call maxcalc(istep)
          write(13,'(I8,3(1x,E24.16),2(1x,I8))') istep,t,dt,cdt,
     >      ivardt,idtstep
#else
          write(*,*) 'imaxc=1 needs a build with various.f'
          stop 2
#endif
        else
c-State-changing part of maxcalc (various.f:739-742, 854-857)
          call horfft(u,1,nz)
          call horfft(v,1,nz)
          call horfft(w,1,nz)
          call horfft(temp,1,nz)
          if (iprobe.eq.1) write(12,'(3(1x,E24.16))')
     >       t,u(ipx,1,kp1),u(ipx,1,kp2)
          call horfft(u,-1,nz)
          call horfft(v,-1,nz)
          call horfft(w,-1,nz)
          call horfft(temp,-1,nz)
        endif
c-main.f:660-664
        if (ivardt.eq.1) idtstep = idtstep+1

        call convec(istep,0,ispecfilternl)

c-main.f:700-777
        ldtadapt = ((ivardt.eq.1).and.(idtstep.lt.2))
        if (istep.lt.2) then
          if (istart.eq.0) then
            fac1 = 2.
            fac2 = -1.
            fac3 = 0.
            bdf1 = 2.
            bdf2 = -0.5
            bdf3 = 0.
          else
            dt1 = dt
            dt2 = dt
            dt3 = cdt*dt
            fac1 = abvar1(dt1,dt2,dt3)
            fac2 = abvar2(dt1,dt2,dt3)
            fac3 = abvar3(dt1,dt2,dt3)
            bdf1 = bdfvar1(dt1,dt2,dt3)
            bdf2 = bdfvar2(dt1,dt2,dt3)
            bdf3 = bdfvar3(dt1,dt2,dt3)
          endif
        else if (.not.ldtadapt) then
          fac1 = 3.
          fac2 = -3.
          fac3 = 1.
          bdf1 = 3.
          bdf2 = -1.5
          bdf3 = 1./3.
        else
          dt1 = dt
          dt2 = dt
          dt3 = cdt*dt
          fac1 = abvar1(dt1,dt2,dt3)
          fac2 = abvar2(dt1,dt2,dt3)
          fac3 = abvar3(dt1,dt2,dt3)
          bdf1 = bdfvar1(dt1,dt2,dt3)
          bdf2 = bdfvar2(dt1,dt2,dt3)
          bdf3 = bdfvar3(dt1,dt2,dt3)
        endif

        call copy(u,su,ntot)
        call copy(v,sv,ntot)
        call copy(w,sw,ntot)
        call copy(temp,st,ntot)

        cpenspecfilter = ((ispecfilterexp.eq.1).and.
     >   (mod(istep,nspecfilterexp).eq.0))
        call calcuvwt(izmode,ispecfilter,irelax,cpenspecfilter)

c-main.f:813-856
        ldtadapt = ((ivardt.eq.1).and.(idtstep.lt.2))
        if (istep.lt.2) then
          if (istart.eq.0) then
            bdfv = 1.5
          else
            bdfv = bdfvarv(dt,dt,cdt*dt)
          endif
        else if (.not.ldtadapt) then
          bdfv = 11./6.
        else
          bdfv = bdfvarv(dt,dt,cdt*dt)
          cdt = 1.
          ivardt = 0
        endif

c-main.f:880-895
        if ((ifilter.eq.1).and.(mod(istep,nstepfilter).eq.0)) then
          if (mod(istep,npostp).eq.0) then
            isgscalc = 1
          else
            isgscalc = 0
          endif
          call filtering(isgscalc)
        endif

        t = t + dt

        if (mod(istep,ncalclegspec).eq.0) then
          call speclegencalc('u',t,u)
          call speclegencalc('v',t,v)
          call speclegencalc('w',t,w)
          call speclegencalc('T',t,temp)
        endif

        if (mod(istep,navgintrfc).eq.0) then
          call interface_avg3d(u,icu)
          call interface_avg3d(v,icv)
          call interface_avg3d(w,icw)
          call interface_avg3d(temp,ict)
          write(12,'(I8,4(1x,I10))') istep,icu,icv,icw,ict
        endif
#ifdef IW_POSTP
        if (mod(istep,npostp).eq.0) then
          call postp(initenergy,t,brunt,tstart)
          call wenergy(istep,t)
        endif
#endif
        call zero(su,ntot)
        if (mod(istep,n3doutput).eq.0) then
          open(3,file='p_out.dat',form='unformatted',status='unknown')
          call plotf(3,iform,t)
          close(3)
        endif
 1000 continue

 900  continue
      close(12)
      if (iwrite.eq.1) then
        open(3,file='restart_out.dat',form='unformatted',
     >       status='unknown')
        call output(3,t)
        close(3)
      endif
      open(11,file=outfile,form='unformatted',access='stream',
     >     status='replace')
      write(11) u,v,w,temp
      write(11) un,vn,wn,tn
      write(11) unm,vnm,wnm,tnm
      write(11) su,sv,sw,st
      write(11) sun,svn,swn,stn
      write(11) sunm,svnm,swnm,stnm
      write(11) t
      if (mode.eq.5) write(11) t11,t12,t13,t22,t23,t33,th1,th2,th3
      close(11)
      end

c-----------------------------------------------------------------------
c  KE_u, KE_v, KE_w, PE of physical fields: rho0/2 q^2 and
c  g^2 rho'^2/(2 rho0 N^2) (CalcE/integration.f:51, 216) averaged over
c  x and y times xlen (times ylen in 3-D), then integrcol per subdomain
c-----------------------------------------------------------------------
      subroutine kepe(u,v,w,temp)
      include 'dim.h'
      include 'comsubd.h'
      include 'comparamphys.h'
      dimension u(nxpp,ny,nz),v(nxpp,ny,nz),w(nxpp,ny,nz)
      dimension temp(nxpp,ny,nz),h(nz),e(4)
      span = xlen
      if (ny.gt.1) span = xlen*ylen
      do iq = 1,4
        c = rho0/2.
        if (iq.eq.4) c = 1./2./rho0*grav**2/brunt**2
        do k = 1,nz
          s = 0.
          do j = 1,ny
            do i = 1,nx
              if (iq.eq.1) f = u(i,j,k)
              if (iq.eq.2) f = v(i,j,k)
              if (iq.eq.3) f = w(i,j,k)
              if (iq.eq.4) f = temp(i,j,k)
              s = s + c*(f*f)
            enddo
          enddo
          h(k) = span*s/float(nx*ny)
        enddo
        e(iq) = 0.
        do ks = 1,nsubd
          call integrcol(h((ks-1)*nzloc+1),outint,nzm,zh(ks))
          e(iq) = e(iq) + outint
        enddo
      enddo
      write(14,'(4(1x,E24.16))') e
      return
      end

c-----------------------------------------------------------------------
c  /energetics/ (comenergy.h) at full precision, fort.60's order
c-----------------------------------------------------------------------
      subroutine wenergy(istep,t)
      include 'comenergy.h'
      write(14,'(I8,17(1x,E24.16))') istep,t,tketot,u1tke,u2tke,
     >  u3tke,prody,prodz,prod,bflux,epsilon,epssgs,epstot,chi,chisgs,
     >  chitot,tkemean,tempvar
      return
      end

#ifndef IW_MAXCALC
c-----------------------------------------------------------------------
c  zero / copy: the two helpers of various.f:2-26 used on the hot path
c-----------------------------------------------------------------------
      subroutine zero(a,n)
      dimension a(*)
      do i=1,n
        a(i) = 0.
      enddo
      return
      end

      subroutine copy(a,b,n)
      dimension a(*), b(*)
      do i=1,n
        b(i) = a(i)
      enddo
      return
      end
#endif
