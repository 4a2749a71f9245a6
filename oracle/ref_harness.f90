!----------------------------------------------------------------------
! Dump harness for the IMPACT-Z reference (test infrastructure only).
!
! Links against the reference's own compiled modules and drives ONE
! space-charge step sequence exactly in the order of the reference's
! segment loop (src/Contrl/AccSimulator.f90:1099-1488, Flagmap=1,
! Flagbc=1, Flagsubstep=0, Flagerr=0; drift, quadrupole, constant-focusing
! and solenoid elements and the RF cavities DTL, CCDTL, CCL, SC and SolRF,
! whose on-axis field comes from rfdataN.in through the reference's own
! read1_Data / read3_Data, as run_AccSimulator reads it, :818-849),
! writing the intermediate quantities the reference never writes itself:
! the particle array after every stage, the conv1st range, the updated
! domain, the charge density and the potential.
!
! Input (list-directed, file named on the command line, default
! harness.in):
!   N Nx Ny Nz
!   curr kinetic_eV mass_eV charge freq phase0
!   flagmap (1 linear maps, 2 nonlinear Lorentz integrator)
!   nelem
!   nelem lines:  length nseg mapstp type v1 ... v14
!                 (the deck line's values = Param(2:15),
!                  src/Contrl/AccSimulator.f90:262-536: type 0: v1 = pipe
!                  radius; type 1: v1 = gradient T/m (or K, 1/m^2, when
!                  v2 < 0), v2 = file ID, v3 = pipe radius; type 2: v1..v3 =
!                  kx0^2, ky0^2, kz0^2, v4 = pipe radius; type 3: v1 = Bz0,
!                  v2 = file ID, v3 = pipe radius; types 101-105: v1 = field
!                  scale, v2 = RF frequency, v3 = phase, v4 = file ID,
!                  v5 = pipe radius, v6.. as the element class lists them)
!   nsteps ndump
!   ndump step indices (1-based) to dump
!   particle file name (stream: int32 N, then Pts1(9,N) doubles)
! Output: dump_<step>.bin (stream) per dumped step, fort.18/24-30 from
! the reference's diagnostic1_Output every step.
!----------------------------------------------------------------------
      program ref_harness
        use mpistub
        use Pgrid2dclass
        use CompDomclass
        use FieldQuantclass
        use BeamLineElemclass
        use BeamBunchclass
        use Timerclass
        use Outputclass
        use PhysConstclass
        use NumConstclass
        use Dataclass
        implicit none
        include 'mpif.h'
        type (Pgrid2d) :: grid2d
        type (CompDom) :: Ageom
        type (BeamBunch) :: Bpts
        type (FieldQuant) :: Potential
        type (DriftTube), target, dimension(:), allocatable :: drifts
        type (Quadrupole), target, dimension(:), allocatable :: quads
        type (ConstFoc), target, dimension(:), allocatable :: cfs
        type (Sol), target, dimension(:), allocatable :: sols
        type (DTL), target, dimension(:), allocatable :: dtls
        type (CCDTL), target, dimension(:), allocatable :: ccdtls
        type (CCL), target, dimension(:), allocatable :: ccls
        type (SC), target, dimension(:), allocatable :: scs
        type (SolRF), target, dimension(:), allocatable :: slrfs
        type (BPM), target, dimension(:), allocatable :: bpms
        ! wakefields (src/Contrl/AccSimulator.f90:627-660,859-875,960-980)
        integer :: flagwake, flagwakeread, flagbtw, iizz, iizz1, kz, ipt
        double precision :: scwk, aawk, ggwk, lengwk, hzwake, ab, tmpwk, rfile, tmp1
        double precision, dimension(:), allocatable :: xwakez, ywakez, densz, &
             exwake, eywake, ezwake
        double precision, dimension(:,:), allocatable :: recvdensz
        type (BeamLineElem), dimension(:), allocatable :: Blnelem
        integer :: N, Nx, Ny, Nz, nelem, nsteps, ndump, ierr, i, j, k
        integer :: bnseg, bmpstp, bitype, ihlf, nstep, idr, iqr, icf, isl
        integer :: Nplocal, Np, nchrg, u
        integer, dimension(:), allocatable :: etype, enseg, empstp, dumpat
        integer, dimension(1) :: nptlist0, idchg
        double precision, dimension(1) :: qmcclist0
        double precision, dimension(:), allocatable :: elen
        double precision, dimension(:,:), allocatable :: ev
        double precision, dimension(25) :: tmprf
        integer :: ifile, tmpfile
        double precision :: curr, kin, mass, chg, freq, ph0, gambet
        double precision :: z, tau1, tau2, blength, piperad, piperad2
        double precision, dimension(21) :: distparam
        double precision, dimension(9) :: tmpdr
        double precision, dimension(9) :: tmpquad
        double precision, dimension(5) :: tmpcf
        double precision, dimension(6) :: ptrange, range
        double precision, dimension(3) :: msize
        integer, dimension(3) :: lcgrid
        double precision, allocatable, dimension(:,:,:) :: chgdens
        character(len=256) :: ctl, pfile, dname
        logical :: dodump
        integer :: n32, flagmap

        ctl = 'harness.in'
        if (command_argument_count() >= 1) call get_command_argument(1, ctl)
        open(unit=11, file=trim(ctl), status='old')
        read(11,*) N, Nx, Ny, Nz
        read(11,*) curr, kin, mass, chg, freq, ph0
        read(11,*) flagmap
        read(11,*) nelem
        allocate(elen(nelem), enseg(nelem), empstp(nelem), etype(nelem))
        allocate(ev(14, nelem))
        do i = 1, nelem
          read(11,*) elen(i), enseg(i), empstp(i), etype(i), ev(1:14, i)
        enddo
        read(11,*) nsteps, ndump
        allocate(dumpat(max(ndump,1)))
        dumpat = -1
        if (ndump > 0) read(11,*) dumpat(1:ndump)
        read(11,'(a)') pfile
        close(11)

        call MPI_INIT(ierr)
        call construct_Timer(0.0d0)
        call construct_Pgrid2d(grid2d, MPI_COMM_WORLD, 1, 1)
        call construct_PhysConst(freq)
        call init_Data()
        distparam = 0.0d0
        distparam(1) = 1.0d-3
        distparam(2) = 1.0d-3
        distparam(8) = 1.0d-3
        distparam(9) = 1.0d-3
        distparam(15) = 1.0d-1
        distparam(16) = 1.0d-3
        gambet = sqrt((kin/mass+1.0d0)**2-1.0d0)
        call construct_CompDom(Ageom, distparam, 21, 19, Nx, Ny, Nz, &
             grid2d, 1, 1, 1, 1.0d0, 1.0d0, 1.0d0, gambet)
        call construct_BeamBunch(Bpts, curr, kin, mass, chg, N, ph0)
        call construct_FieldQuant(Potential, Nx, Ny, Nz, Ageom, grid2d)

        ! particles, exactly as readin_Dist leaves them (9 x N, in place)
        open(unit=12, file=trim(pfile), access='stream', &
             form='unformatted', status='old')
        read(12) n32
        if (n32 /= N) then
          print*, "particle count mismatch", n32, N
          stop
        endif
        allocate(Bpts%Pts1(9, N))
        read(12) Bpts%Pts1
        close(12)
        Bpts%Nptlocal = N
        Nplocal = N
        Np = N

        ! lattice: one target object per element, as init_AccSimulator does
        allocate(drifts(nelem), quads(nelem), Blnelem(nelem))
        allocate(cfs(nelem), sols(nelem))
        allocate(dtls(nelem), ccdtls(nelem), ccls(nelem), scs(nelem), slrfs(nelem))
        allocate(bpms(nelem))
        idr = 0
        iqr = 0
        icf = 0
        isl = 0
        do i = 1, nelem
          if (etype(i) == 0) then
            idr = idr + 1
            call construct_DriftTube(drifts(idr), enseg(i), empstp(i), &
                 etype(i), elen(i))
            tmpdr = 0.0d0
            tmpdr(2) = ev(1,i)
            call setparam_DriftTube(drifts(idr), tmpdr(1:2))
            Blnelem(i) = assign_BeamLineElem(drifts(idr))
          else if (etype(i) == 1) then
            iqr = iqr + 1
            call construct_Quadrupole(quads(iqr), enseg(i), empstp(i), &
                 etype(i), elen(i))
            tmpquad = 0.0d0
            tmpquad(2:4) = ev(1:3,i)
            call setparam_Quadrupole(quads(iqr), tmpquad)
            Blnelem(i) = assign_BeamLineElem(quads(iqr))
          else if (etype(i) == 2) then
            icf = icf + 1
            call construct_ConstFoc(cfs(icf), enseg(i), empstp(i), &
                 etype(i), elen(i))
            tmpcf = 0.0d0
            tmpcf(2:5) = ev(1:4,i)
            call setparam_ConstFoc(cfs(icf), tmpcf)
            Blnelem(i) = assign_BeamLineElem(cfs(icf))
          else if (etype(i) == 3) then
            isl = isl + 1
            call construct_Sol(sols(isl), enseg(i), empstp(i), &
                 etype(i), elen(i))
            tmpquad = 0.0d0
            tmpquad(2:5) = ev(1:4,i)
            call setparam_Sol(sols(isl), tmpquad)
            Blnelem(i) = assign_BeamLineElem(sols(isl))
          else if (etype(i) >= 101 .and. etype(i) <= 105) then
            ! Param(2:15) = v1..v14 (src/Contrl/AccSimulator.f90:375-490)
            tmprf = 0.0d0
            tmprf(2:15) = ev(1:14,i)
            if (etype(i) == 101) then
              call construct_DTL(dtls(i), enseg(i), empstp(i), etype(i), elen(i))
              call setparam_DTL(dtls(i), tmprf)
              Blnelem(i) = assign_BeamLineElem(dtls(i))
            else if (etype(i) == 102) then
              call construct_CCDTL(ccdtls(i), enseg(i), empstp(i), etype(i), elen(i))
              call setparam_CCDTL(ccdtls(i), tmprf)
              Blnelem(i) = assign_BeamLineElem(ccdtls(i))
            else if (etype(i) == 103) then
              call construct_CCL(ccls(i), enseg(i), empstp(i), etype(i), elen(i))
              call setparam_CCL(ccls(i), tmprf)
              Blnelem(i) = assign_BeamLineElem(ccls(i))
            else if (etype(i) == 104) then
              call construct_SC(scs(i), enseg(i), empstp(i), etype(i), elen(i))
              call setparam_SC(scs(i), tmprf)
              Blnelem(i) = assign_BeamLineElem(scs(i))
            else
              call construct_SolRF(slrfs(i), enseg(i), empstp(i), etype(i), elen(i))
              call setparam_SolRF(slrfs(i), tmprf)
              Blnelem(i) = assign_BeamLineElem(slrfs(i))
            endif
          else if (etype(i) == -41) then
            ! v1 = wake scale, v2 = wake file ID, v3 > 0 turns the read-in
            ! wake on (< 10: no transverse wake)
            call construct_BPM(bpms(i), enseg(i), empstp(i), etype(i), elen(i))
            tmprf = 0.0d0
            tmprf(2:15) = ev(1:14,i)
            call setparam_BPM(bpms(i), tmprf)
            Blnelem(i) = assign_BeamLineElem(bpms(i))
          else
            print*, "harness supports element types 0 to 3, 101 to 105 and -41 only"
            stop
          endif
        enddo

        nchrg = 1
        nptlist0(1) = N
        qmcclist0(1) = Bpts%Pts1(7,1)
        allocate(chgdens(1,1,1))
        z = 0.0d0
        nstep = 0
        tmpfile = 0
        flagwake = 0
        flagwakeread = 0
        flagbtw = 0
        scwk = 1.0d0
        allocate(xwakez(Nz), ywakez(Nz), densz(Nz), exwake(Nz), eywake(Nz), &
                 ezwake(Nz), recvdensz(Nz,2))
        exwake = 0.0d0
        eywake = 0.0d0
        ezwake = 0.0d0
        call diagnostic1_Output(z, Bpts, nchrg, nptlist0)

        do i = 1, nelem
          call getparam_BeamLineElem(Blnelem(i), blength, bnseg, bmpstp, &
                                     bitype)
          call getradius_BeamLineElem(Blnelem(i), piperad, piperad2)
          ! the on-axis field of an RF cavity (src/Contrl/AccSimulator.f90:818-849)
          if (bitype > 100) then
            ifile = int(ev(4,i) + 0.1)
            if (ifile /= tmpfile) then
              if (flagmap == 1 .and. bitype /= 105 .and. bitype /= 106) then
                call read1_Data(ifile)
              else
                call read3_Data(ifile)
              endif
              tmpfile = ifile
            endif
          endif
          ! the wake switches of the element loop
          ! (src/Contrl/AccSimulator.f90:759,859-875,960-980)
          if (flagwakeread == 0) flagwake = 0
          if (bitype == 105 .or. bitype == 106) then
            aawk = ev(12,i)
            ggwk = ev(13,i)
            lengwk = ev(14,i)
            if (lengwk > 0.0d0) then
              flagwake = 1
            else
              flagwake = 0
            endif
          endif
          if (bitype == -41) then
            tmpwk = ev(1,i)
            rfile = ev(2,i)
            if (rfile > 0.0d0) call read1wk_Data(int(rfile + 0.1))
            tmp1 = ev(3,i)
            if (tmp1 > 0.0d0) then
              flagwake = 1
              flagwakeread = 1
              if (tmp1 < 10.0) flagbtw = 2
              scwk = tmpwk
            else
              flagwake = 0
              flagwakeread = 0
              scwk = 1.0d0
            endif
          endif
          ! the element's entrance z (src/Contrl/AccSimulator.f90:1025-1026)
          call setparam_BeamLineElem(Blnelem(i), 1, z)
          tau1 = 0.0d0
          if (bitype >= 0) tau1 = 0.5d0*blength/bnseg
          tau2 = 2.0d0*tau1
          ihlf = 0
          do j = 1, bnseg
            if (nstep >= nsteps) goto 900
            dodump = any(dumpat(1:max(ndump,1)) == nstep+1)
            if (dodump) then
              write(dname, '(a,i0,a)') 'dump_', nstep+1, '.bin'
              u = 20
              open(unit=u, file=trim(dname), access='stream', &
                   form='unformatted', status='replace')
              write(u) z
              write(u) Bpts%refptcl
              write(u) Bpts%Current
              call wpts(u)
            endif
            if (flagmap == 1) then
              call push1(i, j)
            else
              call map1_BeamBunch(Bpts, z, tau2)
            endif
            if (dodump) call wpts(u)
            call conv1st_BeamBunch(Bpts, tau2, Nplocal, Np, ptrange, &
                                   1, 1.0d0, piperad, piperad2)
            call chgupdate_BeamBunch(Bpts, nchrg, nptlist0, qmcclist0)
            if (dodump) then
              write(u) ptrange
              write(u) Bpts%Current
              call wpts(u)
            endif
            call update_CompDom(Ageom, ptrange, grid2d, 1)
            call getlcmnum_CompDom(Ageom, lcgrid)
            call setnpt_BeamBunch(Bpts, Nplocal)
            call set_FieldQuant(Potential, Nx, Ny, Nz, Ageom, grid2d, 1, 1)
            deallocate(chgdens)
            allocate(chgdens(lcgrid(1), lcgrid(2), lcgrid(3)))
            call charge_BeamBunch(Bpts, Nplocal, lcgrid(1), lcgrid(2), &
                 lcgrid(3), Ageom, grid2d, chgdens, 1, 1.0d0)
            if (dodump) then
              call getrange_CompDom(Ageom, range)
              call getmsize_CompDom(Ageom, msize)
              write(u) range
              write(u) msize
              write(u) chgdens
            endif
            call update3O_FieldQuant(Potential, chgdens, Ageom, grid2d, &
                 lcgrid(1), lcgrid(2), lcgrid(3), 1, 1, lcgrid(2), lcgrid(3))
            if (dodump) write(u) Potential%FieldQ
            if (flagwake == 1) then
              call wake(u, dodump)
            else if (dodump) then
              densz = 0.0d0
              write(u) densz
              write(u) densz
              write(u) densz
            endif
            call cvbkforth1st_BeamBunch(Bpts)
            call setnpt_BeamBunch(Bpts, Nplocal)
            if (dodump) call wpts(u)
            if (flagmap == 1 .and. flagwake == 1) then
              call kick1wake_BeamBunch(Bpts, tau2, lcgrid(1), lcgrid(2), &
                   lcgrid(3), Potential%FieldQ, Ageom, grid2d, 1, 1.0d0, &
                   exwake, eywake, ezwake, Nz, 1, 1)
            else if (flagmap == 1) then
              call map2_BeamBunch(Bpts, tau2, lcgrid(1), lcgrid(2), &
                   lcgrid(3), Potential%FieldQ, Ageom, grid2d, 1, 1.0d0)
            else if (flagwake == 1) then
              call kick2wake_BeamBunch(Bpts, Blnelem(i), z, tau2, lcgrid(1), &
                   lcgrid(2), lcgrid(3), Potential%FieldQ, Ageom, grid2d, 1, 0, &
                   exwake, eywake, ezwake, Nz, 1, 1)
            else
              call map2_BeamBunch(Bpts, Blnelem(i), z, tau2, lcgrid(1), &
                   lcgrid(2), lcgrid(3), Potential%FieldQ, Ageom, grid2d, 1, 0)
            endif
            if (dodump) call wpts(u)
            if (flagmap == 1) then
              call push1(i, j)
            else
              call map1_BeamBunch(Bpts, z, tau2)
            endif
            if (dodump) then
              call wpts(u)
              write(u) z
              write(u) Bpts%refptcl
              close(u)
            endif
            call diagnostic1_Output(z, Bpts, nchrg, nptlist0)
            nstep = nstep + 1
          enddo
        enddo
900     continue

        ! final particle state of the run
        open(unit=21, file='final.bin', access='stream', &
             form='unformatted', status='replace')
        call wpts(21)
        write(21) z
        write(21) Bpts%refptcl
        write(21) Bpts%Current
        close(21)

      contains
        ! The wake fields of the run loop (src/Contrl/AccSimulator.f90:
        ! 1272-1338, one process): the particles' t-frame z binned linearly
        ! on the Nz mesh lines (unit weight; x and y weighted for the
        ! centroids), the line charge density, then the reference's
        ! wakefield_FieldQuant (analytic) or wakefieldread_FieldQuant (read
        ! in), scaled by scwk.  The dump record (when dumping; zeros without
        ! a wake) is Nz values each of exwake, eywake, ezwake.
        subroutine wake(uu, dd)
          integer, intent(in) :: uu
          logical, intent(in) :: dd
          call getrange_CompDom(Ageom, range)
          hzwake = (range(6)-range(5))/(Nz-1)
          xwakez = 0.0d0
          ywakez = 0.0d0
          densz = 0.0d0
          do ipt = 1, Nplocal
            iizz = (Bpts%Pts1(5,ipt)-range(5))/hzwake + 1
            iizz1 = iizz + 1
            ab = ((range(5)-Bpts%Pts1(5,ipt))+iizz*hzwake)/hzwake
            xwakez(iizz) = xwakez(iizz) + Bpts%Pts1(1,ipt)*ab
            xwakez(iizz1) = xwakez(iizz1) + Bpts%Pts1(1,ipt)*(1.0-ab)
            ywakez(iizz) = ywakez(iizz) + Bpts%Pts1(3,ipt)*ab
            ywakez(iizz1) = ywakez(iizz1) + Bpts%Pts1(3,ipt)*(1.0-ab)
            densz(iizz) = densz(iizz) + ab
            densz(iizz1) = densz(iizz1) + 1.0 -ab
          enddo
          hzwake = hzwake/(-Bpts%refptcl(6))
          do kz = 1, Nz
            recvdensz(kz,1) = densz(kz)*Bpts%Current/Scfreq/Np/(hzwake)*&
                              Bpts%Charge/abs(Bpts%Charge)
          enddo
          do kz = 1, Nz
            if (densz(kz) .ne. 0.0) then
              xwakez(kz) = xwakez(kz)/densz(kz)
              ywakez(kz) = ywakez(kz)/densz(kz)
            else
              xwakez(kz) = 0.0
              ywakez(kz) = 0.0
            endif
            recvdensz(kz,2) = ywakez(kz)
          enddo
          if (flagwakeread == 1) then
            call wakefieldread_FieldQuant(Nz, xwakez, ywakez, recvdensz, exwake, &
                 eywake, ezwake, hzwake, aawk, ggwk, lengwk, flagbtw)
          else
            call wakefield_FieldQuant(Nz, xwakez, ywakez, recvdensz, exwake, &
                 eywake, ezwake, hzwake, aawk, ggwk, lengwk)
          endif
          exwake = scwk*exwake
          eywake = scwk*eywake
          ezwake = scwk*ezwake
          if (dd) then
            write(uu) exwake
            write(uu) eywake
            write(uu) ezwake
          endif
        end subroutine wake

        ! map1_BeamBunch with Flagmap=1 (drift1_BeamBunch).  For the DTL and
        ! SolRF, drift1_BeamBunch's getparam_BeamLineElem(beamln,drange)
        ! copies the element's 25 (DTL) or 15 (SolRF) Param values into its
        ! 12-element drange (src/Appl/BeamBunch.f90:270,323), past the end of
        ! the array; for those two types the harness makes the rest of
        ! drift1's calls itself: the element's maplinear and the 6x6 product
        ! of src/Appl/BeamBunch.f90:436-449, then z += tau.
        subroutine push1(ie, js)
          integer, intent(in) :: ie, js
          double precision, dimension(6,6) :: xm
          double precision, dimension(6) :: tmp
          integer :: ip, jj, kk
          if (bitype == 101 .or. bitype == 105) then
            call maplinear_BeamLineElem(Blnelem(ie), z, tau1, xm, Bpts%refptcl, &
                                        Bpts%Charge, Bpts%Mass)
            do ip = 1, Bpts%Nptlocal
              do jj = 1, 6
                tmp(jj) = 0.0
                do kk = 1, 6
                  tmp(jj) = tmp(jj) + Bpts%Pts1(kk,ip)*xm(jj,kk)
                enddo
              enddo
              do jj = 1, 6
                Bpts%Pts1(jj,ip) = tmp(jj)
              enddo
            enddo
            z = z + tau1
            ihlf = ihlf + 1
          else
            call map1_BeamBunch(Bpts, Blnelem(ie), z, tau1, bitype, bnseg, js, ihlf)
          endif
        end subroutine push1

        subroutine wpts(uu)
          integer, intent(in) :: uu
          integer :: nn
          nn = Bpts%Nptlocal
          write(uu) nn
          write(uu) Bpts%Pts1(1:9, 1:nn)
        end subroutine wpts
      end program ref_harness
