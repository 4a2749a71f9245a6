!----------------------------------------------------------------------------
! izh_lattice -- a Fortran 90 host that marches a particle.in bunch through a
! lattice of drifts, quadrupoles, constant-focusing elements and solenoids
! on the GPU, calling the C ABI through the
! impactz_hip module in the order of run_AccSimulator's segment loop
! (src/Contrl/AccSimulator.f90:1099-1488, Flagbc=1).  Flagmap=1:
!   drift1(tau1) -> conv1st -> update_CompDom -> charge -> update3O ->
!   cvbkforth1st -> kick1(2 tau1) -> drift1(tau1) -> diagnostic1
! Flagmap=2 (the Lorentz integrator): drift2(2 tau1), the same space-charge
! stages, kick2(2 tau1) with the element's field, drift2(2 tau1).
! With fused = .true. the Flagmap=1 segments go through izh_run in one call.
!
! Input: particle.in (readin_Dist format, src/Appl/Distribution.f90:1677-1705)
! and the namelist file named by the first command-line argument; val(1:14,i)
! holds element i's deck values v1..v14 (its Param(2:15)).  An RF cavity
! (bitype 101-105) reads rfdataN.in, N = its file ID v4, as read1_Data
! (rows z Ez Ez' Ez'', bitypes 101-104) or read3_Data (one coefficient per
! line, SolRF) do (src/DataStruct/Data.f90:139-215,348-409).  Output:
! one line per step with z, Npt, the 44 diagnostic1 reductions (37 sums,
! 6 absolute maxima, max x^2+y^2), the reference particle's gamma
! (-refptcl(6), which the Lorentz integrator changes in a longitudinal field)
! and its phase refptcl(5), in the file named by the namelist.
!----------------------------------------------------------------------------
program izh_lattice
  use iso_c_binding
  use impactz_hip
  implicit none
  integer, parameter :: maxelem = 4096
  integer :: nx = 16, ny = 16, nz = 32, precision = 0, device = 0
  real(8) :: current = 0.1d0, kinetic = 150.0d6, mass = 938.272088d6
  real(8) :: charge = 1.0d0, freq = 324.0d6, phase0 = 0.0d0
  integer :: nelem = 0
  real(8) :: elen(maxelem) = 0, val(izh_nval, maxelem) = 0, radius(maxelem) = 1
  integer :: nseg(maxelem) = 1, mapstp(maxelem) = 1, bitype(maxelem) = 0
  logical :: fused = .false.
  integer :: flagmap = 1
  character(len=256) :: particles = 'particle.in', output = 'izh_moments.txt'
  namelist /izh_run_nml/ nx, ny, nz, precision, device, current, kinetic, &
      mass, charge, freq, phase0, nelem, elen, val, radius, nseg, mapstp, &
      bitype, fused, flagmap, particles, output

  character(len=256) :: nmlfile
  real(c_double), allocatable :: pts(:, :), recs(:, :)
  type(izh_segment), allocatable :: segs(:)
  type(c_ptr) :: b, f
  type(izh_compdom) :: dom
  type(izh_bunch_info) :: info
  real(c_double) :: z, zedge, tau1, tau2, ptrange(6), sums(37), amax(6), r2max
  type(c_ptr), allocatable :: rfh(:)
  type(izh_element) :: e
  integer(c_int64_t) :: n, nplc, nptot, nstep, k
  integer :: i, j, u, istep

  call get_command_argument(1, nmlfile)
  open(newunit=u, file=trim(nmlfile), status='old')
  read(u, nml=izh_run_nml)
  close(u)

  ! readin_Dist: N, then one particle (9 values) per line
  open(newunit=u, file=trim(particles), status='old')
  read(u, *) n
  allocate(pts(9, n))
  do k = 1, n
    read(u, *) pts(:, k)
  end do
  close(u)

  call izh_check(izh_init(int(device, c_int)), 'init')
  call izh_check(izh_bunch_create(pts, n, current, kinetic, mass, charge, freq, &
                                  phase0, int(precision, c_int), b), 'bunch_create')
  call izh_check(izh_field_create(int(nx, c_int), int(ny, c_int), int(nz, c_int), &
                                  int(precision, c_int), f), 'field_create')
  allocate(rfh(nelem))
  do i = 1, nelem
    rfh(i) = c_null_ptr
    if (bitype(i) == 105 .and. val(14, i) > 0.0d0) then
      print *, 'izh_lattice: element ', i, ': SolRF wakefields are not run by this host'
      stop 1
    end if
    if (bitype(i) >= 101 .and. bitype(i) <= 105) call read_rfdata(i)
  end do
  dom%nx = nx; dom%ny = ny; dom%nz = nz
  dom%range = 0; dom%msize = 0
  z = 0.0d0

  if (flagmap /= 1 .and. flagmap /= 2) then
    print *, 'izh_lattice: flagmap must be 1 or 2'
    stop 1
  end if
  if (fused .and. flagmap /= 1) then
    print *, 'izh_lattice: izh_run fuses Flagmap=1 segments only'
    stop 1
  end if
  open(newunit=u, file=trim(output), status='replace')
  if (.not. fused) then
    istep = 0
    do i = 1, nelem
      tau1 = 0.5d0 * elen(i) / nseg(i)
      tau2 = 2.0d0 * tau1
      zedge = z
      e = element_gpu(bitype(i), val(:, i), mapstp(i), elen(i), zedge, rfh(i))
      do j = 1, nseg(i)
        if (flagmap == 1) then
          call drift1_gpu(b, z, tau1, e)
        else
          call drift2_gpu(b, z, tau2)
        end if
        if (current >= 1.0e-10) then
          call conv1st_gpu(b, tau2, nplc, nptot, ptrange, 1, radius(i), radius(i))
          call update_compdom_gpu(dom, ptrange, 1)
          call charge_gpu(b, dom, f)
          call update3O_gpu(f, dom)
          call cvbkforth1st_gpu(b)
        end if
        if (flagmap == 1) then
          if (current >= 1.0e-10) call kick1_gpu(b, tau2, f, dom, 1)
          call drift1_gpu(b, z, tau1, e)
        else
          call kick2_gpu(b, z, tau2, e, f, dom, 1)
          call drift2_gpu(b, z, tau2)
        end if
        call diagnostic1_sums_gpu(b, sums, amax, r2max)
        call izh_check(izh_bunch_get_info(b, info), 'bunch_get_info')
        istep = istep + 1
        write(u, '(I8,1X,ES25.17,1X,I12,46(1X,ES25.17))') istep, z, info%npt, &
            sums, amax, r2max, -info%refpt(6), info%refpt(5)
      end do
    end do
  else
    nstep = sum(nseg(1:nelem))
    allocate(segs(nstep), recs(izh_run_record, nstep))
    k = 0
    zedge = 0.0d0
    do i = 1, nelem
      tau1 = 0.5d0 * elen(i) / nseg(i)
      e = element_gpu(bitype(i), val(:, i), mapstp(i), elen(i), zedge, rfh(i))
      do j = 1, nseg(i)
        k = k + 1
        segs(k)%tau1 = tau1
        segs(k)%e = e
        segs(k)%xrad = radius(i)
        segs(k)%yrad = radius(i)
        segs(k)%wake = c_null_ptr
        zedge = (zedge + tau1) + tau1
      end do
    end do
    call izh_check(izh_run(b, f, dom, z, segs, nstep, current, recs), 'run')
    ! z after each step: the lattice's cumulative length
    z = 0.0d0
    k = 0
    do i = 1, nelem
      do j = 1, nseg(i)
        k = k + 1
        z = z + elen(i) / nseg(i)
        write(u, '(I8,1X,ES25.17,1X,I12,46(1X,ES25.17))') k, z, &
            int(recs(45, k), c_int64_t), recs(1:44, k), -recs(47, k), recs(46, k)
      end do
    end do
  end if
  close(u)

  call izh_check(izh_bunch_destroy(b), 'bunch_destroy')
  call izh_check(izh_field_destroy(f), 'field_destroy')
  do i = 1, nelem
    if (c_associated(rfh(i))) call izh_check(izh_rfdata_destroy(rfh(i)), 'rfdata_destroy')
  end do

contains

  ! rfdataN.in of element ie (N = its file ID v4) into an izh_rfdata
  subroutine read_rfdata(ie)
    integer, intent(in) :: ie
    character(len=32) :: fname
    real(c_double), allocatable :: buf(:)
    real(c_double) :: row(4)
    integer :: uu, ios, nrow, ncol, id, kr
    id = int(val(4, ie) + 0.1d0)
    write(fname, '(a,i0,a)') 'rfdata', id, '.in'
    ncol = 4
    if (bitype(ie) == 105) ncol = 1
    open(newunit=uu, file=trim(fname), status='old')
    nrow = 0
    do
      read(uu, *, iostat=ios) row(1:ncol)
      if (ios /= 0) exit
      nrow = nrow + 1
    end do
    rewind(uu)
    allocate(buf(ncol * nrow))
    do kr = 1, nrow
      read(uu, *) buf(ncol * (kr - 1) + 1:ncol * kr)
    end do
    close(uu)
    if (ncol == 4) then
      call izh_check(izh_rfdata_table(buf, int(nrow, c_int), rfh(ie)), 'rfdata_table')
    else
      call izh_check(izh_rfdata_fourier(buf, int(nrow, c_int), rfh(ie)), 'rfdata_fourier')
    end if
    deallocate(buf)
  end subroutine read_rfdata
end program izh_lattice
