!----------------------------------------------------------------------------
! impactz_hip -- Fortran 90 binding of the C ABI in include/impactz_hip.h.
!
! The interfaces bind every entry point with iso_c_binding (handles are
! type(c_ptr), scalars pass by value, arrays by reference).  The wrapper
! subroutines below them keep the argument lists of the reference procedures
! they replace, so a host loop written like run_AccSimulator
! (src/Contrl/AccSimulator.f90:1099-1488) calls them in the same order, and
! turn a non-zero status into the reference's reaction to an error: print
! the message and stop.
!----------------------------------------------------------------------------
module impactz_hip
  use iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: IZH_OK = 0
  integer(c_int), parameter :: IZH_FP64 = 0, IZH_MIXED = 1
  integer(c_int), parameter :: IZH_RHO = 0, IZH_PHI = 1, IZH_EX = 2, &
                               IZH_EY = 3, IZH_EZ = 4

  ! PhysConst scalars (src/DataStruct/PhysConst.f90:18-35)
  type, bind(C) :: izh_const
    real(c_double) :: clight, pi, scxl, scfreq
  end type izh_const

  ! BeamBunch scalars (src/Appl/BeamBunch.f90:23-33)
  type, bind(C) :: izh_bunch_info
    real(c_double) :: current, mass, charge
    integer(c_int64_t) :: npt, nptlocal
    real(c_double) :: refpt(6)
    real(c_double) :: freq
    integer(c_int) :: precision
  end type izh_bunch_info

  ! CompDom for one process (src/Appl/CompDom.f90:19-32)
  ! izh_tuning: how izh_run works on a bunch (no reference counterpart)
  type, bind(C) :: izh_tuning
    integer(c_int) :: resort_period = 0
    integer(c_int) :: reserved(7) = 0
  end type izh_tuning

  type, bind(C) :: izh_compdom
    integer(c_int) :: nx, ny, nz
    real(c_double) :: range(6)
    real(c_double) :: msize(3)
  end type izh_compdom

  ! IZH_NVAL: the deck line's values v1..v14 (Param(2:15))
  integer, parameter :: izh_nval = 14
  ! IZH_RUN_RECORD: values per step from izh_run
  integer, parameter :: izh_run_record = 47

  ! a lattice element: type, RK6 steps, values, length, entrance z (its
  ! Param(1)) and an RF cavity's on-axis field (izh_rfdata, else c_null_ptr)
  type, bind(C) :: izh_element
    integer(c_int) :: bitype
    integer(c_int) :: mapstp
    real(c_double) :: v(izh_nval)
    real(c_double) :: length
    real(c_double) :: zedge
    type(c_ptr) :: rf
  end type izh_element

  ! a wakefield (kind 1 analytic: aa, gg, leng; kind 2 a read-in table with
  ! flagbtw; scale = scwk), src/Contrl/AccSimulator.f90:859-875,960-980
  type, bind(C) :: izh_wake
    integer(c_int) :: kind
    real(c_double) :: aa, gg, leng
    type(c_ptr) :: table
    integer(c_int) :: flagbtw
    real(c_double) :: scale
    real(c_double) :: r0, zoff, blength
  end type izh_wake

  ! one segment of a lattice: the element's half step, the element, the
  ! pipe aperture and the step's izh_wake (c_null_ptr: none)
  type, bind(C) :: izh_segment
    real(c_double) :: tau1
    type(izh_element) :: e
    real(c_double) :: xrad, yrad
    type(c_ptr) :: wake
  end type izh_segment

  interface
    function izh_last_error() bind(C, name="izh_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function izh_init(device) bind(C, name="izh_init") result(rc)
      import :: c_int
      integer(c_int), value :: device
      integer(c_int) :: rc
    end function
    function izh_physconst(freq, c) bind(C, name="izh_physconst") result(rc)
      import :: c_int, c_double, izh_const
      real(c_double), value :: freq
      type(izh_const) :: c
      integer(c_int) :: rc
    end function
    function izh_bunch_create(pts9, n, current, kinetic, mass, charge, freq, &
                              phase0, precision, b) &
        bind(C, name="izh_bunch_create") result(rc)
      import :: c_int, c_int64_t, c_double, c_ptr
      real(c_double), intent(in) :: pts9(*)
      integer(c_int64_t), value :: n
      real(c_double), value :: current, kinetic, mass, charge, freq, phase0
      integer(c_int), value :: precision
      type(c_ptr) :: b
      integer(c_int) :: rc
    end function
    function izh_bunch_destroy(b) bind(C, name="izh_bunch_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: rc
    end function
    function izh_bunch_download(b, pts9, cap, n) &
        bind(C, name="izh_bunch_download") result(rc)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double) :: pts9(*)
      integer(c_int64_t), value :: cap
      integer(c_int64_t) :: n
      integer(c_int) :: rc
    end function
    function izh_bunch_upload(b, pts9, n) bind(C, name="izh_bunch_upload") result(rc)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double), intent(in) :: pts9(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function izh_bunch_get_info(b, info) bind(C, name="izh_bunch_get_info") result(rc)
      import :: c_int, c_ptr, izh_bunch_info
      type(c_ptr), value :: b
      type(izh_bunch_info) :: info
      integer(c_int) :: rc
    end function
    function izh_bunch_set_info(b, info) bind(C, name="izh_bunch_set_info") result(rc)
      import :: c_int, c_ptr, izh_bunch_info
      type(c_ptr), value :: b
      type(izh_bunch_info), intent(in) :: info
      integer(c_int) :: rc
    end function
    function izh_quad_maplinear(c, tau, gradient, mapstp, refpt, charge, mass, xm) &
        bind(C, name="izh_quad_maplinear") result(rc)
      import :: c_int, c_double, izh_const
      type(izh_const), intent(in) :: c
      real(c_double), value :: tau, gradient
      integer(c_int), value :: mapstp
      real(c_double) :: refpt(6)
      real(c_double), value :: charge, mass
      real(c_double) :: xm(36)
      integer(c_int) :: rc
    end function
    function izh_element_maplinear(c, z, tau, e, refpt, charge, mass, xm) &
        bind(C, name="izh_element_maplinear") result(rc)
      import :: c_int, c_double, izh_const, izh_element
      type(izh_const), intent(in) :: c
      real(c_double), value :: z, tau
      type(izh_element), intent(in) :: e
      real(c_double) :: refpt(6)
      real(c_double), value :: charge, mass
      real(c_double) :: xm(36)
      integer(c_int) :: rc
    end function
    function izh_map1(b, z, tau, e) bind(C, name="izh_map1") result(rc)
      import :: c_int, c_double, c_ptr, izh_element
      type(c_ptr), value :: b
      real(c_double) :: z
      real(c_double), value :: tau
      type(izh_element), intent(in) :: e
      integer(c_int) :: rc
    end function
    ! bends: Fpol_Dipole (with the reference phase advance over the bend),
    ! Sector_Dipole, Bpol_Dipole and ISRchicane (src/Appl/Dipole.f90:466-788,
    ! src/Contrl/AccSimulator.f90:1542-1584)
    function izh_fpol_dipole(b, e) bind(C, name="izh_fpol_dipole") result(rc)
      import :: c_int, c_ptr, izh_element
      type(c_ptr), value :: b
      type(izh_element), intent(in) :: e
      integer(c_int) :: rc
    end function
    function izh_sector_dipole(b, z, tau, e) bind(C, name="izh_sector_dipole") result(rc)
      import :: c_int, c_double, c_ptr, izh_element
      type(c_ptr), value :: b
      real(c_double) :: z
      real(c_double), value :: tau
      type(izh_element), intent(in) :: e
      integer(c_int) :: rc
    end function
    function izh_bpol_dipole(b, e) bind(C, name="izh_bpol_dipole") result(rc)
      import :: c_int, c_ptr, izh_element
      type(c_ptr), value :: b
      type(izh_element), intent(in) :: e
      integer(c_int) :: rc
    end function
    function izh_isr_chicane(b, tau, e, rd, nrd) bind(C, name="izh_isr_chicane") result(rc)
      import :: c_int, c_int64_t, c_double, c_ptr, izh_element
      type(c_ptr), value :: b
      real(c_double), value :: tau
      type(izh_element), intent(in) :: e
      type(c_ptr), value :: rd
      integer(c_int64_t), value :: nrd
      integer(c_int) :: rc
    end function
    ! read1_Data's table (rows z, Ez, Ez', Ez'') / read3_Data's coefficients
    function izh_rfdata_table(rows4, n, d) bind(C, name="izh_rfdata_table") result(rc)
      import :: c_int, c_double, c_ptr
      real(c_double), intent(in) :: rows4(*)
      integer(c_int), value :: n
      type(c_ptr) :: d
      integer(c_int) :: rc
    end function
    function izh_rfdata_fourier(coef, n, d) bind(C, name="izh_rfdata_fourier") result(rc)
      import :: c_int, c_double, c_ptr
      real(c_double), intent(in) :: coef(*)
      integer(c_int), value :: n
      type(c_ptr) :: d
      integer(c_int) :: rc
    end function
    function izh_rfdata_destroy(d) bind(C, name="izh_rfdata_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: d
      integer(c_int) :: rc
    end function
    function izh_wakefield(b, f, dom, w) bind(C, name="izh_wakefield") result(rc)
      import :: c_int, c_ptr, izh_compdom, izh_wake
      type(c_ptr), value :: b, f
      type(izh_compdom), intent(in) :: dom
      type(izh_wake), intent(in) :: w
      integer(c_int) :: rc
    end function
    function izh_kick1wake(b, tau, f, dom, flagbc) bind(C, name="izh_kick1wake") result(rc)
      import :: c_int, c_double, c_ptr, izh_compdom
      type(c_ptr), value :: b, f
      real(c_double), value :: tau
      type(izh_compdom), intent(in) :: dom
      integer(c_int), value :: flagbc
      integer(c_int) :: rc
    end function
    function izh_kick2wake(b, z, tau, e, f, dom, flagbc) bind(C, name="izh_kick2wake") &
        result(rc)
      import :: c_int, c_double, c_ptr, izh_element, izh_compdom
      type(c_ptr), value :: b, f
      real(c_double), value :: z, tau
      type(izh_element), intent(in) :: e
      type(izh_compdom), intent(in) :: dom
      integer(c_int), value :: flagbc
      integer(c_int) :: rc
    end function
    function izh_wake_download(f, ex, ey, ez) bind(C, name="izh_wake_download") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: f
      real(c_double) :: ex(*), ey(*), ez(*)
      integer(c_int) :: rc
    end function
    function izh_drift2(b, z, tau) bind(C, name="izh_drift2") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double) :: z
      real(c_double), value :: tau
      integer(c_int) :: rc
    end function
    function izh_kick2(b, z, tau, e, f, dom, flagbc) bind(C, name="izh_kick2") result(rc)
      import :: c_int, c_double, c_ptr, izh_element, izh_compdom
      type(c_ptr), value :: b, f
      real(c_double), value :: z, tau
      type(izh_element), intent(in) :: e
      type(izh_compdom), intent(in) :: dom
      integer(c_int), value :: flagbc
      integer(c_int) :: rc
    end function
    function izh_step2(b, f, dom, z, tau1, e, xrad, yrad, bcurr, wake, moments) &
        bind(C, name="izh_step2") result(rc)
      import :: c_int, c_double, c_ptr, izh_element, izh_compdom
      type(c_ptr), value :: b, f
      type(izh_compdom) :: dom
      real(c_double) :: z
      real(c_double), value :: tau1
      type(izh_element), intent(in) :: e
      real(c_double), value :: xrad, yrad, bcurr
      type(c_ptr), value :: wake
      real(c_double) :: moments(44)
      integer(c_int) :: rc
    end function
    function izh_drift1(b, z, tau, bitype, gradient, mapstp) &
        bind(C, name="izh_drift1") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double) :: z
      real(c_double), value :: tau
      integer(c_int), value :: bitype
      real(c_double), value :: gradient
      integer(c_int), value :: mapstp
      integer(c_int) :: rc
    end function
    function izh_map_linear(b, xm) bind(C, name="izh_map_linear") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double), intent(in) :: xm(36)
      integer(c_int) :: rc
    end function
    function izh_conv1st(b, tau, ptrange, nplc, nptot, xrad, yrad, flagbc) &
        bind(C, name="izh_conv1st") result(rc)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double), value :: tau
      real(c_double) :: ptrange(6)
      integer(c_int64_t) :: nplc, nptot
      real(c_double), value :: xrad, yrad
      integer(c_int), value :: flagbc
      integer(c_int) :: rc
    end function
    function izh_update_compdom(dom, ptrange, flagbc) &
        bind(C, name="izh_update_compdom") result(rc)
      import :: c_int, c_double, izh_compdom
      type(izh_compdom) :: dom
      real(c_double), intent(in) :: ptrange(6)
      integer(c_int), value :: flagbc
      integer(c_int) :: rc
    end function
    function izh_field_create(nx, ny, nz, precision, f) &
        bind(C, name="izh_field_create") result(rc)
      import :: c_int, c_ptr
      integer(c_int), value :: nx, ny, nz, precision
      type(c_ptr) :: f
      integer(c_int) :: rc
    end function
    function izh_field_destroy(f) bind(C, name="izh_field_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: f
      integer(c_int) :: rc
    end function
    function izh_charge(b, dom, f) bind(C, name="izh_charge") result(rc)
      import :: c_int, c_ptr, izh_compdom
      type(c_ptr), value :: b
      type(izh_compdom), intent(in) :: dom
      type(c_ptr), value :: f
      integer(c_int) :: rc
    end function
    function izh_update3O(f, dom) bind(C, name="izh_update3O") result(rc)
      import :: c_int, c_ptr, izh_compdom
      type(c_ptr), value :: f
      type(izh_compdom), intent(in) :: dom
      integer(c_int) :: rc
    end function
    function izh_cvbkforth1st(b) bind(C, name="izh_cvbkforth1st") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: rc
    end function
    function izh_kick1(b, tau, f, dom, flagbc) bind(C, name="izh_kick1") result(rc)
      import :: c_int, c_double, c_ptr, izh_compdom
      type(c_ptr), value :: b
      real(c_double), value :: tau
      type(c_ptr), value :: f
      type(izh_compdom), intent(in) :: dom
      integer(c_int), value :: flagbc
      integer(c_int) :: rc
    end function
    function izh_diagnostic1(b, sums, amax, r2max) &
        bind(C, name="izh_diagnostic1") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: b
      real(c_double) :: sums(37), amax(6), r2max
      integer(c_int) :: rc
    end function
    ! diagnostic1 sums of device rows the caller holds, allocating nothing
    ! (no reference counterpart; for callers that cannot allocate)
    function izh_diagnostic1_rows(rows, ld, n, out, stream) &
        bind(C, name="izh_diagnostic1_rows") result(rc)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: rows, out, stream
      integer(c_int64_t), value :: ld, n
      integer(c_int) :: rc
    end function
    function izh_bunch_set_tuning(b, t) bind(C, name="izh_bunch_set_tuning") result(rc)
      import :: c_int, c_ptr, izh_tuning
      type(c_ptr), value :: b
      type(izh_tuning) :: t
      integer(c_int) :: rc
    end function
    function izh_bunch_get_tuning(b, t) bind(C, name="izh_bunch_get_tuning") result(rc)
      import :: c_int, c_ptr, izh_tuning
      type(c_ptr), value :: b
      type(izh_tuning) :: t
      integer(c_int) :: rc
    end function
    ! a bunch from a resident pool bunch (no reference counterpart)
    function izh_bunch_create_from(src, offset, n, current, precision, b) &
        bind(C, name="izh_bunch_create_from") result(rc)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: src
      integer(c_int64_t), value :: offset, n
      real(c_double), value :: current
      integer(c_int), value :: precision
      type(c_ptr) :: b
      integer(c_int) :: rc
    end function
    ! lattice-scan jobs (no reference counterpart)
    function izh_job_create(b, f, dom, z, period, nperiod, nsteps, bcurr, use_graph, j) &
        bind(C, name="izh_job_create") result(rc)
      import :: c_int, c_int64_t, c_double, c_ptr, izh_compdom, izh_segment
      type(c_ptr), value :: b, f
      type(izh_compdom) :: dom
      real(c_double), value :: z, bcurr
      type(izh_segment), intent(in) :: period(*)
      integer(c_int64_t), value :: nperiod, nsteps
      integer(c_int), value :: use_graph
      type(c_ptr) :: j
      integer(c_int) :: rc
    end function
    function izh_job_advance(j, max_steps, left) bind(C, name="izh_job_advance") result(rc)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: j
      integer(c_int64_t), value :: max_steps
      integer(c_int64_t) :: left
      integer(c_int) :: rc
    end function
    function izh_job_query(j, state) bind(C, name="izh_job_query") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: j
      integer(c_int) :: state
      integer(c_int) :: rc
    end function
    function izh_job_finish(j, z, moments) bind(C, name="izh_job_finish") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: j
      real(c_double) :: z
      real(c_double) :: moments(*)
      integer(c_int) :: rc
    end function
    function izh_job_info(j, info) bind(C, name="izh_job_info") result(rc)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: j
      integer(c_int64_t) :: info(4)
      integer(c_int) :: rc
    end function
    function izh_job_destroy(j) bind(C, name="izh_job_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: j
      integer(c_int) :: rc
    end function
    function izh_chgupdate(b, nchge, qmcclist, counts) &
        bind(C, name="izh_chgupdate") result(rc)
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: b
      integer(c_int), value :: nchge
      real(c_double) :: qmcclist(*)
      integer(c_int64_t) :: counts(*)
      integer(c_int) :: rc
    end function
    function izh_field_download(f, which, host) &
        bind(C, name="izh_field_download") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: f
      integer(c_int), value :: which
      real(c_double) :: host(*)
      integer(c_int) :: rc
    end function
    function izh_field_upload_rho(f, rho) bind(C, name="izh_field_upload_rho") result(rc)
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: f
      real(c_double), intent(in) :: rho(*)
      integer(c_int) :: rc
    end function
    function izh_step(b, f, dom, z, seg, bcurr, moments) &
        bind(C, name="izh_step") result(rc)
      import :: c_int, c_double, c_ptr, izh_compdom, izh_segment
      type(c_ptr), value :: b, f
      type(izh_compdom) :: dom
      real(c_double) :: z
      type(izh_segment), intent(in) :: seg
      real(c_double), value :: bcurr
      real(c_double) :: moments(44)
      integer(c_int) :: rc
    end function
    function izh_run(b, f, dom, z, segs, nseg, bcurr, moments) &
        bind(C, name="izh_run") result(rc)
      import :: c_int, c_int64_t, c_double, c_ptr, izh_compdom, izh_segment
      type(c_ptr), value :: b, f
      type(izh_compdom) :: dom
      real(c_double) :: z
      type(izh_segment), intent(in) :: segs(*)
      integer(c_int64_t), value :: nseg
      real(c_double), value :: bcurr
      real(c_double) :: moments(*)
      integer(c_int) :: rc
    end function
    function izh_run_stats(b, stats) bind(C, name="izh_run_stats") result(rc)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: b
      integer(c_int64_t) :: stats(4)
      integer(c_int) :: rc
    end function
    function izh_probe_copy(nbytes, ms) bind(C, name="izh_probe_copy") result(rc)
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: nbytes
      real(c_double) :: ms
      integer(c_int) :: rc
    end function
    function izh_timer_enable(on) bind(C, name="izh_timer_enable") result(rc)
      import :: c_int
      integer(c_int), value :: on
      integer(c_int) :: rc
    end function
    function izh_timer_read(ms, launches, reset) bind(C, name="izh_timer_read") result(rc)
      import :: c_int, c_int64_t, c_double
      real(c_double) :: ms(*)
      integer(c_int64_t) :: launches(*)
      integer(c_int), value :: reset
      integer(c_int) :: rc
    end function
    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! print + stop on a non-zero status, as the reference reacts to errors
  subroutine izh_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n, i
    character(len=:), allocatable :: text
    if (rc == IZH_OK) return
    p = izh_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    allocate(character(len=n) :: text)
    do i = 1, n
      text(i:i) = msg(i)
    end do
    print *, 'impactz_hip: ', where, ' failed (status ', rc, '): ', text
    stop 1
  end subroutine izh_check

  ! The izh_element of a lattice element: type, values v1..v14 (Param(2:15)),
  ! mapstp, length, entrance z and an RF cavity's field (c_null_ptr if none).
  function element_gpu(bitype, v, mapstp, elen, zedge, rf) result(e)
    integer, intent(in) :: bitype, mapstp
    real(c_double), intent(in) :: v(izh_nval), elen, zedge
    type(c_ptr), intent(in) :: rf
    type(izh_element) :: e
    e%bitype = int(bitype, c_int)
    e%mapstp = int(mapstp, c_int)
    e%v = v
    e%length = elen
    e%zedge = zedge
    e%rf = rf
  end function element_gpu

  ! drift1_BeamBunch(this,beamln,z,tau,bitype,nseg,nst,ihlf)
  ! (src/Appl/BeamBunch.f90:250-258): the element's data enter as an
  ! izh_element (element_gpu).
  subroutine drift1_gpu(b, z, tau, e)
    type(c_ptr), intent(in) :: b
    real(c_double), intent(inout) :: z
    real(c_double), intent(in) :: tau
    type(izh_element), intent(in) :: e
    call izh_check(izh_map1(b, z, tau, e), 'drift1')
  end subroutine drift1_gpu

  ! conv1st_BeamBunch(this,tau,nplc,nptot,ptrange,Flagbc,perd,xrad,yrad)
  ! (src/Appl/BeamBunch.f90:1252-1260)
  subroutine conv1st_gpu(b, tau, nplc, nptot, ptrange, flagbc, xrad, yrad)
    type(c_ptr), intent(in) :: b
    real(c_double), intent(in) :: tau, xrad, yrad
    integer(c_int64_t), intent(out) :: nplc, nptot
    real(c_double), intent(out) :: ptrange(6)
    integer, intent(in) :: flagbc
    call izh_check(izh_conv1st(b, tau, ptrange, nplc, nptot, xrad, yrad, &
                               int(flagbc, c_int)), 'conv1st')
  end subroutine conv1st_gpu

  ! update_CompDom(this,ptrange,grid2d,Flagbc) (src/Appl/CompDom.f90:509-515)
  subroutine update_compdom_gpu(dom, ptrange, flagbc)
    type(izh_compdom), intent(inout) :: dom
    real(c_double), intent(in) :: ptrange(6)
    integer, intent(in) :: flagbc
    call izh_check(izh_update_compdom(dom, ptrange, int(flagbc, c_int)), &
                   'update_CompDom')
  end subroutine update_compdom_gpu

  ! charge_BeamBunch (src/Appl/BeamBunch.f90:3535-3544): the grid charge
  ! stays in the device field f
  subroutine charge_gpu(b, dom, f)
    type(c_ptr), intent(in) :: b, f
    type(izh_compdom), intent(in) :: dom
    call izh_check(izh_charge(b, dom, f), 'charge')
  end subroutine charge_gpu

  ! update3O_FieldQuant (src/Appl/Field.f90:108-117)
  subroutine update3O_gpu(f, dom)
    type(c_ptr), intent(in) :: f
    type(izh_compdom), intent(in) :: dom
    call izh_check(izh_update3O(f, dom), 'update3O')
  end subroutine update3O_gpu

  ! cvbkforth1st_BeamBunch(this) (src/Appl/BeamBunch.f90:1667)
  subroutine cvbkforth1st_gpu(b)
    type(c_ptr), intent(in) :: b
    call izh_check(izh_cvbkforth1st(b), 'cvbkforth1st')
  end subroutine cvbkforth1st_gpu

  ! kick1_BeamBunch(this,tau,innx,inny,innz,temppotent,ptsgeom,grid,Flagbc,perdlen)
  ! (src/Appl/BeamBunch.f90:1707-1716): the potential is the device field f
  subroutine kick1_gpu(b, tau, f, dom, flagbc)
    type(c_ptr), intent(in) :: b, f
    real(c_double), intent(in) :: tau
    type(izh_compdom), intent(in) :: dom
    integer, intent(in) :: flagbc
    call izh_check(izh_kick1(b, tau, f, dom, int(flagbc, c_int)), 'kick1')
  end subroutine kick1_gpu

  ! drift2_BeamBunch(this,z,tau) (src/Appl/BeamBunch.f90:467-509), the
  ! Lorentz integrator's drift (Flagmap=2)
  subroutine drift2_gpu(b, z, tau)
    type(c_ptr), intent(in) :: b
    real(c_double), intent(inout) :: z
    real(c_double), intent(in) :: tau
    call izh_check(izh_drift2(b, z, tau), 'drift2')
  end subroutine drift2_gpu

  ! kick2_BeamBunch(this,beamelem,z,tau,...) (src/Appl/BeamBunch.f90:
  ! 2187-2508): space-charge field plus the element's field (e, with its
  ! entrance zedge), Lorentz push
  subroutine kick2_gpu(b, z, tau, e, f, dom, flagbc)
    type(c_ptr), intent(in) :: b, f
    real(c_double), intent(in) :: z, tau
    type(izh_element), intent(in) :: e
    integer, intent(in) :: flagbc
    type(izh_compdom), intent(in) :: dom
    call izh_check(izh_kick2(b, z, tau, e, f, dom, int(flagbc, c_int)), 'kick2')
  end subroutine kick2_gpu

  ! the reductions of diagnostic1_Output (src/Contrl/Output.f90:119-245)
  subroutine diagnostic1_sums_gpu(b, sums, amax, r2max)
    type(c_ptr), intent(in) :: b
    real(c_double), intent(out) :: sums(37), amax(6), r2max
    call izh_check(izh_diagnostic1(b, sums, amax, r2max), 'diagnostic1')
  end subroutine diagnostic1_sums_gpu

  ! chgupdate_BeamBunch (src/Appl/BeamBunch.f90:688-728): nptlist(j) = live
  ! particles of charge state j
  subroutine chgupdate_gpu(b, nchge, qmcclist, nptlist)
    type(c_ptr), intent(in) :: b
    integer, intent(in) :: nchge
    real(c_double), intent(in) :: qmcclist(nchge)
    integer, intent(out) :: nptlist(nchge)
    integer(c_int64_t) :: counts(nchge)
    call izh_check(izh_chgupdate(b, int(nchge, c_int), qmcclist, counts), 'chgupdate')
    nptlist = int(counts)
  end subroutine chgupdate_gpu

end module impactz_hip
