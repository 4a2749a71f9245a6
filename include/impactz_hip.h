/*
 * impactz_hip.h -- C ABI of the MI355X (gfx950) space-charge step of IMPACT-Z.
 *
 * One entry point per reference procedure on the step, so a Fortran host can
 * bind each with iso_c_binding and keep the reference's call order
 * (src/Contrl/AccSimulator.f90:1099-1488).  Conventions:
 *   - every function returns int status: IZH_OK (0) or an IZH_ERR_* code;
 *     izh_last_error() gives the message of the calling thread's last error;
 *   - no globals carry physics: constants are passed or live in the handles;
 *   - particle arrays crossing the ABI use the reference layout Pts1(9,N)
 *     (9 contiguous doubles per particle: x/Scxl, gamma*beta_x, y/Scxl,
 *     gamma*beta_y, phase, -delta gamma, q/mc^2, charge weight, id);
 *   - grids crossing the ABI use the reference layout rho(Nx,Ny,Nz), x fastest;
 *   - a bunch and its field stay resident on the GPU between calls; host
 *     buffers are only read or written by the upload/download calls.
 * Handles are not thread-safe; one host thread drives a handle at a time.
 */
#ifndef IMPACTZ_HIP_H
#define IMPACTZ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IZH_OK 0
#define IZH_ERR_ARG 1      /* bad argument or unsupported option */
#define IZH_ERR_HIP 2      /* a HIP runtime call failed */
#define IZH_ERR_NOMEM 3    /* device allocation failed */
#define IZH_ERR_STATE 4    /* handle used in the wrong state */

#define IZH_FP64 0         /* every stage in FP64 */
#define IZH_MIXED 1        /* FP32 deposit/FFT/gather, FP64 push, state, moments */

#define IZH_RHO 0          /* field buffers for izh_field_download */
#define IZH_PHI 1
#define IZH_EX 2
#define IZH_EY 3
#define IZH_EZ 4

typedef struct izh_bunch izh_bunch;
typedef struct izh_field izh_field;

/* Constants the reference holds in PhysConst (src/DataStruct/PhysConst.f90:18-35). */
typedef struct {
  double clight, pi, scxl, scfreq;
} izh_const;

/* Host-visible bunch scalars: BeamBunch%Current, %Mass, %Charge, %Npt,
 * %Nptlocal and %refptcl (src/Appl/BeamBunch.f90:23-33). */
typedef struct {
  double current, mass, charge;
  int64_t npt, nptlocal;
  double refpt[6];
  double freq;
  int precision;
} izh_bunch_info;

/* CompDom state for one process (src/Appl/CompDom.f90:19-32):
 * Meshnum, SpatRange and Meshsize. */
typedef struct {
  int nx, ny, nz;
  double range[6];
  double msize[3];
} izh_compdom;

const char *izh_last_error(void);

/* Select the device for the calling thread and warm up the runtime. */
int izh_init(int device);

/* construct_PhysConst (src/DataStruct/PhysConst.f90:25-35). */
int izh_physconst(double freq, izh_const *out);

/* construct_BeamBunch + readin_Dist (src/Appl/BeamBunch.f90:41-59,
 * src/Appl/Distribution.f90:1666-1714): upload n particles (Pts1(9,n)) and
 * keep them device resident in SoA form.  refptcl(1:5)=phase0,
 * refptcl(6)=-(kinetic/mass+1).  precision: IZH_FP64 or IZH_MIXED. */
int izh_bunch_create(const double *pts9, int64_t n, double current,
                     double kinetic, double mass, double charge, double freq,
                     double phase0, int precision, izh_bunch **out);
int izh_bunch_destroy(izh_bunch *b);
/* A new bunch of the n particles that src holds in slots offset ..
 * offset+n-1, copied on the device (no reference counterpart: a server
 * keeps one resident particle pool and gives each request a bunch from it
 * without a host upload).  The beam (mass, charge, frequency, refptcl) is
 * src's; the current is the given one, and each particle's charge weight
 * Pts1(8) = current / (freq n) sign(charge), as the distribution readers
 * set it for n particles.  src must be in the z frame. */
int izh_bunch_create_from(const izh_bunch *src, int64_t offset, int64_t n, double current,
                          int precision, izh_bunch **out);
/* getpts_BeamBunch (src/Appl/BeamBunch.f90:3661-3668): copy the local
 * particles back in the reference's order.  cap = rows of pts9. */
int izh_bunch_download(izh_bunch *b, double *pts9, int64_t cap, int64_t *n);
/* setpts_BeamBunch + setnpt_BeamBunch (src/Appl/BeamBunch.f90:3632-3656). */
int izh_bunch_upload(izh_bunch *b, const double *pts9, int64_t n);
int izh_bunch_get_info(const izh_bunch *b, izh_bunch_info *info);
int izh_bunch_set_info(izh_bunch *b, const izh_bunch_info *info);

/* Element values.  v[0..13] are the deck line's v1..v14, which become the
 * element's Param(2:15) (src/Contrl/AccSimulator.f90:262-536); unused
 * entries are 0:
 *   bitype 0 drift:              pipe radius
 *   bitype 1 quadrupole:         gradient T/m (K in 1/m^2 when v[1] < -1e-6),
 *                                file ID, pipe radius
 *   bitype 2 constant focusing:  kx0^2, ky0^2, kz0^2 (1/m^2), pipe radius
 *   bitype 3 solenoid:           Bz0 (T), file ID, pipe radius, x offset
 *   bitype 101-104 DTL, CCDTL, CCL, SC (src/Appl/DTL.f90:14-45 and alike):
 *                                field scale, RF frequency (Hz), phase
 *                                (degrees), file ID, pipe radius; for the DTL
 *                                also quadrupole 1 length and gradient,
 *                                quadrupole 2 length and gradient
 *   bitype 105 SolRF (src/Appl/SolRF.f90:16-34): field scale, RF frequency,
 *                                phase, file ID, pipe radius, 5 misalignment
 *                                values, Bz0 (T), 3 wakefield values */
#define IZH_NVAL 14

/* The on-axis field of RF cavities, as the reference's Data module holds it
 * (src/DataStruct/Data.f90:41-61) after reading an rfdataN.in file.  Host
 * memory; the library copies the values. */
typedef struct izh_rfdata izh_rfdata;
/* read1_Data (src/DataStruct/Data.f90:139-215): n rows of z, Ez, Ez', Ez''
 * (row-major, 4 doubles each); z is shifted so the first row is at 0.  The
 * DTL, CCDTL, CCL and SC maps interpolate it linearly. */
int izh_rfdata_table(const double *rows4, int n, izh_rfdata **out);
/* read3_Data (src/DataStruct/Data.f90:348-409): the n values of the file
 * (coefficient count, field start, end and period, then the Fourier
 * coefficients), padded with one zero to an odd count.  SolRF's map sums
 * the series. */
int izh_rfdata_fourier(const double *coef, int n, izh_rfdata **out);
int izh_rfdata_destroy(izh_rfdata *d);

/* A lattice element as the push routines see it: its type, RK6 steps, the
 * deck values, its length, the z of its entrance (the Param(1) "zedge"
 * run_AccSimulator sets on entry, src/Contrl/AccSimulator.f90:1025-1026)
 * and, for RF cavities, its on-axis field (else NULL). */
typedef struct {
  int bitype;
  int mapstp;
  double v[IZH_NVAL];
  double length;
  double zedge;
  const izh_rfdata *rf;
} izh_element;

/* maplinear_BeamLineElem (src/Appl/BeamLineElem.f90:776-825) for a
 * constant-gradient quadrupole (bitype 1, file ID <= 1e-5), constant
 * focusing (2), a solenoid (3) and the RF cavities 101-105: the 6x6 linear
 * map of a step tau from z, from RK6 integration over e->mapstp steps
 * (maplinear_Quadrupole, _ConstFoc, _Sol, _DTL, _CCDTL, _CCL, _SC, _SolRF:
 * src/Appl/Quadrupole.f90:120-305, ConstFoc.f90:116-294, Sol.f90:120-321,
 * DTL.f90:136-418, CCDTL.f90:122-373, CCL.f90:122-375, SC.f90:122-373,
 * SolRF.f90:126-479).  Host only.  xm is row-major,
 * xm[6*(j-1)+(k-1)] = xm(j,k); refpt(5:6) are updated (an RF cavity
 * changes the reference energy). */
int izh_element_maplinear(const izh_const *c, double z, double tau, const izh_element *e,
                          double refpt[6], double charge, double mass, double xm[36]);

/* The quadrupole case of izh_element_maplinear, v = {gradient, 0, 0, 0}
 * (maplinear_Quadrupole, src/Appl/Quadrupole.f90:120-305). */
int izh_quad_maplinear(const izh_const *c, double tau, double gradient,
                       int mapstp, double refpt[6], double charge, double mass,
                       double xm[36]);

/* drift1_BeamBunch (src/Appl/BeamBunch.f90:250-465): the exact drift
 * (bitype 0); a quadrupole (1) through its linear map, or through the
 * chromatic K-mode map transfmapK_Quadrupole (src/Appl/Quadrupole.f90:
 * 509-607) when v[1] < -1e-6; constant focusing (2), a solenoid (3) and the
 * RF cavities 101-105 through their linear maps (izh_element_maplinear).
 * *z += tau.  Refused with IZH_ERR_ARG: the undulator map (bitype 1,
 * v[1] < -10), quadrupole fields from a file (v[1] > 1e-5), the ideal
 * RF-cavity branch (bitype >= 2 with Param(5) = v[3] < 0), an RF cavity
 * without its field and other element types (bends: izh_sector_dipole). */
int izh_map1(izh_bunch *b, double *z, double tau, const izh_element *e);

/* Bends (bitype 4).  Their values, as run_AccSimulator reads them
 * (src/Contrl/AccSimulator.f90:1058-1088): v[0] bend angle [rad], v[1] k1,
 * v[2] CSR switch, v[3] half gap [m], v[4] and v[5] the entrance and exit
 * pole-face angles [rad], v[6] and v[7] the pole-face curvatures 1/r1 and
 * 1/r2 [1/m], v[8] the fringe-field integral of both faces.  The reference
 * energy is the bunch's, which no bend map changes.
 * izh_fpol_dipole: the phase advance of the reference particle over the
 * whole bend, refpt[4] += L/(Scxl beta0) (src/Contrl/AccSimulator.f90:1062),
 * then the entrance face Fpol_Dipole (src/Appl/Dipole.f90:466-526). */
int izh_fpol_dipole(izh_bunch *b, const izh_element *e);
/* Sector_Dipole over tau (src/Appl/Dipole.f90:693-788); *z += tau. */
int izh_sector_dipole(izh_bunch *b, double *z, double tau, const izh_element *e);
/* The exit face Bpol_Dipole (src/Appl/Dipole.f90:528-591). */
int izh_bpol_dipole(izh_bunch *b, const izh_element *e);
/* ISRchicane (src/Contrl/AccSimulator.f90:1542-1584), the synchrotron
 * radiation of a bend with current, over tau.  rd: device array of one
 * standard-normal deviate per particle, indexed by the particle's row at
 * izh_bunch_create or the last izh_bunch_upload (the reference draws them
 * with normVec1); nrd: its length, at least the particles of that upload
 * (losses since do not shorten it), else IZH_ERR_ARG. */
int izh_isr_chicane(izh_bunch *b, double tau, const izh_element *e, const double *rd,
                    int64_t nrd);

/* izh_map1 for bitype 0 and 1 with v = {gradient, 0, ...}. */
int izh_drift1(izh_bunch *b, double *z, double tau, int bitype,
               double gradient, int mapstp);
/* The map-apply half of drift1_BeamBunch for any host-computed 6x6 map
 * (src/Appl/BeamBunch.f90:308-320); refpt is the caller's to update. */
int izh_map_linear(izh_bunch *b, const double xm[36]);

/* conv1st_BeamBunch, Flagbc=1 (src/Appl/BeamBunch.f90:1252-1665): z->t
 * frame, phase wrap, range of 0th+1st order positions, loss test and
 * order-preserving compaction, Current rescale.  ptrange may be NULL. */
int izh_conv1st(izh_bunch *b, double tau, double ptrange[6], int64_t *nplc,
                int64_t *nptot, double xrad, double yrad, int flagbc);

/* update_CompDom, one process (src/Appl/CompDom.f90:509-823): new bounding
 * box padded by REAL(4) 1/(N-3) of its extent, h = extent/(N-1). */
int izh_update_compdom(izh_compdom *dom, const double ptrange[6], int flagbc);

/* FieldQuant workspace for one mesh (src/Appl/Field.f90:19-100): charge
 * density, potential, field and the doubled-grid FFT buffers. */
int izh_field_create(int nx, int ny, int nz, int precision, izh_field **out);
int izh_field_destroy(izh_field *f);

/* charge_BeamBunch -> deposit_BeamBunch (src/Appl/BeamBunch.f90:3535-3768):
 * deterministic CIC deposition of the t-frame particles into the field. */
int izh_charge(izh_bunch *b, const izh_compdom *dom, izh_field *f);

/* update3O_FieldQuant -> openBC3D (src/Appl/Field.f90:108-331): open-boundary
 * Poisson solve with the integrated Green function on the doubled grid. */
int izh_update3O(izh_field *f, const izh_compdom *dom);

/* cvbkforth1st_BeamBunch (src/Appl/BeamBunch.f90:1667-1701). */
int izh_cvbkforth1st(izh_bunch *b);

/* kick1_BeamBunch (src/Appl/BeamBunch.f90:1707-2181), Flagbc=1: E from the
 * potential, CIC gather, kick and 0th-order t->z back transform. */
int izh_kick1(izh_bunch *b, double tau, izh_field *f, const izh_compdom *dom,
              int flagbc);

/* diagnostic1_Output reductions (src/Contrl/Output.f90:119-245): the 37 sums
 * in the reference's order, 6 abs-maxima, max of x^2+y^2. */
int izh_diagnostic1(izh_bunch *b, double sums[37], double amax[6],
                    double *r2max);

/* izh_diagnostic1 on particle rows the caller holds, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  rows: device memory, six FP64 rows (x, px, y, py,
 * phi, pt in the Pts1 units) of n particles, row k at rows + k*ld; out:
 * device memory for 44 doubles, the izh_diagnostic1 results in one array
 * (37 sums, 6 abs-maxima, max x^2+y^2), bitwise the same as izh_diagnostic1
 * on a bunch holding those rows in that order.  stream: a hipStream_t (NULL:
 * the null stream); the call only enqueues, the caller synchronises.  The
 * per-block partials live in one static device array per GPU
 * (2048 x 44 doubles, 704 KiB, held until the process ends), so calls that
 * overlap on one GPU race and must be ordered by the caller.  0 <= n <= ld
 * <= IZH_DIAG_ROWS_MAX. */
#define IZH_DIAG_ROWS_MAX (1LL << 40)
int izh_diagnostic1_rows(const double *rows, int64_t ld, int64_t n, double *out,
                         void *stream);

/* chgupdate_BeamBunch (src/Appl/BeamBunch.f90:688-728): counts[j] = live
 * particles of charge state j, a particle counting for state j when
 * |(Pts1(7) - qmcclist[j]) / qmcclist[j]| < 1.0e-4 (the reference's REAL(4)
 * literal); 1 <= nchge <= 100.  The reference keeps these as nptlist for
 * diagnostic1_Output's fort.32. */
int izh_chgupdate(izh_bunch *b, int nchge, const double *qmcclist, int64_t *counts);

/* Parity access to the field (host arrays of nx*ny*nz doubles). */
int izh_field_download(izh_field *f, int which, double *host);
int izh_field_upload_rho(izh_field *f, const double *rho);

/* ---- 1D wakefields (kick1wake_BeamBunch) ---- */

/* The wakefield of a space-charge step, as run_AccSimulator switches it
 * (src/Contrl/AccSimulator.f90:759,859-875,960-980): kind 1 the analytic
 * short-range wake of a SolRF/TWS element with iris radius aa, gap gg and
 * period leng (its Param(13:15), wakefield_FieldQuant); kind 2 the wake
 * functions of a -41 element's file (read1wk_Data rows z, Wz, Wx, Wy, given
 * as an izh_rfdata_table; wakefieldread_FieldQuant) with flagbtw (2, 3: no
 * transverse wake; 4: no longitudinal); scale = scwk.  Kind 3 the coherent
 * synchrotron radiation wake of a bend (csrwakeTrIGF_FieldQuant,
 * src/Appl/Field.f90:1007-1251, as the loop calls it at
 * src/Contrl/AccSimulator.f90:1344-1396): bend radius r0 (m), bend length
 * blength (m), and zoff, the distance (m) the loop adds to range(5)/gamma
 * for the first mesh line's distance from the bend entrance (z - zbleng in
 * the bend, z - zbleng + sentr after it); its transverse wakes are zero and
 * scale is not applied. */
typedef struct {
  int kind;
  double aa, gg, leng;
  const izh_rfdata *table;
  int flagbtw;
  double scale;
  double r0, zoff, blength;
} izh_wake;

/* The wake fields of the bunch after izh_conv1st (its t-frame z binned on
 * the Nz mesh lines, src/Contrl/AccSimulator.f90:1272-1338) and
 * wakefield_FieldQuant / wakefieldread_FieldQuant (src/Appl/Field.f90:
 * 590-1001), kept in the field for izh_kick1wake.  Call it after
 * izh_update3O and before izh_cvbkforth1st, with the domain of the step;
 * Nz <= 1024. */
int izh_wakefield(izh_bunch *b, izh_field *f, const izh_compdom *dom, const izh_wake *w);

/* kick1wake_BeamBunch (src/Appl/BeamBunch.f90:4694-5046): izh_kick1 with the
 * wake fields of the last izh_wakefield added to each z plane of the grid
 * field. */
int izh_kick1wake(izh_bunch *b, double tau, izh_field *f, const izh_compdom *dom,
                  int flagbc);

/* The wake fields exwake, eywake, ezwake (nz doubles each, V/m) of the last
 * izh_wakefield, for parity checks (the reference never writes them). */
int izh_wake_download(izh_field *f, double *ex, double *ey, double *ez);

/* One segment of a lattice: an element's half step and values. */
typedef struct {
  double tau1;            /* half step, L/(2*nseg) */
  izh_element e;          /* the element */
  double xrad, yrad;      /* pipe aperture of the loss test */
  const izh_wake *wake;   /* the step's wakefield (NULL: none) */
} izh_segment;

/* One segment of the reference loop: map1(tau1), conv1st, update_CompDom,
 * charge, update3O, [wakefield,] cvbkforth1st, kick1 or kick1wake (2 tau1),
 * map1(tau1)
 * (src/Contrl/AccSimulator.f90:1099-1488 with Flagmap=1, Flagbc=1,
 * Flagsubstep=0), for every element izh_map1 takes; seg->wake (NULL: none)
 * is the step's wakefield, which a SolRF with Param(15) > 0 must have.
 * moments (NULL or 44 doubles) receives the diagnostic1_Output sums of the
 * state after the step: sums[37], amax[6], r2max.  bcurr is the deck
 * current (no space charge below 1e-10 A). */
int izh_step(izh_bunch *b, izh_field *f, izh_compdom *dom, double *z,
             const izh_segment *seg, double bcurr, double *moments);

/* ---- The nonlinear Lorentz integrator (Flagmap=2) ----
 * Element fields (the izh_element's bitype, values, length and zedge):
 * none for a drift, B = G (y, x, 0) for a quadrupole (file ID <= 0), the
 * constant-focusing and solenoid fields of getfld_ConstFoc / getfld_Sol. */

/* drift2_BeamBunch (src/Appl/BeamBunch.f90:467-509): a drift of tau/2;
 * *z and refpt(5) advance by tau/2. */
int izh_drift2(izh_bunch *b, double *z, double tau);

/* kick2_BeamBunch (src/Appl/BeamBunch.f90:2187-2508), Flagbc=1, at the
 * lattice position z: after izh_cvbkforth1st with current, the space-charge
 * field of the last izh_update3O plus the element's field (scatter2); on a
 * z-frame bunch the element's field alone (scatter20); each applied by the
 * implicit Lorentz push.  refpt(6) follows the on-axis Ez.  The bunch ends
 * in the z frame. */
int izh_kick2(izh_bunch *b, double z, double tau, const izh_element *e, izh_field *f,
              const izh_compdom *dom, int flagbc);

/* kick2wake_BeamBunch (src/Appl/BeamBunch.f90:5048-5410) with
 * scatter2wake (:5412-5938): izh_kick2 with the wake fields of the last
 * izh_wakefield, interpolated linearly in each particle's z and added to
 * its E (not to B).  Without current it is izh_kick2 (scatter20). */
int izh_kick2wake(izh_bunch *b, double z, double tau, const izh_element *e, izh_field *f,
                  const izh_compdom *dom, int flagbc);

/* One segment of the reference loop with Flagmap=2: drift2(2 tau1),
 * conv1st, update_CompDom, charge, update3O, cvbkforth1st, kick2(2 tau1),
 * drift2(2 tau1) (src/Contrl/AccSimulator.f90:1099-1488); with a wake
 * (NULL: none) the wake fields after update3O and kick2wake instead of
 * kick2 (:1446-1454).  Arguments and moments as for izh_step. */
int izh_step2(izh_bunch *b, izh_field *f, izh_compdom *dom, double *z, double tau1,
              const izh_element *e, double xrad, double yrad, double bcurr,
              const izh_wake *wake, double *moments);

/* nseg consecutive izh_step calls (the element/segment loop of
 * run_AccSimulator, src/Contrl/AccSimulator.f90:742-1504; every element of
 * izh_map1, the K-mode quadrupole's chromatic map included) run as one
 * device-resident pipeline: each step's kick, second half push and moment
 * sums are fused with the next step's first half push and z->t transform in
 * one pass over the particles, which are re-sorted into 8x8x8-cell tiles
 * (every izh_tuning.resort_period-th pass, in that same pass) so the deposit
 * accumulates in LDS; the domain update stays
 * on the device, so the host does not wait between steps.  Results equal
 * izh_step's to rounding.  moments (NULL or IZH_RUN_RECORD*nseg doubles)
 * receives per step the 44 diagnostic1_Output values izh_step gives, then
 * Npt, refpt(5) and refpt(6) (RF cavities change the reference energy).  After the call the
 * bunch is in the z frame; its particle order is by tile (izh_bunch_download
 * restores the reference's order). */
#define IZH_RUN_RECORD 47
int izh_run(izh_bunch *b, izh_field *f, izh_compdom *dom, double *z,
            const izh_segment *segs, int64_t nseg, double bcurr,
            double *moments);

/* ---- Lattice-scan jobs (no reference counterpart: the reference runs one
 * scan point per process, all state through files, utilities/Engscan.py:
 * 28-47) ----
 * A job steps one bunch (with its own field and domain) nsteps times
 * through a periodic lattice (period: nperiod segments, repeated) as izh_run
 * does, with space charge (bcurr >= 1e-10 A).  Its passes are enqueued on
 * the bunch's stream without the host waiting; with use_graph, blocks of
 * whole lattice periods are captured once into a HIP graph and replayed.
 * Many jobs run side by side on one GPU (each bunch has its own stream):
 * the host enqueues some steps of each in turn, polls izh_job_query and
 * starts new jobs while others are on the GPU.  The results equal izh_run's
 * on the same lattice to rounding.  Elements: drifts, quadrupoles (gradient
 * or K), constant focusing and solenoids; no RF cavities (they change the
 * reference energy, so the steps do not repeat) and no wakes. */
typedef struct izh_job izh_job;
int izh_job_create(izh_bunch *b, izh_field *f, izh_compdom *dom, double z,
                   const izh_segment *period, int64_t nperiod, int64_t nsteps, double bcurr,
                   int use_graph, izh_job **out);
/* Enqueue up to max_steps more passes (nsteps + 1 in all, the first made by
 * izh_job_create); *left (may be NULL) = passes still to enqueue. */
int izh_job_advance(izh_job *j, int64_t max_steps, int64_t *left);
/* *state: 0 while work enqueued so far is still on the GPU, 1 when it has
 * all run and passes are left to enqueue, 2 when every pass has run.  It
 * does not wait. */
int izh_job_query(izh_job *j, int *state);
/* Enqueue what is left, wait, and return z after the last step and the
 * per-step records (NULL or nsteps x IZH_RUN_RECORD, as izh_run gives them);
 * the bunch and domain then hold the state after the last step. */
int izh_job_finish(izh_job *j, double *z, double *moments);
/* info[0] passes enqueued, [1] steps per graph replay, [2] graph in use,
 * [3] graph launches. */
int izh_job_info(const izh_job *j, int64_t info[4]);
/* Free the job (an unfinished job's bunch is left mid-run: upload it anew
 * before other use). */
int izh_job_destroy(izh_job *j);

/* Tuning of izh_run on a bunch (no reference counterpart; the results do
 * not depend on it beyond the order of the moment sums).  resort_period: a
 * run re-sorts the particles into grid tiles (a counting-sort scatter) at
 * its first and last pass and every resort_period-th pass between; the
 * passes in between write each particle back to its own slot, so the
 * deposit and the gather work on the grouping of the last sort (1: sort
 * every pass; 0: the library's default). */
typedef struct {
  int resort_period;
  int reserved[7];
} izh_tuning;
int izh_bunch_set_tuning(izh_bunch *b, const izh_tuning *t);
int izh_bunch_get_tuning(const izh_bunch *b, izh_tuning *t);

/* Counters of the last izh_run on this bunch (no reference counterpart;
 * for tuning): stats[0] deposit chunks of the grouping the run ends with
 * (the one the next izh_run on this bunch starts from), stats[1] deposited
 * particles over the whole run whose CIC nodes left their chunk's LDS window
 * (they add to global memory directly), stats[2] particles alive,
 * stats[3] sort tiles. */
int izh_run_stats(const izh_bunch *b, int64_t stats[4]);

/* Bandwidth probe (no reference counterpart): one launch of a plain copy
 * kernel of nbytes between two device buffers, the calibration for
 * hardware byte counters.  Returns the kernel time in *ms (HIP events). */
int izh_probe_copy(int64_t nbytes, double *ms);

/* Per-phase kernel timers, the device counterpart of the reference's Timer
 * accumulators (src/Func/Timer.f90:10-146; t_map1, t_rhofas, t_greenf,
 * t_fft2dhpf, t_field, t_force/t_ntrslo, t_diag).  When enabled, each timed
 * kernel launch is bracketed by HIP events on the stream it runs on.
 * izh_timer_read returns, per slot, the summed milliseconds and the number
 * of timed launches (both IZH_T_COUNT long) and optionally resets them. */
#define IZH_T_PUSH 0       /* drift1 / map apply */
#define IZH_T_CONV1ST 1    /* conv1st + range reduction */
#define IZH_T_DEPOSIT 2    /* CIC deposition (charge_BeamBunch) */
#define IZH_T_GREEN 3      /* integrated Green function + its transform */
#define IZH_T_FFT 4        /* rho -> phi convolution passes */
#define IZH_T_EFIELD 5     /* E = -grad phi */
#define IZH_T_KICK 6       /* gather + kick + back transform (scatter1) */
#define IZH_T_CVBK 7       /* cvbkforth1st */
#define IZH_T_DIAG 8       /* diagnostic1 moment reductions */
#define IZH_T_PASS 9       /* izh_run: fused kick/push/moments/z->t/sort pass */
#define IZH_T_FINAL 10     /* izh_run: per-step reductions, domain, scans */
#define IZH_T_WAKE 11      /* 1D wakefield: binning and convolution */
#define IZH_T_COUNT 12
int izh_timer_enable(int on);
int izh_timer_read(double ms[IZH_T_COUNT], int64_t launches[IZH_T_COUNT],
                   int reset);

#ifdef __cplusplus
}
#endif
#endif /* IMPACTZ_HIP_H */
