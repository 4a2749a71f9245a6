/*
 * impactz_oracle.h -- CPU restatement of the IMPACT-Z space-charge step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP library is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it.  Plain C99, FP64, single-threaded, following the
 * reference routines cited per function.
 *
 * Particle arrays use the reference layout Pts1(9,N) (Fortran column-major:
 * the 9 components of a particle are contiguous, pts[9*i + c]).
 * Grid arrays use the reference layout rho(Nx,Ny,Nz), x fastest.
 */
#ifndef IMPACTZ_ORACLE_H
#define IMPACTZ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scalars the reference keeps in module globals (src/DataStruct/PhysConst.f90:25-35). */
typedef struct {
  double clight, pi, scxl, scfreq;
} or_const;

void or_physconst(double freq, or_const *c);

/* Element values: v[0..13] = the deck line's v1..v14, i.e. the element's
 * Param(2:15) (src/Contrl/AccSimulator.f90:262-536).
 *   bitype 1 quadrupole: gradient T/m (K in 1/m^2 when v[1] < -1e-6),
 *            file ID, pipe radius;
 *   bitype 2 constant focusing: kx0^2, ky0^2, kz0^2, pipe radius;
 *   bitype 3 solenoid: Bz0, file ID, pipe radius, x misalignment;
 *   bitype 101-104 DTL, CCDTL, CCL, SC: field scale, RF frequency, phase
 *            (degrees), file ID, pipe radius, (DTL) quadrupole 1 length and
 *            gradient, quadrupole 2 length and gradient;
 *   bitype 105 SolRF: as 103, then 5 misalignment values, Bz0. */
#define OR_NVAL 14

/* The Data module's on-axis field (src/DataStruct/Data.f90:41-61): the
 * table of read1_Data (zdat already shifted to start at 0) or the Fcoef
 * array of read3_Data (n values, already padded to an odd count). */
typedef struct {
  int fourier;
  int n;
  const double *zdat, *edat, *epdat;
  const double *fcoef;
} or_rfdata;

/* An element: type, RK6 steps, values, length, entrance z (its Param(1),
 * set on entry, src/Contrl/AccSimulator.f90:1025-1026), RF field. */
typedef struct {
  int bitype, mapstp;
  double v[OR_NVAL];
  double length, zedge;
  const or_rfdata *rf;
} or_element;

/* The element's 6x6 linear map for a step tau from t: RK6 over mapstp
 * steps of maplinear_Quadrupole (constant gradient), _ConstFoc, _Sol, _DTL,
 * _CCDTL, _CCL, _SC or _SolRF (src/Appl/Quadrupole.f90:120-305,
 * ConstFoc.f90:116-294, Sol.f90:120-321, DTL.f90:136-418, CCDTL.f90:
 * 122-373, CCL.f90:122-375, SC.f90:122-373, SolRF.f90:126-479).  refpt is
 * updated as the reference does.  Returns -1 for anything else. */
int or_element_map(const or_const *c, double t, double tau, const or_element *e,
                   double refpt[6], double charge, double mass, double xm[36]);

/* The quadrupole case of or_element_map with v = {grad, 0, ...}. */
void or_quad_map(const or_const *c, double tau, double grad, int mapstp,
                 double refpt[6], double charge, double mass, double xm[36]);

/* drift1_BeamBunch (src/Appl/BeamBunch.f90:250-465): bitype 0 exact drift,
 * bitype 1 quadrupole (the K-mode chromatic map transfmapK_Quadrupole,
 * src/Appl/Quadrupole.f90:509-607, when v[1] < -1e-6), bitypes 2, 3 and
 * 101-105 through their linear maps.  z is advanced by tau.  Returns -1 for
 * the branches not restated (undulator, ideal RF cavity, field files). */
int or_map1(const or_const *c, double *pts, int64_t n, double refpt[6],
            double *z, double tau, const or_element *e, double charge, double mass);

/* conv1st_BeamBunch, Flagbc=1 (src/Appl/BeamBunch.f90:1285-1352,1657-1663).
 * Compacts lost particles in place; returns the new count; rescales *current. */
int64_t or_conv1st(const or_const *c, double *pts, int64_t n,
                   const double refpt[6], double xrad, double yrad,
                   double ptrange[6], int64_t *npt_total, double *current);

/* update_CompDom, Flagbc=1, one process (src/Appl/CompDom.f90:533-557,614-622). */
void or_update_compdom(const double ptrange[6], int nx, int ny, int nz,
                       double range[6], double msize[3]);

/* deposit_BeamBunch (src/Appl/BeamBunch.f90:3678-3768). */
void or_deposit(const double *pts, int64_t n, int nx, int ny, int nz,
                const double range[6], const double msize[3], double *rho);

/* greenf1tIntnew spatial part (src/Appl/Field.f90:384-428): the integrated
 * Green function g(i,j,k), i<=nx, j<=ny, k<=nz, layout (nx+1,ny+1,nz+1). */
void or_green_igf(int nx, int ny, int nz, double hx, double hy, double hz,
                  double *g);

/* openBC3D (src/Appl/Field.f90:239-331): phi = first octant of the cyclic
 * convolution of zero-padded rho with the even-extended g on (2nx,2ny,2nz),
 * computed with a radix-2 FFT.  nx, ny, nz must be powers of two. */
int or_solve_openbc(const double *rho, int nx, int ny, int nz, double hx,
                    double hy, double hz, double *phi);

/* Same potential by direct summation, O(G^2): phi(i) = sum_j rho(j) g(|i-j|).
 * Independent of any FFT; for tiny grids. */
void or_solve_direct(const double *rho, int nx, int ny, int nz, double hx,
                     double hy, double hz, double *phi);

/* cvbkforth1st_BeamBunch (src/Appl/BeamBunch.f90:1667-1701). */
void or_cvbkforth1st(const or_const *c, double *pts, int64_t n,
                     const double refpt[6]);

/* E = -grad phi as kick1_BeamBunch forms it, one process, Flagbc=1
 * (src/Appl/BeamBunch.f90:1782-1799,1885-1899,1966-2005). */
void or_efield(const double *phi, int nx, int ny, int nz, const double msize[3],
               double *ex, double *ey, double *ez);

/* scatter1_BeamBunch (src/Appl/BeamBunch.f90:2041-2181): gather + kick +
 * 0th-order t->z back-transform. */
void or_scatter1(const or_const *c, double *pts, int64_t n, int nx, int ny,
                 int nz, const double *ex, const double *ey, const double *ez,
                 const double range[6], const double msize[3], double gam,
                 double curr, double tau, double mass, double chrg);

/* kick1_BeamBunch = or_efield + or_scatter1 (skipped below 1e-20 A, :1731). */
void or_kick1(const or_const *c, double *pts, int64_t n, int nx, int ny, int nz,
              const double *phi, const double range[6], const double msize[3],
              const double refpt[6], double curr, double tau, double mass,
              double chrg);

/* ---- the nonlinear Lorentz integrator (Flagmap=2) ---- */

/* drift2_BeamBunch (src/Appl/BeamBunch.f90:467-509): a drift of tau/2;
 * z and refpt(5) advance. */
/* Bends (bitype 4; v[0] angle, v[1] k1, v[3] half gap, v[4]/v[5] pole-face
 * angles, v[6]/v[7] pole-face curvatures, v[8] fringe integral):
 * or_dipole_face: exit = 0 the run loop's refpt[4] += L/(Scxl beta0)
 * (src/Contrl/AccSimulator.f90:1058-1091) and Fpol_Dipole
 * (src/Appl/Dipole.f90:466-526); exit = 1 Bpol_Dipole (:528-591).
 * or_sector_dipole: Sector_Dipole (:693-788) over tau, *z += tau.
 * or_isr_chicane: ISRchicane (src/Contrl/AccSimulator.f90:1542-1584) with
 * the caller's deviates rd[i] for particle i. */
void or_dipole_face(const or_const *c, double *pts, int64_t n, double refpt[6],
                    const or_element *e, int exit, double charge, double mass);
void or_sector_dipole(const or_const *c, double *pts, int64_t n, const double refpt[6],
                      double *z, double tau, const or_element *e, double charge,
                      double mass);
void or_isr_chicane(const or_const *c, double *pts, int64_t n, const double refpt[6],
                    double tau, const or_element *e, double mass, const double *rd);
void or_drift2(const or_const *c, double *pts, int64_t n, double refpt[6], double *z,
               double tau);

/* getfld_BeamLineElem (src/Appl/BeamLineElem.f90:827-870) for the elements
 * with analytic fields: drift (none), quadrupole (file ID <= 0: B = G (y, x, 0)),
 * constant focusing, solenoid (with its linear end ramps) and the RF
 * cavities 101-104 on read3_Data's Fourier coefficients (rf).  pos = (x, y
 * in m, z, t); f = (Ex, Ey, Ez, Bx, By, Bz).  zedge is the element's entrance
 * z.  Returns -1 for anything else. */
int or_elem_field(const or_const *c, int bitype, const double *v, double length,
                  double zedge, const or_rfdata *rf, const double pos[4], double f[6]);

/* kick2_BeamBunch (src/Appl/BeamBunch.f90:2187-2508): the space-charge E of
 * phi (when curr > 0, scatter2) plus the element's field, applied by the
 * implicit Lorentz push; refpt(6) follows the on-axis Ez.  Particles in the
 * 0th-order t frame (curr > 0) or the z frame (scatter20). */
int or_kick2(const or_const *c, double *pts, int64_t n, int nx, int ny, int nz,
             const double *phi, const double range[6], const double msize[3],
             double refpt[6], double curr, double tau, double mass, double chrg,
             int bitype, const double *v, double length, double zedge,
             const or_rfdata *rf, double z, const double *wake);

/* diagnostic1_Output sums (src/Contrl/Output.f90:119-245): 37 sums in the
 * reference's tmplc order, 6 abs-maxima, max of x^2+y^2. */
void or_diag_sums(const double *pts, int64_t n, double sums[37], double amax[6],
                  double *r2max);

/* The derived quantities of one diagnostic row (src/Contrl/Output.f90:247-339),
 * each column scaled as the reference writes it:
 * out[0..4]  : fort.18 columns after z (refpt5, gam, energy, bet, rmax)
 * out[5..10] : fort.24 x0, xrms, px0, pxrms, -xpx/epx, epx
 * out[11..16]: fort.25 (y), out[17..22]: fort.26 (z)
 * out[23..28]: fort.27 maxima, out[29..34]: fort.29, out[35..40]: fort.30 */
void or_diag_row(const or_const *c, const double sums[37], const double amax[6],
                 double r2max, int64_t npt, const double refpt[6], double mass,
                 double out[41]);
/* chgupdate_BeamBunch: particles of each charge state (q/mc^2 within 1e-4). */
void or_chgupdate(const double *pts, int64_t n, int nchge, const double *qmcc,
                  int64_t *counts);

/* A wakefield (src/Contrl/AccSimulator.f90:859-875,960-980): kind 1 the
 * analytic short-range wake of a SolRF/TWS element with iris radius aa, gap
 * gg and period leng (its Param(13:15)); kind 2 the wake functions read by
 * read1wk_Data (n rows: zdat from 0, Wz, Wx, Wy) with flagbtw (2, 3: no
 * transverse wake; 4: no longitudinal); scale = scwk.  Kind 3 the CSR wake
 * of the run loop (src/Contrl/AccSimulator.f90:1344-1396): bend radius r0,
 * bend length blength, and zoff, the distance from the bend entrance the
 * loop adds to range(5)/gamma (z - zbleng inside the bend, z - zbleng +
 * sentr after it); its transverse wakes are zero and scale is not used. */
typedef struct {
  int kind;
  double aa, gg, leng;
  int n;
  const double *zdat, *edat, *epdat, *eppdat;
  int flagbtw;
  double scale;
  double r0, zoff, blength;
} or_wake;

/* The wake fields ex, ey, ez (nz each) of the particles after conv1st. */
int or_wakefield(const or_const *c, const double *pts, int64_t n, const double range[6],
                 int nz, double current, int64_t np, double gam, double charge,
                 const or_wake *w, double *ex, double *ey, double *ez);

/* csrwakeTrIGF_FieldQuant (src/Appl/Field.f90:1007-1251): the 1D CSR wake
 * ezwake (V/m) on nx mesh lines from the line density rhonew, for a bend of
 * radius r0 and length blength (m); ptmin is the first line's distance from
 * the bend entrance and hx the line spacing, both in the lab frame (m). */
void or_csr_wake(int nx, double r0, double ptmin, double hx, double blength,
                 const double *rhonew, double gam, double *ezwake);

/* kick1wake_BeamBunch (src/Appl/BeamBunch.f90:4694-5046): kick1 with the
 * wake fields added to the grid field plane by plane. */
void or_kick1wake(const or_const *c, double *pts, int64_t n, int nx, int ny, int nz,
                  const double *phi, const double range[6], const double msize[3],
                  const double refpt[6], double curr, double tau, double mass,
                  double chrg, const double *ex, const double *ey, const double *ez);

/* One segment of the reference loop (src/Contrl/AccSimulator.f90:1099-1488)
 * for an element or_map1 supports:
 * map1(tau1) -> conv1st -> update_CompDom -> deposit -> solve -> cvbkforth1st
 * -> kick1(2 tau1) -> map1(tau1).  Workspace is allocated internally.
 * Returns the particle count after the step (<0 on error). */
typedef struct {
  double bcurr;                   /* deck current; below 1e-10 A no SC block */
  double current, mass, charge;   /* BeamBunch%Current, %Mass, %Charge */
  int64_t npt;                    /* BeamBunch%Npt */
  double refpt[6];                /* BeamBunch%refptcl */
  double z;
} or_bunch_state;

int64_t or_step(const or_const *c, double *pts, int64_t n, or_bunch_state *st,
                int nx, int ny, int nz, double tau1, const or_element *e,
                double pipe_radius, const or_wake *wake, double *rho_out,
                double *phi_out);

/* One Flagmap=2 segment: drift2(2 tau1) -> conv1st -> update_CompDom ->
 * deposit -> solve -> cvbkforth1st -> kick2(2 tau1) -> drift2(2 tau1). */
int64_t or_step2(const or_const *c, double *pts, int64_t n, or_bunch_state *st,
                 int nx, int ny, int nz, double tau1, const or_element *e,
                 double pipe_radius, const or_wake *wake, double *rho_out,
                 double *phi_out);

#ifdef __cplusplus
}
#endif
#endif
