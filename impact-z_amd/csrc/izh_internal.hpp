// izh_internal.hpp -- handles, error plumbing and launch helpers shared by the
// HIP translation units of libimpactz_hip.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "impactz_hip.h"

namespace izh {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const char *fmt, ...);

#define IZH_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return ::izh::fail(IZH_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__,    \
                         #call, hipGetErrorString(e_));                      \
  } while (0)

#define IZH_CHECK(expr)          \
  do {                               \
    int rc_ = (expr);                \
    if (rc_ != IZH_OK) return rc_;   \
  } while (0)

// ---------------------------------------------------------- constants
constexpr int kBlock = 256;          // threads per block for particle passes
constexpr int kMaxRedBlocks = 2048;  // fixed partial count => deterministic sums
constexpr int kMoments = 44;         // 37 sums + 6 abs-max + r^2 max
constexpr int kRunRecord = kMoments + 1;  // + Npt, per step of izh_run (device)
static_assert(IZH_RUN_RECORD == kRunRecord + 2, "izh_run record: + refpt(5:6)");
constexpr int kConvRed = 8;          // 3 min, 3 max, lost, alive
constexpr int kPartStride = kMoments + kConvRed;
constexpr int kChunk = 16384;        // particles per deposit chunk
constexpr int kWindowCells = 4096;   // LDS deposit window (int64 cells)
constexpr int kTileShift = 3;        // sort tiles of 8x8x8 cells
constexpr uint32_t kDeadKey = 0xffffffffu;
constexpr int kSinkBytes = 64 * 8 * 16;  // one 8-byte word per lane, 16 rows
// Passes of a fused run between two counting-sort scatters when the bunch's
// tuning leaves it open (izh_tuning.resort_period = 0).
constexpr int kDefaultResortPeriod = 1;

// Frame of the particle coordinates held by a bunch; the reference's call
// order moves the bunch through these (src/Contrl/AccSimulator.f90:1110-1429).
enum Frame : int {
  kFrameZ = 0,     // z as independent variable: x/Scxl, px, y/Scxl, py, phi, pt
  kFrameT1 = 1,    // after conv1st: 1st-order t frame, metres
  kFrameT0 = 2,    // after cvbkforth1st: 0th-order t frame, metres
};

// Computational domain as the kernels read it (one per field, in device
// memory, so a run of steps can update it without the host).
// The grid a grouping of sort keys was computed in: lower corner and cell
// size per axis.  Key tile t of an axis starts at node t << shift of it.
struct KeyGrid {
  double lo[3], h[3];
};

struct DevDom {
  double range[6];    // xmin, xmax, ymin, ymax, zmin, zmax (SpatRange)
  double msize[3];    // hx, hy, hz (Meshsize)
  double hinv[3];     // 1/h
  double ptrange[6];  // the conv1st range it was built from
  KeyGrid kgrid[2];   // grid of the keys of each of the two groupings in flight
};

// First node, in the grid with lower corner lo and 1/h hinv, of the key tile
// that starts at node t0 of key grid (klo, kh).  The domain moves and
// rescales between steps, so a tile's nodes shift by a few cells.
__device__ inline int tile_first_node(int t0, double klo, double kh, double lo,
                                      double hinv) {
  const double u = floor((klo + (double)t0 * kh - lo) * hinv);
  return (int)fmin(fmax(u, -1.0e6), 1.0e6);  // callers clamp to the grid
}

// Bunch scalars a run of steps keeps on the device.
struct DevScal {
  double current;     // BeamBunch%Current
  double npt;         // BeamBunch%Npt (as a double for the rescale)
  double xycon, tcon; // scatter1 constants of the current step
  int64_t nalive;     // particles alive after the last conv1st
  int64_t nslots;     // slots written by the last pass (alive + just lost)
  int64_t nchunks[2]; // deposit chunks of the two groupings in flight
  int64_t overflow;   // deposited particles whose nodes left their LDS window
  int64_t err;        // scatter slots out of range (must stay 0)
  int64_t rec;        // next record of the run's per-step moments (the finalize
                      // kernels count it, so a replayed step graph writes in turn)
};

// A run of at most kChunk consecutive slots of one sort tile.
struct Chunk {
  uint32_t start, len, tile;
};

// Sort tiles are 2^s nodes per axis (s <= kTileShift, 2^s <= n).
inline int tile_shift(int n) {
  int s = 0;
  while ((1 << (s + 1)) <= n && s < kTileShift) s++;
  return s;
}

}  // namespace izh

// Device-resident bunch in structure-of-arrays form.  The six phase-space
// rows and the particle's original index move with the particle (sorting,
// compaction); q/m, charge weight and id are immutable per particle and are
// read through the original index (or held as one value when uniform).
struct izh_bunch {
  int device = 0;
  int precision = IZH_FP64;
  hipStream_t stream = nullptr;
  int64_t cap = 0;       // allocated rows
  int64_t n0 = 0;        // particles at upload (size of the immutable rows)
  int64_t n = 0;         // Nptlocal: live rows in r[]
  int64_t npt = 0;       // Npt (one process: == n after conv1st)
  double current = 0, mass = 0, charge = 0, freq = 0;
  double refpt[6] = {0, 0, 0, 0, 0, 0};
  izh_const c{};
  int frame = izh::kFrameZ;
  double wabs_sum = 0;   // sum |Pts1(8,:)|, bounds every cell of the grid charge
  double wabs_max = 0;   // max |Pts1(8,:)|, bounds every particle's contribution
  // moving rows: x, px, y, py, phi, pt and the original index
  double *r[6] = {nullptr};
  double *r_alt[6] = {nullptr};
  uint32_t *orig = nullptr, *orig_alt = nullptr;
  uint32_t *key = nullptr, *key_alt = nullptr;   // sort keys of the fused run
  void *tpos = nullptr;  // t-frame positions for the deposit: 3*cap of T
  void *sink = nullptr;  // kSinkBytes the fused pass's empty lanes store into
  // immutable rows, indexed by the original index
  double *qm0 = nullptr, *w0 = nullptr, *id0 = nullptr;
  int w_uniform = 0;
  double w_value = 0;
  int qm_uniform = 0;    // every particle has q/mc^2 == qm_value
  double qm_value = 0;
  unsigned long long *d_chg = nullptr, *h_chg = nullptr;  // charge-state counts
  // reduction scratch
  double *d_part = nullptr;     // kMaxRedBlocks * kPartStride
  double *d_out = nullptr;      // final reductions (kPartStride + 8)
  double *h_out = nullptr;      // pinned mirror of d_out
  // fused-run scratch
  izh::DevScal *d_scal = nullptr;
  izh::DevScal *h_scal = nullptr;          // pinned mirror
  uint32_t *hist[2] = {nullptr, nullptr};   // tile histograms (ping-pong)
  uint32_t *cursor = nullptr;               // scatter cursors per tile
  izh::Chunk *chunks = nullptr;             // deposit work list
  int64_t ntiles_cap = 0, nchunks_cap = 0;
  double *d_moments = nullptr;              // per-step moments of a run
  int64_t moments_cap = 0;
  int64_t run_stats[4] = {0, 0, 0, 0};      // izh_run_stats
  int resort_period = 0;                    // izh_tuning (0: the default)
  // Sort keys, their scan (cursor) and first deposit chunks left by the last
  // izh_run for the next one; any other change to the particles clears this.
  int keys_valid = 0;
  const izh_field *keys_field = nullptr;
  int keys_grid[3] = {0, 0, 0};
  int keys_po = 0;  // chunk slot holding the next run's first deposit grouping
};

// Device-resident FieldQuant + solver workspace for one mesh.
struct izh_field {
  int device = 0;
  int precision = IZH_FP64;
  hipStream_t stream = nullptr;
  int nx = 0, ny = 0, nz = 0;
  int64_t g = 0;                // nx*ny*nz
  unsigned long long *rho_fx = nullptr;  // fixed-point deposit accumulator
  void *rho = nullptr;          // T[g]   charge density (T = double or float)
  void *phi = nullptr;          // T[g]   potential
  void *e = nullptr;            // T4[g]  (Ex, Ey, Ez, 0) interleaved
  int64_t ldx = 0;              // x pitch of the half spectra: nx+1 rounded up
                                // to whole 128-byte rows
  void *spec_a = nullptr;       // complex T[ldx*ny*nz]
  void *spec_b = nullptr;       // complex T[ldx*2ny*nz]
  void *green = nullptr;        // T[(ny+1)][(nz+1)][ldx] transformed Green fn
  void *green_t = nullptr;      // T[(nx+1)*(ny+1)*(nz+1)] work
  double *corner = nullptr;     // double[(nx+2)*(ny+2)*(nz+2)] IGF antiderivative
  void *tw_x = nullptr, *tw_y = nullptr, *tw_z = nullptr;  // twiddles per length
  izh::DevDom *d_dom = nullptr; // domain the kernels read
  izh::DevDom *h_dom = nullptr; // pinned staging for host-set domains
  izh::Chunk *chunks = nullptr; // uniform chunks for unsorted deposits
  int64_t nchunks_cap = 0;
  double rho_scale = 1.0;       // grid value = physical charge * rho_scale
  int have_rho = 0, have_phi = 0;
  double fixed_scale = 0;       // 2^e used by the fixed-point deposit
  double *d_wake = nullptr;     // exwake, eywake, ezwake, dEx, dEy, dEz (nz each), bins
  int have_wake = 0;
};

namespace izh {

// Element push of a half step, applied per particle inside the kernels.
struct Push {
  int kind;           // 0 none, 1 exact drift, 2 linear 6x6 map
  double tau, gam0, ibeta0, scxl;
  double xm[36];
};

// A bend map applied in Transport coordinates (x, x', y, y', z, delta) by
// k_dipole: mode 0 the second-order sector map Sector_Dipole
// (src/Appl/Dipole.f90:693-788), 1 the entrance pole face Fpol_Dipole
// (:466-526), 2 the exit pole face Bpol_Dipole (:528-591).  gam0, beta0 and
// gambet are the reference energy as each routine forms it; qm0 is the beam's
// Charge/Mass.
enum { kDipoleSector = 0, kDipoleEntrance = 1, kDipoleExit = 2 };
struct DipoleMap {
  int mode;
  double scxl, gam0, beta0, gambet, qm0;
  // pole faces
  double h0, h1, k1, psi, tanphi, tanphi2, secphi2, secphi3, tanphib, secphib2;
  // sector: first- and second-order transport coefficients
  double len, R11, R12, R16, R21, R22, R26, R51, R52, R56;
  double T111, T112, T116, T122, T126, T166, T144, T216, T222, T266, T244, T314,
      T324, T346, T522, T544, T526;
};

// The inputs of the wake kernels (izh_wake.hip): the wake (kind 1 analytic
// aa, gg, leng; kind 2 a device table of ntab rows: z, Wz, Wx, Wy), the
// scale scwk, the reference energy of the step and the bunch constants;
// current, npt, xycon and tcon come from a DevScal when the launch gives one.
struct WakeParams {
  int kind, nz, flagbtw, ntab;
  double aa, gg, leng, scale;
  double r0, zoff, blength;   // kind 3: the CSR wake of a bend
  const double *tab;
  double gam, mass, charge, scfreq, clight;
  double current, npt, xycon, tcon;
};
constexpr int kWakeMaxNz = 1024;   // LDS of the one-workgroup convolution

int launch_wake(izh_bunch *b, izh_field *f, const WakeParams &p, const DevScal *S,
                const Chunk *chunks, const int64_t *d_nchunks, int64_t max_chunks);
int launch_wake_add(izh_field *f);
// WakeParams of a wake for the bunch's present reference energy (izh_api.hip).
int wake_params_of(const izh_bunch *b, const izh_field *f, const izh_wake *w,
                   WakeParams *p);
int wake_table_device(const izh_rfdata *d, const double **out);

// Called by every operation that changes the particles or their order: the
// sort state a run of steps leaves for the next one no longer applies.
inline void touch_particles(izh_bunch *b) { b->keys_valid = 0; }

// Launchers implemented in the kernel translation units.
int launch_drift(izh_bunch *b, double tau, double gam0, double beta0);
int launch_quad_kmode(izh_bunch *b, double tau, double kq, double qmass, double gambet0,
                      double beta0);
int launch_dipole(izh_bunch *b, const DipoleMap &m);
int launch_isr(izh_bunch *b, double gam0, double p1, double p2, const double *rd);
// The model drift1_BeamBunch applies to an element (src/Appl/BeamBunch.f90:
// 276-449): kMap1Drift, kMap1Linear (a host 6x6 map) or kMap1QuadK; -1 with
// *why set for the branches not on this path.
enum { kMap1Drift = 0, kMap1Linear = 1, kMap1QuadK = 2 };
inline bool is_rf_cavity(int bitype) { return bitype >= 101 && bitype <= 105; }
inline int map1_kind(int bitype, const double v[IZH_NVAL], const char **why) {
  *why = nullptr;
  if (bitype == 0) return kMap1Drift;
  if (bitype == 1) {
    // src/Appl/BeamBunch.f90:299-307: the file-ID slot selects the model
    if (v[1] < -10.0) { *why = "undulator transfer map"; return -1; }
    if (v[1] < -1.0e-6) return kMap1QuadK;
    if (v[1] > 1.0e-5) { *why = "quadrupole fringe field from a file"; return -1; }
    return kMap1Linear;
  }
  if (bitype == 2 || bitype == 3 || is_rf_cavity(bitype)) {
    // Param(5) < 0 selects the ideal RF cavity model (:324-330)
    if (v[3] < 0.0) { *why = "Param(5) < 0 (ideal RF cavity branch)"; return -1; }
    return kMap1Linear;
  }
  *why = "element type not on this path (0 drift, 1 quadrupole, 2 constant "
         "focusing, 3 solenoid, 101-105 RF cavities)";
  return -1;
}

}  // namespace izh
// The on-axis field of the Data module (src/DataStruct/Data.f90:41-61):
// read1_Data's table (z from 0, Ez, Ez', Ez'') or read3_Data's coefficients.
struct izh_rfdata {
  int fourier;                 // 0: table (read1_Data), 1: coefficients (read3_Data)
  std::vector<double> z, ez, ezp, ezpp;
  std::vector<double> coef;    // Fcoef(1..Ndata)
  mutable double *d_table = nullptr;  // device copy of the four columns (wake tables)
};
namespace izh {

// ---------------------------------------- nonlinear integrator (Flagmap=2)

// An element's analytic field for the Lorentz integrator: getfld_DriftTube,
// getfld_Quadrupole (file ID <= 0), getfld_ConstFoc and getfld_Sol
// (src/Appl/DriftTube.f90:290-305, Quadrupole.f90:369-402, ConstFoc.f90:
// 297-322, Sol.f90:388-435).  pos = (x, y in m, z, t); f = E (V/m), B (T).
struct ElemField {
  int type;
  double v[9];          // Param(2:10)
  double length, zedge;
  // RF cavities 101-104: the on-axis field and its first three derivatives
  // at the kick's z (the same for every particle: scatter2 evaluates the
  // field at pos(3) = z), the angular frequency in units of the reference
  // one, Scxl/ww and the phase theta0 (src/Appl/DTL.f90:780-838)
  double ax[4];
  double ww, xl, theta0;
  // SolRF (src/Appl/SolRF.f90:482-620): which sections cover the kick's z
  // (bit 0 the RF one, bit 1 the solenoid one; 0 outside the element), the
  // solenoid section's Bz and derivatives (unscaled), its scale Param(12)
  // and 2*pi*Scfreq of the time conversion; xl holds ww/clite there
  int win;
  double bz[4];
  double bscale, tdiv;
};

__host__ __device__ inline void elem_field(const ElemField &e, double scxl, double clight,
                                           const double pos[4], double f[6]) {
  for (int k = 0; k < 6; k++) f[k] = 0.0;
  if (e.type == 1) {
    f[3] = e.v[0] * pos[1];
    f[4] = e.v[0] * pos[0];
  } else if (e.type == 2) {
    // the reference's placeholders: qmcc = gambetz = 1
    const double qmcc = 1.0, pz = 1.0;
    f[2] = e.v[2] * pos[3] * pz * pz * pz * scxl / qmcc;
    f[3] = -e.v[0] * pos[1] * pz / qmcc / clight;
    f[4] = e.v[1] * pos[0] * pz / qmcc / clight;
  } else if (e.type == 3) {
    // uniform Bz with linear ramps over 2*Param(4) at both ends
    const double x02 = 2 * e.v[2], zz = pos[2] - e.zedge, b0 = e.v[0];
    if (zz >= 0.0 && zz < x02) {
      const double bb = b0 / x02;
      f[3] = -bb / 2 * pos[0];
      f[4] = -bb / 2 * pos[1];
      f[5] = bb * zz;
    } else if (zz >= x02 && zz <= e.length - x02) {
      f[5] = b0;
    } else if (zz > e.length - x02 && zz <= e.length) {
      const double bb = -b0 / x02, zshift = e.length - x02;
      f[3] = -bb / 2 * pos[0];
      f[4] = -bb / 2 * pos[1];
      f[5] = b0 + bb * (zz - zshift);
    }
  } else if (e.type == 105) {
    const double clite = (double)299792458.e0f;
    const double r2 = pos[0] * pos[0] + pos[1] * pos[1];
    if (e.win & 1) {
      const double ez1 = e.ax[0], ezp1 = e.ax[1], ezpp1 = e.ax[2], ezppp = e.ax[3];
      const double xlrep = e.xl;
      const double tt = pos[3] / e.tdiv;
      const double tmpcos = cos(e.ww * tt + e.theta0), tmpsin = sin(e.ww * tt + e.theta0);
      const double f1 = -(ezpp1 + ez1 * xlrep * xlrep) / 4;
      const double f1p = -(ezppp + ezp1 * xlrep * xlrep) / 4;
      f[0] += -pos[0] * (ezp1 / 2 + f1p * r2 / 4) * tmpcos;
      f[1] += -pos[1] * (ezp1 / 2 + f1p * r2 / 4) * tmpcos;
      f[2] += (ez1 + f1 * r2) * tmpcos;
      f[3] += pos[1] * xlrep / clite * (ez1 / 2 + f1 * r2 / 4) * tmpsin;
      f[4] += -pos[0] * xlrep / clite * (ez1 / 2 + f1 * r2 / 4) * tmpsin;
    }
    if (e.win & 2) {
      const double bs = e.bscale;
      f[3] = f[3] - bs * e.bz[1] / 2 * pos[0] + bs * e.bz[3] * pos[0] * r2 / 16;
      f[4] = f[4] - bs * e.bz[1] / 2 * pos[1] + bs * e.bz[3] * pos[1] * r2 / 16;
      f[5] = f[5] + bs * e.bz[0] - bs * e.bz[2] * r2 / 4;
    }
  } else if (e.type >= 101 && e.type <= 104) {
    // getfld_DTL / _CCDTL / _CCL / _SC (src/Appl/DTL.f90:769-855): the DTL's
    // drift-tube quadrupoles, then the paraxial expansion of the on-axis
    // field with the REAL(4) speed of light of the reference
    if (e.type == 101) {
      const double zz = pos[2] - e.zedge;
      const double bgrad = zz < e.v[5] ? e.v[6] : (zz < e.length - e.v[7] ? 0.0 : e.v[8]);
      f[3] = bgrad * pos[1];
      f[4] = bgrad * pos[0];
    }
    const double clite = (double)299792458.e0f;
    const double ez1 = e.ax[0], ezp1 = e.ax[1], ezpp1 = e.ax[2], ezppp = e.ax[3];
    const double tmpcos = cos(e.ww * pos[3] + e.theta0);
    const double tmpsin = sin(e.ww * pos[3] + e.theta0);
    const double f1 = -(ezpp1 + ez1 / e.xl / e.xl) / 4;
    const double f1p = -(ezppp + ezp1 / e.xl / e.xl) / 4;
    const double r2 = pos[0] * pos[0] + pos[1] * pos[1];
    f[0] += -pos[0] * (ezp1 / 2 + f1p * r2 / 4) * tmpcos;
    f[1] += -pos[1] * (ezp1 / 2 + f1p * r2 / 4) * tmpcos;
    f[2] += (ez1 + f1 * r2) * tmpcos;
    f[3] += pos[1] / (e.xl * clite) * (ez1 / 2 + f1 * r2 / 4) * tmpsin;
    f[4] += -pos[0] / (e.xl * clite) * (ez1 / 2 + f1 * r2 / 4) * tmpsin;
  }
}

// The element types with an analytic field on this path, or a message.
inline const char *elem_field_unsupported(int bitype, const double v[4]) {
  if (bitype == 0 || bitype == 2 || bitype == 3) return nullptr;
  if (bitype >= 101 && bitype <= 105) return nullptr;   // with Fourier coefficients
  if (bitype == 1)
    return v[1] > 0.0 ? "quadrupole field profile from a file (file ID > 0)" : nullptr;
  return "element type not on this path (0 drift, 1 quadrupole, 2 constant focusing, "
         "3 solenoid, 101-105 RF cavities)";
}

// Implicit Lorentz push of scatter2/scatter20 (src/Appl/BeamBunch.f90:
// 3026-3066): the momenta solve a 3x3 system with the longitudinal momentum
// averaged over the step, in two iterations.
__device__ inline void lorentz_push(double &px, double &py, double &pt, double q, double ex,
                                    double ey, double ez, double bx, double by, double bz,
                                    double tau, double clight, double tmpgamma0,
                                    double gam, double qmcc, double ez0) {
  const double tmp12 = -tau * clight * bz * q / 2;
  const double tmp13 = tau * ex * q / 2;
  const double tmp23 = tau * ey * q / 2;
  const double tmp1 = tmpgamma0 * ex * tau * q;
  const double tmp2 = clight * by * tau * q;
  const double tmp3 = tmpgamma0 * ey * tau * q;
  const double tmp4 = clight * bx * tau * q;
  const double tmp5 = (ez0 * qmcc - ez * q) * tau;
  double pz = sqrt((gam - pt) * (gam - pt) - 1.0 - px * px - py * py);
  const double p1 = px, p2 = py, p3 = pt;
  const double tmpgamma00 = gam + tau * qmcc * ez0;
  for (int ii = 0; ii < 2; ii++) {
    const double a12 = tmp12 / pz, a13 = tmp13 / pz, a23 = tmp23 / pz;
    const double b1 = p1 - a12 * p2 - a13 * p3 + tmp1 / pz - tmp2;
    const double b2 = a12 * p1 + p2 - a23 * p3 + tmp3 / pz + tmp4;
    const double b3 = -a13 * p1 - a23 * p2 + p3 + tmp5;
    const double det = 1 + a12 * a12 - a13 * a13 - a23 * a23;
    const double s11 = 1 - a23 * a23, s12 = -a12 + a13 * a23, s13 = a12 * a23 - a13;
    const double s21 = a12 + a13 * a23, s22 = 1 - a13 * a13, s23 = -a23 - a13 * a12;
    const double s31 = -a12 * a23 - a13, s32 = -a23 + a12 * a13, s33 = 1 + a12 * a12;
    px = (s11 * b1 + s12 * b2 + s13 * b3) / det;
    py = (s21 * b1 + s22 * b2 + s23 * b3) / det;
    pt = (s31 * b1 + s32 * b2 + s33 * b3) / det;
    const double tmpsq2 = (tmpgamma00 - pt) * (tmpgamma00 - pt) - 1.0 - px * px - py * py;
    if (tmpsq2 > 0.0) pz = 0.5 * pz + 0.5 * sqrt(tmpsq2);
  }
}

// Constants of one kick2 call, set on the host.
struct Kick2 {
  ElemField e;
  double gam, beta0_c, xk, xl, clight, tg, z, tau, tmpscale, tmpgamma0, qmcc, ez0;
  // kick2wake (src/Appl/BeamBunch.f90:5048-5410): exwake, eywake, ezwake of
  // the field (nz each, V/m) interpolated in z and added to E, or null
  const double *wake;
  int nzw;
};

int launch_drift2(izh_bunch *b, double tau, double gamma0, double beta0);
int launch_kick2(izh_bunch *b, izh_field *f, const Kick2 &k, bool space_charge);

int launch_map_linear(izh_bunch *b, const double xm[36]);
// Charge states of chgupdate_BeamBunch: their q/mc^2 and the relative
// tolerance a particle's value must be within.
constexpr int kMaxChargeStates = 100;   // the reference's list length
struct ChgList {
  int n;
  double eps;
  double qmcc[kMaxChargeStates];
};
int launch_chg_count(izh_bunch *b, const ChgList &cl, int64_t *out);
int launch_conv1st(izh_bunch *b, double xrad, double yrad, double ptrange[6],
                   int64_t *nlost);
int launch_compact(izh_bunch *b, double xrad, double yrad, int64_t nkeep);
int launch_cvbk(izh_bunch *b);
int launch_moments(izh_bunch *b, double out[kMoments]);
// izh_diagnostic1_rows: the moments of caller-held rows into out (device),
// partials in the library's static per-device scratch; allocates nothing.
int launch_moments_rows(const double *rows, int64_t ld, int64_t n, double *out,
                        hipStream_t st);
// Looks up the scratch array's address on `device` (current) once.
int scratch_setup(int device);
int launch_clone(const izh_bunch *src, izh_bunch *dst, int64_t offset, int64_t n, double w);
int field_set_domain(izh_field *f, const izh_compdom *dom);
int launch_deposit(izh_bunch *b, izh_field *f, const Chunk *chunks, const KeyGrid *kgrid,
                   const int64_t *d_nchunks, int64_t max_chunks);
int launch_deposit_granular(izh_bunch *b, izh_field *f);
int launch_solve(izh_field *f);
// The two halves of launch_solve: G^ of the current domain on stream gs,
// then the convolution on the field's stream (after G^ is complete).
int launch_green(izh_field *f, hipStream_t gs);
int launch_convolve(izh_field *f);
int launch_efield(izh_field *f);
int launch_gather_kick(izh_bunch *b, izh_field *f, double tau, double xycon,
                       double tcon);
int launch_probe_copy(int64_t nbytes, double *ms);
int run_steps(izh_bunch *b, izh_field *f, izh_compdom *dom, double *z,
              const izh_segment *segs, int64_t nseg, double *moments);
double fixed_point_scale(double wabs_sum);
int bunch_alloc_alt(izh_bunch *b);

// Kernel timers (izh_timer_*): begin/end record events on the launch stream
// when timing is enabled; both are no-ops otherwise.
bool timing_on();
void timer_begin(int slot, hipStream_t s);
void timer_end(int slot, hipStream_t s);

struct ScopedTimer {
  int slot;
  hipStream_t s;
  ScopedTimer(int slot_, hipStream_t s_) : slot(slot_), s(s_) { timer_begin(slot, s); }
  ~ScopedTimer() { timer_end(slot, s); }
};

inline int grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > kMaxRedBlocks) g = kMaxRedBlocks;
  return (int)g;
}

// ------------------------------------------------------------ device math
// REAL(4) pi of conv1st: pilc = 2.0*asin(1.0) (src/Appl/BeamBunch.f90:1268).
__host__ __device__ inline double pilc_f32() { return (double)3.14159274101257324f; }

// Phase wrap of conv1st (src/Appl/BeamBunch.f90:1296-1298).
__device__ inline double wrap_phase(double p5) {
  const double pilc = pilc_f32();
  int ntmp5 = (int)(p5 / pilc);
  double tmp5 = p5 - ntmp5 * pilc;
  return tmp5 - (ntmp5 % 2) * pilc;
}

// Loss predicate of conv1st (src/Appl/BeamBunch.f90:1340-1343) on the
// transformed coordinates.
__device__ inline bool lost_t1(double xt, double yt, double px, double py,
                               double pt, double gam, double rad) {
  double rcp = 1.0 / (-pt + gam);
  double tmpbet = 1.0 - rcp * rcp * (1.0 + px * px + py * py);
  double radtest = sqrt(xt * xt + yt * yt);
  return radtest >= rad || tmpbet <= 0.0;
}

// Half-step push of one particle (drift1_BeamBunch, src/Appl/BeamBunch.f90:
// 276-320): exact drift or the 6x6 map product.
__device__ inline void apply_push(const Push &p, double &x, double &px, double &y,
                                  double &py, double &ph, double &pt) {
  if (p.kind == 1) {
    double d = pt - p.gam0;
    double tmph = sqrt(d * d - 1.0 - px * px - py * py);
    x = x + px * p.tau / tmph / p.scxl;
    y = y + py * p.tau / tmph / p.scxl;
    ph = ph - (p.ibeta0 + d / tmph) * p.tau / p.scxl;
  } else if (p.kind == 2) {
    const double v[6] = {x, px, y, py, ph, pt};
    double t[6];
#pragma unroll
    for (int j = 0; j < 6; j++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 6; k++) s += v[k] * p.xm[6 * j + k];
      t[j] = s;
    }
    x = t[0]; px = t[1]; y = t[2]; py = t[3]; ph = t[4]; pt = t[5];
  }
}

template <typename T>
struct alignas(sizeof(T) * 4) E4 {
  T x, y, z, w;
};

}  // namespace izh
