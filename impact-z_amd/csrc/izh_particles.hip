// izh_particles.hip -- per-particle passes of the procedure-by-procedure API.
//
// Particles live in structure-of-arrays form (one FP64 array per phase-space
// row), so every pass streams only the rows it touches with coalesced loads.
// Reductions (conv1st range, diagnostic moments) go through a fixed number of
// per-block partials and a single-block tree in LDS: the summation order
// depends only on N, so results are bitwise reproducible run to run.
#include <cmath>
#include <cstring>
#include <mutex>

#include "izh_device.hpp"

namespace izh {

// Single-block partial reduction (declared in izh_device.hpp).
__global__ void __launch_bounds__(1024)
k_reduce_partials(const double *__restrict__ part, int nparts, int nq,
                  unsigned long long min_mask, unsigned long long max_mask,
                  double *__restrict__ out) {
  __shared__ double sh[1024];
  for (int q = 0; q < nq; q++) {
    const int op = ((min_mask >> q) & 1ull) ? 1 : (((max_mask >> q) & 1ull) ? 2 : 0);
    double acc = op == 1 ? 1e300 : (op == 2 ? -1e300 : 0.0);
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
      double v = part[(int64_t)b * nq + q];
      acc = op == 1 ? fmin(acc, v) : (op == 2 ? fmax(acc, v) : acc + v);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) {
        double a = sh[threadIdx.x], c = sh[threadIdx.x + s];
        sh[threadIdx.x] = op == 1 ? fmin(a, c) : (op == 2 ? fmax(a, c) : a + c);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh[0];
    __syncthreads();
  }
}

int launch_reduce(hipStream_t s, const double *part, int nparts, int nq,
                  unsigned long long min_mask, unsigned long long max_mask,
                  double *out) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, part, nparts,
                     nq, min_mask, max_mask, out);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

namespace {

__global__ void k_zero_u64_small(unsigned long long *__restrict__ p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// ------------------------------------------------------------------- push
__global__ void __launch_bounds__(kBlock)
k_push(double *__restrict__ x, double *__restrict__ px, double *__restrict__ y,
       double *__restrict__ py, double *__restrict__ ph, double *__restrict__ pt,
       int64_t n, Push p) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double a = x[i], b = px[i], c = y[i], d = py[i], e = ph[i], f = pt[i];
    apply_push(p, a, b, c, d, e, f);
    x[i] = a; px[i] = b; y[i] = c; py[i] = d; ph[i] = e; pt[i] = f;
  }
}

// ------------------------------------------------------- K-mode quadrupole
// transfmapK_Quadrupole (src/Appl/Quadrupole.f90:509-607): each particle's
// own strength kstr = (q/m)/(q/m)_0 * gambet0/|gamma beta| * K gives a
// cos/sin (focusing) and cosh/sinh (defocusing) pair of 2x2 maps; the phase
// slips by (1/beta_z - 1/beta0) tau/Scxl.  In place, one particle per lane.
__global__ void __launch_bounds__(kBlock)
k_quad_kmode(double *__restrict__ X, double *__restrict__ PX, double *__restrict__ Y,
             double *__restrict__ PY, double *__restrict__ PH, const double *__restrict__ PT,
             const uint32_t *__restrict__ orig, const double *__restrict__ qm0, int64_t n,
             double tau, double kq, double qmass, double gam0, double gambet0, double beta0,
             double scxl) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double x = X[i], px = PX[i], y = Y[i], py = PY[i];
    const double gam = gam0 - PT[i];
    double gambetz = sqrt(gam * gam - 1.0 - px * px - py * py);
    const double betaz = gambetz / gam;
    gambetz = sqrt(gam * gam - 1.0);
    const double kstr = qm0[orig[i]] / qmass * gambet0 / gambetz * kq;
    const double rtkstr = sqrt(fabs(kstr));
    const double rtkstrzz = rtkstr * tau;
    double xm11, xm12, xm21, xm22, xm33, xm34, xm43, xm44;
    if (kstr != 0.0) {
      const double co = cos(rtkstrzz), si = sin(rtkstrzz);
      const double t = exp(rtkstrzz);
      const double ch = (t * t + 1.0) / (t + t), sh = (t * t - 1.0) / (t + t);
      // focusing plane: x when kstr > 0, y when kstr < 0
      const double fc = co, fs = si, dc = ch, ds = sh;
      if (kstr > 0.0) {
        xm11 = fc; xm12 = fs / rtkstr; xm21 = -rtkstr * fs; xm22 = fc;
        xm33 = dc; xm34 = ds / rtkstr; xm43 = ds * rtkstr; xm44 = dc;
      } else {
        xm11 = dc; xm12 = ds / rtkstr; xm21 = rtkstr * ds; xm22 = dc;
        xm33 = fc; xm34 = fs / rtkstr; xm43 = -fs * rtkstr; xm44 = fc;
      }
    } else {
      xm11 = 1.0; xm12 = tau; xm21 = 0.0; xm22 = 1.0;
      xm33 = 1.0; xm34 = tau; xm43 = 0.0; xm44 = 1.0;
    }
    X[i] = xm11 * x + xm12 * px / gambetz / scxl;
    PX[i] = gambetz * scxl * xm21 * x + xm22 * px;
    Y[i] = xm33 * y + xm34 * py / gambetz / scxl;
    PY[i] = gambetz * scxl * xm43 * y + xm44 * py;
    PH[i] = PH[i] + (1.0 / betaz - 1.0 / beta0) * tau / scxl;
  }
}

// ------------------------------------------------------------------ bends
// Sector_Dipole, Fpol_Dipole and Bpol_Dipole (src/Appl/Dipole.f90:466-788):
// each particle goes to Transport coordinates x, x' = px/(gamma beta_z),
// y, y', z = -phi beta0 Scxl, delta = -pt/(beta0 gambet) - dq with dq the
// particle's relative charge-to-mass offset, through the second-order map
// and back with gamma beta_z = |gamma beta| / sqrt(1 + x'^2 + y'^2).
// In place, one particle per lane.
template <int MODE>
__global__ void __launch_bounds__(kBlock)
k_dipole(double *__restrict__ X, double *__restrict__ PX, double *__restrict__ Y,
         double *__restrict__ PY, double *__restrict__ PH, double *__restrict__ PT,
         const uint32_t *__restrict__ orig, const double *__restrict__ qm0, int64_t n,
         DipoleMap m) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double px = PX[i], py = PY[i], pt = PT[i];
    const double dq = (qm0[orig[i]] - m.qm0) / m.qm0;
    const double gam = m.gam0 - pt;
    const double gbtot = sqrt(gam * gam - 1.0);
    const double gbz = MODE == kDipoleSector ? sqrt(gbtot * gbtot - px * px - py * py)
                                             : sqrt(gam * gam - 1.0 - px * px - py * py);
    const double x = X[i] * m.scxl, xp = px / gbz, y = Y[i] * m.scxl, yp = py / gbz;
    const double z = -PH[i] * m.beta0 * m.scxl;
    const double eta = -pt / m.beta0 / m.gambet - dq;
    double o1, o2, o3, o4, o5;
    if (MODE == kDipoleSector) {
      o1 = m.R11 * x + m.R12 * xp + m.R16 * eta + m.T111 * x * x + m.T112 * x * xp +
           m.T116 * x * eta + m.T122 * xp * xp + m.T126 * xp * eta + m.T166 * eta * eta +
           m.T144 * yp * yp;
      o2 = m.R21 * x + m.R22 * xp + m.R26 * eta + m.T216 * x * eta + m.T222 * xp * xp +
           m.T266 * eta * eta + m.T244 * yp * yp;
      o3 = y + m.len * yp + m.T314 * x * yp + m.T324 * xp * yp + m.T346 * yp * eta;
      o4 = yp;
      o5 = m.R51 * x + m.R52 * xp + z + m.R56 * eta + m.T522 * xp * xp + m.T544 * yp * yp +
           m.T526 * xp * eta;
    } else {
      const double h0 = m.h0, t = m.tanphi, t2 = m.tanphi2, s2 = m.secphi2;
      const double a = 0.5 * h0 * m.h1 * m.secphi3 + m.k1 * t;
      const double yeta = h0 * t - h0 * m.psi * m.secphib2;
      if (MODE == kDipoleEntrance) {
        o1 = x - 0.5 * h0 * t2 * x * x + 0.5 * h0 * s2 * y * y;
        o2 = h0 * t * x + xp + h0 * t2 * x * xp + a * x * x -
             (a - h0 * h0 * t * t2 - 0.5 * h0 * h0 * t) * y * y - h0 * t2 * y * yp -
             h0 * t * x * eta;
        o3 = y + h0 * t2 * x * y;
        o4 = -h0 * m.tanphib * y + yp - h0 * t2 * x * yp - h0 * s2 * xp * y -
             (h0 * m.h1 * m.secphi3 + 2.0 * m.k1 * t) * x * y + yeta * y * eta;
      } else {
        o1 = x + 0.5 * h0 * t2 * x * x - 0.5 * h0 * s2 * y * y;
        o2 = h0 * t * x + xp - h0 * t2 * x * xp + (a - h0 * h0 * t2 * t / 2.0) * x * x -
             (a + h0 * h0 * t * t2 / 2.0) * y * y + h0 * t2 * y * yp - h0 * t * x * eta;
        o3 = y - h0 * t2 * x * y;
        o4 = -h0 * m.tanphib * y + yp + h0 * t2 * x * yp + h0 * s2 * xp * y -
             (h0 * m.h1 * m.secphi3 + 2.0 * m.k1 * t - h0 * h0 * s2 * t) * x * y +
             yeta * y * eta;
      }
      o5 = z;
    }
    const double gbz2 = gbtot / sqrt(o2 * o2 + o4 * o4 + 1.0);
    X[i] = o1 / m.scxl;
    PX[i] = o2 * gbz2;
    Y[i] = o3 / m.scxl;
    PY[i] = o4 * gbz2;
    PH[i] = -o5 / m.scxl / m.beta0;
    PT[i] = -m.beta0 * m.gambet * (eta + dq);
  }
}

// ISRchicane (src/Contrl/AccSimulator.f90:1542-1584): the mean loss p1 and
// the fluctuation p2 * rd of incoherent synchrotron radiation remove
// (gamma beta)^2 (p1 + p2 rd) from each particle's momentum along its
// direction.  rd holds one deviate per particle, by original index.
__global__ void __launch_bounds__(kBlock)
k_isr(double *__restrict__ PX, double *__restrict__ PY, double *__restrict__ PT,
      const uint32_t *__restrict__ orig, const double *__restrict__ rd, int64_t n,
      double gam0, double p1, double p2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double gam = gam0 - PT[i];
    const double gambet = sqrt(gam * gam - 1.0);
    const double ptot = gambet * gambet * (p1 + p2 * rd[orig[i]]);
    PX[i] = PX[i] - ptot * PX[i] / gambet;
    PY[i] = PY[i] - ptot * PY[i] / gambet;
    const double gambet2 = gambet - ptot;
    PT[i] = gam0 - sqrt(1.0 + gambet2 * gambet2);
  }
}

// -------------------------------------------------------------- conv1st
// conv1st_BeamBunch, Flagbc=1 (src/Appl/BeamBunch.f90:1285-1352), in place.
// Partials per block: 3 minima, 3 maxima, lost count.
__global__ void __launch_bounds__(kBlock)
k_conv1st(double *__restrict__ x, const double *__restrict__ px,
          double *__restrict__ y, const double *__restrict__ py,
          double *__restrict__ ph, const double *__restrict__ pt, int64_t n,
          double gam, double xl, double rad, double *__restrict__ part) {
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  double lost = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double p5 = wrap_phase(ph[i]);
    double p2 = px[i], p4 = py[i], p6 = pt[i];
    double rcp = 1.0 / (-p6 + gam);
    double tmpbet = 1.0 - rcp * rcp * (1.0 + p2 * p2 + p4 * p4);
    double betai = sqrt(fabs(tmpbet));
    double x0 = x[i], y0 = y[i];
    double tmpx = x0 * xl;
    double xt = (x0 - p2 * p5 * rcp) * xl;
    double tmpy = y0 * xl;
    double yt = (y0 - p4 * p5 * rcp) * xl;
    double zt = -gam * betai * p5 * xl;
    mn[0] = fmin(mn[0], fmin(tmpx, xt));
    mx[0] = fmax(mx[0], fmax(tmpx, xt));
    mn[1] = fmin(mn[1], fmin(tmpy, yt));
    mx[1] = fmax(mx[1], fmax(tmpy, yt));
    mn[2] = fmin(mn[2], zt);
    mx[2] = fmax(mx[2], zt);
    x[i] = xt;
    y[i] = yt;
    ph[i] = zt;
    if (lost_t1(xt, yt, p2, p4, p6, gam, rad)) lost += 1.0;
  }
  __shared__ double sh[kBlock / 64][7];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double v[7];
  for (int k = 0; k < 3; k++) { v[k] = wave_min(mn[k]); v[3 + k] = wave_max(mx[k]); }
  v[6] = wave_sum(lost);
  if (lane == 0)
    for (int k = 0; k < 7; k++) sh[wv][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 7) {
    int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < kBlock / 64; w++)
      r = k < 3 ? fmin(r, sh[w][k]) : (k < 6 ? fmax(r, sh[w][k]) : r + sh[w][k]);
    part[blockIdx.x * 7 + k] = r;
  }
}

// ------------------------------------------------------------- compaction
// Order-preserving removal of lost particles (src/Appl/BeamBunch.f90:1290,
// 1345-1347): chunks of kCompactChunk consecutive particles; kept counts, an
// exclusive scan over chunks, then a stable in-chunk scatter.
constexpr int kCompactChunk = kBlock * 8;

__global__ void __launch_bounds__(kBlock)
k_compact_count(const double *__restrict__ x, const double *__restrict__ px,
                const double *__restrict__ y, const double *__restrict__ py,
                const double *__restrict__ pt, int64_t n, double gam,
                double rad, int64_t *__restrict__ cnt) {
  int64_t base = (int64_t)blockIdx.x * kCompactChunk;
  int kept = 0;
  for (int k = 0; k < kCompactChunk / kBlock; k++) {
    int64_t i = base + k * kBlock + threadIdx.x;
    if (i < n && !lost_t1(x[i], y[i], px[i], py[i], pt[i], gam, rad)) kept++;
  }
  __shared__ int sh[kBlock / 64];
  int s = kept;
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kBlock / 64; w++) t += sh[w];
    cnt[blockIdx.x] = t;
  }
}

__global__ void k_scan_counts(int64_t *__restrict__ cnt, int64_t nchunks) {
  // one thread: nchunks is N / 2048 and the path runs only after losses
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t acc = 0;
  for (int64_t c = 0; c < nchunks; c++) {
    int64_t v = cnt[c];
    cnt[c] = acc;
    acc += v;
  }
  cnt[nchunks] = acc;
}

struct Rows7 {
  double *r[6];
  uint32_t *orig;
};

__global__ void __launch_bounds__(kBlock)
k_compact_scatter(Rows7 src, Rows7 dst, int64_t n, double gam, double rad,
                  const int64_t *__restrict__ offs) {
  int64_t base = (int64_t)blockIdx.x * kCompactChunk;
  int64_t out = offs[blockIdx.x];
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < kCompactChunk / kBlock; k++) {
    int64_t i = base + k * kBlock + threadIdx.x;
    bool keep = i < n && !lost_t1(src.r[0][i], src.r[2][i], src.r[1][i],
                                  src.r[3][i], src.r[5][i], gam, rad);
    unsigned long long ballot = __ballot(keep);
    int before_in_wave = __popcll(ballot & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wv] = __popcll(ballot);
    __syncthreads();
    int before = before_in_wave;
    int total = 0;
    for (int w = 0; w < kBlock / 64; w++) {
      if (w < wv) before += wsum[w];
      total += wsum[w];
    }
    if (keep) {
      int64_t j = out + before;
      for (int r = 0; r < 6; r++) dst.r[r][j] = src.r[r][i];
      dst.orig[j] = src.orig[i];
    }
    out += total;
    __syncthreads();
  }
}

// --------------------------------------------------------- cvbkforth1st
// src/Appl/BeamBunch.f90:1679-1699; both reference loops are per particle,
// so they are applied back to back per particle.
__global__ void __launch_bounds__(kBlock)
k_cvbk(double *__restrict__ x, const double *__restrict__ px,
       double *__restrict__ y, const double *__restrict__ py,
       double *__restrict__ ph, const double *__restrict__ pt, int64_t n,
       double gamma0, double xl) {
  const double xk = 1.0 / xl;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double p2 = px[i], p4 = py[i];
    double rcp = 1.0 / (-pt[i] + gamma0);
    double betai = sqrt(1.0 - rcp * rcp * (1.0 + p2 * p2 + p4 * p4));
    double p5 = ph[i] * xk / (-gamma0 * betai);
    double p1 = x[i] * xk + p2 * p5 * rcp;
    double p3 = y[i] * xk + p4 * p5 * rcp;
    x[i] = p1 * xl;
    y[i] = p3 * xl;
    ph[i] = -gamma0 * betai * p5 * xl;
  }
}

// ---------------------------------------------------------- gather + kick
// scatter1_BeamBunch (src/Appl/BeamBunch.f90:2041-2181): CIC weights of the
// lower node (ab, cd, ef), 8-node gather in the reference's term order, kick,
// 0th-order t->z back transform.  T is the field type (double or float).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gather_kick(double *__restrict__ x, double *__restrict__ px,
              double *__restrict__ y, double *__restrict__ py,
              double *__restrict__ ph, double *__restrict__ pt, int64_t n,
              const E4<T> *__restrict__ e, const DevDom *__restrict__ dom,
              int nx, int ny, int nz, double gam, double xk, double tau_xycon,
              double tau_tcon) {
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  const double xmin = dom->range[0], ymin = dom->range[2], zmin = dom->range[4];
  const double hx = dom->msize[0], hy = dom->msize[1], hz = dom->msize[2];
  const double hxi = dom->hinv[0], hyi = dom->hinv[1], hzi = dom->hinv[2];
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double xi = x[i], yi = y[i], zi = ph[i];
    int ix = (int)((xi - xmin) * hxi + 1.0);
    double ab = ((xmin - xi) + ix * hx) * hxi;
    int jx = (int)((yi - ymin) * hyi + 1.0);
    double cd = ((ymin - yi) + jx * hy) * hyi;
    int kx = (int)((zi - zmin) * hzi + 1.0);
    double ef = ((zmin - zi) + kx * hz) * hzi;
    // the reference reads the grid unchecked; a particle outside the padded
    // box would be a corrupt state, so clamp rather than read out of bounds
    ix = min(max(ix, 1), nx - 1);
    jx = min(max(jx, 1), ny - 1);
    kx = min(max(kx, 1), nz - 1);
    int64_t o = (ix - 1) + sy * (jx - 1) + sz * (kx - 1);
    T gx, gy, gz;
    cic8_e4<T>(e, o, sy, sz, (T)ab, (T)cd, (T)ef, gx, gy, gz);
    double exn = (double)gx * gam, eyn = (double)gy * gam, ezn = (double)gz;
    double p2 = px[i], p4 = py[i], p6 = pt[i];
    double rcp = 1.0 / (-p6 + gam);
    double betai = sqrt(1.0 - rcp * rcp * (1.0 + p2 * p2 + p4 * p4));
    x[i] = xi * xk;
    px[i] = p2 + tau_xycon * exn;
    y[i] = yi * xk;
    py[i] = p4 + tau_xycon * eyn;
    ph[i] = zi * xk / (-gam * betai);
    pt[i] = p6 - tau_tcon * ezn;
  }
}

// ------------------------------------------------- Lorentz integrator
// drift2_BeamBunch (src/Appl/BeamBunch.f90:467-509): a drift of tau/2.
__global__ void __launch_bounds__(kBlock)
k_drift2(double *__restrict__ X, const double *__restrict__ PX, double *__restrict__ Y,
         const double *__restrict__ PY, double *__restrict__ PH, const double *__restrict__ PT,
         int64_t n, double tau, double gamma0, double beta0, double xl) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double px = PX[i], py = PY[i], g = gamma0 - PT[i];
    const double pz = sqrt(g * g - 1.0 - px * px - py * py);
    X[i] = X[i] + 0.5 * tau * px / (xl * pz);
    Y[i] = Y[i] + 0.5 * tau * py / (xl * pz);
    PH[i] = PH[i] + 0.5 * tau * (g / (xl * pz) - 1.0 / (beta0 * xl));
  }
}

// scatter2_BeamBunch (src/Appl/BeamBunch.f90:2726-3067) with SC: the space-
// charge E gathered at the 0th-order t-frame position (scaled back by
// tmpscale), the back transform to the z frame, the element's field at the
// particle and the implicit Lorentz push.  Without SC it is scatter20
// (:3436-3527): the element's field only, on z-frame particles.
template <typename T, bool SC>
__global__ void __launch_bounds__(kBlock)
k_kick2(double *__restrict__ X, double *__restrict__ PX, double *__restrict__ Y,
        double *__restrict__ PY, double *__restrict__ PH, double *__restrict__ PT,
        const uint32_t *__restrict__ orig, const double *__restrict__ qm0, int64_t n,
        const E4<T> *__restrict__ e, const DevDom *__restrict__ dom, int nx, int ny, int nz,
        Kick2 k) {
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double x = X[i], y = Y[i], ph = PH[i], px = PX[i], py = PY[i], pt = PT[i];
    double exn = 0.0, eyn = 0.0, ezn = 0.0;
    if constexpr (SC) {
      const double xmin = dom->range[0], ymin = dom->range[2], zmin = dom->range[4];
      const double hx = dom->msize[0], hy = dom->msize[1], hz = dom->msize[2];
      const double hxi = dom->hinv[0], hyi = dom->hinv[1], hzi = dom->hinv[2];
      int ix = (int)((x - xmin) * hxi + 1.0);
      const double ab = ((xmin - x) + ix * hx) * hxi;
      int jx = (int)((y - ymin) * hyi + 1.0);
      const double cd = ((ymin - y) + jx * hy) * hyi;
      int kx = (int)((ph - zmin) * hzi + 1.0);
      const double ef = ((zmin - ph) + kx * hz) * hzi;
      ix = min(max(ix, 1), nx - 1);
      jx = min(max(jx, 1), ny - 1);
      kx = min(max(kx, 1), nz - 1);
      T gx, gy, gz;
      cic8_e4<T>(e, (ix - 1) + sy * (jx - 1) + sz * (kx - 1), sy, sz, (T)ab, (T)cd, (T)ef,
                 gx, gy, gz);
      exn = (double)gx * k.gam;
      eyn = (double)gy * k.gam;
      ezn = (double)gz;
      const double rcp = 1.0 / (-pt + k.gam);
      const double betai = sqrt(1.0 - rcp * rcp * (1 + px * px + py * py));
      ph = ph * k.xk / (-k.gam * betai);
      x = x * k.xk;
      y = y * k.xk;
    }
    const double pos[4] = {x * k.xl, y * k.xl, k.z, ph + k.tg};
    double f[6];
    elem_field(k.e, k.xl, k.clight, pos, f);
    double ex = k.tmpscale * exn + f[0], ey = k.tmpscale * eyn + f[1];
    double ez = k.tmpscale * ezn + f[2];
    if (SC && k.wake) {
      // scatter2wake: the wake of the particle's z, linear in its CIC weight
      // (src/Appl/BeamBunch.f90:5634-5636,5886-5888)
      const double zmin = dom->range[4], hz = dom->msize[2], hzi = dom->hinv[2];
      const double phz = PH[i];
      int kx = (int)((phz - zmin) * hzi + 1.0);
      const double ef = ((zmin - phz) + kx * hz) * hzi;
      kx = min(max(kx, 1), nz - 1);
      const double *w = k.wake;
      ex = ex + (w[kx - 1] * ef + w[kx] * (1.0 - ef));
      ey = ey + (w[k.nzw + kx - 1] * ef + w[k.nzw + kx] * (1.0 - ef));
      ez = ez + (w[2 * k.nzw + kx - 1] * ef + w[2 * k.nzw + kx] * (1.0 - ef));
    }
    const double bx = -k.beta0_c * k.tmpscale * eyn + f[3];
    const double by = k.beta0_c * k.tmpscale * exn + f[4];
    lorentz_push(px, py, pt, qm0[orig[i]], ex, ey, ez, bx, by, f[5], k.tau, k.clight,
                 k.tmpgamma0, k.gam, k.qmcc, k.ez0);
    X[i] = x; Y[i] = y; PH[i] = ph;
    PX[i] = px; PY[i] = py; PT[i] = pt;
  }
}

// ----------------------------------------------------------------- moments
__global__ void __launch_bounds__(kBlock)
k_moments(const double *__restrict__ X, const double *__restrict__ PX,
          const double *__restrict__ Y, const double *__restrict__ PY,
          const double *__restrict__ Z, const double *__restrict__ PZ, int64_t n,
          double *__restrict__ part) {
  Mom a;
  mom_zero(a);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    mom_add(a, X[i], PX[i], Y[i], PY[i], Z[i], PZ[i]);
  __shared__ double sh[kBlock / 64][kMoments];
  mom_block_store(a, sh, part + (int64_t)blockIdx.x * kMoments);
}

// Plain streaming copy, 16 bytes per lane per iteration: the reference
// kernel for calibrating hardware byte counters.
__global__ void __launch_bounds__(kBlock)
k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

// ================================================================ launchers

int launch_probe_copy(int64_t nbytes, double *ms) {
  const int64_t n = nbytes / 16;
  if (n < 1) return fail(IZH_ERR_ARG, "izh_probe_copy: at least 16 bytes");
  uint4 *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  IZH_HIP(hipMalloc((void **)&a, n * 16));
  IZH_HIP(hipMalloc((void **)&b, n * 16));
  IZH_HIP(hipMemset(a, 1, n * 16));
  IZH_HIP(hipEventCreate(&e0));
  IZH_HIP(hipEventCreate(&e1));
  int64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  IZH_HIP(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(kBlock), 0, 0, a, b, n);
  IZH_HIP(hipEventRecord(e1, 0));
  IZH_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  IZH_HIP(hipEventElapsedTime(&t, e0, e1));
  if (ms) *ms = t;
  IZH_HIP(hipEventDestroy(e0));
  IZH_HIP(hipEventDestroy(e1));
  IZH_HIP(hipFree(a));
  IZH_HIP(hipFree(b));
  return IZH_OK;
}

static int launch_push(izh_bunch *b, const Push &p) {
  if (b->n == 0) return IZH_OK;
  ScopedTimer tm_(IZH_T_PUSH, b->stream);
  hipLaunchKernelGGL(k_push, dim3(grid_for(b->n)), dim3(kBlock), 0, b->stream,
                     b->r[0], b->r[1], b->r[2], b->r[3], b->r[4], b->r[5], b->n, p);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_drift(izh_bunch *b, double tau, double gam0, double beta0) {
  touch_particles(b);
  Push p{};
  p.kind = 1;
  p.tau = tau;
  p.gam0 = gam0;
  p.ibeta0 = 1.0 / beta0;
  p.scxl = b->c.scxl;
  return launch_push(b, p);
}

int launch_quad_kmode(izh_bunch *b, double tau, double kq, double qmass, double gambet0,
                      double beta0) {
  touch_particles(b);
  if (b->n == 0) return IZH_OK;
  ScopedTimer tm_(IZH_T_PUSH, b->stream);
  hipLaunchKernelGGL(k_quad_kmode, dim3(grid_for(b->n)), dim3(kBlock), 0, b->stream,
                     b->r[0], b->r[1], b->r[2], b->r[3], b->r[4], b->r[5], b->orig, b->qm0,
                     b->n, tau, kq, qmass, -b->refpt[5], gambet0, beta0, b->c.scxl);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_dipole(izh_bunch *b, const DipoleMap &m) {
  touch_particles(b);
  if (b->n == 0) return IZH_OK;
  ScopedTimer tm_(IZH_T_PUSH, b->stream);
  const dim3 g(grid_for(b->n)), t(kBlock);
  if (m.mode == kDipoleSector)
    hipLaunchKernelGGL(k_dipole<kDipoleSector>, g, t, 0, b->stream, b->r[0], b->r[1],
                       b->r[2], b->r[3], b->r[4], b->r[5], b->orig, b->qm0, b->n, m);
  else if (m.mode == kDipoleEntrance)
    hipLaunchKernelGGL(k_dipole<kDipoleEntrance>, g, t, 0, b->stream, b->r[0], b->r[1],
                       b->r[2], b->r[3], b->r[4], b->r[5], b->orig, b->qm0, b->n, m);
  else
    hipLaunchKernelGGL(k_dipole<kDipoleExit>, g, t, 0, b->stream, b->r[0], b->r[1],
                       b->r[2], b->r[3], b->r[4], b->r[5], b->orig, b->qm0, b->n, m);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_isr(izh_bunch *b, double gam0, double p1, double p2, const double *rd) {
  touch_particles(b);
  if (b->n == 0) return IZH_OK;
  ScopedTimer tm_(IZH_T_PUSH, b->stream);
  hipLaunchKernelGGL(k_isr, dim3(grid_for(b->n)), dim3(kBlock), 0, b->stream, b->r[1],
                     b->r[3], b->r[5], b->orig, rd, b->n, gam0, p1, p2);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_map_linear(izh_bunch *b, const double xm[36]) {
  touch_particles(b);
  Push p{};
  p.kind = 2;
  std::memcpy(p.xm, xm, sizeof p.xm);
  return launch_push(b, p);
}

int launch_conv1st(izh_bunch *b, double xrad, double yrad, double ptrange[6],
                   int64_t *nlost) {
  touch_particles(b);
  const double gam = -b->refpt[5];
  const double rad = (xrad + yrad) / 2.0;
  const int nb = grid_for(b->n);
  {
    ScopedTimer tm_(IZH_T_CONV1ST, b->stream);
    hipLaunchKernelGGL(k_conv1st, dim3(nb), dim3(kBlock), 0, b->stream, b->r[0],
                       b->r[1], b->r[2], b->r[3], b->r[4], b->r[5], b->n, gam,
                       b->c.scxl, rad, b->d_part);
    IZH_CHECK(launch_reduce(b->stream, b->d_part, nb, 7, 0x7ull, 0x38ull, b->d_out));
  }
  IZH_HIP(hipGetLastError());
  IZH_HIP(hipMemcpyAsync(b->h_out, b->d_out, 7 * sizeof(double),
                         hipMemcpyDeviceToHost, b->stream));
  IZH_HIP(hipStreamSynchronize(b->stream));
  // ptrange(1..6) = xmin, xmax, ymin, ymax, zmin, zmax
  ptrange[0] = b->h_out[0]; ptrange[1] = b->h_out[3];
  ptrange[2] = b->h_out[1]; ptrange[3] = b->h_out[4];
  ptrange[4] = b->h_out[2]; ptrange[5] = b->h_out[5];
  *nlost = (int64_t)b->h_out[6];
  return IZH_OK;
}

int launch_compact(izh_bunch *b, double xrad, double yrad, int64_t nkeep) {
  touch_particles(b);
  const double gam = -b->refpt[5];
  const double rad = (xrad + yrad) / 2.0;
  const int64_t nchunks = (b->n + kCompactChunk - 1) / kCompactChunk;
  int64_t *cnt = nullptr;
  IZH_HIP(hipMalloc((void **)&cnt, (nchunks + 1) * sizeof(int64_t)));
  hipLaunchKernelGGL(k_compact_count, dim3(nchunks), dim3(kBlock), 0, b->stream,
                     b->r[0], b->r[1], b->r[2], b->r[3], b->r[5], b->n, gam,
                     rad, cnt);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(64), 0, b->stream, cnt, nchunks);
  Rows7 src, dst;
  for (int r = 0; r < 6; r++) { src.r[r] = b->r[r]; dst.r[r] = b->r_alt[r]; }
  src.orig = b->orig;
  dst.orig = b->orig_alt;
  hipLaunchKernelGGL(k_compact_scatter, dim3(nchunks), dim3(kBlock), 0,
                     b->stream, src, dst, b->n, gam, rad, cnt);
  IZH_HIP(hipGetLastError());
  int64_t total = -1;
  IZH_HIP(hipMemcpyAsync(&total, cnt + nchunks, sizeof(int64_t),
                         hipMemcpyDeviceToHost, b->stream));
  IZH_HIP(hipStreamSynchronize(b->stream));
  IZH_HIP(hipFree(cnt));
  if (total != nkeep)
    return fail(IZH_ERR_STATE, "compaction kept %lld particles, conv1st %lld",
                (long long)total, (long long)nkeep);
  for (int r = 0; r < 6; r++) std::swap(b->r[r], b->r_alt[r]);
  std::swap(b->orig, b->orig_alt);
  b->n = nkeep;
  return IZH_OK;
}

int launch_cvbk(izh_bunch *b) {
  touch_particles(b);
  if (b->n == 0) return IZH_OK;
  ScopedTimer tm_(IZH_T_CVBK, b->stream);
  hipLaunchKernelGGL(k_cvbk, dim3(grid_for(b->n)), dim3(kBlock), 0, b->stream,
                     b->r[0], b->r[1], b->r[2], b->r[3], b->r[4], b->r[5], b->n,
                     -b->refpt[5], b->c.scxl);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_gather_kick(izh_bunch *b, izh_field *f, double tau, double xycon,
                       double tcon) {
  touch_particles(b);
  if (b->n == 0) return IZH_OK;
  const double gam = -b->refpt[5];
  const double xk = 1.0 / b->c.scxl;
  // the grid holds charge * rho_scale, so the field carries the same factor
  const double txy = tau * xycon / f->rho_scale, ttc = tau * tcon / f->rho_scale;
  ScopedTimer tm_(IZH_T_KICK, b->stream);
  if (f->precision == IZH_FP64)
    hipLaunchKernelGGL(k_gather_kick<double>, dim3(grid_for(b->n)), dim3(kBlock),
                       0, b->stream, b->r[0], b->r[1], b->r[2], b->r[3], b->r[4],
                       b->r[5], b->n, (const E4<double> *)f->e, f->d_dom, f->nx,
                       f->ny, f->nz, gam, xk, txy, ttc);
  else
    hipLaunchKernelGGL(k_gather_kick<float>, dim3(grid_for(b->n)), dim3(kBlock),
                       0, b->stream, b->r[0], b->r[1], b->r[2], b->r[3], b->r[4],
                       b->r[5], b->n, (const E4<float> *)f->e, f->d_dom, f->nx,
                       f->ny, f->nz, gam, xk, txy, ttc);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_drift2(izh_bunch *b, double tau, double gamma0, double beta0) {
  touch_particles(b);
  if (b->n == 0) return IZH_OK;
  ScopedTimer tm_(IZH_T_PUSH, b->stream);
  hipLaunchKernelGGL(k_drift2, dim3(grid_for(b->n)), dim3(kBlock), 0, b->stream, b->r[0],
                     b->r[1], b->r[2], b->r[3], b->r[4], b->r[5], b->n, tau, gamma0, beta0,
                     b->c.scxl);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_kick2(izh_bunch *b, izh_field *f, const Kick2 &k, bool space_charge) {
  touch_particles(b);
  if (b->n == 0) return IZH_OK;
  ScopedTimer tm_(IZH_T_KICK, b->stream);
  if (!space_charge) {
    hipLaunchKernelGGL((k_kick2<double, false>), dim3(grid_for(b->n)), dim3(kBlock), 0,
                       b->stream, b->r[0], b->r[1], b->r[2], b->r[3], b->r[4], b->r[5],
                       b->orig, b->qm0, b->n, (const E4<double> *)nullptr, nullptr, 2, 2, 2,
                       k);
  } else if (f->precision == IZH_FP64) {
    hipLaunchKernelGGL((k_kick2<double, true>), dim3(grid_for(b->n)), dim3(kBlock), 0,
                       b->stream, b->r[0], b->r[1], b->r[2], b->r[3], b->r[4], b->r[5],
                       b->orig, b->qm0, b->n, (const E4<double> *)f->e, f->d_dom, f->nx,
                       f->ny, f->nz, k);
  } else {
    hipLaunchKernelGGL((k_kick2<float, true>), dim3(grid_for(b->n)), dim3(kBlock), 0,
                       b->stream, b->r[0], b->r[1], b->r[2], b->r[3], b->r[4], b->r[5],
                       b->orig, b->qm0, b->n, (const E4<float> *)f->e, f->d_dom, f->nx,
                       f->ny, f->nz, k);
  }
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

namespace {
// chgupdate_BeamBunch (src/Appl/BeamBunch.f90:688-728): live particles per
// charge state, a particle counting for state j when its q/mc^2 lies within
// relative eps of qmcc[j].  Counts are integers, so the atomics' order does
// not matter.
__global__ void __launch_bounds__(kBlock)
k_chg_count(const uint32_t *__restrict__ orig, const double *__restrict__ qm0, int64_t n,
            ChgList cl, unsigned long long *__restrict__ counts) {
  __shared__ unsigned int lc[kMaxChargeStates];
  for (int j = threadIdx.x; j < cl.n; j += blockDim.x) lc[j] = 0u;
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double q = qm0[orig[i]];
    for (int j = 0; j < cl.n; j++)
      if (fabs((q - cl.qmcc[j]) / cl.qmcc[j]) < cl.eps) atomicAdd(&lc[j], 1u);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < cl.n; j += blockDim.x)
    if (lc[j]) atomicAdd(&counts[j], (unsigned long long)lc[j]);
}
}  // namespace

int launch_chg_count(izh_bunch *b, const ChgList &cl, int64_t *out) {
  if (!b->d_chg) {
    IZH_HIP(hipMalloc((void **)&b->d_chg, kMaxChargeStates * sizeof(unsigned long long)));
    IZH_HIP(hipHostMalloc((void **)&b->h_chg, kMaxChargeStates * sizeof(unsigned long long),
                          hipHostMallocDefault));
  }
  hipLaunchKernelGGL(k_zero_u64_small, dim3(1), dim3(kMaxChargeStates), 0, b->stream,
                     b->d_chg, cl.n);
  if (b->n > 0)
    hipLaunchKernelGGL(k_chg_count, dim3(grid_for(b->n)), dim3(kBlock), 0, b->stream,
                       b->orig, b->qm0, b->n, cl, b->d_chg);
  IZH_HIP(hipGetLastError());
  IZH_HIP(hipMemcpyAsync(b->h_chg, b->d_chg, cl.n * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, b->stream));
  IZH_HIP(hipStreamSynchronize(b->stream));
  for (int j = 0; j < cl.n; j++) out[j] = (int64_t)b->h_chg[j];
  return IZH_OK;
}

// ------------------------------------------------------------------ clone
namespace {
// Slots offset..offset+n-1 of a bunch into slots 0..n-1 of another: the six
// rows, and the immutable rows through the source's original index; the new
// bunch's original index is the identity and its charge weights are w.
__global__ void __launch_bounds__(kBlock)
k_clone(const double *__restrict__ src, int64_t src_cap, const uint32_t *__restrict__ sorig,
        const double *__restrict__ sqm, const double *__restrict__ sid, double *__restrict__ dst,
        int64_t dst_cap, uint32_t *__restrict__ dorig, double *__restrict__ dqm,
        double *__restrict__ dw, double *__restrict__ did, int64_t offset, int64_t n, double w) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = offset + i;
#pragma unroll
    for (int k = 0; k < 6; k++) dst[k * dst_cap + i] = src[k * src_cap + j];
    const uint32_t o = sorig[j];
    dorig[i] = (uint32_t)i;
    dqm[i] = sqm[o];
    did[i] = sid[o];
    dw[i] = w;
  }
}
}  // namespace

int launch_clone(const izh_bunch *src, izh_bunch *dst, int64_t offset, int64_t n, double w) {
  if (n == 0) return IZH_OK;
  hipLaunchKernelGGL(k_clone, dim3(grid_for(n)), dim3(kBlock), 0, dst->stream, src->r[0],
                     src->cap, src->orig, src->qm0, src->id0, dst->r[0], dst->cap, dst->orig,
                     dst->qm0, dst->w0, dst->id0, offset, n, w);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

// ------------------------------------------------ allocation-free moments
// The scratch of izh_diagnostic1_rows: the per-block partials of the moment
// sums, kMaxRedBlocks x kMoments doubles (704 KiB).  Each GPU holds its own
// copy for the life of the process.  Host code never hands `scratch` itself
// to a kernel (on the host it names a stand-in, not device memory): its
// address on each device comes from hipGetSymbolAddress, kept per device.
static __device__ __align__(256) double scratch[kMaxRedBlocks * kMoments];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kScratchDevices] = {};
}  // namespace

int scratch_setup(int device) {
  if (device < 0 || device >= kScratchDevices)
    return fail(IZH_ERR_ARG, "scratch: device %d beyond the %d the library tracks", device,
                kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return IZH_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(IZH_ERR_HIP, "scratch: hipGetSymbolAddress on device %d: %s", device,
                hipGetErrorString(e));
  g_scratch_ptr[device] = (double *)p;
  return IZH_OK;
}

int launch_moments_rows(const double *rows, int64_t ld, int64_t n, double *out,
                        hipStream_t st) {
  int dev = 0;
  IZH_HIP(hipGetDevice(&dev));
  IZH_CHECK(scratch_setup(dev));   // set up by izh_init; here for a device it skipped
  double *part = g_scratch_ptr[dev];
  if (n == 0) {
    hipLaunchKernelGGL(k_zero_u64_small, dim3(1), dim3(64), 0, st,
                       (unsigned long long *)out, kMoments);
    IZH_HIP(hipGetLastError());
    return IZH_OK;
  }
  // the kernels of launch_moments, on the scratch partials
  const int nb = grid_for(n);
  hipLaunchKernelGGL(k_moments, dim3(nb), dim3(kBlock), 0, st, rows, rows + ld,
                     rows + 2 * ld, rows + 3 * ld, rows + 4 * ld, rows + 5 * ld, n, part);
  IZH_CHECK(launch_reduce(st, part, nb, kMoments, 0ull, kMomentsMaxMask, out));
  return IZH_OK;
}

int launch_moments(izh_bunch *b, double out[kMoments]) {
  const int nb = grid_for(b->n);
  {
    ScopedTimer tm_(IZH_T_DIAG, b->stream);
    hipLaunchKernelGGL(k_moments, dim3(nb), dim3(kBlock), 0, b->stream, b->r[0],
                       b->r[1], b->r[2], b->r[3], b->r[4], b->r[5], b->n,
                       b->d_part);
    IZH_CHECK(launch_reduce(b->stream, b->d_part, nb, kMoments, 0ull,
                            kMomentsMaxMask, b->d_out));
  }
  IZH_HIP(hipMemcpyAsync(b->h_out, b->d_out, kMoments * sizeof(double),
                         hipMemcpyDeviceToHost, b->stream));
  IZH_HIP(hipStreamSynchronize(b->stream));
  std::memcpy(out, b->h_out, kMoments * sizeof(double));
  return IZH_OK;
}

}  // namespace izh
