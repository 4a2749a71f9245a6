// izh_particles.hip -- per-particle passes of the space-charge step on gfx950.
//
// Particles live in structure-of-arrays form (one FP64 array per Pts1 row),
// so every pass streams only the rows it touches with fully coalesced loads.
// Reductions (conv1st range, diagnostic moments) go through a fixed number of
// per-block partials and a single-block final pass: the summation order
// depends only on N, so results are bitwise reproducible run to run.
#include <cmath>
#include <cstring>

#include "izh_internal.hpp"

namespace izh {
namespace {

// REAL(4) pi of conv1st: pilc = 2.0*asin(1.0) (src/Appl/BeamBunch.f90:1268).
__device__ __host__ inline double pilc_f32() { return (double)3.14159274101257324f; }

// ------------------------------------------------------------ wave reductions
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}
__device__ inline double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
  return v;
}

// ------------------------------------------------------------------- drift
// drift1_BeamBunch, bitype 0 (src/Appl/BeamBunch.f90:276-298).
__global__ void __launch_bounds__(kBlock)
k_drift(double *__restrict__ x, const double *__restrict__ px,
        double *__restrict__ y, const double *__restrict__ py,
        double *__restrict__ ph, const double *__restrict__ pt, int64_t n,
        double tau, double gam0, double beta0, double scxl) {
  const double ib = 1.0 / beta0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double tpx = px[i], tpy = py[i], tpt = pt[i];
    double d = tpt - gam0;
    double tmph = sqrt(d * d - 1.0 - tpx * tpx - tpy * tpy);
    x[i] = x[i] + tpx * tau / tmph / scxl;
    y[i] = y[i] + tpy * tau / tmph / scxl;
    ph[i] = ph[i] - (ib + d / tmph) * tau / scxl;
  }
}

struct Map6 {
  double m[36];
};

// The 6x6 map product of drift1_BeamBunch (src/Appl/BeamBunch.f90:308-320).
__global__ void __launch_bounds__(kBlock)
k_map_linear(double *__restrict__ x, double *__restrict__ px,
             double *__restrict__ y, double *__restrict__ py,
             double *__restrict__ ph, double *__restrict__ pt, int64_t n,
             Map6 xm) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double p[6] = {x[i], px[i], y[i], py[i], ph[i], pt[i]};
    double t[6];
#pragma unroll
    for (int j = 0; j < 6; j++) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 6; k++) s += p[k] * xm.m[6 * j + k];
      t[j] = s;
    }
    x[i] = t[0]; px[i] = t[1]; y[i] = t[2]; py[i] = t[3]; ph[i] = t[4]; pt[i] = t[5];
  }
}

// -------------------------------------------------------------- conv1st
// Loss predicate of conv1st (src/Appl/BeamBunch.f90:1340-1343), evaluated on
// the transformed coordinates so the compaction pass can re-evaluate it.
__device__ inline bool lost_t1(double xt, double yt, double px, double py,
                               double pt, double gam, double rad) {
  double rcp = 1.0 / (-pt + gam);
  double tmpbet = 1.0 - rcp * rcp * (1.0 + px * px + py * py);
  double radtest = sqrt(xt * xt + yt * yt);
  return radtest >= rad || tmpbet <= 0.0;
}

// conv1st_BeamBunch, Flagbc=1 (src/Appl/BeamBunch.f90:1285-1352), in place.
// Partials per block: 3 minima, 3 maxima, lost count.
__global__ void __launch_bounds__(kBlock)
k_conv1st(double *__restrict__ x, const double *__restrict__ px,
          double *__restrict__ y, const double *__restrict__ py,
          double *__restrict__ ph, const double *__restrict__ pt, int64_t n,
          double gam, double xl, double rad, double *__restrict__ part) {
  const double pilc = pilc_f32();
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  double lost = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double p5 = ph[i];
    int ntmp5 = (int)(p5 / pilc);
    double tmp5 = p5 - ntmp5 * pilc;
    p5 = tmp5 - (ntmp5 % 2) * pilc;
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
  __shared__ double sh[4][7];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double v[7];
  for (int k = 0; k < 3; k++) { v[k] = wave_min(mn[k]); v[3 + k] = wave_max(mx[k]); }
  v[6] = wave_sum(lost);
  if (lane == 0)
    for (int k = 0; k < 7; k++) sh[wv][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[7];
    for (int k = 0; k < 7; k++) r[k] = sh[0][k];
    for (int w = 1; w < kBlock / 64; w++) {
      for (int k = 0; k < 3; k++) { r[k] = fmin(r[k], sh[w][k]); r[3 + k] = fmax(r[3 + k], sh[w][3 + k]); }
      r[6] += sh[w][6];
    }
    for (int k = 0; k < 7; k++) part[blockIdx.x * 7 + k] = r[k];
  }
}

__global__ void k_conv1st_final(const double *__restrict__ part, int nparts,
                                double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double r[7] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300, 0.0};
  for (int b = 0; b < nparts; b++) {
    for (int k = 0; k < 3; k++) {
      r[k] = fmin(r[k], part[b * 7 + k]);
      r[3 + k] = fmax(r[3 + k], part[b * 7 + 3 + k]);
    }
    r[6] += part[b * 7 + 6];
  }
  for (int k = 0; k < 7; k++) out[k] = r[k];
}

// ------------------------------------------------------------- compaction
// Order-preserving removal of lost particles (src/Appl/BeamBunch.f90:1290,
// 1345-1347): chunks of kChunk consecutive particles; kept counts, an
// exclusive scan over chunks, then a stable in-chunk scatter.
constexpr int kChunk = kBlock * 8;

__global__ void __launch_bounds__(kBlock)
k_compact_count(const double *__restrict__ x, const double *__restrict__ px,
                const double *__restrict__ y, const double *__restrict__ py,
                const double *__restrict__ pt, int64_t n, double gam,
                double rad, int64_t *__restrict__ cnt) {
  int64_t base = (int64_t)blockIdx.x * kChunk;
  int kept = 0;
  for (int k = 0; k < kChunk / kBlock; k++) {
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
  // single thread: nchunks is N / 2048, small enough for a serial scan of a
  // path that runs only when particles were lost
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t acc = 0;
  for (int64_t c = 0; c < nchunks; c++) {
    int64_t v = cnt[c];
    cnt[c] = acc;
    acc += v;
  }
  cnt[nchunks] = acc;
}

struct Rows9 {
  double *r[9];
};

__global__ void __launch_bounds__(kBlock)
k_compact_scatter(Rows9 src, Rows9 dst, int64_t n, double gam, double rad,
                  const int64_t *__restrict__ offs) {
  int64_t base = (int64_t)blockIdx.x * kChunk;
  int64_t out = offs[blockIdx.x];
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < kChunk / kBlock; k++) {
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
      for (int r = 0; r < 9; r++) dst.r[r][j] = src.r[r][i];
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
__device__ inline T cic8(const T *__restrict__ e, int64_t o, int64_t sy,
                         int64_t sz, T ab, T cd, T ef) {
  const T one = (T)1;
  return e[o] * ab * cd * ef + e[o + sy] * ab * (one - cd) * ef +
         e[o + sy + sz] * ab * (one - cd) * (one - ef) +
         e[o + sz] * ab * cd * (one - ef) +
         e[o + 1 + sz] * (one - ab) * cd * (one - ef) +
         e[o + 1 + sy + sz] * (one - ab) * (one - cd) * (one - ef) +
         e[o + 1 + sy] * (one - ab) * (one - cd) * ef +
         e[o + 1] * (one - ab) * cd * ef;
}

struct Dom {
  double xmin, ymin, zmin, hx, hy, hz, hxi, hyi, hzi;
  int nx, ny, nz;
};

template <typename T>
__global__ void __launch_bounds__(kBlock)
k_gather_kick(double *__restrict__ x, double *__restrict__ px,
              double *__restrict__ y, double *__restrict__ py,
              double *__restrict__ ph, double *__restrict__ pt, int64_t n,
              const T *__restrict__ ex, const T *__restrict__ ey,
              const T *__restrict__ ez, Dom d, double gam, double xk,
              double tau_xycon, double tau_tcon) {
  const int64_t sy = d.nx, sz = (int64_t)d.nx * d.ny;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double xi = x[i], yi = y[i], zi = ph[i];
    int ix = (int)((xi - d.xmin) * d.hxi + 1.0);
    double ab = ((d.xmin - xi) + ix * d.hx) * d.hxi;
    int jx = (int)((yi - d.ymin) * d.hyi + 1.0);
    double cd = ((d.ymin - yi) + jx * d.hy) * d.hyi;
    int kx = (int)((zi - d.zmin) * d.hzi + 1.0);
    double ef = ((d.zmin - zi) + kx * d.hz) * d.hzi;
    // the reference reads the grid unchecked; a particle outside the padded
    // box would be a corrupt state, so clamp rather than read out of bounds
    ix = min(max(ix, 1), d.nx - 1);
    jx = min(max(jx, 1), d.ny - 1);
    kx = min(max(kx, 1), d.nz - 1);
    int64_t o = (ix - 1) + sy * (jx - 1) + sz * (kx - 1);
    T tab = (T)ab, tcd = (T)cd, tef = (T)ef;
    double exn = (double)cic8<T>(ex, o, sy, sz, tab, tcd, tef) * gam;
    double eyn = (double)cic8<T>(ey, o, sy, sz, tab, tcd, tef) * gam;
    double ezn = (double)cic8<T>(ez, o, sy, sz, tab, tcd, tef);
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

// ----------------------------------------------------------------- moments
// diagnostic1_Output sums (src/Contrl/Output.f90:119-199), tmplc order.
__global__ void __launch_bounds__(kBlock)
k_moments(const double *__restrict__ X, const double *__restrict__ PX,
          const double *__restrict__ Y, const double *__restrict__ PY,
          const double *__restrict__ Z, const double *__restrict__ PZ, int64_t n,
          double *__restrict__ part) {
  double s[37];
#pragma unroll
  for (int k = 0; k < 37; k++) s[k] = 0.0;
  double m[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double x = X[i], px = PX[i], y = Y[i], py = PY[i], z = Z[i], pz = PZ[i];
    double x2 = x * x, px2 = px * px, y2 = y * y, py2 = py * py, z2 = z * z, pz2 = pz * pz;
    s[0] += x; s[6] += x2; s[15] += x2 * x; s[16] += x2 * x2;
    s[12] += x * px; s[27] += x * px2; s[28] += x * px2 * px;
    s[29] += x2 * px; s[30] += x2 * px2; s[31] += x2 * x * px;
    s[1] += px; s[7] += px2; s[17] += px2 * px; s[18] += px2 * px2;
    s[2] += y; s[8] += y2; s[19] += y2 * y; s[20] += y2 * y2;
    s[13] += y * py; s[32] += y * py2; s[33] += y * py2 * py;
    s[34] += y2 * py; s[35] += y2 * py2; s[36] += y2 * y * py;
    s[3] += py; s[9] += py2; s[21] += py2 * py; s[22] += py2 * py2;
    s[4] += z; s[10] += z2; s[23] += fabs(z2 * z); s[24] += z2 * z2;
    s[14] += z * pz;
    s[5] += pz; s[11] += pz2; s[25] += pz2 * pz; s[26] += pz2 * pz2;
    m[0] = fmax(m[0], fabs(x)); m[1] = fmax(m[1], fabs(px));
    m[2] = fmax(m[2], fabs(y)); m[3] = fmax(m[3], fabs(py));
    m[4] = fmax(m[4], fabs(z)); m[5] = fmax(m[5], fabs(pz));
    m[6] = fmax(m[6], x2 + y2);
  }
  __shared__ double sh[kBlock / 64][kMoments];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 37; k++) {
    double v = wave_sum(s[k]);
    if (lane == 0) sh[wv][k] = v;
  }
#pragma unroll
  for (int k = 0; k < 7; k++) {
    double v = wave_max(m[k]);
    if (lane == 0) sh[wv][37 + k] = v;
  }
  __syncthreads();
  if (threadIdx.x < kMoments) {
    int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < kBlock / 64; w++)
      r = (k < 37) ? r + sh[w][k] : fmax(r, sh[w][k]);
    part[blockIdx.x * kMoments + k] = r;
  }
}

__global__ void k_moments_final(const double *__restrict__ part, int nparts,
                                double *__restrict__ out) {
  int k = threadIdx.x;
  if (k >= kMoments) return;
  double r = part[k];
  for (int b = 1; b < nparts; b++)
    r = (k < 37) ? r + part[b * kMoments + k] : fmax(r, part[b * kMoments + k]);
  out[k] = r;
}

}  // namespace

// ================================================================ launchers

int launch_drift(izh_bunch *b, double tau, double gam0, double beta0) {
  if (b->n == 0) return IZH_OK;
  {
    ScopedTimer tm_(IZH_T_PUSH, b->stream);
    hipLaunchKernelGGL(k_drift, dim3(grid_for(b->n)), dim3(kBlock), 0, b->stream,
                       b->s[0], b->s[1], b->s[2], b->s[3], b->s[4], b->s[5], b->n,
                       tau, gam0, beta0, b->c.scxl);
  }
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_map_linear(izh_bunch *b, const double xm[36]) {
  if (b->n == 0) return IZH_OK;
  Map6 m;
  std::memcpy(m.m, xm, sizeof m.m);
  {
    ScopedTimer tm_(IZH_T_PUSH, b->stream);
    hipLaunchKernelGGL(k_map_linear, dim3(grid_for(b->n)), dim3(kBlock), 0,
                       b->stream, b->s[0], b->s[1], b->s[2], b->s[3], b->s[4],
                       b->s[5], b->n, m);
  }
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_conv1st(izh_bunch *b, double xrad, double yrad, double ptrange[6],
                   int64_t *nlost) {
  const double gam = -b->refpt[5];
  const double rad = (xrad + yrad) / 2.0;
  const int nb = grid_for(b->n);
  {
    ScopedTimer tm_(IZH_T_CONV1ST, b->stream);
    hipLaunchKernelGGL(k_conv1st, dim3(nb), dim3(kBlock), 0, b->stream, b->s[0],
                       b->s[1], b->s[2], b->s[3], b->s[4], b->s[5], b->n, gam,
                       b->c.scxl, rad, b->d_part);
    IZH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_conv1st_final, dim3(1), dim3(64), 0, b->stream,
                       b->d_part, nb, b->d_out);
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
  const double gam = -b->refpt[5];
  const double rad = (xrad + yrad) / 2.0;
  const int64_t nchunks = (b->n + kChunk - 1) / kChunk;
  int64_t *cnt = nullptr;
  IZH_HIP(hipMalloc((void **)&cnt, (nchunks + 1) * sizeof(int64_t)));
  hipLaunchKernelGGL(k_compact_count, dim3(nchunks), dim3(kBlock), 0, b->stream,
                     b->s[0], b->s[1], b->s[2], b->s[3], b->s[5], b->n, gam,
                     rad, cnt);
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(64), 0, b->stream, cnt, nchunks);
  Rows9 src, dst;
  for (int r = 0; r < 9; r++) { src.r[r] = b->s[r]; dst.r[r] = b->alt[r]; }
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
  for (int r = 0; r < 9; r++) std::swap(b->s[r], b->alt[r]);
  b->n = nkeep;
  return IZH_OK;
}

int launch_cvbk(izh_bunch *b) {
  if (b->n == 0) return IZH_OK;
  {
    ScopedTimer tm_(IZH_T_CVBK, b->stream);
    hipLaunchKernelGGL(k_cvbk, dim3(grid_for(b->n)), dim3(kBlock), 0, b->stream,
                       b->s[0], b->s[1], b->s[2], b->s[3], b->s[4], b->s[5], b->n,
                       -b->refpt[5], b->c.scxl);
  }
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_gather_kick(izh_bunch *b, izh_field *f, const izh_compdom *dom,
                       double tau, double xycon, double tcon) {
  if (b->n == 0) return IZH_OK;
  Dom d;
  d.xmin = dom->range[0]; d.ymin = dom->range[2]; d.zmin = dom->range[4];
  d.hx = dom->msize[0]; d.hy = dom->msize[1]; d.hz = dom->msize[2];
  d.hxi = 1.0 / d.hx; d.hyi = 1.0 / d.hy; d.hzi = 1.0 / d.hz;
  d.nx = f->nx; d.ny = f->ny; d.nz = f->nz;
  const double gam = -b->refpt[5];
  const double xk = 1.0 / b->c.scxl;
  // the grid holds charge * rho_scale, so the field carries the same factor
  const double txy = tau * xycon / f->rho_scale, ttc = tau * tcon / f->rho_scale;
  {
    ScopedTimer tm_(IZH_T_KICK, b->stream);
    if (f->precision == IZH_FP64) {
      const double *e = (const double *)f->e;
      hipLaunchKernelGGL(k_gather_kick<double>, dim3(grid_for(b->n)), dim3(kBlock),
                         0, b->stream, b->s[0], b->s[1], b->s[2], b->s[3], b->s[4],
                         b->s[5], b->n, e, e + f->g, e + 2 * f->g, d, gam, xk,
                         txy, ttc);
    } else {
      const float *e = (const float *)f->e;
      hipLaunchKernelGGL(k_gather_kick<float>, dim3(grid_for(b->n)), dim3(kBlock),
                         0, b->stream, b->s[0], b->s[1], b->s[2], b->s[3], b->s[4],
                         b->s[5], b->n, e, e + f->g, e + 2 * f->g, d, gam, xk,
                         txy, ttc);
    }
  }
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

int launch_moments(izh_bunch *b, double out[kMoments]) {
  const int nb = grid_for(b->n);
  {
    ScopedTimer tm_(IZH_T_DIAG, b->stream);
    hipLaunchKernelGGL(k_moments, dim3(nb), dim3(kBlock), 0, b->stream, b->s[0],
                       b->s[1], b->s[2], b->s[3], b->s[4], b->s[5], b->n,
                       b->d_part);
    IZH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(64), 0, b->stream, b->d_part,
                       nb, b->d_out);
  }
  IZH_HIP(hipGetLastError());
  IZH_HIP(hipMemcpyAsync(b->h_out, b->d_out, kMoments * sizeof(double),
                         hipMemcpyDeviceToHost, b->stream));
  IZH_HIP(hipStreamSynchronize(b->stream));
  std::memcpy(out, b->h_out, kMoments * sizeof(double));
  return IZH_OK;
}

}  // namespace izh
