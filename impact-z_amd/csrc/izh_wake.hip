// izh_wake.hip -- 1D wakefields of the space-charge step: the line charge
// density and centroids along z (src/Contrl/AccSimulator.f90:1272-1338),
// the short-range wake convolutions of wakefield_FieldQuant and
// wakefieldread_FieldQuant (src/Appl/Field.f90:590-1001), the CSR wake of
// csrwakeTrIGF_FieldQuant (:1007-1251, one workgroup per mesh line), and the wake's
// share of the grid field kick1wake_BeamBunch adds plane by plane before
// scatter1 (src/Appl/BeamBunch.f90:4694-5046).
//
// Sizes: Nz <= kWakeMaxNz mesh lines.  The binning reads the t-frame z (and
// x, y) of every particle once: 3 reals per particle, the only HBM traffic
// worth counting; the convolution is O(Nz^2) in one workgroup, and the field
// update is one read-modify-write of the E grid.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "izh_device.hpp"

namespace izh {

namespace {

__global__ void k_zero_f64(double *__restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// Bins of one pass: rows 0..Nz of unit weight, x-weighted and y-weighted
// sums.  A particle on the last mesh line (iz = Nz, ab = 1) adds its zero
// upper weight to the extra row Nz, where the reference's loop writes it one
// past densz(Nz) (src/Contrl/AccSimulator.f90:1284-1294); the domain's pad
// keeps every particle of a Flagbc=1 step below that line, so the clamp only
// guards memory.
__device__ inline void wake_bin(double x, double y, double z, double zmin, double hz,
                                int nz, double *lds) {
  int iz = (int)((z - zmin) / hz + 1);
  const double ab = ((zmin - z) + iz * hz) / hz;
  iz = min(max(iz, 1), nz);
  const int s = nz + 1;
  atomicAdd(&lds[iz - 1], ab);
  atomicAdd(&lds[iz], 1.0 - ab);
  atomicAdd(&lds[s + iz - 1], x * ab);
  atomicAdd(&lds[s + iz], x * (1.0 - ab));
  atomicAdd(&lds[2 * s + iz - 1], y * ab);
  atomicAdd(&lds[2 * s + iz], y * (1.0 - ab));
}

__device__ inline void wake_flush(const double *lds, int nz, double *bins) {
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * (nz + 1); q += blockDim.x)
    if (lds[q] != 0.0) atomicAdd(&bins[q], lds[q]);
}

// The live rows of a bunch in the t frame after conv1st (x, y, z in m).
__global__ void __launch_bounds__(kBlock)
k_wake_bins_rows(const double *__restrict__ X, const double *__restrict__ Y,
                 const double *__restrict__ Z, int64_t n, const DevDom *__restrict__ dom,
                 int nz, double *__restrict__ bins) {
  extern __shared__ double lds[];
  for (int q = threadIdx.x; q < 3 * (nz + 1); q += blockDim.x) lds[q] = 0.0;
  __syncthreads();
  const double zmin = dom->range[4], hz = (dom->range[5] - dom->range[4]) / (nz - 1);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    wake_bin(X[i], Y[i], Z[i], zmin, hz, nz, lds);
  wake_flush(lds, nz, bins);
}

// The fused run's t positions, chunk by chunk as the deposit reads them
// (lost particles hold NaN).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_wake_bins_chunks(const T *__restrict__ X, const T *__restrict__ Y, const T *__restrict__ Z,
                   const Chunk *__restrict__ chunks, const int64_t *__restrict__ d_nchunks,
                   const DevDom *__restrict__ dom, int nz, double *__restrict__ bins) {
  extern __shared__ double lds[];
  if (blockIdx.x >= *d_nchunks) return;
  for (int q = threadIdx.x; q < 3 * (nz + 1); q += blockDim.x) lds[q] = 0.0;
  __syncthreads();
  const Chunk ch = chunks[blockIdx.x];
  const double zmin = dom->range[4], hz = (dom->range[5] - dom->range[4]) / (nz - 1);
  for (uint32_t t = threadIdx.x; t < ch.len; t += blockDim.x) {
    const int64_t i = (int64_t)ch.start + t;
    const T x = X[i];
    if (!(x == x)) continue;
    wake_bin((double)x, (double)Y[i], (double)Z[i], zmin, hz, nz, lds);
  }
  wake_flush(lds, nz, bins);
}

// A wake function at distance zz behind the source (plane 0 longitudinal,
// 1 horizontal, 2 vertical): Bane's short-range fits with the reference's
// REAL(4) constants (src/Appl/Field.f90:627-642,692-706), or the read-in
// table with the reference's interpolation weights (ab = iwk*hwk - zwk, not
// divided by hwk; :820-830), rows past the table reading 0.
__device__ inline double wake_fn(const WakeParams &p, int plane, double zz) {
  if (p.kind == 1) {
    const double pilc = 2 * asin(1.0), z0 = 120 * pilc;
    const double aa = p.aa, gg = p.gg, leng = p.leng;
    if (plane == 0) {
      const double zz00 = (double)0.41f * pow(aa / leng, (double)0.8f) *
                          pow(gg / leng, (double)1.6f) * aa;
      return z0 * p.clight * exp(-sqrt(zz / zz00)) / (pilc * aa * aa);
    }
    const double zz00 = 1.0 * (double)0.169f * aa * pow(aa / leng, (double)1.17f) *
                        pow(gg / aa, (double)0.38f);
    const double coef1 = 4 * z0 * p.clight * zz00 / (pilc * aa * aa * aa * aa);
    return coef1 * (1.0 - (1. + sqrt(zz / zz00)) * exp(-sqrt(zz / zz00)));
  }
  const double *zt = p.tab, *col = p.tab + (size_t)(plane + 1) * p.ntab;
  const double hwk = zt[1] - zt[0], zwk = zz - zt[0];
  const int iwk = (int)(zwk / hwk + 1), iwk1 = iwk + 1;
  const double ab = iwk * hwk - zwk;
  const double e1 = iwk >= 1 && iwk <= p.ntab ? col[iwk - 1] : 0.0;
  const double e2 = iwk1 >= 1 && iwk1 <= p.ntab ? col[iwk1 - 1] : 0.0;
  return e1 * ab + e2 * (1.0 - ab);
}

// One workgroup: line density and centroids from the bins, the wake
// functions on the Nz+1 distances, the convolutions (for line k: its own
// bin at half weight for the longitudinal wake, and every bin j > k at
// distance (j-k)*hz/gamma, which is the cyclic convolution the reference
// takes by FFT on 2*Nz points), then the field increments of
// kick1wake_BeamBunch.  out: exwake, eywake, ezwake, then dEx, dEy, dEz.
__global__ void __launch_bounds__(kBlock)
k_wake_field(const double *__restrict__ bins, WakeParams p, const DevDom *__restrict__ dom,
             const DevScal *__restrict__ S, double *__restrict__ out) {
  extern __shared__ double lds[];
  const int nz = p.nz, s = nz + 1;
  double *rho = lds, *xb = lds + nz, *yb = lds + 2 * nz, *g = lds + 3 * nz;  // g: 3*(nz+1)
  const double current = S ? S->current : p.current, npt = S ? S->npt : p.npt;
  const double xycon = S ? S->xycon : p.xycon, tcon = S ? S->tcon : p.tcon;
  const double hzw = (dom->range[5] - dom->range[4]) / (nz - 1) / p.gam;
  const double sgn = p.charge / fabs(p.charge);
  for (int k = threadIdx.x; k < nz; k += blockDim.x) {
    const double d = bins[k];
    rho[k] = d * current / p.scfreq / npt / hzw * sgn;
    xb[k] = d != 0.0 ? bins[s + k] / d : 0.0;
    yb[k] = d != 0.0 ? bins[2 * s + k] / d : 0.0;
  }
  for (int q = threadIdx.x; q < 3 * s; q += blockDim.x) {
    const int plane = q / s, m = q - plane * s;
    g[q] = wake_fn(p, p.kind == 1 && plane == 2 ? 1 : plane, m * hzw);
  }
  __syncthreads();
  const bool lon = !(p.kind == 2 && p.flagbtw == 4);
  const bool tra = !(p.kind == 2 && (p.flagbtw == 2 || p.flagbtw == 3));
  const double bet = sqrt(p.gam * p.gam - 1.0) / p.gam;
  for (int k = threadIdx.x; k < nz; k += blockDim.x) {
    double sz = lon ? 0.5 * g[0] * rho[k] : 0.0, sx = 0.0, sy = 0.0;
    for (int j = k + 1; j < nz; j++) {
      const int m = j - k;
      if (lon) sz += g[m] * rho[j];
      if (tra) {
        sx += g[s + m] * rho[j] * xb[j];
        sy += g[2 * s + m] * rho[j] * yb[j];
      }
    }
    const double ez = p.scale * (-sz * hzw), ex = p.scale * (sx * hzw),
                 ey = p.scale * (sy * hzw);
    out[k] = ex;
    out[nz + k] = ey;
    out[2 * nz + k] = ez;
    // divided by the constants scatter1 multiplies the grid field with
    const bool kick = xycon != 0.0 && tcon != 0.0;
    out[3 * nz + k] = kick ? p.charge * ex / xycon / p.mass / bet / p.gam : 0.0;
    out[4 * nz + k] = kick ? p.charge * ey / xycon / p.mass / bet / p.gam : 0.0;
    out[5 * nz + k] = kick ? p.charge * ez / tcon / p.mass : 0.0;
  }
}

// ------------------------------------------------------------- CSR wake
// The integrated CSR kernels of the four regimes of Saldin et al. that
// csrwakeTrIGF_FieldQuant sums (src/Appl/Field.f90:1253-1301): entrance
// transient (A), steady state (B), exit transients with the source inside
// (C) or past (D) the bend, where D's retarded angle is the root of a
// quartic found by Newton-Raphson with a bisection safeguard (root and
// funcd, :1305-1384, with their tolerance and iteration limit).
__device__ inline double csr_a(double phih, double ssh, double xk2) {
  const double p3 = phih * phih * phih, p4 = phih * phih * phih * phih;
  const double bb = 2 * phih + p3 / 3 - 2 * ssh;
  const double cc = phih * phih + p4 / 12 - 2 * ssh * phih;
  const double yh = (-bb + sqrt(bb * bb - 4 * cc)) / 2;
  const double phpy = phih + yh;
  return xk2 * (-(2 * phpy + p3) / (phpy * phpy + p4 / 4) + 1.0 / ssh);
}

__device__ inline double csr_b(double ssh, double xk2) {
  const double aa = sqrt(64.0 + 144.0 * (ssh * ssh));
  const double uuh = pow(aa + 12 * ssh, 1.0 / 3.0) - pow(aa - 12 * ssh, 1.0 / 3.0);
  const double u2 = uuh * uuh;
  return xk2 * (-4 * uuh * (u2 + 8) / ((u2 + 4.0) * (u2 + 12.0)));
}

__device__ inline double csr_c(double phimh, double xxbarh, double ssh, double xk2) {
  const double s = phimh + xxbarh;
  const double bb = 2 * s - 2 * ssh + phimh * phimh * phimh / 3 + phimh * phimh * xxbarh;
  const double cc = s * s + phimh * phimh * (phimh * phimh + 4 * phimh * xxbarh) / 12 -
                    2 * ssh * s;
  const double yh = (-bb + sqrt(bb * bb - 4 * cc)) / 2;
  const double pa = s + yh, pb = phimh * xxbarh + phimh * phimh / 2.0;
  return xk2 * (-2.0 * (pa + phimh * pb) / (pa * pa + pb * pb) + 1.0 / ssh);
}

__device__ inline void csr_funcd(double psi, double xxh, double ds, double &f, double &df) {
  f = psi * psi * psi * psi / 12 + xxh * psi * psi * psi / 3 + psi * psi +
      (2 * xxh - 2 * ds) * psi - 2 * ds * xxh + xxh * xxh;
  df = psi * psi * psi / 3 + xxh * psi * psi + 2 * psi + 2 * xxh - 2 * ds;
}

__device__ inline double csr_root(double x1, double x2, double xacc, int maxit, double zeta,
                                  double xxh) {
  double fl, fh, df, f;
  csr_funcd(x1, xxh, zeta, fl, df);
  csr_funcd(x2, xxh, zeta, fh, df);
  if (fabs(fl) < xacc) return x1;
  if (fabs(fh) < xacc) return x2;
  double xl = fl < 0.0 ? x1 : x2, xh = fl < 0.0 ? x2 : x1;
  double rts = 0.5 * (x1 + x2), dxold = fabs(x2 - x1), dx = dxold;
  csr_funcd(rts, xxh, zeta, f, df);
  for (int j = 1; j <= maxit; j++) {
    if (((rts - xh) * df - f) * ((rts - xl) * df - f) > 0.0 ||
        fabs(2.0 * f) > fabs(dxold * df)) {
      dxold = dx;
      dx = 0.5 * (xh - xl);
      rts = xl + dx;
      if (xl == rts) return rts;
    } else {
      dxold = dx;
      dx = f / df;
      const double temp = rts;
      rts = rts - dx;
      if (temp == rts) return rts;
    }
    if (fabs(dx) < xacc) return rts;
    csr_funcd(rts, xxh, zeta, f, df);
    if (f < 0.0) xl = rts; else xh = rts;
  }
  return rts;
}

constexpr double kCsrTol = 2.0e-9;   // the reference's epstol
constexpr int kCsrMaxIt = 100;       // and Nmax

__device__ inline double csr_d(double xxbarh, double psimax, double ssh, double xk2) {
  const double psi = csr_root(-kCsrTol, psimax * (1.0 + kCsrTol), kCsrTol, kCsrMaxIt, ssh,
                              xxbarh);
  const double pa = psi + xxbarh, pb = psi * (xxbarh + psi / 2.0);
  return xk2 * (-2.0 * (pa + psi * pb) / (pa * pa + pb * pb) + 1.0 / ssh);
}

// One workgroup per mesh line i (1-based i = blockIdx.x + 1): the CSR wake
// of csrwakeTrIGF_FieldQuant at line i, with the interior sums over the
// lines behind it spread over the threads (each term the line density
// times the kernel integrated over that line's cell) and the end-cell
// corrections taken by thread 0; then the kick increment of kick1wake
// (no transverse wake).  The line density is the wake's (the bins times
// Current/Scfreq/Npt/(hz/gamma)); the first line sits range(5)/gamma + zoff
// from the bend entrance (src/Contrl/AccSimulator.f90:1344-1396).
__global__ void __launch_bounds__(kBlock)
k_csr_lines(const double *__restrict__ bins, WakeParams p, const DevDom *__restrict__ dom,
            const DevScal *__restrict__ S, double *__restrict__ out) {
  const int nz = p.nz, i = (int)blockIdx.x + 1, tid = threadIdx.x;
  const double current = S ? S->current : p.current, npt = S ? S->npt : p.npt;
  const double xycon = S ? S->xycon : p.xycon, tcon = S ? S->tcon : p.tcon;
  const double gam = p.gam, r0 = p.r0, blength = p.blength;
  const double hx = (dom->range[5] - dom->range[4]) / (nz - 1) / gam;
  const double ptmin = dom->range[4] / gam + p.zoff;
  const double rscale = current / p.scfreq / npt / hx * (p.charge / fabs(p.charge));
  const double pilc = 2 * asin(1.0), phim = blength / r0;
  const double xconst = 1. / (4 * pilc * 8.854187817e-12);
  const double xk2 = gam / r0, g3 = gam * gam * gam;
  const double xx = ptmin + (i - 1) * hx;
  auto rho = [&](int j) { return bins[j - 1] * rscale; };
  auto ssh = [&](double xp) { return (xx - xp) * g3 / r0; };
  auto edge = [&](int j, double half) { return ptmin + (j - 1) * hx + half; };
  double part = 0.0, ends = 0.0;
  if (i >= 2 && xx > 0) {
    if (xx <= blength) {
      const double xr = xx / r0, xxl = xr * xr * xr * r0 / 24;
      const double xslp = (xx / r0) * r0 / 2 / gam / gam + xxl;
      const int islp1 = (int)((xx - xslp - ptmin) / hx) + 1;
      const double phih = xx / r0 * gam;
      for (int j = 2 + tid; j <= islp1 - 1; j += blockDim.x)
        part += rho(j) * (csr_a(phih, ssh(edge(j, -hx / 2)), xk2) -
                          csr_a(phih, ssh(edge(j, hx / 2)), xk2));
      const int jstart = islp1 > 1 ? islp1 : 1;
      for (int j = jstart + 2 + tid; j <= i - 1; j += blockDim.x)
        part += rho(j) * (csr_b(ssh(edge(j, -hx / 2)), xk2) - csr_b(ssh(edge(j, hx / 2)), xk2));
      if (tid == 0) {
        if (islp1 >= 2) {
          ends += rho(islp1) * (csr_a(phih, ssh(edge(islp1, -hx / 2)), xk2) -
                                csr_a(phih, ssh(xx - xslp), xk2));
          ends += rho(1) * (csr_a(phih, ssh(ptmin), xk2) - csr_a(phih, ssh(ptmin + hx / 2), xk2));
        } else if (islp1 == 1) {
          ends += rho(1) * (csr_a(phih, ssh(ptmin), xk2) - csr_a(phih, ssh(xx - xslp), xk2));
        }
        if (jstart <= i - 2) {
          ends += rho(i) * (csr_b(ssh(edge(i, -hx / 2)), xk2) - csr_b(ssh(edge(i, 0.0)), xk2));
          if (islp1 < 1) {
            ends += rho(1) * (csr_b(ssh(ptmin), xk2) - csr_b(ssh(ptmin + hx / 2), xk2));
          } else {
            const int j = islp1 + 1;
            ends += rho(j) * (csr_b(ssh(xx - xslp), xk2) - csr_b(ssh(edge(j, hx / 2)), xk2));
          }
        } else if (jstart == i - 1) {
          ends += rho(i) * (csr_b(ssh(xx - xslp), xk2) - csr_b(ssh(edge(i, 0.0)), xk2));
        }
      }
    } else {
      const double xxbar = xx - blength;
      const double xslp = (r0 * phim + xxbar) / 2 / gam / gam +
                          r0 * phim * phim * phim / 24 * (r0 * phim + 4 * xxbar) /
                              (r0 * phim + xxbar);
      const int islp1 = (int)((xx - xslp - ptmin) / hx) + 1;
      const double xslpn = xxbar / 2.0 / gam / gam;
      const int islpn1 = (int)((xx - xslpn - ptmin) / hx) + 1;
      const double xxbarh = xxbar * gam / r0, phimh = phim * gam, psimax = phim * gam;
      for (int j = 2 + tid; j <= islp1 - 1; j += blockDim.x)
        part += rho(j) * (csr_c(phimh, xxbarh, ssh(edge(j, -hx / 2)), xk2) -
                          csr_c(phimh, xxbarh, ssh(edge(j, hx / 2)), xk2));
      const int jstart = islp1 > 0 ? islp1 : 0;
      for (int j = jstart + 2 + tid; j <= islpn1 - 1; j += blockDim.x)
        part += rho(j) * (csr_d(xxbarh, psimax, ssh(edge(j, -hx / 2)), xk2) -
                          csr_d(xxbarh, psimax, ssh(edge(j, hx / 2)), xk2));
      if (tid == 0) {
        if (islp1 >= 2) {
          ends += rho(islp1) * (csr_c(phimh, xxbarh, ssh(edge(islp1, -hx / 2)), xk2) -
                                csr_c(phimh, xxbarh, ssh(xx - xslp), xk2));
          ends += rho(1) * (csr_c(phimh, xxbarh, ssh(ptmin), xk2) -
                            csr_c(phimh, xxbarh, ssh(ptmin + hx / 2), xk2));
        } else if (islp1 == 1) {
          ends += rho(1) * (csr_c(phimh, xxbarh, ssh(ptmin), xk2) -
                            csr_c(phimh, xxbarh, ssh(xx - xslp), xk2));
        }
        if (islpn1 >= jstart + 2) {
          ends += rho(islpn1) * (csr_d(xxbarh, psimax, ssh(edge(islpn1, -hx / 2)), xk2) -
                                 csr_d(xxbarh, psimax, xslpn * g3 / r0, xk2));
          ends += rho(jstart + 1) * (csr_d(xxbarh, psimax, xslp * g3 / r0, xk2) -
                                     csr_d(xxbarh, psimax, ssh(ptmin + jstart * hx + hx / 2),
                                           xk2));
        } else if (islpn1 == jstart + 1) {
          ends += rho(jstart + 1) * (csr_d(xxbarh, psimax, xslp * g3 / r0, xk2) -
                                     csr_d(xxbarh, psimax, xslpn * g3 / r0, xk2));
        }
      }
    }
  }
  // fixed-order block sum of the interior terms, then the end terms
  __shared__ double red[kBlock / 64];
  const double w = wave_sum(part);
  if ((tid & 63) == 0) red[tid >> 6] = w;
  __syncthreads();
  if (tid == 0) {
    double e = 0.0;
    for (int k = 0; k < kBlock / 64; k++) e += red[k];
    const double ez = i >= 2 ? (e + ends) * xconst : 0.0;
    const int k = i - 1;
    const bool kick = xycon != 0.0 && tcon != 0.0;
    out[k] = 0.0;
    out[nz + k] = 0.0;
    out[2 * nz + k] = ez;
    out[3 * nz + k] = 0.0;
    out[4 * nz + k] = 0.0;
    out[5 * nz + k] = kick ? p.charge * ez / tcon / p.mass : 0.0;
  }
}

// E of every node of plane k += the plane's wake increment (times the grid
// scaling of the mixed form).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_wake_add(E4<T> *__restrict__ e, int64_t plane, int nz, const double *__restrict__ de,
           double rho_scale) {
  const int64_t n = plane * nz;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(c / plane);
    E4<T> v = e[c];
    v.x = (T)((double)v.x + de[k] * rho_scale);
    v.y = (T)((double)v.y + de[nz + k] * rho_scale);
    v.z = (T)((double)v.z + de[2 * nz + k] * rho_scale);
    e[c] = v;
  }
}

}  // namespace

int wake_alloc(izh_field *f) {
  if (f->d_wake) return IZH_OK;
  const size_t nz = (size_t)f->nz;
  IZH_HIP(hipMalloc((void **)&f->d_wake, (6 * nz + 3 * (nz + 1)) * sizeof(double)));
  return IZH_OK;
}

int wake_table_device(const izh_rfdata *d, const double **out) {
  if (!d->d_table) {
    const size_t n = d->z.size();
    std::vector<double> h(4 * n);
    for (size_t i = 0; i < n; i++) {
      h[i] = d->z[i];
      h[n + i] = d->ez[i];
      h[2 * n + i] = d->ezp[i];
      h[3 * n + i] = d->ezpp[i];
    }
    double *p = nullptr;
    IZH_HIP(hipMalloc((void **)&p, h.size() * sizeof(double)));
    IZH_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    d->d_table = p;
  }
  *out = d->d_table;
  return IZH_OK;
}

int launch_wake(izh_bunch *b, izh_field *f, const WakeParams &p0, const DevScal *S,
                const Chunk *chunks, const int64_t *d_nchunks, int64_t max_chunks) {
  ScopedTimer tm_(IZH_T_WAKE, f->stream);
  IZH_CHECK(wake_alloc(f));
  const int nz = f->nz;
  WakeParams p = p0;
  p.nz = nz;
  double *bins = f->d_wake + 6 * (size_t)nz;
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(kBlock), 0, f->stream, bins,
                     (int64_t)3 * (nz + 1));
  const size_t lds_bins = 3 * (size_t)(nz + 1) * sizeof(double);
  if (chunks) {
    if (f->precision == IZH_MIXED) {
      const float *tp = (const float *)b->tpos;
      hipLaunchKernelGGL(k_wake_bins_chunks<float>, dim3((unsigned)max_chunks), dim3(kBlock),
                         lds_bins, f->stream, tp, tp + b->cap, tp + 2 * b->cap, chunks,
                         d_nchunks, f->d_dom, nz, bins);
    } else {
      const double *tp = (const double *)b->tpos;
      hipLaunchKernelGGL(k_wake_bins_chunks<double>, dim3((unsigned)max_chunks),
                         dim3(kBlock), lds_bins, f->stream, tp, tp + b->cap, tp + 2 * b->cap,
                         chunks, d_nchunks, f->d_dom, nz, bins);
    }
  } else if (b->n > 0) {
    hipLaunchKernelGGL(k_wake_bins_rows, dim3(grid_for(b->n)), dim3(kBlock), lds_bins,
                       f->stream, b->r[0], b->r[2], b->r[4], b->n, f->d_dom, nz, bins);
  }
  if (p.kind == 3) {
    hipLaunchKernelGGL(k_csr_lines, dim3(nz), dim3(kBlock), 0, f->stream, bins, p, f->d_dom,
                       S, f->d_wake);
  } else {
    const size_t lds_field = (3 * (size_t)nz + 3 * (size_t)(nz + 1)) * sizeof(double);
    hipLaunchKernelGGL(k_wake_field, dim3(1), dim3(kBlock), lds_field, f->stream, bins, p,
                       f->d_dom, S, f->d_wake);
  }
  IZH_HIP(hipGetLastError());
  f->have_wake = 1;
  return IZH_OK;
}

int launch_wake_add(izh_field *f) {
  const int64_t plane = (int64_t)f->nx * f->ny;
  const double *de = f->d_wake + 3 * (size_t)f->nz;
  if (f->precision == IZH_FP64)
    hipLaunchKernelGGL(k_wake_add<double>, dim3(grid_for(f->g)), dim3(kBlock), 0, f->stream,
                       (E4<double> *)f->e, plane, f->nz, de, f->rho_scale);
  else
    hipLaunchKernelGGL(k_wake_add<float>, dim3(grid_for(f->g)), dim3(kBlock), 0, f->stream,
                       (E4<float> *)f->e, plane, f->nz, de, f->rho_scale);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

}  // namespace izh
