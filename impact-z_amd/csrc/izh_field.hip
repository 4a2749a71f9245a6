// izh_field.hip -- charge deposition, open-boundary Poisson solve and E field.
//
// Deposition: CIC weights as deposit_BeamBunch forms them, accumulated as
// fixed-point integers (scale 2^e from the total |charge|), so the sum is
// exact and independent of the order atomics land in: bitwise
// deterministic.
//
// Poisson: the Hockney doubled-grid convolution of openBC3D
// (src/Appl/Field.f90:239-331), factored into batched 1D FFTs over lines:
//   fwd-x (real, half zero) -> fwd-y (half zero) -> z: FFT * G^ * IFFT fused
//   in LDS, keeping only the first half -> inv-y (keep half) -> inv-x (C2R).
// Only the first octant of the result is ever stored.  G^ of the even
// extended integrated Green function is real and even, so it is built and
// stored on the (nx+1)(ny+1)(nz+1) half-grid and folded on use.
// Each block transforms B lines of length L (B*L = 2048 points) in LDS with
// in-place radix-2 stages: decimation in frequency for forward transforms
// (natural in, bit-reversed out) and decimation in time for the inverse of
// the fused z pass (bit-reversed in, natural out), so no reorder pass is
// needed between them.
#include <cmath>
#include <cstring>
#include <vector>

#include "izh_internal.hpp"

namespace izh {
namespace {

template <typename T>
struct cx {
  T re, im;
};

template <typename T>
__device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

constexpr int kFftThreads = 256;
constexpr int kFftPoints = 2048;  // points per block in LDS

enum Load { kLoadReal = 0, kLoadCplx = 1, kLoadEven = 2, kLoadHerm = 3 };
enum Store { kStoreCplx = 0, kStoreReal = 1 };

struct Lines {
  const void *src;
  void *dst;
  int64_t nf, ns;          // lines = nf * ns, f fastest
  int64_t sf, ss, se;      // source strides (elements)
  int64_t df, ds, de;      // destination strides
  int L, log2L, B, Lin, Lout;
  double out_scale;
  // fused z pass: multiply by the half-grid G^ folded in the line's y index
  const void *green;
  int64_t gkx, gky;        // G^ strides: (nx+1), (nx+1)*(ny+1)
  int my;                  // doubled y length, for folding the line's y index
};

__device__ inline int bitrev(int v, int bits) {
  return (int)(__brev((unsigned)v) >> (32 - bits));
}

// In-place radix-2 decimation in frequency: natural in, bit-reversed out.
template <typename T>
__device__ void fft_dif(cx<T> *buf, const Lines &p, const cx<T> *__restrict__ tw,
                        bool inverse) {
  const int L = p.L, nbf = p.B * (L >> 1);
  for (int half = L >> 1; half >= 1; half >>= 1) {
    const int tstride = L / (2 * half);
    for (int t = threadIdx.x; t < nbf; t += kFftThreads) {
      int lb = t / (L >> 1), j = t - lb * (L >> 1);
      int k = j & (half - 1), g = j / half;
      int i0 = lb * L + g * 2 * half + k, i1 = i0 + half;
      cx<T> w = tw[k * tstride];
      if (inverse) w.im = -w.im;
      cx<T> u = buf[i0], v = buf[i1];
      buf[i0] = {u.re + v.re, u.im + v.im};
      buf[i1] = cmul<T>({u.re - v.re, u.im - v.im}, w);
    }
    __syncthreads();
  }
}

// In-place radix-2 decimation in time: bit-reversed in, natural out.
template <typename T>
__device__ void fft_dit(cx<T> *buf, const Lines &p, const cx<T> *__restrict__ tw,
                        bool inverse) {
  const int L = p.L, nbf = p.B * (L >> 1);
  for (int half = 1; half < L; half <<= 1) {
    const int tstride = L / (2 * half);
    for (int t = threadIdx.x; t < nbf; t += kFftThreads) {
      int lb = t / (L >> 1), j = t - lb * (L >> 1);
      int k = j & (half - 1), g = j / half;
      int i0 = lb * L + g * 2 * half + k, i1 = i0 + half;
      cx<T> w = tw[k * tstride];
      if (inverse) w.im = -w.im;
      cx<T> u = buf[i0], v = cmul<T>(buf[i1], w);
      buf[i0] = {u.re + v.re, u.im + v.im};
      buf[i1] = {u.re - v.re, u.im - v.im};
    }
    __syncthreads();
  }
}

// Thread t of a block -> (line in block, element) with coalesced addresses:
// unit element stride -> consecutive threads walk one line; otherwise they
// walk across neighbouring lines (unit line stride).
__device__ inline void map_thread(int t, int64_t estride, const Lines &p, int &lb,
                                  int &m) {
  if (estride == 1) {
    lb = t >> p.log2L;
    m = t & (p.L - 1);
  } else {
    m = t / p.B;
    lb = t - m * p.B;
  }
}

template <typename T, int LOAD, int STORE, bool CONV>
__global__ void __launch_bounds__(kFftThreads)
k_fft_lines(Lines p, const cx<T> *__restrict__ tw, int inverse) {
  __shared__ cx<T> buf[kFftPoints];
  const int64_t nlines = p.nf * p.ns;
  const int64_t line0 = (int64_t)blockIdx.x * p.B;
  const int tot = p.B * p.L;
  // ---- load (zero padding / even extension / Hermitian completion)
  for (int t = threadIdx.x; t < tot; t += kFftThreads) {
    int lb, m;
    map_thread(t, p.se, p, lb, m);
    int64_t line = line0 + lb;
    cx<T> v = {(T)0, (T)0};
    if (line < nlines) {
      int64_t f = line % p.nf, s = line / p.nf;
      int64_t base = f * p.sf + s * p.ss;
      if (LOAD == kLoadReal) {
        if (m < p.Lin) v.re = ((const T *)p.src)[base + m * p.se];
      } else if (LOAD == kLoadCplx) {
        if (m < p.Lin) v = ((const cx<T> *)p.src)[base + m * p.se];
      } else if (LOAD == kLoadEven) {
        int mm = m <= (p.L >> 1) ? m : p.L - m;
        v.re = ((const T *)p.src)[base + mm * p.se];
      } else {  // kLoadHerm: first L/2+1 values stored
        if (m <= (p.L >> 1)) {
          v = ((const cx<T> *)p.src)[base + m * p.se];
        } else {
          v = ((const cx<T> *)p.src)[base + (p.L - m) * p.se];
          v.im = -v.im;
        }
      }
    }
    buf[lb * p.L + m] = v;
  }
  __syncthreads();
  fft_dif<T>(buf, p, tw, inverse != 0);
  if (CONV) {
    // forward z spectrum is bit-reversed: position q holds frequency bitrev(q)
    const T *G = (const T *)p.green;
    for (int t = threadIdx.x; t < tot; t += kFftThreads) {
      int lb = t >> p.log2L, q = t & (p.L - 1);
      int64_t line = line0 + lb;
      if (line < nlines) {
        int64_t i = line % p.nf, j2 = line / p.nf;
        int64_t jf = j2 <= p.my / 2 ? j2 : p.my - j2;
        int kz = bitrev(q, p.log2L);
        int kf = kz <= (p.L >> 1) ? kz : p.L - kz;
        T g = G[i + p.gkx * jf + p.gky * kf];
        cx<T> v = buf[t];
        buf[t] = {v.re * g, v.im * g};
      }
    }
    __syncthreads();
    fft_dit<T>(buf, p, tw, true);
  }
  // ---- store the kept part of each line
  const int totout = p.B * p.L;
  for (int t = threadIdx.x; t < totout; t += kFftThreads) {
    int lb, m;
    map_thread(t, p.de, p, lb, m);
    int64_t line = line0 + lb;
    if (line >= nlines || m >= p.Lout) continue;
    int64_t f = line % p.nf, s = line / p.nf;
    int64_t o = f * p.df + s * p.ds + m * p.de;
    cx<T> v = CONV ? buf[lb * p.L + m] : buf[lb * p.L + bitrev(m, p.log2L)];
    if (STORE == kStoreCplx) {
      ((cx<T> *)p.dst)[o] = {(T)(v.re * p.out_scale), (T)(v.im * p.out_scale)};
    } else {
      ((T *)p.dst)[o] = (T)(v.re * p.out_scale);
    }
  }
}

// ------------------------------------------------------ Green function
// IGF antiderivative at cell corners (src/Appl/Field.f90:384-418), FP64.
__global__ void k_igf_corner(double *__restrict__ out, int tx, int ty, int tz,
                             double hx, double hy, double hz) {
  int64_t n = (int64_t)tx * ty * tz;
  const double vol = hx * hy * hz;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(c % tx);
    int j = (int)((c / tx) % ty);
    int k = (int)(c / ((int64_t)tx * ty));
    double x = i * hx - hx / 2.0, y = j * hy - hy / 2.0, z = k * hz - hz / 2.0;
    double rr = sqrt(x * x + y * y + z * z);
    double aa = -z * z * atan(x * y / (z * rr)) / 2.0;
    double bb = -y * y * atan(x * z / (y * rr)) / 2.0;
    double cc = -x * x * atan(y * z / (x * rr)) / 2.0;
    double dd = y * z * log(x + rr);
    double ee = x * z * log(y + rr);
    double ff = x * y * log(z + rr);
    out[c] = (aa + bb + cc + dd + ee + ff) / vol;
  }
}

// 8-corner difference (src/Appl/Field.f90:420-428) -> g on (nx+1,ny+1,nz+1).
template <typename T>
__global__ void k_igf_diff(const double *__restrict__ t, T *__restrict__ g,
                           int nx, int ny, int nz) {
  const int tx = nx + 2, ty = ny + 2;
  const int gx = nx + 1, gy = ny + 1, gz = nz + 1;
  int64_t n = (int64_t)gx * gy * gz;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(c % gx);
    int j = (int)((c / gx) % gy);
    int k = (int)(c / ((int64_t)gx * gy));
#define TT(a, b, cc) t[(a) + (int64_t)tx * ((b) + (int64_t)ty * (cc))]
    double v = TT(i + 1, j + 1, k + 1) - TT(i, j + 1, k + 1) - TT(i + 1, j, k + 1) +
               TT(i, j, k + 1) - TT(i + 1, j + 1, k) + TT(i, j + 1, k) +
               TT(i + 1, j, k) - TT(i, j, k);
#undef TT
    g[c] = (T)v;
  }
}

// --------------------------------------------------------------- deposit
struct Dom {
  double xmin, ymin, zmin, hx, hy, hz, hxi, hyi, hzi;
  int nx, ny, nz;
};

__global__ void k_zero_u64(unsigned long long *__restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// deposit_BeamBunch (src/Appl/BeamBunch.f90:3737-3766) in fixed point.
__global__ void __launch_bounds__(kBlock)
k_deposit_fx(const double *__restrict__ X, const double *__restrict__ Y,
             const double *__restrict__ Z, const double *__restrict__ W,
             int64_t n, Dom d, double scale, unsigned long long *__restrict__ rho) {
  const int64_t sy = d.nx, sz = (int64_t)d.nx * d.ny;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double x = X[i], y = Y[i], z = Z[i];
    int ix = (int)((x - d.xmin) * d.hxi + 1.0);
    double ab = ((d.xmin - x) + ix * d.hx) * d.hxi;
    int jx = (int)((y - d.ymin) * d.hyi + 1.0);
    double cd = ((d.ymin - y) + jx * d.hy) * d.hyi;
    int kx = (int)((z - d.zmin) * d.hzi + 1.0);
    double ef = ((d.zmin - z) + kx * d.hz) * d.hzi;
    if (ix < 1 || ix + 1 > d.nx || jx < 1 || jx + 1 > d.ny || kx < 1 ||
        kx + 1 > d.nz)
      continue;
    double w = W[i] * scale;
    int64_t o = (ix - 1) + sy * (jx - 1) + sz * (kx - 1);
    auto add = [&](int64_t idx, double c) {
      atomicAdd(&rho[idx], (unsigned long long)__double2ll_rn(c));
    };
    add(o, ab * cd * ef * w);
    add(o + sy, ab * (1.0 - cd) * ef * w);
    add(o + sy + sz, ab * (1.0 - cd) * (1.0 - ef) * w);
    add(o + sz, ab * cd * (1.0 - ef) * w);
    add(o + 1 + sz, (1.0 - ab) * cd * (1.0 - ef) * w);
    add(o + 1 + sy + sz, (1.0 - ab) * (1.0 - cd) * (1.0 - ef) * w);
    add(o + 1 + sy, (1.0 - ab) * (1.0 - cd) * ef * w);
    add(o + 1, (1.0 - ab) * cd * ef * w);
  }
}

template <typename T>
__global__ void k_fx_to_grid(const unsigned long long *__restrict__ fx,
                             T *__restrict__ rho, int64_t n, double inv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    rho[i] = (T)((double)(long long)fx[i] * inv);
}

// ------------------------------------------------------------- E field
// kick1_BeamBunch one-process Flagbc=1 stencils (src/Appl/BeamBunch.f90:
// 1782-1794, 1885-1899, 1966-2005): central differences, one-sided at edges.
template <typename T>
__global__ void k_efield(const T *__restrict__ phi, T *__restrict__ ex,
                         T *__restrict__ ey, T *__restrict__ ez, int nx, int ny,
                         int nz, double hxi, double hyi, double hzi) {
  const int64_t sy = nx, sz = (int64_t)nx * ny, n = sz * nz;
  const T hx = (T)hxi, hy = (T)hyi, hz = (T)hzi;
  const T hx2 = (T)(0.5 * hxi), hy2 = (T)(0.5 * hyi), hz2 = (T)(0.5 * hzi);
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n;
       c += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(c % nx);
    int j = (int)((c / nx) % ny);
    int k = (int)(c / sz);
    T vx, vy, vz;
    if (i == 0) vx = hx * (phi[c] - phi[c + 1]);
    else if (i == nx - 1) vx = hx * (phi[c - 1] - phi[c]);
    else vx = hx2 * (phi[c - 1] - phi[c + 1]);
    if (j == 0) vy = hy * (phi[c] - phi[c + sy]);
    else if (j == ny - 1) vy = hy * (phi[c - sy] - phi[c]);
    else vy = hy2 * (phi[c - sy] - phi[c + sy]);
    if (k == 0) vz = hz * (phi[c] - phi[c + sz]);
    else if (k == nz - 1) vz = hz * (phi[c - sz] - phi[c]);
    else vz = hz2 * (phi[c - sz] - phi[c + sz]);
    ex[c] = vx;
    ey[c] = vy;
    ez[c] = vz;
  }
}

int ilog2(int v) {
  int r = 0;
  while ((1 << r) < v) r++;
  return r;
}

template <typename T, int LOAD, int STORE, bool CONV>
int run_lines(izh_field *f, Lines p, const void *tw, bool inverse) {
  p.log2L = ilog2(p.L);
  p.B = kFftPoints / p.L;
  int64_t nlines = p.nf * p.ns;
  int64_t nblk = (nlines + p.B - 1) / p.B;
  hipLaunchKernelGGL((k_fft_lines<T, LOAD, STORE, CONV>), dim3((unsigned)nblk),
                     dim3(kFftThreads), 0, f->stream, p, (const cx<T> *)tw,
                     inverse ? 1 : 0);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

template <typename T>
int solve_t(izh_field *f, const izh_compdom *dom) {
  const int nx = f->nx, ny = f->ny, nz = f->nz;
  const int mx = 2 * nx, my = 2 * ny, mz = 2 * nz, kx = nx + 1;
  const double hx = dom->msize[0], hy = dom->msize[1], hz = dom->msize[2];
  // ---- G^: IGF corners, 8-corner difference, 3 even real transforms
  {
    ScopedTimer tm_(IZH_T_GREEN, f->stream);
    int64_t nc = (int64_t)(nx + 2) * (ny + 2) * (nz + 2);
    hipLaunchKernelGGL(k_igf_corner, dim3(grid_for(nc)), dim3(kBlock), 0,
                       f->stream, f->corner, nx + 2, ny + 2, nz + 2, hx, hy, hz);
    int64_t ng = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    hipLaunchKernelGGL(k_igf_diff<T>, dim3(grid_for(ng)), dim3(kBlock), 0,
                       f->stream, f->corner, (T *)f->green_t, nx, ny, nz);
    IZH_HIP(hipGetLastError());
    Lines p{};
    // x: lines (j<=ny, k<=nz), contiguous; green_t -> green
    p.src = f->green_t; p.dst = f->green;
    p.nf = ny + 1; p.ns = nz + 1;
    p.sf = kx; p.ss = (int64_t)kx * (ny + 1); p.se = 1;
    p.df = kx; p.ds = (int64_t)kx * (ny + 1); p.de = 1;
    p.L = mx; p.Lin = p.L; p.Lout = nx + 1; p.out_scale = 1.0;
    IZH_CHECK((run_lines<T, kLoadEven, kStoreReal, false>(f, p, f->tw_x, false)));
    // y: lines (i<=nx, k<=nz); green -> green_t
    p.src = f->green; p.dst = f->green_t;
    p.nf = kx; p.ns = nz + 1;
    p.sf = 1; p.ss = (int64_t)kx * (ny + 1); p.se = kx;
    p.df = 1; p.ds = (int64_t)kx * (ny + 1); p.de = kx;
    p.L = my; p.Lin = p.L; p.Lout = ny + 1;
    IZH_CHECK((run_lines<T, kLoadEven, kStoreReal, false>(f, p, f->tw_y, false)));
    // z: lines (i<=nx, j<=ny); green_t -> green, scaled by 1/(mx my mz)
    p.src = f->green_t; p.dst = f->green;
    p.nf = kx; p.ns = ny + 1;
    p.sf = 1; p.ss = kx; p.se = (int64_t)kx * (ny + 1);
    p.df = 1; p.ds = kx; p.de = (int64_t)kx * (ny + 1);
    p.L = mz; p.Lin = p.L; p.Lout = nz + 1;
    p.out_scale = 1.0 / ((double)mx * my * mz);
    IZH_CHECK((run_lines<T, kLoadEven, kStoreReal, false>(f, p, f->tw_z, false)));
  }
  // ---- forward x: rho (nx,ny,nz) real -> A (kx, ny, nz)
  ScopedTimer tm_(IZH_T_FFT, f->stream);
  Lines p{};
  p.src = f->rho; p.dst = f->spec_a;
  p.nf = ny; p.ns = nz;
  p.sf = nx; p.ss = (int64_t)nx * ny; p.se = 1;
  p.df = kx; p.ds = (int64_t)kx * ny; p.de = 1;
  p.L = mx; p.Lin = nx; p.Lout = kx; p.out_scale = 1.0;
  IZH_CHECK((run_lines<T, kLoadReal, kStoreCplx, false>(f, p, f->tw_x, false)));
  // ---- forward y: A -> B (kx, my, nz)
  p.src = f->spec_a; p.dst = f->spec_b;
  p.nf = kx; p.ns = nz;
  p.sf = 1; p.ss = (int64_t)kx * ny; p.se = kx;
  p.df = 1; p.ds = (int64_t)kx * my; p.de = kx;
  p.L = my; p.Lin = ny; p.Lout = my;
  IZH_CHECK((run_lines<T, kLoadCplx, kStoreCplx, false>(f, p, f->tw_y, false)));
  // ---- z: FFT, * G^, IFFT, keep k < nz; in place in B
  p.src = f->spec_b; p.dst = f->spec_b;
  p.nf = kx; p.ns = my;
  p.sf = 1; p.ss = kx; p.se = (int64_t)kx * my;
  p.df = 1; p.ds = kx; p.de = (int64_t)kx * my;
  p.L = mz; p.Lin = nz; p.Lout = nz;
  p.green = f->green; p.gkx = kx; p.gky = (int64_t)kx * (ny + 1); p.my = my;
  IZH_CHECK((run_lines<T, kLoadCplx, kStoreCplx, true>(f, p, f->tw_z, false)));
  // ---- inverse y: B -> A, keep j < ny
  p.green = nullptr;
  p.src = f->spec_b; p.dst = f->spec_a;
  p.nf = kx; p.ns = nz;
  p.sf = 1; p.ss = (int64_t)kx * my; p.se = kx;
  p.df = 1; p.ds = (int64_t)kx * ny; p.de = kx;
  p.L = my; p.Lin = my; p.Lout = ny;
  IZH_CHECK((run_lines<T, kLoadCplx, kStoreCplx, false>(f, p, f->tw_y, true)));
  // ---- inverse x (C2R): A -> phi, keep i < nx
  p.src = f->spec_a; p.dst = f->phi;
  p.nf = ny; p.ns = nz;
  p.sf = kx; p.ss = (int64_t)kx * ny; p.se = 1;
  p.df = nx; p.ds = (int64_t)nx * ny; p.de = 1;
  p.L = mx; p.Lin = mx; p.Lout = nx;
  IZH_CHECK((run_lines<T, kLoadHerm, kStoreReal, false>(f, p, f->tw_x, true)));
  return IZH_OK;
}

template <typename T>
void make_twiddles(std::vector<cx<T>> &tw, int L) {
  tw.resize(L / 2);
  for (int t = 0; t < L / 2; t++) {
    long double a = -2.0L * 3.141592653589793238462643383279502884L * t / L;
    tw[t] = {(T)cosl(a), (T)sinl(a)};
  }
}

template <typename T>
int upload_twiddles(izh_field *f) {
  std::vector<cx<T>> tw;
  void **dst[3] = {&f->tw_x, &f->tw_y, &f->tw_z};
  int len[3] = {2 * f->nx, 2 * f->ny, 2 * f->nz};
  for (int a = 0; a < 3; a++) {
    make_twiddles<T>(tw, len[a]);
    IZH_HIP(hipMalloc(dst[a], tw.size() * sizeof(cx<T>)));
    IZH_HIP(hipMemcpy(*dst[a], tw.data(), tw.size() * sizeof(cx<T>),
                      hipMemcpyHostToDevice));
  }
  return IZH_OK;
}

}  // namespace

// ================================================================ launchers

int field_alloc(izh_field *f) {
  const size_t es = f->precision == IZH_FP64 ? sizeof(double) : sizeof(float);
  const int64_t g = f->g, kx = f->nx + 1;
  const int64_t ghalf = kx * (f->ny + 1) * (f->nz + 1);
  IZH_HIP(hipMalloc((void **)&f->rho_fx, g * sizeof(unsigned long long)));
  IZH_HIP(hipMalloc(&f->rho, g * es));
  IZH_HIP(hipMalloc(&f->phi, g * es));
  IZH_HIP(hipMalloc(&f->e, 3 * g * es));
  IZH_HIP(hipMalloc(&f->spec_a, kx * f->ny * f->nz * 2 * es));
  IZH_HIP(hipMalloc(&f->spec_b, kx * 2 * f->ny * f->nz * 2 * es));
  IZH_HIP(hipMalloc(&f->green, ghalf * es));
  IZH_HIP(hipMalloc(&f->green_t, ghalf * es));
  IZH_HIP(hipMalloc((void **)&f->corner,
                    (size_t)(f->nx + 2) * (f->ny + 2) * (f->nz + 2) * sizeof(double)));
  if (f->precision == IZH_FP64) return upload_twiddles<double>(f);
  return upload_twiddles<float>(f);
}

int launch_deposit(izh_bunch *b, const izh_compdom *dom, izh_field *f) {
  Dom d;
  d.xmin = dom->range[0]; d.ymin = dom->range[2]; d.zmin = dom->range[4];
  d.hx = dom->msize[0]; d.hy = dom->msize[1]; d.hz = dom->msize[2];
  d.hxi = 1.0 / d.hx; d.hyi = 1.0 / d.hy; d.hzi = 1.0 / d.hz;
  d.nx = f->nx; d.ny = f->ny; d.nz = f->nz;
  // every cell holds at most sum|w| (CIC weights of a particle sum to one),
  // so 2^e * sum|w| < 2^61 leaves headroom for the +-0.5 rounding per term
  double wsum = b->wabs_sum > 0 ? b->wabs_sum : 1.0;
  int e = 61 - (int)std::ceil(std::log2(wsum));
  f->fixed_scale = std::ldexp(1.0, e);
  hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(f->g)), dim3(kBlock), 0, f->stream,
                     f->rho_fx, f->g);
  {
    ScopedTimer tm_(IZH_T_DEPOSIT, f->stream);
    if (b->n > 0)
      hipLaunchKernelGGL(k_deposit_fx, dim3(grid_for(b->n)), dim3(kBlock), 0,
                         f->stream, b->s[0], b->s[2], b->s[4], b->s[7], b->n, d,
                         f->fixed_scale, f->rho_fx);
  }
  // FP64: the grid holds charge itself; mixed: charge / sum|w| (O(1) values)
  f->rho_scale = f->precision == IZH_FP64 ? 1.0 : 1.0 / wsum;
  const double inv = f->rho_scale / f->fixed_scale;
  if (f->precision == IZH_FP64)
    hipLaunchKernelGGL(k_fx_to_grid<double>, dim3(grid_for(f->g)), dim3(kBlock), 0,
                       f->stream, f->rho_fx, (double *)f->rho, f->g, inv);
  else
    hipLaunchKernelGGL(k_fx_to_grid<float>, dim3(grid_for(f->g)), dim3(kBlock), 0,
                       f->stream, f->rho_fx, (float *)f->rho, f->g, inv);
  IZH_HIP(hipGetLastError());
  f->have_rho = 1;
  f->have_phi = 0;
  return IZH_OK;
}

int launch_solve(izh_field *f, const izh_compdom *dom) {
  int rc = f->precision == IZH_FP64 ? solve_t<double>(f, dom) : solve_t<float>(f, dom);
  if (rc == IZH_OK) f->have_phi = 1;
  return rc;
}

int launch_efield(izh_field *f, const izh_compdom *dom) {
  const double hxi = 1.0 / dom->msize[0], hyi = 1.0 / dom->msize[1],
               hzi = 1.0 / dom->msize[2];
  {
    ScopedTimer tm_(IZH_T_EFIELD, f->stream);
    if (f->precision == IZH_FP64) {
      double *e = (double *)f->e;
      hipLaunchKernelGGL(k_efield<double>, dim3(grid_for(f->g)), dim3(kBlock), 0,
                         f->stream, (const double *)f->phi, e, e + f->g,
                         e + 2 * f->g, f->nx, f->ny, f->nz, hxi, hyi, hzi);
    } else {
      float *e = (float *)f->e;
      hipLaunchKernelGGL(k_efield<float>, dim3(grid_for(f->g)), dim3(kBlock), 0,
                         f->stream, (const float *)f->phi, e, e + f->g,
                         e + 2 * f->g, f->nx, f->ny, f->nz, hxi, hyi, hzi);
    }
  }
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

}  // namespace izh
