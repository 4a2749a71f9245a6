// izh_device.hpp -- device helpers shared by the particle kernels of the
// procedure-by-procedure API (izh_particles.hip) and of the fused run
// (izh_fused.hip): wavefront reductions, the diagnostic moment accumulator
// and the 8-node CIC gather.
#pragma once

#include "izh_internal.hpp"

namespace izh {

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

// diagnostic1_Output sums (src/Contrl/Output.f90:119-199) in the reference's
// tmplc order, then |x|,|px|,|y|,|py|,|z|,|pz| maxima and max(x^2+y^2).
struct Mom {
  double s[37];
  double m[7];
};

__device__ inline void mom_zero(Mom &a) {
#pragma unroll
  for (int k = 0; k < 37; k++) a.s[k] = 0.0;
#pragma unroll
  for (int k = 0; k < 7; k++) a.m[k] = 0.0;
}

__device__ inline void mom_add(Mom &a, double x, double px, double y, double py,
                               double z, double pz) {
  double x2 = x * x, px2 = px * px, y2 = y * y, py2 = py * py, z2 = z * z,
         pz2 = pz * pz;
  double *s = a.s;
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
  double *m = a.m;
  m[0] = fmax(m[0], fabs(x)); m[1] = fmax(m[1], fabs(px));
  m[2] = fmax(m[2], fabs(y)); m[3] = fmax(m[3], fabs(py));
  m[4] = fmax(m[4], fabs(z)); m[5] = fmax(m[5], fabs(pz));
  m[6] = fmax(m[6], x2 + y2);
}

// Block-reduce a per-thread accumulator and store kMoments values at out
// (sums, then maxima).  sh: [blockDim/64][kMoments] in LDS.
__device__ inline void mom_block_store(const Mom &a, double (*sh)[kMoments],
                                       double *out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < 37; k++) {
    double v = wave_sum(a.s[k]);
    if (lane == 0) sh[wv][k] = v;
  }
#pragma unroll
  for (int k = 0; k < 7; k++) {
    double v = wave_max(a.m[k]);
    if (lane == 0) sh[wv][37 + k] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < kMoments) {
    int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < nw; w++) r = (k < 37) ? r + sh[w][k] : fmax(r, sh[w][k]);
    out[k] = r;
  }
}

constexpr unsigned long long kMomentsMaxMask =
    ((1ull << kMoments) - 1ull) & ~((1ull << 37) - 1ull);

// 8-node CIC gather of interleaved (Ex, Ey, Ez, 0) at lower-node weights
// ab, cd, ef, summed in the term order of scatter1_BeamBunch
// (src/Appl/BeamBunch.f90:2130-2155).
template <typename T>
__device__ inline void cic8_e4(const E4<T> *__restrict__ e, int64_t o, int64_t sy,
                               int64_t sz, T ab, T cd, T ef, T &gx, T &gy, T &gz) {
  const T one = (T)1;
  const int64_t idx[8] = {o, o + sy, o + sy + sz, o + sz, o + 1 + sz,
                          o + 1 + sy + sz, o + 1 + sy, o + 1};
  const T w[8] = {ab * cd * ef,
                  ab * (one - cd) * ef,
                  ab * (one - cd) * (one - ef),
                  ab * cd * (one - ef),
                  (one - ab) * cd * (one - ef),
                  (one - ab) * (one - cd) * (one - ef),
                  (one - ab) * (one - cd) * ef,
                  (one - ab) * cd * ef};
  T sx = (T)0, syy = (T)0, szz = (T)0;
#pragma unroll
  for (int c = 0; c < 8; c++) {
    E4<T> v = e[idx[c]];
    sx += v.x * w[c];
    syy += v.y * w[c];
    szz += v.z * w[c];
  }
  gx = sx;
  gy = syy;
  gz = szz;
}

// Single-block reduction of per-block partials (nq <= 64 quantities each) in
// a fixed tree order; bit q of min_mask / max_mask selects min / max for
// quantity q, sums otherwise.
__global__ void __launch_bounds__(1024)
k_reduce_partials(const double *__restrict__ part, int nparts, int nq,
                  unsigned long long min_mask, unsigned long long max_mask,
                  double *__restrict__ out);

int launch_reduce(hipStream_t s, const double *part, int nparts, int nq,
                  unsigned long long min_mask, unsigned long long max_mask,
                  double *out);

}  // namespace izh
