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

// Exchange a double with the neighbouring lane of the pair (lane ^ 1) through
// a DPP quad permutation [1,0,3,2]; every lane of the wave must be active.
__device__ inline double swap_lane_pair(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int slo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);
  const int shi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
  return __hiloint2double(shi, slo);
}

// The same diagnostic1_Output sums accumulated by lane pairs, with half the
// registers of Mom: both lanes of a pair see both particles; the even lane
// sums the x plane and the powers of phi, the odd lane the y plane and the
// powers of pt.  Plane sums are u, u^2, u^3, u^4, u pu, u pu^2, u pu^3,
// u^2 pu, u^2 pu^2, u^3 pu, pu, pu^2, pu^3, pu^4 with (u, pu) = (x, px) or
// (y, py), matching the 14 x- and y-plane entries of Mom::s.
struct PairMom {
  double p[14];
  double w[4];   // w, w^2, w^3 (|phi^3| for phi), w^4
  double zp;     // phi*pt of the lane's own particle
  double m[4];   // max |u|, |pu|, |w| over the pair; max x^2+y^2 (own)
};

__device__ inline void pmom_zero(PairMom &a) {
#pragma unroll
  for (int k = 0; k < 14; k++) a.p[k] = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) a.w[k] = 0.0;
  a.zp = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) a.m[k] = 0.0;
}

__device__ inline void pmom_one(PairMom &a, double u, double pu, double w, bool absw) {
  const double u2 = u * u, pu2 = pu * pu, w2 = w * w;
  double *p = a.p;
  p[0] += u; p[1] += u2; p[2] += u2 * u; p[3] += u2 * u2;
  p[4] += u * pu; p[5] += u * pu2; p[6] += u * pu2 * pu;
  p[7] += u2 * pu; p[8] += u2 * pu2; p[9] += u2 * u * pu;
  p[10] += pu; p[11] += pu2; p[12] += pu2 * pu; p[13] += pu2 * pu2;
  const double w3 = w2 * w;
  a.w[0] += w; a.w[1] += w2; a.w[2] += absw ? fabs(w3) : w3; a.w[3] += w2 * w2;
  a.m[0] = fmax(a.m[0], fabs(u));
  a.m[1] = fmax(a.m[1], fabs(pu));
  a.m[2] = fmax(a.m[2], fabs(w));
}

// Add one particle per lane (zeros for a lane without one).  Call with every
// lane of the wave active.
__device__ inline void pmom_add(PairMom &a, double x, double px, double y, double py,
                                double ph, double pt) {
  const bool odd = threadIdx.x & 1;
  const double q0 = swap_lane_pair(odd ? x : y);
  const double q1 = swap_lane_pair(odd ? px : py);
  const double q2 = swap_lane_pair(odd ? ph : pt);
  pmom_one(a, odd ? y : x, odd ? py : px, odd ? pt : ph, !odd);
  pmom_one(a, q0, q1, q2, !odd);
  a.zp += ph * pt;
  a.m[3] = fmax(a.m[3], x * x + y * y);
}

// Block-reduce PairMom into the kMoments values of Mom's layout, value k at
// out[k * ostride].  sh: [blockDim/64][kMoments] in LDS.
__device__ inline void pmom_block_store(const PairMom &a, double (*sh)[kMoments],
                                        double *out, int64_t ostride) {
  constexpr int kx[14] = {0, 6, 15, 16, 12, 27, 28, 29, 30, 31, 1, 7, 17, 18};
  constexpr int ky[14] = {2, 8, 19, 20, 13, 32, 33, 34, 35, 36, 3, 9, 21, 22};
  constexpr int kz[4] = {4, 10, 23, 24}, kp[4] = {5, 11, 25, 26};
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  const bool odd = lane & 1;
#pragma unroll
  for (int k = 0; k < 14; k++) {
    const double e = wave_sum(odd ? 0.0 : a.p[k]);
    const double o = wave_sum(odd ? a.p[k] : 0.0);
    if (lane == 0) { sh[wv][kx[k]] = e; sh[wv][ky[k]] = o; }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const double e = wave_sum(odd ? 0.0 : a.w[k]);
    const double o = wave_sum(odd ? a.w[k] : 0.0);
    if (lane == 0) { sh[wv][kz[k]] = e; sh[wv][kp[k]] = o; }
  }
  {
    const double v = wave_sum(a.zp);
    if (lane == 0) sh[wv][14] = v;
  }
  constexpr int kme[3] = {37, 38, 41}, kmo[3] = {39, 40, 42};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double e = wave_max(odd ? 0.0 : a.m[k]);
    const double o = wave_max(odd ? a.m[k] : 0.0);
    if (lane == 0) { sh[wv][kme[k]] = e; sh[wv][kmo[k]] = o; }
  }
  {
    const double v = wave_max(a.m[3]);
    if (lane == 0) sh[wv][43] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < kMoments) {
    const int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < nw; w++) r = (k < 37) ? r + sh[w][k] : fmax(r, sh[w][k]);
    out[k * ostride] = r;
  }
}

// 8-node CIC gather at lower-node weights ab, cd, ef, summed in the term
// order of scatter1_BeamBunch (src/Appl/BeamBunch.f90:2130-2155).  ld(q, x,
// y, z) loads the field of node q; o is the lower node, sy and sz the node
// strides in y and z.
template <typename T, typename Load>
__device__ inline void cic8(Load ld, int64_t o, int64_t sy, int64_t sz, T ab, T cd,
                            T ef, T &gx, T &gy, T &gz) {
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
    T vx, vy, vz;
    ld(idx[c], vx, vy, vz);
    sx += vx * w[c];
    syy += vy * w[c];
    szz += vz * w[c];
  }
  gx = sx;
  gy = syy;
  gz = szz;
}

// The gather from interleaved (Ex, Ey, Ez, 0) in global memory.
template <typename T>
__device__ inline void cic8_e4(const E4<T> *__restrict__ e, int64_t o, int64_t sy,
                               int64_t sz, T ab, T cd, T ef, T &gx, T &gy, T &gz) {
  cic8<T>([e](int64_t q, T &x, T &y, T &z) {
            const E4<T> v = e[q];
            x = v.x; y = v.y; z = v.z;
          },
          o, sy, sz, ab, cd, ef, gx, gy, gz);
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
