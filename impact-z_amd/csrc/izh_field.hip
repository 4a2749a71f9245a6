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
// Each block transforms B lines of length L (B*L = 2048 points, 4096 in the
// fused z pass) in LDS with in-place radix-2 stages: decimation in
// frequency for forward transforms (natural in, bit-reversed out) and
// decimation in time for the inverse of the fused z pass (bit-reversed in,
// natural out), so no reorder pass is needed between them.
#include <climits>
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

// Threads per line-kernel block: 512 puts four waves per SIMD on a CU that
// holds two 4096-point z-pass blocks in LDS (256 left two, and the solve
// 8% slower).
constexpr int kFftThreads = 512;
constexpr int kFftPoints = 2048;  // points per block in LDS
constexpr int kFftPointsZ = 4096; // the fused z pass: 8 lines of 512 at 128x128x256
constexpr int kFftMinLine = 8;    // shortest line (a mesh of 4 nodes doubled)

enum Load { kLoadReal = 0, kLoadCplx = 1, kLoadEven = 2, kLoadHerm = 3 };
enum Store { kStoreCplx = 0, kStoreReal = 1 };

struct Lines {
  const void *src;
  void *dst;
  int64_t nf, ns;          // lines = nf * ns, f fastest
  int64_t nfv;             // lines with f >= nfv are padding: not read or written
  int64_t sf, ss, se;      // source strides (elements)
  int64_t df, ds, de;      // destination strides
  int L, log2L, B, log2B, Lin, Lout;
  uint64_t nf_mul;         // ceil(2^40 / nf): line / nf by a multiply
  double out_scale;
  // fused z pass: multiply by the half-grid G^ folded in the line's y index
  const void *green;
  int64_t gkx, gky;        // G^ strides of ky and kz: ldx*(nz+1), ldx
  int my;                  // doubled y length, for folding the line's y index
};

// Logical block of physical block p of n: the hardware deals blocks to the
// 8 XCDs in turn (block p to XCD p mod 8), each XCD with its own L2.  XCD x
// takes the x-th contiguous run of logical blocks, one after another, so
// blocks whose lines share 128-byte lines of a spectrum or of G^ (FP32 rows,
// G^'s runs of 8 values) read them through one L2.  A bijection of [0, n)
// for any n (the first n mod 8 XCDs take one block more).
__device__ inline int xcd_block(int p, int n) {
  const int x = p & 7, base = n >> 3, extra = n & 7;
  return x * base + min(x, extra) + (p >> 3);
}

__device__ inline int bitrev(int v, int bits) {
  return (int)(__brev((unsigned)v) >> (32 - bits));
}

// In-place radix-2 decimation in frequency (natural in, bit-reversed out) and
// decimation in time (bit-reversed in, natural out) over the B lines of a
// block.  Up to three consecutive radix-2 stages run in registers on groups
// of 8 elements: the elements a run of stages combines form closed groups
// (stride = the run's smallest half), so each group is read from LDS once per
// three stages.  Every element sees the same butterflies with the same
// twiddles, in the same order, as the stage-by-stage transform.  tw: L/2
// twiddles exp(-2 pi i k / L) in LDS.
// Position of point m in its line's LDS row.  Without it the radix-8 groups
// of the later runs put the lanes of a wavefront on a few banks (group
// stride 1: 16-byte points 128 bytes apart); XOR-ing the low 4 bits with
// m >> 3 spreads every access pattern of the kernel over all banks.  mask
// is 15 for lines of 64 points or more, else 0 (short lines do no better).
__device__ inline int lds_pos(int m, int mask) { return m ^ ((m >> 3) & mask); }

// The K stages of one run on a group held in registers: v[j] is position
// m0 + j * 2^ls of its line, r = m0 mod 2^ls.
template <typename T, bool DIT, int K>
__device__ __attribute__((always_inline)) inline void run_regs(cx<T> (&v)[1 << K], int ls,
                                                               int r, int log2L,
                                                               const cx<T> *tw, bool inverse) {
  constexpr int span = 1 << K;
#pragma unroll
  for (int st = 0; st < K; st++) {
    const int e = DIT ? st : K - 1 - st;  // pair distance 2^e in the group
    const int lt = log2L - 1 - ls - e;    // log2 of the twiddle stride
#pragma unroll
    for (int j = 0; j < span; j++) {
      if ((j >> e) & 1) continue;
      const int k = r + ((j & ((1 << e) - 1)) << ls);
      cx<T> w = tw[k << lt];
      if (inverse) w.im = -w.im;
      const cx<T> u = v[j], x = v[j + (1 << e)];
      if (DIT) {
        const cx<T> y = cmul<T>(x, w);
        v[j] = {u.re + y.re, u.im + y.im};
        v[j + (1 << e)] = {u.re - y.re, u.im - y.im};
      } else {
        v[j] = {u.re + x.re, u.im + x.im};
        v[j + (1 << e)] = cmul<T>({u.re - x.re, u.im - x.im}, w);
      }
    }
  }
}

template <typename T, bool DIT, int K>
__device__ inline void fft_run(cx<T> *buf, int log2L, int LP, int lines, const cx<T> *tw,
                               bool inverse, int done, int mask) {
  constexpr int span = 1 << K;
  // DIF runs halves L/2 .. 1 downwards, DIT 1 .. L/2 upwards; every size
  // here is a power of two, so the group index splits with shifts and masks
  const int ls = DIT ? done : log2L - done - K;  // log2 of the group stride
  const int s = 1 << ls;
  const int lpl = log2L - K;                     // log2 of the groups per line
  const int ngroups = lines << lpl;
  for (int g = threadIdx.x; g < ngroups; g += kFftThreads) {
    const int lb = g >> lpl, q = g & ((1 << lpl) - 1);
    const int blk = q >> ls, r = q & (s - 1);
    const int row = lb * LP, m0 = (blk << (ls + K)) + r;
    cx<T> v[span];
#pragma unroll
    for (int j = 0; j < span; j++) v[j] = buf[row + lds_pos(m0 + j * s, mask)];
    run_regs<T, DIT, K>(v, ls, r, log2L, tw, inverse);
#pragma unroll
    for (int j = 0; j < span; j++) buf[row + lds_pos(m0 + j * s, mask)] = v[j];
  }
  __syncthreads();
}

// The fused z pass's middle: the forward transform's last run (stages
// log2L-3 .. log2L-1, groups of 8 neighbouring positions) and the inverse
// transform's first run (stages 0 .. 2, the same groups) with the multiply by
// G^ between them, in registers: two LDS round trips and two barriers fewer
// than running them apart.  Thread t takes the group of positions 8q .. 8q+7
// of line lb, lb = t mod B and q = t / B (lanes walk across the block's lines,
// so their G^ values, gv[j] for position 8q + j, were read in runs of
// neighbouring x lines); one group per thread (PTS / 8 == kFftThreads).
template <typename T>
__device__ __attribute__((always_inline)) inline void fft_conv_middle(
    cx<T> *buf, int LP, int log2L, int log2B, const cx<T> *tw, bool inverse,
    const T (&gv)[8], int mask) {
  const int lb = threadIdx.x & ((1 << log2B) - 1), q = threadIdx.x >> log2B;
  const int row = lb * LP, m0 = q << 3;
  cx<T> v[8];
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = buf[row + lds_pos(m0 + j, mask)];
  run_regs<T, false, 3>(v, 0, 0, log2L, tw, inverse);
#pragma unroll
  for (int j = 0; j < 8; j++) v[j] = {v[j].re * gv[j], v[j].im * gv[j]};
  run_regs<T, true, 3>(v, 0, 0, log2L, tw, true);
#pragma unroll
  for (int j = 0; j < 8; j++) buf[row + lds_pos(m0 + j, mask)] = v[j];
  __syncthreads();
}

// Stages [from, to) of the transform (all of them by default), in runs of
// up to three.
template <typename T, bool DIT>
__device__ __attribute__((always_inline)) inline void fft_stages(cx<T> *buf, int LP, int log2L, int lines, const cx<T> *tw,
                           bool inverse, int mask, int from = 0, int to = -1) {
  if (to < 0) to = log2L;
  int done = from;
  while (done < to) {
    const int k = min(3, to - done);
    if (k == 3)
      fft_run<T, DIT, 3>(buf, log2L, LP, lines, tw, inverse, done, mask);
    else if (k == 2)
      fft_run<T, DIT, 2>(buf, log2L, LP, lines, tw, inverse, done, mask);
    else
      fft_run<T, DIT, 1>(buf, log2L, LP, lines, tw, inverse, done, mask);
    done += k;
  }
}

// line = f + nf * s for f < nf: s by a multiply with nf_mul = ceil(2^40 / nf),
// exact for line < 2^24 and nf < 2^16 (run_lines checks both).
__device__ inline void line_fs(int line, const Lines &p, int &f, int &s) {
  s = (int)(((uint64_t)(uint32_t)line * p.nf_mul) >> 40);
  f = line - s * (int)p.nf;
}

// Thread t of a block -> (line in block, element) with coalesced addresses:
// unit element stride -> consecutive threads walk one line; otherwise they
// walk across neighbouring lines (unit line stride).
__device__ inline void map_thread(int t, int64_t estride, int log2L, int log2B, int &lb,
                                  int &m) {
  if (estride == 1) {
    lb = t >> log2L;
    m = t & ((1 << log2L) - 1);
  } else {
    m = t >> log2B;
    lb = t & ((1 << log2B) - 1);
  }
}

// PAIR: two real lines share one complex transform (line 2j in the real
// part, line 2j+1 in the imaginary part).  Real even input (kLoadEven) gives
// a real even spectrum each, so the parts separate directly; real input
// (kLoadReal) is separated from Z(k) and conj Z(L-k) when stored; two
// Hermitian spectra (kLoadHerm) are combined as X + iY, whose inverse has
// the two real outputs as its parts.
// PTS points of B = PTS/L lines per block; each line is held in LDS at a
// pitch of L+1 points, so lanes that walk across lines (a line stride
// other than 1, or the fused z pass's multiply) do not all hit one bank.
// LL: log2 of the line length when it is fixed at compile time (the common
// lengths, so every index split and LDS position folds), 0 to take it from p.
template <typename T, int LOAD, int STORE, bool CONV, bool PAIR, int PTS, int LL>
__global__ void __launch_bounds__(kFftThreads)
k_fft_lines(Lines p, const cx<T> *__restrict__ twg, int inverse) {
  __shared__ cx<T> buf[PTS + PTS / kFftMinLine];
  static_assert(PTS == 2048 || PTS == 4096, "points per block");
  constexpr int kLog2Pts = PTS == 4096 ? 12 : 11;
  const int log2L = LL ? LL : p.log2L;
  const int L = 1 << log2L;
  const int log2B = kLog2Pts - log2L;
  const int B = 1 << log2B;
  const int LP = L + 1;
  const int mask = L >= 64 ? 15 : 0;
  extern __shared__ __align__(16) unsigned char fft_dyn[];  // L/2 twiddles
  cx<T> *tw = reinterpret_cast<cx<T> *>(fft_dyn);
  // twiddles are written to LDS together with the line data below
  constexpr int kTwPer = PTS / 2 / kFftThreads;  // L/2 <= PTS/2
  cx<T> twv[kTwPer];
#pragma unroll
  for (int it = 0; it < kTwPer; it++) {
    const int t = threadIdx.x + it * kFftThreads;
    T re = (T)0, im = (T)0;
    if (t < (L >> 1)) {
      const cx<T> v = twg[t];
      re = v.re; im = v.im;
    }
    twv[it] = {re, im};
  }
  const int nf = (int)p.nf;
  const int nlines = nf * (int)p.ns;
  const int line0 = xcd_block((int)blockIdx.x, (int)gridDim.x) * B;
  // fused z pass: fetch the G^ values of the thread's group of the fused
  // middle run first, so their latency hides behind the forward transform.
  // Lanes walk across the block's lines first (neighbouring x indices,
  // neighbouring addresses of G^); the forward spectrum is bit-reversed, so
  // position m of a line holds natural frequency kz = bitrev(m).
  static_assert(!CONV || PTS / 8 == kFftThreads, "one middle-run group per thread");
  T gv[8];
  if (CONV) {
    const T *G = (const T *)p.green;
    const int lb = threadIdx.x & (B - 1), q = threadIdx.x >> log2B;
    const int line = line0 + lb;
    int i, j2;
    line_fs(line, p, i, j2);
    const bool live = line < nlines && i < p.nfv;
    const int jf = j2 <= p.my / 2 ? j2 : p.my - j2;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int kz = bitrev((q << 3) + j, log2L);
      const int kf = kz <= (L >> 1) ? kz : L - kz;
      T g = (T)0;
      if (live) g = G[i + p.gkx * jf + p.gky * kf];
      gv[j] = g;
    }
  }
  // Complex lines walked across (unit line stride, element stride other
  // than one: the y passes and the fused z pass) skip the load's LDS round
  // trip: thread t loads the 8 points of its group of the first radix-8
  // run, positions q + j L/8 of line lb (lb = t mod B, q = t / B: lanes walk
  // across the block's neighbouring lines, so each run of B points is one
  // 128-byte line), straight from global memory and runs those three
  // stages in registers before its first LDS write.
  // The y passes (complex lines in and out, both walked across) run as a
  // decimation in time: the first radix-8 run takes its groups of
  // bit-reversed positions 8q .. 8q+7, points bitrev(q) + bitrev3(j) L/8,
  // straight from global memory, and the last leaves natural positions
  // q + j L/8, stored from registers, so the data passes through LDS only
  // between runs.
  constexpr bool kFuseLoad = !PAIR && LOAD == kLoadCplx;
  const bool fuse = kFuseLoad && p.se != 1 && p.sf == 1 && log2L >= 6;
  const bool dit = fuse && !CONV && STORE == kStoreCplx && p.df == 1;
  int first = 0;  // stages run before the data reaches LDS
  if (fuse) {
    const int lb = threadIdx.x & (B - 1), q = threadIdx.x >> log2B;
    const bool act = threadIdx.x < (PTS >> 3);  // one group per thread
    const int line = line0 + lb;
    int f = 0, sl = 0;
    line_fs(line, p, f, sl);
    const bool live = act && line < nlines && f < p.nfv;
    const int64_t base = f * p.sf + sl * p.ss;
    const int s8 = L >> 3;
    const int qr = dit ? bitrev(q, log2L - 3) : q;
    cx<T> v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int m = qr + (dit ? bitrev(j, 3) : j) * s8;
      T re = (T)0, im = (T)0;
      if (live && m < p.Lin) {
        const cx<T> x = ((const cx<T> *)p.src)[base + m * p.se];
        re = x.re; im = x.im;
      }
      v[j] = {re, im};
    }
#pragma unroll
    for (int it = 0; it < kTwPer; it++) {
      const int t = threadIdx.x + it * kFftThreads;
      if (t < (L >> 1)) tw[t] = twv[it];
    }
    __syncthreads();
    if (act && dit) {
      run_regs<T, true, 3>(v, 0, 0, log2L, tw, inverse != 0);
#pragma unroll
      for (int j = 0; j < 8; j++) buf[lb * LP + lds_pos((q << 3) + j, mask)] = v[j];
    } else if (act) {
      run_regs<T, false, 3>(v, log2L - 3, q, log2L, tw, inverse != 0);
#pragma unroll
      for (int j = 0; j < 8; j++) buf[lb * LP + lds_pos(q + j * s8, mask)] = v[j];
    }
    __syncthreads();
    first = 3;
  } else {
    // ---- load (zero padding / even extension / Hermitian completion): all of
    // a thread's loads are issued before the first LDS write
    constexpr int kPer = PTS / kFftThreads;  // B * L == PTS
    cx<T> ld[kPer];
    int slot[kPer];
#pragma unroll
    for (int it = 0; it < kPer; it++) {
      const int t = threadIdx.x + it * kFftThreads;
      int lb, m;
      map_thread(t, p.se, log2L, log2B, lb, m);
      const int line = line0 + lb;
      slot[it] = lb * LP + lds_pos(m, mask);
      T re = (T)0, im = (T)0;
      if (PAIR) {
        const int l0 = 2 * line, l1 = l0 + 1;
        const T *src = (const T *)p.src;
        const cx<T> *csrc = (const cx<T> *)p.src;
        if (l0 < nlines) {
          int f0, s0, f1, s1;
          line_fs(l0, p, f0, s0);
          line_fs(l1, p, f1, s1);
          const int64_t b0 = f0 * p.sf + s0 * p.ss;
          const bool has1 = l1 < nlines;
          const int64_t b1 = has1 ? f1 * p.sf + s1 * p.ss : b0;
          if (LOAD == kLoadEven) {
            const int mm = m <= (L >> 1) ? m : L - m;
            re = src[b0 + mm * p.se];
            im = has1 ? src[b1 + mm * p.se] : (T)0;
          } else if (LOAD == kLoadReal) {
            if (m < p.Lin) {
              re = src[b0 + m * p.se];
              im = has1 ? src[b1 + m * p.se] : (T)0;
            }
          } else {  // kLoadHerm: Z = X + iY of the completed spectra
            const bool lo = m <= (L >> 1);
            const int mm = lo ? m : L - m;
            cx<T> x = csrc[b0 + mm * p.se];
            cx<T> y = has1 ? csrc[b1 + mm * p.se] : cx<T>{(T)0, (T)0};
            if (!lo) { x.im = -x.im; y.im = -y.im; }
            re = x.re - y.im;
            im = x.im + y.re;
          }
        }
      } else if (line < nlines) {
        int f, s;
        line_fs(line, p, f, s);
        const int64_t base = f * p.sf + s * p.ss;
        if (f >= p.nfv) {
          // a padding line: zeros
        } else if (LOAD == kLoadReal) {
          if (m < p.Lin) re = ((const T *)p.src)[base + m * p.se];
        } else if (LOAD == kLoadCplx) {
          if (m < p.Lin) {
            const cx<T> v = ((const cx<T> *)p.src)[base + m * p.se];
            re = v.re; im = v.im;
          }
        } else if (LOAD == kLoadEven) {
          const int mm = m <= (L >> 1) ? m : L - m;
          re = ((const T *)p.src)[base + mm * p.se];
        } else {  // kLoadHerm: first L/2+1 values stored
          const bool lo = m <= (L >> 1);
          const cx<T> v = ((const cx<T> *)p.src)[base + (lo ? m : L - m) * p.se];
          re = v.re; im = lo ? v.im : -v.im;
        }
      }
      ld[it] = {re, im};
    }
#pragma unroll
    for (int it = 0; it < kPer; it++) buf[slot[it]] = ld[it];
#pragma unroll
    for (int it = 0; it < kTwPer; it++) {
      const int t = threadIdx.x + it * kFftThreads;
      if (t < (L >> 1)) tw[t] = twv[it];
    }
    __syncthreads();
  }
  // The last run of a decimation in time leaves positions q + j L/8 of
  // each line in natural order: with lanes walking across the lines, the
  // thread stores its group from registers (runs of B neighbouring x
  // lines), skipping the store's LDS round trip.
  auto last_run_store = [&](bool inv) {
    const int lb = threadIdx.x & (B - 1), q = threadIdx.x >> log2B;
    if (threadIdx.x >= (PTS >> 3)) return;  // one group per thread
    const int s8 = L >> 3;
    cx<T> v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = buf[lb * LP + lds_pos(q + j * s8, mask)];
    run_regs<T, true, 3>(v, log2L - 3, q, log2L, tw, inv);
    const int line = line0 + lb;
    if (line >= nlines) return;
    int f, sl;
    line_fs(line, p, f, sl);
    if (f >= p.nfv) return;
    const int64_t o = f * p.df + sl * p.ds;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int m = q + j * s8;
      if (m < p.Lout)
        ((cx<T> *)p.dst)[o + m * p.de] = {(T)(v[j].re * p.out_scale),
                                          (T)(v[j].im * p.out_scale)};
    }
  };
  if (dit) {
    fft_stages<T, true>(buf, LP, log2L, B, tw, inverse != 0, mask, 3, log2L - 3);
    last_run_store(inverse != 0);
    return;
  }
  if (CONV) {
    fft_stages<T, false>(buf, LP, log2L, B, tw, inverse != 0, mask, first, log2L - 3);
    fft_conv_middle<T>(buf, LP, log2L, log2B, tw, inverse != 0, gv, mask);
    if (log2L >= 6 && p.df == 1 && STORE == kStoreCplx) {
      fft_stages<T, true>(buf, LP, log2L, B, tw, true, mask, 3, log2L - 3);
      last_run_store(true);
      return;
    }
    fft_stages<T, true>(buf, LP, log2L, B, tw, true, mask, 3);
  } else {
    fft_stages<T, false>(buf, LP, log2L, B, tw, inverse != 0, mask, first);
  }
  // ---- store the kept part of each line
  const int totout = B * L;
  for (int t = threadIdx.x; t < totout; t += kFftThreads) {
    int lb, m;
    map_thread(t, p.de, log2L, log2B, lb, m);
    if (m >= p.Lout) continue;
    if (PAIR) {
      const int l0 = 2 * (line0 + lb), l1 = l0 + 1;
      if (l0 >= nlines) continue;
      const bool has1 = l1 < nlines;
      int f0, s0, f1, s1;
      line_fs(l0, p, f0, s0);
      line_fs(l1, p, f1, s1);
      const int64_t o0 = f0 * p.df + s0 * p.ds + m * p.de;
      const int64_t o1 = has1 ? f1 * p.df + s1 * p.ds + m * p.de : o0;
      const cx<T> z = buf[lb * LP + lds_pos(bitrev(m, log2L), mask)];
      if (STORE == kStoreReal) {
        ((T *)p.dst)[o0] = (T)(z.re * p.out_scale);
        if (has1) ((T *)p.dst)[o1] = (T)(z.im * p.out_scale);
      } else {
        // X(m) = (Z(m) + conj Z(L-m)) / 2,  Y(m) = (Z(m) - conj Z(L-m)) / 2i
        const cx<T> w = buf[lb * LP + lds_pos(bitrev((L - m) & (L - 1), log2L), mask)];
        const T h = (T)(0.5 * p.out_scale);
        ((cx<T> *)p.dst)[o0] = {(z.re + w.re) * h, (z.im - w.im) * h};
        if (has1) ((cx<T> *)p.dst)[o1] = {(z.im + w.im) * h, (w.re - z.re) * h};
      }
      continue;
    }
    int line = line0 + lb;
    if (line >= nlines) continue;
    int f, s;
    line_fs(line, p, f, s);
    if (f >= p.nfv) continue;
    int64_t o = f * p.df + s * p.ds + m * p.de;
    cx<T> v = buf[lb * LP + lds_pos(CONV ? m : bitrev(m, log2L), mask)];
    if (STORE == kStoreCplx) {
      ((cx<T> *)p.dst)[o] = {(T)(v.re * p.out_scale), (T)(v.im * p.out_scale)};
    } else {
      ((T *)p.dst)[o] = (T)(v.re * p.out_scale);
    }
  }
}

// ------------------------------------------------------ Green function
// IGF antiderivative at cell corners (src/Appl/Field.f90:384-418), FP64,
// mesh sizes from the device domain.
__global__ void k_igf_corner(double *__restrict__ out, int tx, int ty, int tz,
                             const DevDom *__restrict__ dom) {
  const double hx = dom->msize[0], hy = dom->msize[1], hz = dom->msize[2];
  const double vol = hx * hy * hz;
  const int n = tx * ty * tz;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += gridDim.x * blockDim.x) {
    int i = c % tx;
    int j = (c / tx) % ty;
    int k = c / (tx * ty);
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
  const int n = gx * gy * gz;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += gridDim.x * blockDim.x) {
    int i = c % gx;
    int j = (c / gx) % gy;
    int k = c / (gx * gy);
#define TT(a, b, cc) t[(a) + tx * ((b) + ty * (cc))]
    double v = TT(i + 1, j + 1, k + 1) - TT(i, j + 1, k + 1) - TT(i + 1, j, k + 1) +
               TT(i, j, k + 1) - TT(i + 1, j + 1, k) + TT(i, j + 1, k) +
               TT(i + 1, j, k) - TT(i, j, k);
#undef TT
    g[c] = (T)v;
  }
}

// --------------------------------------------------------------- deposit
__global__ void k_zero_u64(unsigned long long *__restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// Lower-node cell (0-based) and CIC weights of deposit_BeamBunch
// (src/Appl/BeamBunch.f90:3737-3744) in the position type P.
template <typename P>
struct Cic {
  int i, j, k;
  P ab, cd, ef;
};

template <typename P>
__device__ inline Cic<P> cic_of(P x, P y, P z, P xmin, P ymin, P zmin, P hx, P hy,
                                P hz, P hxi, P hyi, P hzi) {
  Cic<P> c;
  int ix = (int)((x - xmin) * hxi + (P)1);
  c.ab = ((xmin - x) + ix * hx) * hxi;
  int jx = (int)((y - ymin) * hyi + (P)1);
  c.cd = ((ymin - y) + jx * hy) * hyi;
  int kx = (int)((z - zmin) * hzi + (P)1);
  c.ef = ((zmin - z) + kx * hz) * hzi;
  c.i = ix - 1;
  c.j = jx - 1;
  c.k = kx - 1;
  return c;
}

// x rounded to the nearest integer (ties to even), as __double2ll_rn.
// FAST: valid for |x| < 2^50 (the launch checks that every contribution is
// below it): adding 1.5 * 2^52 leaves round(x) in the low mantissa bits, one
// add and one integer subtract instead of the split conversion.
template <bool FAST>
__device__ inline long long fx_round(double x) {
  if constexpr (FAST) {
    const double t = __dadd_rn(x, 6755399441055744.0);
    return (long long)(__double_as_longlong(t) - 0x4338000000000000LL);
  } else {
    return __double2ll_rn(x);
  }
}

// Fixed-point contributions of one particle, in the reference's node order.
template <typename P, bool FAST = false>
__device__ inline void cic_contrib(const Cic<P> &c, double wq, long long v[8]) {
  // each product is rounded to FP64 before it is rounded to an integer: no
  // multiply may fuse with the short form's add
#pragma clang fp contract(off)
  const P one = (P)1;
  const P a = c.ab, b = c.cd, e = c.ef;
  v[0] = fx_round<FAST>((double)(a * b * e) * wq);
  v[1] = fx_round<FAST>((double)(a * (one - b) * e) * wq);
  v[2] = fx_round<FAST>((double)(a * (one - b) * (one - e)) * wq);
  v[3] = fx_round<FAST>((double)(a * b * (one - e)) * wq);
  v[4] = fx_round<FAST>((double)((one - a) * b * (one - e)) * wq);
  v[5] = fx_round<FAST>((double)((one - a) * (one - b) * (one - e)) * wq);
  v[6] = fx_round<FAST>((double)((one - a) * (one - b) * e) * wq);
  v[7] = fx_round<FAST>((double)((one - a) * b * e) * wq);
}

// Node offsets (di, dj, dk) in the order of cic_contrib.
__constant__ int kNode[8][3] = {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1},
                                {1, 0, 1}, {1, 1, 1}, {1, 1, 0}, {1, 0, 0}};

__device__ inline int wave_min_i(int v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_down(v, off, 64));
  return v;
}
__device__ inline int wave_max_i(int v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_down(v, off, 64));
  return v;
}

// deposit_BeamBunch (src/Appl/BeamBunch.f90:3678-3768) for one chunk of
// consecutive particles per block.  The chunk's cells are bounded first; if
// the box (plus the upper CIC node) fits kWindowCells, the block accumulates
// in LDS and adds only the box to the grid, otherwise it adds straight to the
// grid.  Values are fixed point (int64, scale 2^e), so the result is exact
// and independent of the order of the additions.  Particles with a NaN x
// (lost in the fused run) and nodes outside the grid are skipped, as the
// reference skips them (:3746-3747).
template <typename P>
__global__ void __launch_bounds__(kBlock)
k_deposit_chunks(const P *__restrict__ X, const P *__restrict__ Y,
                 const P *__restrict__ Z, const uint32_t *__restrict__ orig,
                 const double *__restrict__ w0, double w_uniform, int uniform,
                 int64_t n_flat, const DevDom *__restrict__ dom, int nx, int ny,
                 int nz, double scale, unsigned long long *__restrict__ rho) {
  __shared__ unsigned long long win[kWindowCells];
  __shared__ int red[kBlock / 64][6];
  __shared__ int box[6];
  const int64_t start = (int64_t)blockIdx.x * kChunk;
  if (start >= n_flat) return;
  const int64_t len = min((int64_t)kChunk, n_flat - start);
  const P xmin = (P)dom->range[0], ymin = (P)dom->range[2], zmin = (P)dom->range[4];
  const P hx = (P)dom->msize[0], hy = (P)dom->msize[1], hz = (P)dom->msize[2];
  const P hxi = (P)dom->hinv[0], hyi = (P)dom->hinv[1], hzi = (P)dom->hinv[2];
  // ---- bounding box of the chunk's lower nodes
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (int64_t t = threadIdx.x; t < len; t += kBlock) {
    int64_t i = start + t;
    P x = X[i];
    if (!(x == x)) continue;
    Cic<P> c = cic_of<P>(x, Y[i], Z[i], xmin, ymin, zmin, hx, hy, hz, hxi, hyi, hzi);
    if (c.i < 0 || c.i + 1 >= nx || c.j < 0 || c.j + 1 >= ny || c.k < 0 ||
        c.k + 1 >= nz)
      continue;
    lo[0] = min(lo[0], c.i); lo[1] = min(lo[1], c.j); lo[2] = min(lo[2], c.k);
    hi[0] = max(hi[0], c.i); hi[1] = max(hi[1], c.j); hi[2] = max(hi[2], c.k);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int v6[6];
  for (int a = 0; a < 3; a++) { v6[a] = wave_min_i(lo[a]); v6[3 + a] = wave_max_i(hi[a]); }
  if (lane == 0)
    for (int a = 0; a < 6; a++) red[wv][a] = v6[a];
  __syncthreads();
  if (threadIdx.x < 6) {
    int a = threadIdx.x, r = red[0][a];
    for (int w = 1; w < kBlock / 64; w++) r = a < 3 ? min(r, red[w][a]) : max(r, red[w][a]);
    box[a] = r;
  }
  __syncthreads();
  if (box[0] == INT_MAX) return;  // nothing on the grid in this chunk
  const int bx = box[0], by = box[1], bz = box[2];
  const int wx = box[3] - bx + 2, wy = box[4] - by + 2, wz = box[5] - bz + 2;
  const int64_t wcells = (int64_t)wx * wy * wz;
  const bool local = wcells <= kWindowCells;
  if (local) {
    for (int c = threadIdx.x; c < wcells; c += kBlock) win[c] = 0ull;
    __syncthreads();
  }
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  for (int64_t t = threadIdx.x; t < len; t += kBlock) {
    int64_t i = start + t;
    P x = X[i];
    if (!(x == x)) continue;
    Cic<P> c = cic_of<P>(x, Y[i], Z[i], xmin, ymin, zmin, hx, hy, hz, hxi, hyi, hzi);
    if (c.i < 0 || c.i + 1 >= nx || c.j < 0 || c.j + 1 >= ny || c.k < 0 ||
        c.k + 1 >= nz)
      continue;
    const double w = uniform ? w_uniform : w0[orig[i]];
    long long v[8];
    cic_contrib<P>(c, w * scale, v);
    if (local) {
      const int l0 = (c.i - bx) + wx * ((c.j - by) + wy * (c.k - bz));
#pragma unroll
      for (int q = 0; q < 8; q++)
        atomicAdd(&win[l0 + kNode[q][0] + wx * (kNode[q][1] + wy * kNode[q][2])],
                  (unsigned long long)v[q]);
    } else {
      const int64_t g0 = c.i + sy * c.j + sz * c.k;
#pragma unroll
      for (int q = 0; q < 8; q++)
        atomicAdd(&rho[g0 + kNode[q][0] + sy * kNode[q][1] + sz * kNode[q][2]],
                  (unsigned long long)v[q]);
    }
  }
  if (!local) return;
  __syncthreads();
  for (int c = threadIdx.x; c < wcells; c += kBlock) {
    unsigned long long v = win[c];
    if (v == 0ull) continue;
    int lx = c % wx, ly = (c / wx) % wy, lz = c / (wx * wy);
    atomicAdd(&rho[(bx + lx) + sy * (by + ly) + sz * (bz + lz)], v);
  }
}

// The fused run's deposit.  Each block takes one chunk of the sorted bunch:
// at most kChunk slots grouped under one sort tile (the tile of the
// particles' previous position, in the grid *kgrid that key was computed
// in).  The block accumulates in an LDS window of the tile's nodes, mapped
// into this step's grid and widened on every side (plus the upper CIC node)
// by kTileMarginT nodes across the beam and kTileMarginL along it, so no
// pass over the positions is needed to bound them; particles that moved
// further add to the grid directly and are counted in *spill.  The margins
// follow how far particles move in cells between the key and the deposit
// (one or two steps): transversely several cells in a drift step, along the
// bunch (the t-frame phase) about one.  Same fixed-point arithmetic and
// node order as k_deposit_chunks.
constexpr int kTileMarginT = 6;
constexpr int kTileMarginL = 2;
constexpr int kTileWinEdgeT = (1 << kTileShift) + 1 + 2 * kTileMarginT;
constexpr int kTileWinEdgeL = (1 << kTileShift) + 1 + 2 * kTileMarginL;
constexpr int kTileWinCells = kTileWinEdgeT * kTileWinEdgeT * kTileWinEdgeL;

template <typename P, bool FAST>
__global__ void __launch_bounds__(kBlock)
k_deposit_tiles(const P *__restrict__ X, const P *__restrict__ Y,
                const P *__restrict__ Z, const uint32_t *__restrict__ orig,
                const double *__restrict__ w0, double w_uniform, int uniform,
                const Chunk *__restrict__ chunks, const int64_t *__restrict__ d_nchunks,
                const KeyGrid *__restrict__ kgrid, const DevDom *__restrict__ dom, int nx,
                int ny, int nz, int s0, int s1, int s2, int tnx, int tny, double scale,
                unsigned long long *__restrict__ rho, int64_t *__restrict__ spill) {
  __shared__ unsigned long long win[kTileWinCells];
  if (blockIdx.x >= *d_nchunks) return;
  const Chunk ch = chunks[blockIdx.x];
  const int tz = (int)(ch.tile / ((uint32_t)tnx * (uint32_t)tny));
  const int rem = (int)(ch.tile % ((uint32_t)tnx * (uint32_t)tny));
  const int ty = rem / tnx, tx = rem % tnx;
  const int wx = min((1 << s0) + 1 + 2 * kTileMarginT, nx);
  const int wy = min((1 << s1) + 1 + 2 * kTileMarginT, ny);
  const int wz = min((1 << s2) + 1 + 2 * kTileMarginL, nz);
  const KeyGrid kg = *kgrid;
  const int bx = min(max(tile_first_node(tx << s0, kg.lo[0], kg.h[0], dom->range[0],
                                         dom->hinv[0]) - kTileMarginT, 0), nx - wx);
  const int by = min(max(tile_first_node(ty << s1, kg.lo[1], kg.h[1], dom->range[2],
                                         dom->hinv[1]) - kTileMarginT, 0), ny - wy);
  const int bz = min(max(tile_first_node(tz << s2, kg.lo[2], kg.h[2], dom->range[4],
                                         dom->hinv[2]) - kTileMarginL, 0), nz - wz);
  const int wcells = wx * wy * wz;
  for (int c = threadIdx.x; c < wcells; c += kBlock) win[c] = 0ull;
  __syncthreads();
  const P xmin = (P)dom->range[0], ymin = (P)dom->range[2], zmin = (P)dom->range[4];
  const P hx = (P)dom->msize[0], hy = (P)dom->msize[1], hz = (P)dom->msize[2];
  const P hxi = (P)dom->hinv[0], hyi = (P)dom->hinv[1], hzi = (P)dom->hinv[2];
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  uint32_t nspill = 0;
  // kDepBatch slots per thread per round: all their loads are issued first
  constexpr int kDepBatch = 4;
  for (uint32_t t0 = threadIdx.x; t0 < ch.len; t0 += kDepBatch * kBlock) {
    P px[kDepBatch], py[kDepBatch], pz[kDepBatch];
    double pw[kDepBatch];
#pragma unroll
    for (int u = 0; u < kDepBatch; u++) {
      const uint32_t t = t0 + u * kBlock;
      const int64_t i = (int64_t)ch.start + (t < ch.len ? t : ch.len - 1);
      px[u] = t < ch.len ? X[i] : (P)NAN;
      py[u] = Y[i];
      pz[u] = Z[i];
      pw[u] = uniform ? w_uniform : w0[orig[i]];
    }
#pragma unroll
    for (int u = 0; u < kDepBatch; u++) {
      const P x = px[u];
      if (!(x == x)) continue;
      const Cic<P> c = cic_of<P>(x, py[u], pz[u], xmin, ymin, zmin, hx, hy, hz, hxi,
                                 hyi, hzi);
      if (c.i < 0 || c.i + 1 >= nx || c.j < 0 || c.j + 1 >= ny || c.k < 0 ||
          c.k + 1 >= nz)
        continue;
      long long v[8];
      cic_contrib<P, FAST>(c, pw[u] * scale, v);
      const int li = c.i - bx, lj = c.j - by, lk = c.k - bz;
      if (li >= 0 && li + 1 < wx && lj >= 0 && lj + 1 < wy && lk >= 0 && lk + 1 < wz) {
        const int l0 = li + wx * (lj + wy * lk);
#pragma unroll
        for (int q = 0; q < 8; q++)
          atomicAdd(&win[l0 + kNode[q][0] + wx * (kNode[q][1] + wy * kNode[q][2])],
                    (unsigned long long)v[q]);
      } else {
        const int64_t g0 = c.i + sy * c.j + sz * c.k;
#pragma unroll
        for (int q = 0; q < 8; q++)
          atomicAdd(&rho[g0 + kNode[q][0] + sy * kNode[q][1] + sz * kNode[q][2]],
                    (unsigned long long)v[q]);
        nspill++;
      }
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < wcells; c += kBlock) {
    const unsigned long long v = win[c];
    if (v == 0ull) continue;
    const int lx = c % wx, ly = (c / wx) % wy, lz = c / (wx * wy);
    atomicAdd(&rho[(bx + lx) + sy * (by + ly) + sz * (bz + lz)], v);
  }
  if (spill) {
    int tot = (int)nspill;
    for (int off = 32; off > 0; off >>= 1) tot += __shfl_down(tot, off, 64);
    if ((threadIdx.x & 63) == 0 && tot > 0)
      atomicAdd((unsigned long long *)spill, (unsigned long long)tot);
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
// 1782-1794, 1885-1899, 1966-2005): central differences, one-sided at edges,
// written interleaved (Ex, Ey, Ez, 0) for the 8-node gather.
template <typename T>
__global__ void k_efield(const T *__restrict__ phi, E4<T> *__restrict__ e,
                         int nx, int ny, int nz, const DevDom *__restrict__ dom) {
  const int sy = nx, sz = nx * ny, n = sz * nz;
  const T hx = (T)dom->hinv[0], hy = (T)dom->hinv[1], hz = (T)dom->hinv[2];
  const T hx2 = (T)(0.5 * dom->hinv[0]), hy2 = (T)(0.5 * dom->hinv[1]),
          hz2 = (T)(0.5 * dom->hinv[2]);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += gridDim.x * blockDim.x) {
    int i = c % nx;
    int j = (c / nx) % ny;
    int k = c / sz;
    E4<T> v;
    if (i == 0) v.x = hx * (phi[c] - phi[c + 1]);
    else if (i == nx - 1) v.x = hx * (phi[c - 1] - phi[c]);
    else v.x = hx2 * (phi[c - 1] - phi[c + 1]);
    if (j == 0) v.y = hy * (phi[c] - phi[c + sy]);
    else if (j == ny - 1) v.y = hy * (phi[c - sy] - phi[c]);
    else v.y = hy2 * (phi[c - sy] - phi[c + sy]);
    if (k == 0) v.z = hz * (phi[c] - phi[c + sz]);
    else if (k == nz - 1) v.z = hz * (phi[c - sz] - phi[c]);
    else v.z = hz2 * (phi[c - sz] - phi[c + sz]);
    v.w = (T)0;
    e[c] = v;
  }
}

int ilog2(int v) {
  int r = 0;
  while ((1 << r) < v) r++;
  return r;
}

template <typename T, int LOAD, int STORE, bool CONV, bool PAIR = false>
int run_lines(hipStream_t stream, Lines p, const void *tw, bool inverse) {
  // the fused z pass takes twice the lines per block: its rows of the
  // spectrum and of G^ are read across the lines
  constexpr int pts = CONV ? kFftPointsZ : kFftPoints;
  p.log2L = ilog2(p.L);
  if (p.nfv == 0) p.nfv = p.nf;
  if (p.L < kFftMinLine || p.L > pts || (1 << p.log2L) != p.L)
    return fail(IZH_ERR_ARG, "FFT line of %d points: not a power of two in %d..%d", p.L,
                kFftMinLine, pts);
  p.B = pts / p.L;
  p.log2B = ilog2(p.B);
  // line / nf is a multiply by ceil(2^40 / nf), exact below these bounds
  if (p.nf < 1 || p.nf >= (1 << 16) || 2 * p.nf * p.ns >= (1 << 24))
    return fail(IZH_ERR_ARG, "FFT of %lld x %lld lines outside the kernel's index range",
                (long long)p.nf, (long long)p.ns);
  p.nf_mul = ((1ull << 40) + (uint64_t)p.nf - 1) / (uint64_t)p.nf;
  const int64_t nlines = p.nf * p.ns;
  const int64_t nfft = PAIR ? (nlines + 1) / 2 : nlines;
  const int64_t nblk = (nfft + p.B - 1) / p.B;
  const size_t tw_bytes = (size_t)(p.L / 2) * sizeof(cx<T>);
  const dim3 g((unsigned)nblk), b(kFftThreads);
  const cx<T> *twc = (const cx<T> *)tw;
  const int inv = inverse ? 1 : 0;
  // the line lengths of the 64^3 .. 128x128x256 grids as compile-time sizes
  switch (p.log2L) {
    case 7:
      hipLaunchKernelGGL((k_fft_lines<T, LOAD, STORE, CONV, PAIR, pts, 7>), g, b, tw_bytes,
                         stream, p, twc, inv);
      break;
    case 8:
      hipLaunchKernelGGL((k_fft_lines<T, LOAD, STORE, CONV, PAIR, pts, 8>), g, b, tw_bytes,
                         stream, p, twc, inv);
      break;
    case 9:
      hipLaunchKernelGGL((k_fft_lines<T, LOAD, STORE, CONV, PAIR, pts, 9>), g, b, tw_bytes,
                         stream, p, twc, inv);
      break;
    default:
      hipLaunchKernelGGL((k_fft_lines<T, LOAD, STORE, CONV, PAIR, pts, 0>), g, b, tw_bytes,
                         stream, p, twc, inv);
      break;
  }
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

// G^ of the field's current domain: IGF corners, 8-corner difference and
// three even real transforms, on stream gs.
template <typename T>
int green_t(izh_field *f, hipStream_t gs) {
  const int nx = f->nx, ny = f->ny, nz = f->nz;
  const int mx = 2 * nx, my = 2 * ny, mz = 2 * nz, kx = nx + 1;
  {
    ScopedTimer tm_(IZH_T_GREEN, gs);
    int nc = (nx + 2) * (ny + 2) * (nz + 2);
    hipLaunchKernelGGL(k_igf_corner, dim3(grid_for(nc)), dim3(kBlock), 0,
                       gs, f->corner, nx + 2, ny + 2, nz + 2, f->d_dom);
    int ng = (nx + 1) * (ny + 1) * (nz + 1);
    hipLaunchKernelGGL(k_igf_diff<T>, dim3(grid_for(ng)), dim3(kBlock), 0,
                       gs, f->corner, (T *)f->green_t, nx, ny, nz);
    IZH_HIP(hipGetLastError());
    Lines p{};
    // x: lines (j<=ny, k<=nz), contiguous; green_t -> green
    p.src = f->green_t; p.dst = f->green;
    p.nf = ny + 1; p.ns = nz + 1;
    p.sf = kx; p.ss = (int64_t)kx * (ny + 1); p.se = 1;
    p.df = kx; p.ds = (int64_t)kx * (ny + 1); p.de = 1;
    p.L = mx; p.Lin = p.L; p.Lout = nx + 1; p.out_scale = 1.0;
    IZH_CHECK((run_lines<T, kLoadEven, kStoreReal, false, true>(gs, p, f->tw_x, false)));
    // y: lines (i<=nx, k<=nz); green -> green_t
    p.src = f->green; p.dst = f->green_t;
    p.nf = kx; p.ns = nz + 1;
    p.sf = 1; p.ss = (int64_t)kx * (ny + 1); p.se = kx;
    p.df = 1; p.ds = (int64_t)kx * (ny + 1); p.de = kx;
    p.L = my; p.Lin = p.L; p.Lout = ny + 1;
    IZH_CHECK((run_lines<T, kLoadEven, kStoreReal, false, true>(gs, p, f->tw_y, false)));
    // z: lines (i<=nx, j<=ny); green_t -> green, scaled by 1/(mx my mz).
    // G^ is stored [j][k][i] with i rows at the spectra's pitch ldx: the
    // fused z pass of a block of x lines then reads, for its one y index,
    // aligned runs of neighbouring i at neighbouring k, one contiguous
    // region of ldx*(nz+1) values instead of one value run per z plane.
    const int64_t ld = f->ldx;
    p.src = f->green_t; p.dst = f->green;
    p.nf = kx; p.ns = ny + 1;
    p.sf = 1; p.ss = kx; p.se = (int64_t)kx * (ny + 1);
    p.df = 1; p.ds = ld * (nz + 1); p.de = ld;
    p.L = mz; p.Lin = p.L; p.Lout = nz + 1;
    p.out_scale = 1.0 / ((double)mx * my * mz);
    IZH_CHECK((run_lines<T, kLoadEven, kStoreReal, false, true>(gs, p, f->tw_z, false)));
  }
  return IZH_OK;
}

// The convolution with the G^ in place: forward x, forward y, the fused z
// pass, inverse y, inverse x, on the field's stream.
template <typename T>
int convolve_t(izh_field *f) {
  const int nx = f->nx, ny = f->ny, nz = f->nz;
  const int mx = 2 * nx, my = 2 * ny, mz = 2 * nz, kx = nx + 1;
  const hipStream_t fs = f->stream;
  ScopedTimer tm_(IZH_T_FFT, fs);
  // The half spectra hold x rows of ld = f->ldx >= kx values.  The y and z
  // passes run over all ld x lines, so a block's run of neighbouring lines
  // starts on a 128-byte boundary; lines kx..ld-1 are padding.
  const int64_t ld = f->ldx;
  // ---- forward x: rho (nx,ny,nz) real -> A (kx of ld, ny, nz)
  Lines p{};
  p.src = f->rho; p.dst = f->spec_a;
  p.nf = ny; p.ns = nz;
  p.sf = nx; p.ss = (int64_t)nx * ny; p.se = 1;
  p.df = ld; p.ds = ld * ny; p.de = 1;
  p.L = mx; p.Lin = nx; p.Lout = kx; p.out_scale = 1.0;
  IZH_CHECK((run_lines<T, kLoadReal, kStoreCplx, false, true>(fs, p, f->tw_x, false)));
  // ---- forward y: A -> B (kx of ld, my, nz)
  p.src = f->spec_a; p.dst = f->spec_b;
  p.nf = ld; p.nfv = kx; p.ns = nz;
  p.sf = 1; p.ss = ld * ny; p.se = ld;
  p.df = 1; p.ds = ld * my; p.de = ld;
  p.L = my; p.Lin = ny; p.Lout = my;
  IZH_CHECK((run_lines<T, kLoadCplx, kStoreCplx, false>(fs, p, f->tw_y, false)));
  // ---- z: FFT, * G^, IFFT, keep k < nz; in place in B
  p.src = f->spec_b; p.dst = f->spec_b;
  p.nf = ld; p.nfv = kx; p.ns = my;
  p.sf = 1; p.ss = ld; p.se = ld * my;
  p.df = 1; p.ds = ld; p.de = ld * my;
  p.L = mz; p.Lin = nz; p.Lout = nz;
  p.green = f->green; p.gkx = ld * (nz + 1); p.gky = ld; p.my = my;
  IZH_CHECK((run_lines<T, kLoadCplx, kStoreCplx, true>(fs, p, f->tw_z, false)));
  // ---- inverse y: B -> A, keep j < ny
  p.green = nullptr;
  p.src = f->spec_b; p.dst = f->spec_a;
  p.nf = ld; p.nfv = kx; p.ns = nz;
  p.sf = 1; p.ss = ld * my; p.se = ld;
  p.df = 1; p.ds = ld * ny; p.de = ld;
  p.L = my; p.Lin = my; p.Lout = ny;
  IZH_CHECK((run_lines<T, kLoadCplx, kStoreCplx, false>(fs, p, f->tw_y, true)));
  // ---- inverse x (C2R): A -> phi, keep i < nx
  p.src = f->spec_a; p.dst = f->phi;
  p.nf = ny; p.nfv = ny; p.ns = nz;
  p.sf = ld; p.ss = ld * ny; p.se = 1;
  p.df = nx; p.ds = (int64_t)nx * ny; p.de = 1;
  p.L = mx; p.Lin = mx; p.Lout = nx;
  IZH_CHECK((run_lines<T, kLoadHerm, kStoreReal, false, true>(fs, p, f->tw_x, true)));
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

double fixed_point_scale(double wabs_sum) {
  // every cell holds at most sum|w| (CIC weights of a particle sum to one),
  // so 2^e * sum|w| <= 2^61 leaves headroom for the +-0.5 rounding per term
  double wsum = wabs_sum > 0 ? wabs_sum : 1.0;
  int e = 61 - (int)std::ceil(std::log2(wsum));
  return std::ldexp(1.0, e);
}

int field_alloc(izh_field *f) {
  const size_t es = f->precision == IZH_FP64 ? sizeof(double) : sizeof(float);
  const int64_t g = f->g, kx = f->nx + 1;
  const int64_t ghalf = kx * (f->ny + 1) * (f->nz + 1);
  // rows of whole 128-byte lines: the y and z passes read and write runs of
  // neighbouring x lines, which then start on a line boundary
  const int64_t row = 128 / (2 * (int64_t)es);
  f->ldx = (kx + row - 1) / row * row;
  IZH_HIP(hipMalloc((void **)&f->rho_fx, g * sizeof(unsigned long long)));
  IZH_HIP(hipMalloc(&f->rho, g * es));
  IZH_HIP(hipMalloc(&f->phi, g * es));
  IZH_HIP(hipMalloc(&f->e, 4 * g * es));
  IZH_HIP(hipMalloc(&f->spec_a, f->ldx * f->ny * f->nz * 2 * es));
  IZH_HIP(hipMalloc(&f->spec_b, f->ldx * 2 * f->ny * f->nz * 2 * es));
  // G^ at x pitch ldx ([j][k][i], written by its z pass; the x pass's
  // output at pitch kx fits as well)
  IZH_HIP(hipMalloc(&f->green, f->ldx * (f->ny + 1) * (f->nz + 1) * es));
  IZH_HIP(hipMalloc(&f->green_t, ghalf * es));
  IZH_HIP(hipMalloc((void **)&f->corner,
                    (size_t)(f->nx + 2) * (f->ny + 2) * (f->nz + 2) * sizeof(double)));
  IZH_HIP(hipMalloc((void **)&f->d_dom, sizeof(DevDom)));
  IZH_HIP(hipHostMalloc((void **)&f->h_dom, sizeof(DevDom), hipHostMallocDefault));
  IZH_HIP(hipMemset(f->d_dom, 0, sizeof(DevDom)));
  if (f->precision == IZH_FP64) return upload_twiddles<double>(f);
  return upload_twiddles<float>(f);
}

int field_set_domain(izh_field *f, const izh_compdom *dom) {
  // the pinned staging buffer may still feed an earlier copy
  IZH_HIP(hipStreamSynchronize(f->stream));
  DevDom d{};
  for (int k = 0; k < 6; k++) d.range[k] = dom->range[k];
  for (int k = 0; k < 3; k++) {
    d.msize[k] = dom->msize[k];
    d.hinv[k] = 1.0 / dom->msize[k];
  }
  *f->h_dom = d;
  IZH_HIP(hipMemcpyAsync(f->d_dom, f->h_dom, sizeof(DevDom), hipMemcpyHostToDevice,
                         f->stream));
  return IZH_OK;
}

// Zero the accumulator, deposit, convert to the grid type.  With a chunk
// list (the fused run) the positions are the t-frame buffer (T) of the
// tile-sorted bunch; without one, the bunch's own rows (t frame, FP64) in
// chunks of kChunk consecutive particles.
int launch_deposit(izh_bunch *b, izh_field *f, const Chunk *chunks, const KeyGrid *kgrid,
                   const int64_t *d_nchunks, int64_t max_chunks) {
  f->fixed_scale = fixed_point_scale(b->wabs_sum);
  // every contribution is at most |w| * scale: below 2^50 the fused
  // deposit rounds with the short form
  const bool fast = b->wabs_max * f->fixed_scale < 0x1p50;
  hipLaunchKernelGGL(k_zero_u64, dim3(grid_for(f->g)), dim3(kBlock), 0, f->stream,
                     f->rho_fx, f->g);
  if (max_chunks > 0) {
    ScopedTimer tm_(IZH_T_DEPOSIT, f->stream);
    if (chunks) {
      // the fused run: chunks of the tile-sorted bunch, t positions in T
      if (!kgrid) return fail(IZH_ERR_ARG, "deposit: chunk list without its key grid");
      const int s0 = tile_shift(f->nx), s1 = tile_shift(f->ny), s2 = tile_shift(f->nz);
      const int tnx = f->nx >> s0, tny = f->ny >> s1;
      int64_t *spill = b->d_scal ? &b->d_scal->overflow : nullptr;
      const auto go = [&](auto kern, const auto *tp) {
        hipLaunchKernelGGL(kern, dim3((unsigned)max_chunks), dim3(kBlock), 0, f->stream, tp,
                           tp + b->cap, tp + 2 * b->cap, b->orig, b->w0, b->w_value,
                           b->w_uniform, chunks, d_nchunks, kgrid, f->d_dom, f->nx, f->ny,
                           f->nz, s0, s1, s2, tnx, tny, f->fixed_scale, f->rho_fx, spill);
      };
      if (f->precision == IZH_MIXED) {
        const float *tp = (const float *)b->tpos;
        if (fast) go(k_deposit_tiles<float, true>, tp);
        else go(k_deposit_tiles<float, false>, tp);
      } else {
        const double *tp = (const double *)b->tpos;
        if (fast) go(k_deposit_tiles<double, true>, tp);
        else go(k_deposit_tiles<double, false>, tp);
      }
    } else {
      // one call's bunch in the t frame, in its own (unsorted) order
      hipLaunchKernelGGL(k_deposit_chunks<double>, dim3((unsigned)max_chunks),
                         dim3(kBlock), 0, f->stream, b->r[0], b->r[2], b->r[4],
                         b->orig, b->w0, b->w_value, b->w_uniform, b->n, f->d_dom,
                         f->nx, f->ny, f->nz, f->fixed_scale, f->rho_fx);
    }
    IZH_HIP(hipGetLastError());
  }
  // FP64: the grid holds charge itself; mixed: charge / sum|w| (O(1) values)
  double wsum = b->wabs_sum > 0 ? b->wabs_sum : 1.0;
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

int launch_deposit_granular(izh_bunch *b, izh_field *f) {
  const int64_t nch = (b->n + kChunk - 1) / kChunk;
  return launch_deposit(b, f, nullptr, nullptr, nullptr, nch);
}

int launch_green(izh_field *f, hipStream_t gs) {
  return f->precision == IZH_FP64 ? green_t<double>(f, gs) : green_t<float>(f, gs);
}

int launch_convolve(izh_field *f) {
  int rc = f->precision == IZH_FP64 ? convolve_t<double>(f) : convolve_t<float>(f);
  if (rc == IZH_OK) f->have_phi = 1;
  return rc;
}

int launch_solve(izh_field *f) {
  const int rc = launch_green(f, f->stream);
  return rc == IZH_OK ? launch_convolve(f) : rc;
}

int launch_efield(izh_field *f) {
  ScopedTimer tm_(IZH_T_EFIELD, f->stream);
  if (f->precision == IZH_FP64)
    hipLaunchKernelGGL(k_efield<double>, dim3(grid_for(f->g)), dim3(kBlock), 0,
                       f->stream, (const double *)f->phi, (E4<double> *)f->e,
                       f->nx, f->ny, f->nz, f->d_dom);
  else
    hipLaunchKernelGGL(k_efield<float>, dim3(grid_for(f->g)), dim3(kBlock), 0,
                       f->stream, (const float *)f->phi, (E4<float> *)f->e, f->nx,
                       f->ny, f->nz, f->d_dom);
  IZH_HIP(hipGetLastError());
  return IZH_OK;
}

}  // namespace izh
