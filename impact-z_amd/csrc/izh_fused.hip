// izh_fused.hip -- a run of space-charge steps as one device-resident pipeline.
//
// Per step s the particles make ONE streaming pass (k_pass):
//   [step s-1] 0th-order t-frame position, 8-node gather of E(s-1), kick,
//              back transform, second half push, diagnostic moment sums;
//   [step s]   first half push, phase wrap, 1st-order z->t transform, range,
//              loss test, tile key of the new position.
// The pass writes each particle to its slot in the other ping-pong buffer,
// grouped by the tile key computed one step earlier (a counting sort whose
// histogram and scan were done by the previous k_finalize): neighbours in
// memory stay neighbours in space, so the gather hits cache and the deposit
// (izh_field.hip, k_deposit_chunks) accumulates each chunk in LDS.
// k_finalize (one block) then reduces the pass's partials into the step's
// moments, the new domain (update_CompDom), particle count, Current rescale
// and kick constants, scans the new tile histogram into scatter cursors and
// the next deposit's chunk list.  The host only launches kernels.
//
// Reference call order per segment (src/Contrl/AccSimulator.f90:1099-1488):
// drift1 -> conv1st -> update_CompDom -> charge -> update3O -> cvbkforth1st ->
// kick1 -> drift1 -> diagnostic1.  cvbkforth1st(conv1st(P)) is the 0th-order
// transform of the wrapped z-frame state, so the pass gathers there directly.
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "izh_device.hpp"

namespace izh {
namespace {

// One half-step push in the form the pass applies it.  The quadrupole map of
// maplinear_Quadrupole is block diagonal (src/Appl/Quadrupole.f90:179-196),
// so only its three 2x2 blocks travel.
struct HalfPush {
  double tau_xl, gam0, phase;  // exact drift: tau/Scxl, gamma0, tau/(Scxl beta0)
  // kPushMap: the (x,px), (y,py), (phi,pt) 2x2 blocks, row major (m[0..11]);
  // kPushMapC: the 4x4 transverse block, row major (m[0..15]), then the
  // (phi,pt) block (m[16..19]);
  // kPushQuadK: m[0] tau, m[1] K, m[2] the beam's q/mc^2, m[3] gamma0,
  // m[4] gamma0 beta0, m[5] beta0, m[6] Scxl
  double m[20];
};

// kPushMapC is a map whose transverse planes are coupled (the solenoid);
// kPushQuadK the chromatic map of a quadrupole given by K, per particle.
enum PushKind : int {
  kPushNone = 0, kPushDrift = 1, kPushMap = 2, kPushMapC = 3, kPushQuadK = 4
};

// What a pass does with the slots: kPassScatter writes each particle to its
// counting-sort slot by the tile key the previous pass computed (and moves
// its original index); without it the pass writes every particle back to
// the slot it read (coalesced, no atomics, the original index stays).
// kPassKeys computes the tile key of each particle's new position and its
// histogram, for a scatter in the next pass.
enum PassMode : int { kPassInPlace = 0, kPassScatter = 1, kPassKeys = 2 };

struct PassCfg {
  const double *in;         // six rows of cap doubles, row k at in + k*cap
  double *out;              // (== in for a pass without kPassScatter)
  const uint32_t *in_orig;  // original index of the input slots
  uint32_t *out_orig;       // kPassScatter: original index of the output slots
  const uint32_t *in_key;   // kPassScatter: tile key of the input slots
  uint32_t *out_key;        // kPassKeys: tile key of the output slots
  void *tpos;               // t-frame positions (T), 3 blocks of cap
  int64_t cap;
  const int64_t *d_nslots;  // input slots (device)
  uint32_t *cursor;         // kPassScatter: scatter cursor per tile of the input keys
  uint32_t *hist_new;       // kPassKeys: histogram of the output keys
  const DevDom *dom;        // domain of step s-1: gather and key geometry
  const void *e;            // E4<T> of step s-1
  const DevScal *scal;      // current and kick constants of step s-1
  DevScal *err;             // the same record, for the slot-overflow flag
  int nx, ny, nz;
  int tshift[3], tnx, tny, ntiles;
  int kb, ka;               // push kinds of pb and pa (the same for every lane)
  int need_orig;            // a K-mode push needs each particle's own q/mc^2
  const double *qm0;        // q/mc^2 by original index
  double qm_value;          // the q/mc^2 of every particle (without need_orig)
  double gam, xl, rad;      // -refpt(6), Scxl, pipe radius of step s
  double gam_k;             // -refpt(6) of step s-1 (its kick; RF cavities change it)
  double tau2_prev;         // 2*tau1 of step s-1
  double inv_rho_scale;     // undo the grid-charge scaling in the kick
  HalfPush pb, pa;          // second half of step s-1, first half of step s
  double *part;             // kMoments moment values, kConvRed conv, per block
  void *sink;               // kSinkBytes for the stores of empty lanes
  const Chunk *in_chunks;   // slot ranges of the input grouping, by tile
  const int64_t *in_nchunks;
  const KeyGrid *kgrid;     // grid the tiles of in_chunks were computed in
};

// drift1_BeamBunch (src/Appl/BeamBunch.f90:276-320) for one particle.  The
// map rows keep the reference's product order: the zero entries of the 6x6
// product add nothing.
template <int K>
__device__ inline void half_push(const HalfPush &p, double &x, double &px,
                                 double &y, double &py, double &ph, double &pt) {
  if constexpr (K == kPushDrift) {
    // drift1_BeamBunch bitype 0 (src/Appl/BeamBunch.f90:290-297) with the
    // common factor (tau/Scxl)/tmph taken once: one division instead of six
    const double d = pt - p.gam0;
    const double r = p.tau_xl / sqrt(d * d - 1.0 - px * px - py * py);
    x = fma(px, r, x);
    y = fma(py, r, y);
    ph = ph - fma(d, r, p.phase);
  } else if constexpr (K == kPushMap) {
    const double *m = p.m;
    const double x1 = fma(px, m[1], x * m[0]), px1 = fma(px, m[3], x * m[2]);
    const double y1 = fma(py, m[5], y * m[4]), py1 = fma(py, m[7], y * m[6]);
    const double ph1 = fma(pt, m[9], ph * m[8]), pt1 = fma(pt, m[11], ph * m[10]);
    x = x1; px = px1; y = y1; py = py1; ph = ph1; pt = pt1;
  } else if constexpr (K == kPushMapC) {
    const double *m = p.m;
    double t[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
      t[j] = fma(py, m[4 * j + 3], fma(y, m[4 * j + 2], fma(px, m[4 * j + 1], x * m[4 * j])));
    const double ph1 = fma(pt, m[17], ph * m[16]), pt1 = fma(pt, m[19], ph * m[18]);
    x = t[0]; px = t[1]; y = t[2]; py = t[3]; ph = ph1; pt = pt1;
  }
}

// transfmapK_Quadrupole (src/Appl/Quadrupole.f90:509-607) for one particle of
// q/mc^2 qm: its own strength kstr = qm/qm0 * gambet0/|gamma beta| * K gives a
// cos/sin (focusing) and cosh/sinh (defocusing) pair of 2x2 maps; the phase
// slips by (1/beta_z - 1/beta0) tau/Scxl.  The arithmetic of k_quad_kmode.
__device__ inline void half_push_kmode(const HalfPush &p, double qm, double &x, double &px,
                                       double &y, double &py, double &ph, double pt) {
  const double tau = p.m[0], kq = p.m[1], qmass = p.m[2], gam0 = p.m[3];
  const double gambet0 = p.m[4], beta0 = p.m[5], scxl = p.m[6];
  const double gam = gam0 - pt;
  double gambetz = sqrt(gam * gam - 1.0 - px * px - py * py);
  const double betaz = gambetz / gam;
  gambetz = sqrt(gam * gam - 1.0);
  const double kstr = qm / qmass * gambet0 / gambetz * kq;
  const double rtkstr = sqrt(fabs(kstr));
  const double rtkstrzz = rtkstr * tau;
  double xm11, xm12, xm21, xm22, xm33, xm34, xm43, xm44;
  if (kstr != 0.0) {
    const double co = cos(rtkstrzz), si = sin(rtkstrzz);
    const double t = exp(rtkstrzz);
    const double ch = (t * t + 1.0) / (t + t), sh = (t * t - 1.0) / (t + t);
    if (kstr > 0.0) {
      xm11 = co; xm12 = si / rtkstr; xm21 = -rtkstr * si; xm22 = co;
      xm33 = ch; xm34 = sh / rtkstr; xm43 = sh * rtkstr; xm44 = ch;
    } else {
      xm11 = ch; xm12 = sh / rtkstr; xm21 = rtkstr * sh; xm22 = ch;
      xm33 = co; xm34 = si / rtkstr; xm43 = -si * rtkstr; xm44 = co;
    }
  } else {
    xm11 = 1.0; xm12 = tau; xm21 = 0.0; xm22 = 1.0;
    xm33 = 1.0; xm34 = tau; xm43 = 0.0; xm44 = 1.0;
  }
  const double x1 = xm11 * x + xm12 * px / gambetz / scxl;
  const double px1 = gambetz * scxl * xm21 * x + xm22 * px;
  const double y1 = xm33 * y + xm34 * py / gambetz / scxl;
  const double py1 = gambetz * scxl * xm43 * y + xm44 * py;
  ph = ph + (1.0 / betaz - 1.0 / beta0) * tau / scxl;
  x = x1; px = px1; y = y1; py = py1;
}

// The push of a kind chosen at launch (the same for every lane, so the
// branches are scalar).
__device__ inline void half_push_any(int kind, const HalfPush &p, double qm, double &x,
                                     double &px, double &y, double &py, double &ph,
                                     double &pt) {
  if (kind == kPushDrift) half_push<kPushDrift>(p, x, px, y, py, ph, pt);
  else if (kind == kPushMap) half_push<kPushMap>(p, x, px, y, py, ph, pt);
  else if (kind == kPushMapC) half_push<kPushMapC>(p, x, px, y, py, ph, pt);
  else if (kind == kPushQuadK) half_push_kmode(p, qm, x, px, y, py, ph, pt);
}

// The push of kind K fixed at compile time (K >= 0), or of the launch's
// kind read at run time (K = kPushAnyKind).  The pass is instantiated for
// the drift/map/none combinations of the FODO-type lattices, so their
// launches carry neither the other kinds' code nor their coefficients as
// live uniform values; every other combination takes the run-time form.
constexpr int kPushAnyKind = -1;
template <int K>
__device__ inline void half_push_k(int kind, const HalfPush &p, double qm, double &x,
                                   double &px, double &y, double &py, double &ph,
                                   double &pt) {
  if constexpr (K == kPushAnyKind) half_push_any(kind, p, qm, x, px, y, py, ph, pt);
  else if constexpr (K == kPushDrift || K == kPushMap || K == kPushMapC)
    half_push<K>(p, x, px, y, py, ph, pt);
  else if constexpr (K == kPushQuadK) half_push_kmode(p, qm, x, px, y, py, ph, pt);
}

// Counting-sort slot of each participating lane, in two halves so the
// atomics' latency overlaps other work.  wave_rank_issue groups the lanes by
// key with ballots (no memory traffic), then makes ONE atomic instruction in
// which each group's first lane adds the group's size; wave_rank_slot later
// combines that old counter value with the lane's rank in its group.
struct RankTicket {
  uint32_t ret;   // the atomic's old value (group leaders)
  int leader;     // lane holding this lane's group atomic
  uint32_t rank;  // position of this lane within its group
};

__device__ inline RankTicket wave_rank_issue(uint32_t *counters, uint32_t key, bool take) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(take);
  RankTicket t{0u, lane, 0u};
  uint32_t count = 0;
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
    const unsigned long long same = __ballot(take && key == lk) & pending;
    if ((same >> lane) & 1ull) {
      t.leader = leader;
      t.rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
    }
    if (lane == leader) count = (uint32_t)__popcll(same);
    pending &= ~same;
  }
  if (take && t.leader == lane) t.ret = atomicAdd(&counters[key], count);
  return t;
}

__device__ inline uint32_t wave_rank_slot(const RankTicket &t) {
  return (uint32_t)__shfl((int)t.ret, t.leader, 64) + t.rank;
}

// Block-staged counting-sort scatter.  The kBlock slots a block takes per
// group are grouped by key in an LDS table, so the cursor atomic is made
// once per key per block instead of once per key per wavefront, and the
// particles are staged in LDS in the order of their new slots: each
// wavefront then stores consecutive staged particles, whose slots form runs
// merged across the block's wavefronts, instead of the few short runs of its
// own particles (short runs cost the vector memory path a write request per
// partial line).  The table is double buffered by the slot loop's two
// register sets, so filling one group's table needs no barrier against the
// clearing of the previous group's.
constexpr bool kStagedScatter = true;
template <typename T, bool ON>
struct ScatterStage {};
template <typename T>
struct ScatterStage<T, true> {
  uint32_t key[2][kBlock];     // the group's distinct keys, hashed (kDeadKey: empty)
  uint32_t cnt[2][kBlock];     // particles of each key
  uint32_t base[kBlock];       // cursor value the block's run of a key starts at
  uint32_t start[kBlock + 1];  // staged position of a key's run; [kBlock]: the total
  double row[6][kBlock];
  T tp[3][kBlock];
  uint32_t orig[kBlock], nkey[kBlock], slot[kBlock];
};

struct StageTicket {
  int h;         // table entry of the lane's key
  uint32_t pos;  // position of the lane's particle in the block's run of that key
};

// Call with every lane of the wave active.  The lanes are grouped by key
// with ballots; each group's first lane finds (or makes) its key's entry
// with LDS compare-and-swap and adds the group's size to its count, whose
// old value places the group within the block's run of that key.  The table
// has kBlock entries, so a group's at most kBlock distinct keys always fit.
template <typename S>
__device__ inline StageTicket stage_insert(S &sg, int par, uint32_t key, bool take) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(take);
  int leader = lane;
  uint32_t rank = 0, count = 0;
  while (pending) {
    const int ld = __ffsll((long long)pending) - 1;
    const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, ld);
    const unsigned long long same = __ballot(take && key == lk) & pending;
    if ((same >> lane) & 1ull) {
      leader = ld;
      rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
    }
    if (lane == ld) count = (uint32_t)__popcll(same);
    pending &= ~same;
  }
  int h = 0;
  uint32_t off = 0;
  if (take && leader == lane) {
    h = (int)(key & (kBlock - 1));
    for (;;) {
      const uint32_t old = atomicCAS(&sg.key[par][h], kDeadKey, key);
      if (old == kDeadKey || old == key) break;
      h = (h + 1) & (kBlock - 1);
    }
    off = atomicAdd(&sg.cnt[par][h], count);
  }
  StageTicket t;
  t.h = __shfl(h, leader, 64);
  t.pos = (uint32_t)__shfl((int)off, leader, 64) + rank;
  return t;
}

// Histogram update: the lanes are grouped by key with ballots, then one
// atomic instruction adds each group's size at its first lane.
__device__ inline void wave_count_add(uint32_t *counters, uint32_t key, bool take) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(take);
  uint32_t count = 0;
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint32_t lk = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
    const unsigned long long same = __ballot(take && key == lk) & pending;
    if (lane == leader) count = (uint32_t)__popcll(same);
    pending &= ~same;
  }
  if (count) atomicAdd(&counters[key], count);
}

// Slot range of a block: contiguous, and the blocks one XCD receives (block
// b runs on XCD b mod 8) cover adjacent ranges, so each XCD's L2 holds the
// field of its own part of the sorted bunch.  gridDim.x is a multiple of 8.
__device__ inline void block_range(int64_t n, int64_t &lo, int64_t &hi) {
  const int nb = gridDim.x;
  const int vb = (int)(blockIdx.x & 7) * (nb >> 3) + (int)(blockIdx.x >> 3);
  const int64_t per = ((n + nb - 1) / nb + kBlock - 1) / kBlock * kBlock;
  lo = (int64_t)vb * per;
  hi = lo + per < n ? lo + per : n;
}

// E of a neighbourhood of the block's current sort tile (kGBoxT nodes
// across the beam, kGBoxL along it), staged in LDS so the 8-node gather of
// the dense core reads LDS instead of the vector L1 (whose tag rate bounds a
// gather over 8 scattered nodes).  Particles whose stencil leaves the box
// read global memory; both paths sum the same values in the same order.
constexpr int kGBoxT = 12;
constexpr int kGBoxL = 12;
constexpr int kGBoxCells = kGBoxT * kGBoxT * kGBoxL;
constexpr int kGBoxLowT = 2;  // nodes below the tile's first node, x and y
constexpr int kGBoxLowL = 2;  // and z

template <typename T> struct GatherLds;
template <> struct GatherLds<double> {
  double2 xy[kGBoxCells];
  double z[kGBoxCells];
};
template <> struct GatherLds<float> {
  float4 e[kGBoxCells];
};

__device__ inline void box_store(GatherLds<double> &g, int q, const E4<double> &v) {
  g.xy[q] = make_double2(v.x, v.y);
  g.z[q] = v.z;
}
__device__ inline void box_store(GatherLds<float> &g, int q, const E4<float> &v) {
  g.e[q] = make_float4(v.x, v.y, v.z, 0.0f);
}
__device__ inline void box_load(const GatherLds<double> &g, int64_t q, double &x,
                                double &y, double &z) {
  const double2 v = g.xy[q];
  x = v.x; y = v.y; z = g.z[q];
}
__device__ inline void box_load(const GatherLds<float> &g, int64_t q, float &x,
                                float &y, float &z) {
  const float4 v = g.e[q];
  x = v.x; y = v.y; z = v.z;
}

// One slot of the pass's input, read an iteration before it is processed.
// The loaded values are kept as loaded (a lane past the range holds its
// block's last slot): nothing between the loads and the next group's first
// use reads them, so the waits for them fall where that group starts.
struct SlotIn {
  uint32_t key, orig;
  uint32_t k0;       // sort tile of the group's first slot (block uniform)
  bool in;           // the lane has a slot of the block's range
  double r[6];
};

// 8-node gather in two batches of four nodes: each batch's loads are issued
// before any of them is used.  ld(q, x, y, z) loads node q.  Same weights and
// summation order as cic8.
template <typename T, typename Load>
__device__ inline void gather_batched(Load ld, int64_t o, int64_t sy, int64_t sz, T ab,
                                      T cd, T ef, T &gx, T &gy, T &gz) {
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
  for (int h = 0; h < 2; h++) {
    T vx[4], vy[4], vz[4];
#pragma unroll
    for (int q = 0; q < 4; q++) ld(idx[4 * h + q], vx[q], vy[q], vz[q]);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      sx += vx[q] * w[4 * h + q];
      syy += vy[q] * w[4 * h + q];
      szz += vz[q] * w[4 * h + q];
    }
  }
  gx = sx;
  gy = syy;
  gz = szz;
}

// One pass finishes step s-1 and starts step s.  pb (kind kb) is the second
// half push of step s-1 (kPushNone on the first pass of a run, which then has
// no kick and no moments); pa (kind ka) the first half push of step s
// (kPushNone on the last pass, which then has no z->t transform).  MODE is
// a set of PassMode bits.  The slot loop is unrolled by two over two
// register sets, so the loads of the next slot group land in the set not
// being worked on.
// Two blocks (wavefronts) per SIMD; three for an in-place pass with an FP32
// grid and compile-time push kinds, which then spills at most 4 VGPRs and
// gains from the extra wavefront (measured: 3%).  FP64 would spill 70 VGPRs
// to scratch at three (32% slower); a scatter pass's LDS (70 KB) allows two.
template <typename T, int MODE, int KB>
constexpr int pass_blocks_per_cu(int ka) {
  return ((MODE & kPassScatter) == 0 && sizeof(T) == 4 && KB != kPushAnyKind &&
          ka != kPushAnyKind) ? 3 : 2;
}

template <typename T, int MODE, int KB, int KA>
__global__ void __launch_bounds__(kBlock, (pass_blocks_per_cu<T, MODE, KB>(KA))) k_pass(PassCfg c) {
  constexpr bool kScatter = (MODE & kPassScatter) != 0, kKeys = (MODE & kPassKeys) != 0;
  const bool has_kick = KB == kPushAnyKind ? c.kb != kPushNone : KB != kPushNone;
  const bool has_conv = KA == kPushAnyKind ? c.ka != kPushNone : KA != kPushNone;
  const int64_t nslots = *c.d_nslots;
  int64_t lo, hi;
  block_range(nslots, lo, hi);
  const int64_t cap = c.cap;
  const double xk = 1.0 / c.xl;
  const double xmin = c.dom->range[0], ymin = c.dom->range[2], zmin = c.dom->range[4];
  const double hx = c.dom->msize[0], hy = c.dom->msize[1], hz = c.dom->msize[2];
  const double hxi = c.dom->hinv[0], hyi = c.dom->hinv[1], hzi = c.dom->hinv[2];
  const bool kick = has_kick && c.scal->current >= (double)1.0e-20f;
  const double txy = c.tau2_prev * c.scal->xycon * c.inv_rho_scale;
  const double ttc = c.tau2_prev * c.scal->tcon * c.inv_rho_scale;
  const int sy = c.nx, sz = c.nx * c.ny;
  const E4<T> *E = (const E4<T> *)c.e;
  T *tpx = (T *)c.tpos;
  __shared__ GatherLds<T> gbox;
  constexpr bool kStage = kScatter && kStagedScatter;
  __shared__ ScatterStage<T, kStage> sg;
  if constexpr (kStage) {
    for (int par = 0; par < 2; par++) {
      sg.key[par][threadIdx.x] = kDeadKey;
      sg.cnt[par][threadIdx.x] = 0u;
    }
    __syncthreads();
  }
  const int bdx = min(kGBoxT, c.nx), bdy = min(kGBoxT, c.ny), bdz = min(kGBoxL, c.nz);
  uint32_t box_key = kDeadKey;   // tile the staged box was built for
  int ox = 0, oy = 0, oz = 0;    // its first node
  PairMom mom;
  pmom_zero(mom);
  double mn[3] = {1e300, 1e300, 1e300}, mxv[3] = {-1e300, -1e300, -1e300};
  uint32_t nlost = 0, nalive = 0, nerr = 0;
  // The input slots are grouped by the tiles of the keys of the last
  // scatter; in_chunks lists those slot ranges in slot order.  The gather
  // box follows the tile of each group's first slot: the block walks the
  // list forward from the chunk holding its first slot.
  // The list is not written during the pass, so it is read through the
  // constant address space: scalar loads, which no vector-memory wait counts.
  using ConstChunk = __attribute__((address_space(4))) const Chunk;
  const ConstChunk *chs = (const ConstChunk *)c.in_chunks;
  int64_t nch = 0, ci = 0, nxt_start = INT64_MAX;
  uint32_t cur_tile = kDeadKey;
  if (has_kick) {
    nch = *(__attribute__((address_space(4))) const int64_t *)c.in_nchunks;
    int64_t a = 0, b = nch;  // last chunk starting at or before lo
    while (b - a > 1) {
      const int64_t m = (a + b) >> 1;
      if ((int64_t)chs[m].start <= lo) a = m; else b = m;
    }
    ci = a;
    if (nch > 0) {
      cur_tile = chs[ci].tile;
      nxt_start = ci + 1 < nch ? (int64_t)chs[ci + 1].start : INT64_MAX;
    }
  }

  // Loads of the slot group starting at base (called while lo < hi).  Every
  // lane loads (an index past the range is clamped to its last slot and the
  // lane is marked empty), so no branch skips the loads and the waits for
  // them and for the atomics before them can be counted exactly.
  auto load = [&](int64_t base, SlotIn &sl) {
    const int64_t j = base + threadIdx.x;
    const int64_t jj = j < hi ? j : hi - 1;
    uint32_t k0 = kDeadKey;
    if (has_kick && base < hi) {
      // the next chunk's start is held, so only a crossing loads
      while (nxt_start <= base) {
        ci++;
        cur_tile = chs[ci].tile;
        nxt_start = ci + 1 < nch ? (int64_t)chs[ci + 1].start : INT64_MAX;
      }
      k0 = cur_tile;
    }
    uint32_t key = 0u, og = 0u;
    if constexpr (kScatter) {
      key = c.in_key[jj];
      og = c.in_orig[jj];
    } else {
      if (c.need_orig) og = c.in_orig[jj];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) sl.r[k] = c.in[k * cap + jj];
    sl.in = j < hi;
    sl.k0 = k0;
    sl.key = key;
    sl.orig = og;
  };

  auto step = [&](SlotIn &sl, uint32_t k0, int64_t base, int par, auto &&prefetch) {
    double &x = sl.r[0], &px = sl.r[1], &y = sl.r[2], &py = sl.r[3], &ph = sl.r[4],
           &pt = sl.r[5];
    // a scatter pass takes the particles the previous pass keyed; a pass in
    // place every particle not marked lost (x = NaN).  A key past the tile
    // count, which must not happen, is counted as an error and the particle
    // dropped, so no cursor outside the scan is touched.
    bool valid;
    if constexpr (kScatter) {
      const bool bad = sl.in && sl.key != kDeadKey && sl.key >= (uint32_t)c.ntiles;
      nerr += bad ? 1u : 0u;
      valid = sl.in && !bad && sl.key != kDeadKey;
    } else {
      valid = sl.in && x == x;
    }
    const uint32_t kin = valid ? sl.key : kDeadKey;
    nalive += valid ? 1u : 0u;
    // counting-sort scatter by the key of the previous pass; the next slot
    // group's loads are issued after the atomic, so waiting for its result
    // does not wait for them
    RankTicket tk{0u, 0, 0u};
    StageTicket stk{0, 0u};
    if constexpr (kStage) stk = stage_insert(sg, par, kin, valid);
    else if constexpr (kScatter) tk = wave_rank_issue(c.cursor, kin, valid);
    prefetch();
    // staged scatter: once every wavefront's keys are in the table, thread t
    // takes the cursor run of entry t (the result is first used after the
    // group's arithmetic) and wavefront 0 places the entries' runs in the
    // staged order (an exclusive scan of their counts in table order)
    uint32_t run_base = 0u;
    bool run_owner = false;
    if constexpr (kStage) {
      __syncthreads();
      const int t = threadIdx.x;
      const uint32_t ek = sg.key[par][t];
      run_owner = ek != kDeadKey;
      if (run_owner) run_base = atomicAdd(&c.cursor[ek], sg.cnt[par][t]);
      if (t < 64) {
        uint32_t a[4], loc = 0u;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          a[u] = sg.cnt[par][4 * t + u];
          loc += a[u];
        }
        uint32_t inc = loc;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const uint32_t o = (uint32_t)__shfl_up((int)inc, d, 64);
          if (t >= d) inc += o;
        }
        uint32_t e = inc - loc;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          sg.start[4 * t + u] = e;
          e += a[u];
        }
        if (t == 63) sg.start[kBlock] = inc;
      }
    }
    const double qm = c.need_orig ? c.qm0[sl.orig] : c.qm_value;
    if (has_kick) {
      // restage the gather box when the slot group starts another tile
      if (kick && k0 != kDeadKey && k0 != box_key) {
        const uint32_t txy = (uint32_t)c.tnx * (uint32_t)c.tny;
        const int tz = (int)(k0 / txy), rem = (int)(k0 % txy);
        const int ty = rem / c.tnx, tx = rem % c.tnx;
        // the tile's first node in the grid of this E field
        const KeyGrid &kg = *c.kgrid;
        ox = min(max(tile_first_node(tx << c.tshift[0], kg.lo[0], kg.h[0], xmin, hxi) -
                     kGBoxLowT, 0), c.nx - bdx);
        oy = min(max(tile_first_node(ty << c.tshift[1], kg.lo[1], kg.h[1], ymin, hyi) -
                     kGBoxLowT, 0), c.ny - bdy);
        oz = min(max(tile_first_node(tz << c.tshift[2], kg.lo[2], kg.h[2], zmin, hzi) -
                     kGBoxLowL, 0), c.nz - bdz);
        __syncthreads();  // the previous box is no longer read
        const int bxy = bdx * bdy, nb = bxy * bdz;
        for (int q0 = threadIdx.x; q0 < nb; q0 += 2 * kBlock) {
          E4<T> ev[2];
#pragma unroll
          for (int u = 0; u < 2; u++) {
            const int q = q0 + u * kBlock;
            E4<T> v{};
            if (q < nb) {
              const int lz = q / bxy, r2 = q - lz * bxy, ly = r2 / bdx, lx = r2 - ly * bdx;
              v = E[(int64_t)(ox + lx) + (int64_t)sy * (oy + ly) + (int64_t)sz * (oz + lz)];
            }
            ev[u] = v;
          }
#pragma unroll
          for (int u = 0; u < 2; u++)
            if (q0 + u * kBlock < nb) box_store(gbox, q0 + u * kBlock, ev[u]);
        }
        __syncthreads();
        box_key = k0;
      }
      if (kick) {
        // cvbkforth1st + scatter1 (src/Appl/BeamBunch.f90:1667-1701,2041-2181).
        // Computed on every lane, so no branch separates the slot's loads from
        // their first use; only valid lanes gather (the node indices are
        // clamped to the grid) and keep the result.
        const double rcp = 1.0 / (-pt + c.gam_k);
        const double betai = sqrt(1.0 - rcp * rcp * (1.0 + px * px + py * py));
        const double X0 = x * c.xl, Y0 = y * c.xl, Z0 = -c.gam_k * betai * ph * c.xl;
        int ix = (int)((X0 - xmin) * hxi + 1.0);
        const double ab = ((xmin - X0) + ix * hx) * hxi;
        int jx = (int)((Y0 - ymin) * hyi + 1.0);
        const double cd = ((ymin - Y0) + jx * hy) * hyi;
        int kx = (int)((Z0 - zmin) * hzi + 1.0);
        const double ef = ((zmin - Z0) + kx * hz) * hzi;
        ix = min(max(ix, 1), c.nx - 1);
        jx = min(max(jx, 1), c.ny - 1);
        kx = min(max(kx, 1), c.nz - 1);
        T gx = (T)0, gy = (T)0, gz = (T)0;
        const int li = ix - 1 - ox, lj = jx - 1 - oy, lk = kx - 1 - oz;
        const bool inbox = box_key != kDeadKey && li >= 0 && li + 1 < bdx && lj >= 0 &&
                           lj + 1 < bdy && lk >= 0 && lk + 1 < bdz;
        if (valid && inbox) {
          gather_batched<T>(
              [](int64_t q, T &ex, T &ey, T &ez) { box_load(gbox, q, ex, ey, ez); },
              (int64_t)li + (int64_t)bdx * (lj + bdy * lk), bdx, bdx * bdy, (T)ab, (T)cd,
              (T)ef, gx, gy, gz);
        } else if (valid) {
          const int64_t o = (int64_t)(ix - 1) + (int64_t)sy * (jx - 1) + (int64_t)sz * (kx - 1);
          gather_batched<T>(
              [E](int64_t q, T &ex, T &ey, T &ez) {
                const E4<T> v = E[q];
                ex = v.x; ey = v.y; ez = v.z;
              },
              o, sy, sz, (T)ab, (T)cd, (T)ef, gx, gy, gz);
        }
        const double exn = (double)gx * c.gam_k, eyn = (double)gy * c.gam_k, ezn = (double)gz;
        if (valid) {
          x = X0 * xk;
          y = Y0 * xk;
          ph = Z0 * xk / (-c.gam_k * betai);
          px = px + txy * exn;
          py = py + txy * eyn;
          pt = pt - ttc * ezn;
        }
      }
      half_push_k<KB>(c.kb, c.pb, qm, x, px, y, py, ph, pt);
      // diagnostic1 sums of the state after step s-1 (all lanes take part)
      pmom_add(mom, valid ? x : 0.0, valid ? px : 0.0, valid ? y : 0.0,
               valid ? py : 0.0, valid ? ph : 0.0, valid ? pt : 0.0);
    }
    uint32_t knew = valid ? 0u : kDeadKey;
    T tx = (T)NAN, ty = (T)NAN, tz = (T)NAN;
    if (has_conv) {
      if (valid) {
        half_push_k<KA>(c.ka, c.pa, qm, x, px, y, py, ph, pt);
        // conv1st_BeamBunch, Flagbc=1 (src/Appl/BeamBunch.f90:1285-1352)
        ph = wrap_phase(ph);
        const double rcp = 1.0 / (-pt + c.gam);
        const double tmpbet = 1.0 - rcp * rcp * (1.0 + px * px + py * py);
        const double betai = sqrt(fabs(tmpbet));
        const double tmpx = x * c.xl, xt = (x - px * ph * rcp) * c.xl;
        const double tmpy = y * c.xl, yt = (y - py * ph * rcp) * c.xl;
        const double zt = -c.gam * betai * ph * c.xl;
        mn[0] = fmin(mn[0], fmin(tmpx, xt));
        mxv[0] = fmax(mxv[0], fmax(tmpx, xt));
        mn[1] = fmin(mn[1], fmin(tmpy, yt));
        mxv[1] = fmax(mxv[1], fmax(tmpy, yt));
        mn[2] = fmin(mn[2], zt);
        mxv[2] = fmax(mxv[2], zt);
        const double radtest = sqrt(xt * xt + yt * yt);
        if (radtest >= c.rad || tmpbet <= 0.0) {
          // lost: marked by x = NaN for the passes that keep it in its slot
          nlost++;
          knew = kDeadKey;
          x = (double)NAN;
        } else {
          if constexpr (kKeys) {
            const int ci = min(max((int)((xt - xmin) * hxi), 0), c.nx - 1);
            const int cj = min(max((int)((yt - ymin) * hyi), 0), c.ny - 1);
            const int ck = min(max((int)((zt - zmin) * hzi), 0), c.nz - 1);
            knew = (uint32_t)((ci >> c.tshift[0]) +
                              c.tnx * ((cj >> c.tshift[1]) + c.tny * (ck >> c.tshift[2])));
          }
          tx = (T)xt;
          ty = (T)yt;
          tz = (T)zt;
        }
      }
    } else if (kKeys && valid) {
      // last pass: the tile of the 0th-order t-frame position the next run
      // starts from (k_prologue_keys' position, on the current domain), so
      // the next run needs no prologue
      const double rcp = 1.0 / (-pt + c.gam);
      const double betai = sqrt(fabs(1.0 - rcp * rcp * (1.0 + px * px + py * py)));
      const double X0 = x * c.xl, Y0 = y * c.xl;
      const double Z0 = -c.gam * betai * wrap_phase(ph) * c.xl;
      const int ci = min(max((int)((X0 - xmin) * hxi), 0), c.nx - 1);
      const int cj = min(max((int)((Y0 - ymin) * hyi), 0), c.ny - 1);
      const int ck = min(max((int)((Z0 - zmin) * hzi), 0), c.nz - 1);
      knew = (uint32_t)((ci >> c.tshift[0]) +
                        c.tnx * ((cj >> c.tshift[1]) + c.tny * (ck >> c.tshift[2])));
    }
    // histogram of the new keys for the next pass's scatter
    if constexpr (kKeys) wave_count_add(c.hist_new, knew, knew != kDeadKey);
    if constexpr (kStage) {
      // every key's cursor run is known; stage each particle at its place
      if (run_owner) sg.base[threadIdx.x] = run_base;
      __syncthreads();
      if (valid) {
        const uint32_t sp = sg.start[stk.h] + stk.pos;
        sg.row[0][sp] = x;
        sg.row[1][sp] = px;
        sg.row[2][sp] = y;
        sg.row[3][sp] = py;
        sg.row[4][sp] = ph;
        sg.row[5][sp] = pt;
        if (has_conv) {
          sg.tp[0][sp] = tx;
          sg.tp[1][sp] = ty;
          sg.tp[2][sp] = tz;
        }
        sg.orig[sp] = sl.orig;
        sg.nkey[sp] = knew;
        sg.slot[sp] = sg.base[stk.h] + stk.pos;
      }
      __syncthreads();
      // thread t stores the t-th staged particle; the group's table is
      // cleared for the group after next
      const int t = threadIdx.x;
      const bool have = (uint32_t)t < sg.start[kBlock];
      const uint32_t q = have ? sg.slot[t] : 0u;
      const bool ok = have && q < (uint32_t)cap;
      nerr += (have && !ok) ? 1u : 0u;
      const int lane = threadIdx.x & 63;
      double *sd = (double *)c.sink + lane;
      uint32_t *su = (uint32_t *)((double *)c.sink + 8 * 64) + lane;
      T *stt = (T *)((double *)c.sink + 10 * 64) + lane;
#pragma unroll
      for (int k = 0; k < 6; k++) (ok ? c.out[k * cap + q] : sd[64 * k]) = sg.row[k][t];
      (ok ? c.out_orig[q] : su[0]) = sg.orig[t];
      if constexpr (kKeys) (ok ? c.out_key[q] : su[64]) = sg.nkey[t];
      if (has_conv) {
        (ok ? tpx[q] : stt[0]) = sg.tp[0][t];
        (ok ? tpx[cap + q] : stt[64]) = sg.tp[1][t];
        (ok ? tpx[2 * cap + q] : stt[128]) = sg.tp[2][t];
      }
      sg.key[par][t] = kDeadKey;
      sg.cnt[par][t] = 0u;
      return;
    }
    uint32_t slot;
    if constexpr (kScatter) slot = wave_rank_slot(tk);
    else slot = (uint32_t)(base + threadIdx.x);
    // Every lane stores: a lane without a particle (or with a slot past the
    // buffer, which must not happen and is counted) writes into the sink, so
    // no branch skips the stores and later waits can be counted exactly.  A
    // pass in place leaves the slots of particles lost earlier as they are.
    const bool ok = valid && slot < (uint32_t)cap;
    nerr += (valid && !ok) ? 1u : 0u;
    const int lane = threadIdx.x & 63;
    double *sd = (double *)c.sink + lane;
    uint32_t *su = (uint32_t *)((double *)c.sink + 8 * 64) + lane;
    T *st = (T *)((double *)c.sink + 10 * 64) + lane;
    (ok ? c.out[slot] : sd[0]) = x;
    (ok ? c.out[cap + slot] : sd[64]) = px;
    (ok ? c.out[2 * cap + slot] : sd[128]) = y;
    (ok ? c.out[3 * cap + slot] : sd[192]) = py;
    (ok ? c.out[4 * cap + slot] : sd[256]) = ph;
    (ok ? c.out[5 * cap + slot] : sd[320]) = pt;
    if constexpr (kScatter) (ok ? c.out_orig[slot] : su[0]) = sl.orig;
    // In place, the key of EVERY slot of the block's range is rewritten: the
    // slot of a particle lost before this pass takes kDeadKey, so the next
    // scatter, which trusts the keys, drops it.
    if constexpr (kKeys) {
      const bool kok = kScatter ? ok : base + (int64_t)threadIdx.x < hi;
      (kok ? c.out_key[slot] : su[64]) = knew;
    }
    if (has_conv) {
      (ok ? tpx[slot] : st[0]) = tx;
      (ok ? tpx[cap + slot] : st[64]) = ty;
      (ok ? tpx[2 * cap + slot] : st[128]) = tz;
    }
  };

  // The first group is peeled off, so both paths into the loop end the same
  // way (a group's stores after the next group's loads) and the wait at the
  // top of the loop can be counted exactly.
  SlotIn sa, sb;
  if (lo < hi) {
    load(lo, sa);
    step(sa, (uint32_t)__builtin_amdgcn_readfirstlane((int)sa.k0), lo, 0,
         [&] { load(lo + kBlock, sb); });
    for (int64_t base = lo + kBlock; base < hi; base += 2 * kBlock) {
      step(sb, (uint32_t)__builtin_amdgcn_readfirstlane((int)sb.k0), base, 1,
           [&] { load(base + kBlock, sa); });
      if (base + kBlock >= hi) break;
      step(sa, (uint32_t)__builtin_amdgcn_readfirstlane((int)sa.k0), base + kBlock, 0,
           [&] { load(base + 2 * kBlock, sb); });
    }
  }

  if (nerr) atomicAdd((unsigned long long *)&c.err->err, (unsigned long long)nerr);
  __shared__ double sh[kBlock / 64][kMoments];
  // partials are quantity-major: value q of block b at part[q * kMaxRedBlocks + b]
  double *out = c.part + blockIdx.x;
  if (has_kick) {
    pmom_block_store(mom, sh, out, kMaxRedBlocks);
  } else if (threadIdx.x < kMoments) {
    out[(int64_t)threadIdx.x * kMaxRedBlocks] = 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double v[kConvRed];
  for (int k = 0; k < 3; k++) { v[k] = wave_min(mn[k]); v[3 + k] = wave_max(mxv[k]); }
  v[6] = wave_sum((double)nlost);
  v[7] = wave_sum((double)nalive);
  if (lane == 0)
    for (int k = 0; k < kConvRed; k++) sh[wv][k] = v[k];
  __syncthreads();
  if (threadIdx.x < kConvRed) {
    int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < kBlock / 64; w++)
      r = k < 3 ? fmin(r, sh[w][k]) : (k < 6 ? fmax(r, sh[w][k]) : r + sh[w][k]);
    out[(int64_t)(kMoments + k) * kMaxRedBlocks] = r;
  }
}

template <typename T, int KB, int KA>
void launch_pass_kinds(int mode, int nb, hipStream_t st, const PassCfg &pc) {
  const dim3 g(nb), b(kBlock);
  switch (mode) {
    case kPassInPlace:
      hipLaunchKernelGGL((k_pass<T, kPassInPlace, KB, KA>), g, b, 0, st, pc);
      break;
    case kPassKeys: hipLaunchKernelGGL((k_pass<T, kPassKeys, KB, KA>), g, b, 0, st, pc); break;
    case kPassScatter:
      hipLaunchKernelGGL((k_pass<T, kPassScatter, KB, KA>), g, b, 0, st, pc);
      break;
    default:
      hipLaunchKernelGGL((k_pass<T, kPassScatter | kPassKeys, KB, KA>), g, b, 0, st, pc);
      break;
  }
}

// Drift, map and none as compile-time kinds (3 x 3); the run-time form
// for the rest (the coupled map, the K-mode quadrupole).
template <typename T, int KB>
void launch_pass_b(int mode, int nb, hipStream_t st, const PassCfg &pc) {
  switch (pc.ka) {
    case kPushNone: launch_pass_kinds<T, KB, kPushNone>(mode, nb, st, pc); break;
    case kPushDrift: launch_pass_kinds<T, KB, kPushDrift>(mode, nb, st, pc); break;
    default: launch_pass_kinds<T, KB, kPushMap>(mode, nb, st, pc); break;
  }
}

template <typename T>
void launch_pass(int mode, int nb, hipStream_t st, const PassCfg &pc) {
  const auto simple = [](int k) { return k == kPushNone || k == kPushDrift || k == kPushMap; };
  if (!simple(pc.kb) || !simple(pc.ka) || pc.need_orig) {
    launch_pass_kinds<T, kPushAnyKind, kPushAnyKind>(mode, nb, st, pc);
    return;
  }
  switch (pc.kb) {
    case kPushNone: launch_pass_b<T, kPushNone>(mode, nb, st, pc); break;
    case kPushDrift: launch_pass_b<T, kPushDrift>(mode, nb, st, pc); break;
    default: launch_pass_b<T, kPushMap>(mode, nb, st, pc); break;
  }
}

// ------------------------------------------------------------ prologue
// Initial tile keys: the 0th-order t-frame position of the uploaded z-frame
// particles, binned on a domain that bounds them.
__global__ void __launch_bounds__(kBlock)
k_prologue_range(const double *__restrict__ X, const double *__restrict__ PX,
                 const double *__restrict__ Y, const double *__restrict__ PY,
                 const double *__restrict__ PH, const double *__restrict__ PT,
                 int64_t n, double gam, double xl, double *__restrict__ part) {
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double px = PX[i], py = PY[i];
    double rcp = 1.0 / (-PT[i] + gam);
    double betai = sqrt(fabs(1.0 - rcp * rcp * (1.0 + px * px + py * py)));
    double p[3] = {X[i] * xl, Y[i] * xl, -gam * betai * wrap_phase(PH[i]) * xl};
    for (int a = 0; a < 3; a++) { mn[a] = fmin(mn[a], p[a]); mx[a] = fmax(mx[a], p[a]); }
  }
  __shared__ double sh[kBlock / 64][6];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double v[6];
  for (int a = 0; a < 3; a++) { v[a] = wave_min(mn[a]); v[3 + a] = wave_max(mx[a]); }
  if (lane == 0)
    for (int k = 0; k < 6; k++) sh[wv][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    double r = sh[0][k];
    for (int w = 1; w < kBlock / 64; w++) r = k < 3 ? fmin(r, sh[w][k]) : fmax(r, sh[w][k]);
    part[blockIdx.x * 6 + k] = r;
  }
}

__global__ void __launch_bounds__(kBlock)
k_prologue_keys(const double *__restrict__ X, const double *__restrict__ PX,
                const double *__restrict__ Y, const double *__restrict__ PY,
                const double *__restrict__ PH, const double *__restrict__ PT,
                int64_t n, double gam, double xl, const double *__restrict__ box,
                int nx, int ny, int nz, int s0, int s1, int s2, int tnx, int tny,
                uint32_t *__restrict__ key, uint32_t *__restrict__ hist) {
  // box = reduced (min x, min y, min z, max x, max y, max z)
  const double h[3] = {(box[3] - box[0]) / nx, (box[4] - box[1]) / ny,
                       (box[5] - box[2]) / nz};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    uint32_t k = kDeadKey;
    if (i < n) {
      double px = PX[i], py = PY[i];
      double rcp = 1.0 / (-PT[i] + gam);
      double betai = sqrt(fabs(1.0 - rcp * rcp * (1.0 + px * px + py * py)));
      double p[3] = {X[i] * xl, Y[i] * xl, -gam * betai * wrap_phase(PH[i]) * xl};
      int ci[3];
      const int nn[3] = {nx, ny, nz};
      for (int a = 0; a < 3; a++) {
        double u = h[a] > 0 ? (p[a] - box[a]) / h[a] : 0.0;
        ci[a] = min(max((int)u, 0), nn[a] - 1);
      }
      k = (uint32_t)((ci[0] >> s0) + tnx * ((ci[1] >> s1) + tny * (ci[2] >> s2)));
      key[i] = k;
    }
    wave_count_add(hist, k, k != kDeadKey);
  }
}

// Provisional domain of the initial keys: the padded box of the bunch
// (update_CompDom's padding, src/Appl/CompDom.f90:547-557,614-622).
__global__ void k_prologue_dom(const double *__restrict__ box, DevDom *__restrict__ d,
                               int nx, int ny, int nz) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int nn[3] = {nx, ny, nz};
  for (int a = 0; a < 3; a++) {
    const double eps = (double)(1.0f / ((float)nn[a] - 3.0f));
    const double lo = box[a], hi = box[3 + a];
    d->range[2 * a] = lo - eps * (hi - lo);
    d->range[2 * a + 1] = hi + eps * (hi - lo);
    d->msize[a] = (d->range[2 * a + 1] - d->range[2 * a]) / (nn[a] - 1);
    d->hinv[a] = 1.0 / d->msize[a];
    d->ptrange[2 * a] = lo;
    d->ptrange[2 * a + 1] = hi;
  }
}

// ------------------------------------------------------------ finalize
struct FinCfg {
  const double *part;
  int nparts;
  int mode;                 // 0 prologue, 1 step, 2 final
  DevScal *scal;
  DevDom *dom;
  double *moments;          // the run's records (kRunRecord doubles per step)
  int want_moments;         // write the record of the step just finished
  const uint32_t *hist;     // histogram of the keys just written
  uint32_t *cursor;         // -> exclusive offsets of hist
  Chunk *chunks;            // -> deposit chunks of that grouping
  int64_t *nchunks;         // -> their number
  uint32_t *hist_zero;      // histogram the next pass fills
  int scan;                 // the pass computed keys: scan their histogram
  int inplace;              // the pass kept every particle in its slot
  int kg;                   // grouping whose keys were just written (dom->kgrid)
  int ntiles;
  int nx, ny, nz;
  double gam, scfreq, clight, pi, scxl, mass, charge;
  const double *prologue_box;  // mode 0: the box k_prologue_keys used
};

__global__ void __launch_bounds__(1024) k_finalize(FinCfg f) {
  __shared__ double red[kPartStride];
  __shared__ uint32_t ssum[1024];
  __shared__ uint32_t csum[1024];
  const int tid = threadIdx.x;
  // ---- reduce the pass partials: one wavefront per quantity, fixed order
  if (f.mode != 0) {
    const int lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
    for (int q = wv; q < kPartStride; q += nw) {
      const bool isconv = q >= kMoments;
      const int cq = q - kMoments;
      const int op = !isconv ? (q < 37 ? 0 : 2) : (cq < 3 ? 1 : (cq < 6 ? 2 : 0));
      const double ident = op == 1 ? 1e300 : (op == 2 ? -1e300 : 0.0);
      double acc = ident;
      // eight loads in flight per lane, combined in slot order
      constexpr int kU = 8;
      const double *src = f.part + (int64_t)q * kMaxRedBlocks;
      for (int b0 = lane; b0 < f.nparts; b0 += 64 * kU) {
        double v[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
          const int b = b0 + 64 * u;
          v[u] = b < f.nparts ? src[b] : ident;
        }
#pragma unroll
        for (int u = 0; u < kU; u++)
          acc = op == 1 ? fmin(acc, v[u]) : (op == 2 ? fmax(acc, v[u]) : acc + v[u]);
      }
      acc = op == 1 ? wave_min(acc) : (op == 2 ? wave_max(acc) : wave_sum(acc));
      if (lane == 0) red[q] = acc;
    }
    __syncthreads();
  }
  if (tid == 0) {
    DevScal &S = *f.scal;
    // the keys the pass just wrote were computed in the domain as it stands
    // before this step's update; k_prologue_keys used the unpadded box
    if (f.scan) {
      DevDom &D = *f.dom;
      const int nn[3] = {f.nx, f.ny, f.nz};
      for (int a = 0; a < 3; a++) {
        const double *pb = f.prologue_box;
        D.kgrid[f.kg].lo[a] = f.mode == 0 ? pb[a] : D.range[2 * a];
        D.kgrid[f.kg].h[a] = f.mode == 0 ? (pb[3 + a] - pb[a]) / nn[a] : D.msize[a];
      }
    }
    if (f.mode != 0 && f.want_moments) {
      double *o = f.moments + S.rec * kRunRecord;
      for (int k = 0; k < kMoments; k++) o[k] = red[k];
      o[kMoments] = red[kMoments + 7];  // Npt of those moments
      S.rec++;
    }
    if (f.mode == 1) {
      const double *cv = red + kMoments;
      const int64_t alive_in = (int64_t)cv[7];
      const int64_t nnew = alive_in - (int64_t)cv[6];
      // a scatter writes one slot per particle it took; in place the slots stay
      if (!f.inplace) S.nslots = alive_in;
      S.nalive = nnew;
      // conv1st: Current = Current*nptot/Npt (src/Appl/BeamBunch.f90:1657-1663)
      if (S.npt > 0) S.current = S.current * (double)nnew / S.npt;
      S.npt = (double)nnew;
      if (nnew > 0) {
        // update_CompDom, Flagbc=1 (src/Appl/CompDom.f90:547-557,614-622)
        const double pr[6] = {cv[0], cv[3], cv[1], cv[4], cv[2], cv[5]};
        const int nn[3] = {f.nx, f.ny, f.nz};
        DevDom &D = *f.dom;
        for (int a = 0; a < 3; a++) {
          const double eps = (double)(1.0f / ((float)nn[a] - 3.0f));
          const double lo = pr[2 * a], hi = pr[2 * a + 1];
          D.range[2 * a] = lo - eps * (hi - lo);
          D.range[2 * a + 1] = hi + eps * (hi - lo);
          D.msize[a] = (D.range[2 * a + 1] - D.range[2 * a]) / (nn[a] - 1);
          D.hinv[a] = 1.0 / D.msize[a];
          D.ptrange[2 * a] = lo;
          D.ptrange[2 * a + 1] = hi;
        }
      }
      // scatter1 constants (src/Appl/BeamBunch.f90:2065-2084)
      const double gam = f.gam, curr = S.current;
      const double tmpscale = curr / f.scfreq;
      const double twopi = 2.0 * f.pi;
      const double fpei = f.clight * f.clight * (double)1.0e-7f;
      const double xk = 1.0 / f.scxl;
      const double bet = sqrt(gam * gam - 1.0) / gam;
      const double bbyk = bet / xk;
      const double gambet = gam * bet;
      const double brho = gambet / (f.clight * f.charge) * f.mass;
      const double vz0 = bet * f.clight;
      const double perv0 = 2.0 * curr * fpei / (brho * vz0 * vz0 * gam * gam);
      S.xycon = curr > 0 ? 0.5 * perv0 * gambet * bbyk * twopi / tmpscale : 0.0;
      S.tcon = bet * S.xycon * gam * gam;
    } else if (f.mode == 2) {
      const int64_t alive_in = (int64_t)red[kMoments + 7];
      S.nslots = alive_in;
      S.nalive = alive_in;
    } else {
      // prologue: the input slots are the uploaded particles
      S.nslots = S.nalive;
    }
  }
  __syncthreads();
  if (!f.scan) return;
  // ---- exclusive scan of the tile histogram and the chunk list (after the
  // last pass: of the keys the next run starts from)
  const int per = (f.ntiles + blockDim.x - 1) / blockDim.x;
  const int t0 = tid * per, t1 = min(t0 + per, f.ntiles);
  uint32_t cnt = 0, nch = 0;
  for (int t = t0; t < t1; t++) {
    uint32_t h = f.hist[t];
    cnt += h;
    nch += (h + kChunk - 1) / kChunk;
  }
  ssum[tid] = cnt;
  csum[tid] = nch;
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    uint32_t a = tid >= off ? ssum[tid - off] : 0u;
    uint32_t b2 = tid >= off ? csum[tid - off] : 0u;
    __syncthreads();
    ssum[tid] += a;
    csum[tid] += b2;
    __syncthreads();
  }
  uint32_t off = ssum[tid] - cnt, coff = csum[tid] - nch;
  for (int t = t0; t < t1; t++) {
    uint32_t h = f.hist[t];
    f.cursor[t] = off;
    for (uint32_t s = 0; s < h; s += kChunk) {
      f.chunks[coff].start = off + s;
      f.chunks[coff].len = min((uint32_t)kChunk, h - s);
      f.chunks[coff].tile = (uint32_t)t;
      coff++;
    }
    off += h;
  }
  if (tid == blockDim.x - 1) *f.nchunks = csum[tid];
  __syncthreads();  // hist_zero may be hist itself
  for (int t = tid; t < f.ntiles; t += blockDim.x) f.hist_zero[t] = 0u;
}

__global__ void k_zero_u32(uint32_t *__restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

}  // namespace

int bunch_alloc_fused(izh_bunch *b, int ntiles, int64_t nseg) {
  const size_t ps = b->precision == IZH_FP64 ? sizeof(double) : sizeof(float);
  if (!b->key) {
    IZH_HIP(hipMalloc((void **)&b->key, (size_t)b->cap * sizeof(uint32_t)));
    IZH_HIP(hipMalloc((void **)&b->key_alt, (size_t)b->cap * sizeof(uint32_t)));
    IZH_HIP(hipMalloc(&b->tpos, (size_t)3 * b->cap * ps));
    IZH_HIP(hipMalloc(&b->sink, kSinkBytes));
    IZH_HIP(hipMalloc((void **)&b->d_scal, sizeof(DevScal)));
  }
  if (!b->r_alt[0]) IZH_CHECK(bunch_alloc_alt(b));
  if (!b->h_scal)
    IZH_HIP(hipHostMalloc((void **)&b->h_scal, sizeof(DevScal), hipHostMallocDefault));
  const int64_t nch = b->cap / kChunk + ntiles + 1;
  if (ntiles > b->ntiles_cap || nch > b->nchunks_cap) {
    for (int k = 0; k < 2; k++)
      if (b->hist[k]) (void)hipFree(b->hist[k]);
    if (b->cursor) (void)hipFree(b->cursor);
    if (b->chunks) (void)hipFree(b->chunks);
    for (int k = 0; k < 2; k++)
      IZH_HIP(hipMalloc((void **)&b->hist[k], (size_t)ntiles * sizeof(uint32_t)));
    IZH_HIP(hipMalloc((void **)&b->cursor, (size_t)ntiles * sizeof(uint32_t)));
    IZH_HIP(hipMalloc((void **)&b->chunks, (size_t)2 * nch * sizeof(Chunk)));
    b->ntiles_cap = ntiles;
    b->nchunks_cap = nch;
  }
  if (nseg > b->moments_cap) {
    if (b->d_moments) (void)hipFree(b->d_moments);
    IZH_HIP(hipMalloc((void **)&b->d_moments, (size_t)nseg * kRunRecord * sizeof(double)));
    b->moments_cap = nseg;
  }
  return IZH_OK;
}

// Host-side descriptor of a segment's half-step push; refpt and z follow the
// reference's bookkeeping (drift1_BeamBunch and the element's maplinear).
static int segment_push(izh_bunch *b, const izh_segment &sg, HalfPush &p, int &kind,
                        double *z) {
  p = HalfPush{};
  const char *why = nullptr;
  const int m1 = map1_kind(sg.e.bitype, sg.e.v, &why);
  if (m1 == kMap1Drift) {
    const double gam0 = -b->refpt[5];
    const double beta0 = std::sqrt(gam0 * gam0 - 1.0) / gam0;
    kind = kPushDrift;
    p.tau_xl = sg.tau1 / b->c.scxl;
    p.gam0 = gam0;
    p.phase = sg.tau1 / (b->c.scxl * beta0);
    b->refpt[4] = b->refpt[4] + sg.tau1 / (b->c.scxl * beta0);
  } else if (m1 == kMap1Linear) {
    double xm[36];
    IZH_CHECK(izh_element_maplinear(&b->c, *z, sg.tau1, &sg.e, b->refpt, b->charge,
                                    b->mass, xm));
    // the longitudinal plane is never coupled to the transverse ones
    for (int j = 0; j < 6; j++)
      for (int k = 0; k < 6; k++)
        if ((j < 4) != (k < 4) && xm[6 * j + k] != 0.0)
          return fail(IZH_ERR_STATE, "izh_run: element map couples phi/pt to x/y");
    bool coupled = false;
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 4; k++)
        if (j / 2 != k / 2 && xm[6 * j + k] != 0.0) coupled = true;
    if (coupled) {
      kind = kPushMapC;
      for (int j = 0; j < 4; j++)
        for (int k = 0; k < 4; k++) p.m[4 * j + k] = xm[6 * j + k];
      p.m[16] = xm[6 * 4 + 4];
      p.m[17] = xm[6 * 4 + 5];
      p.m[18] = xm[6 * 5 + 4];
      p.m[19] = xm[6 * 5 + 5];
    } else {
      kind = kPushMap;
      for (int blk = 0; blk < 3; blk++) {
        const int r = 2 * blk;
        p.m[4 * blk + 0] = xm[6 * r + r];
        p.m[4 * blk + 1] = xm[6 * r + r + 1];
        p.m[4 * blk + 2] = xm[6 * (r + 1) + r];
        p.m[4 * blk + 3] = xm[6 * (r + 1) + r + 1];
      }
    }
  } else if (m1 == kMap1QuadK) {
    // transfmapK_Quadrupole (src/Appl/Quadrupole.f90:509-607), per particle
    const double r6 = b->refpt[5];
    kind = kPushQuadK;
    p.m[0] = sg.tau1;
    p.m[1] = sg.e.v[0];
    p.m[2] = b->charge / b->mass;
    p.m[3] = -r6;
    p.m[4] = std::sqrt(r6 * r6 - 1.0);
    p.m[5] = std::sqrt(1.0 - 1.0 / (r6 * r6));
    p.m[6] = b->c.scxl;
    b->refpt[4] = b->refpt[4] + sg.tau1 / p.m[5] / b->c.scxl;
  } else {
    return fail(IZH_ERR_ARG, "izh_run: bitype %d: %s", sg.e.bitype, why);
  }
  *z = *z + sg.tau1;
  return IZH_OK;
}

// A run of steps on one bunch and field, pass by pass: begin() sets up the
// device scalars and the initial sort keys, pass() makes one pass (and, but
// for the last, the charge, field solve and E of the step it starts), end()
// copies the results back.  izh_run drives it through all its passes; a
// lattice-scan job replays captured graphs of its passes (dry passes keep the
// host bookkeeping in step with the replays).
struct Runner {
  izh_bunch *b = nullptr;
  izh_field *f = nullptr;
  hipStream_t st = nullptr, fstream = nullptr;
  // G^ of a step needs only the step's domain: with fork_green it runs on
  // the field's own stream (idle during the run) beside the deposit, forked
  // after the pass and joined before the convolution.  Lattice-scan jobs
  // keep to one stream: with many jobs in flight the device is already
  // full, and a second stream per job only adds contention for the
  // hardware queues (served scan p50 13.5 -> 17.9 s with it).
  bool fork_green = true;
  hipEvent_t ev_dom = nullptr, ev_green = nullptr;
  int nx = 0, ny = 0, nz = 0, s0 = 0, s1 = 0, s2 = 0, tnx = 0, tny = 0, ntiles = 0;
  int nb = 0, period = 1;
  int g_cur = 1, g_next = 0;
  int64_t total = 0;   // steps of the run
  int64_t s = 0;       // index of the next pass
  PassCfg pc{};
  FinCfg fc{};
  HalfPush pb_prev{};
  int kb_prev = kPushNone;
  double tau2_prev = 0.0;
  double gam_sc = 0.0;             // -refpt(6) during the space-charge stages of step s-1
  std::vector<double> refpt_after; // refpt(5:6) after each step
  bool began = false;

  bool scatters(int64_t k) const { return k == 0 || k == total || k % period == 0; }

  int begin(izh_bunch *b_, izh_field *f_, int64_t nsteps) {
    b = b_;
    f = f_;
    total = nsteps;
    nx = f->nx; ny = f->ny; nz = f->nz;
    s0 = tile_shift(nx); s1 = tile_shift(ny); s2 = tile_shift(nz);
    tnx = nx >> s0; tny = ny >> s1;
    const int tnz = nz >> s2;
    ntiles = tnx * tny * tnz;
    IZH_CHECK(bunch_alloc_fused(b, ntiles, total));
    st = b->stream;
    // everything runs on the bunch's stream, after work queued on the field's
    {
      hipEvent_t ev;
      IZH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      IZH_HIP(hipEventRecord(ev, f->stream));
      IZH_HIP(hipStreamWaitEvent(st, ev, 0));
      IZH_HIP(hipEventDestroy(ev));
    }
    fstream = f->stream;
    f->stream = st;
    began = true;
    IZH_HIP(hipEventCreateWithFlags(&ev_dom, hipEventDisableTiming));
    IZH_HIP(hipEventCreateWithFlags(&ev_green, hipEventDisableTiming));
    const double gam = -b->refpt[5];
    gam_sc = gam;
    const double wsum = b->wabs_sum > 0 ? b->wabs_sum : 1.0;
    const double rho_scale = f->precision == IZH_FP64 ? 1.0 : 1.0 / wsum;
    nb = (grid_for(b->n) + 7) / 8 * 8;  // multiple of 8 for block_range
    period = b->resort_period > 0 ? b->resort_period : kDefaultResortPeriod;
    refpt_after.assign(2 * (size_t)total, 0.0);
    // Two chunk lists (with their key grids and counts) describe groupings
    // of the slots: g_cur the order the particles are in, g_next the order
    // the keys of the last scan put them in at the next scatter.  The
    // previous run on this bunch and field left its last scan (keys,
    // cursors, chunks) in slot b->keys_po: start from it.
    const bool warm = b->keys_valid && b->keys_field == f && b->keys_grid[0] == nx &&
                      b->keys_grid[1] == ny && b->keys_grid[2] == nz;
    g_next = warm ? b->keys_po : 0;
    g_cur = 1 - g_next;
    b->keys_valid = 0;
    uint32_t *hist = b->hist[0];
    // ---- device scalars
    DevScal hs{};
    if (warm) {
      hs.nchunks[0] = b->h_scal->nchunks[0];
      hs.nchunks[1] = b->h_scal->nchunks[1];
    }
    hs.current = b->current;
    hs.npt = (double)b->npt;
    hs.nalive = b->n;
    hs.nslots = b->n;
    *b->h_scal = hs;
    IZH_HIP(hipMemcpyAsync(b->d_scal, b->h_scal, sizeof hs, hipMemcpyHostToDevice, st));
    fc = FinCfg{};
    fc.part = b->d_part;
    fc.nparts = nb;
    fc.scal = b->d_scal;
    fc.dom = f->d_dom;
    fc.moments = b->d_moments;
    fc.ntiles = ntiles;
    fc.nx = nx; fc.ny = ny; fc.nz = nz;
    fc.gam = gam; fc.scfreq = b->c.scfreq; fc.clight = b->c.clight; fc.pi = b->c.pi;
    fc.scxl = b->c.scxl; fc.mass = b->mass; fc.charge = b->charge;
    fc.cursor = b->cursor;
    fc.hist = hist;
    fc.hist_zero = hist;
    // ---- prologue: initial keys, their histogram, cursors and chunks
    if (!warm) {
      hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(ntiles)), dim3(kBlock), 0, st, hist,
                         (int64_t)ntiles);
      hipLaunchKernelGGL(k_prologue_range, dim3(nb), dim3(kBlock), 0, st, b->r[0],
                         b->r[1], b->r[2], b->r[3], b->r[4], b->r[5], b->n, gam,
                         b->c.scxl, b->d_part);
      IZH_CHECK(launch_reduce(st, b->d_part, nb, 6, 0x7ull, 0x38ull, b->d_out));
      hipLaunchKernelGGL(k_prologue_keys, dim3(nb), dim3(kBlock), 0, st, b->r[0],
                         b->r[1], b->r[2], b->r[3], b->r[4], b->r[5], b->n, gam,
                         b->c.scxl, b->d_out, nx, ny, nz, s0, s1, s2, tnx, tny,
                         b->key, hist);
      hipLaunchKernelGGL(k_prologue_dom, dim3(1), dim3(64), 0, st, b->d_out, f->d_dom,
                         nx, ny, nz);
      IZH_HIP(hipGetLastError());
      fc.mode = 0;
      fc.scan = 1;
      fc.inplace = 0;
      fc.kg = g_next;
      fc.prologue_box = b->d_out;
      fc.chunks = b->chunks + g_next * b->nchunks_cap;
      fc.nchunks = &b->d_scal->nchunks[g_next];
      {
        ScopedTimer tm_(IZH_T_FINAL, st);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, fc);
      }
      IZH_HIP(hipGetLastError());
    }
    pc = PassCfg{};
    pc.cap = b->cap;
    pc.d_nslots = &b->d_scal->nslots;
    pc.cursor = b->cursor;
    pc.hist_new = hist;
    pc.dom = f->d_dom;
    pc.e = f->e;
    pc.scal = b->d_scal;
    pc.err = b->d_scal;
    pc.nx = nx; pc.ny = ny; pc.nz = nz;
    pc.tshift[0] = s0; pc.tshift[1] = s1; pc.tshift[2] = s2;
    pc.tnx = tnx; pc.tny = tny;
    pc.ntiles = ntiles;
    pc.gam = gam;
    pc.gam_k = gam;
    pc.xl = b->c.scxl;
    pc.inv_rho_scale = 1.0 / rho_scale;
    pc.tpos = b->tpos;
    pc.part = b->d_part;
    pc.sink = b->sink;
    pc.qm0 = b->qm0;
    pc.qm_value = b->qm_value;
    s = 0;
    kb_prev = kPushNone;
    pb_prev = HalfPush{};
    tau2_prev = 0.0;
    return IZH_OK;
  }

  // Pass s: it finishes step s-1 and starts step s (seg = the segment of
  // step s, null for the last pass).  Passes 0 and total and every
  // period-th pass between scatter the particles into tile order; the others
  // write them back in place.  A pass before a scatter computes the keys it
  // scatters by.  dry: the host bookkeeping only (a replayed graph launched
  // the kernels).
  int pass(const izh_segment *seg, double *z, bool dry = false) {
    const bool last = seg == nullptr;
    const bool scatter = scatters(s);
    const bool keys = last || scatters(s + 1);
    HalfPush pa{};
    int ka = kPushNone;
    if (!last) IZH_CHECK(segment_push(b, *seg, pa, ka, z));  // first half of step s
    // the kick of step s-1 and the z->t transform of step s each use the
    // reference energy of their own step (after its first half push)
    pc.gam_k = gam_sc;
    pc.gam = -b->refpt[5];
    gam_sc = pc.gam;
    fc.gam = pc.gam;
    pc.in = b->r[0];
    pc.out = scatter ? b->r_alt[0] : b->r[0];
    pc.in_orig = b->orig;
    pc.out_orig = b->orig_alt;
    pc.in_key = b->key;
    pc.out_key = scatter ? b->key_alt : b->key;
    // the input slots are grouped as g_cur says (the last scatter's keys)
    pc.in_chunks = b->chunks + g_cur * b->nchunks_cap;
    pc.in_nchunks = &b->d_scal->nchunks[g_cur];
    pc.kgrid = &f->d_dom->kgrid[g_cur];
    pc.kb = kb_prev;
    pc.ka = ka;
    pc.pb = pb_prev;
    pc.pa = pa;
    pc.need_orig = (kb_prev == kPushQuadK || ka == kPushQuadK) && !b->qm_uniform;
    pc.tau2_prev = tau2_prev;
    pc.rad = last ? 0.0 : 0.5 * (seg->xrad + seg->yrad);
    const int mode = (scatter ? kPassScatter : 0) | (keys ? kPassKeys : 0);
    if (!dry) {
      ScopedTimer tm_(IZH_T_PASS, st);
      if (f->precision == IZH_FP64)
        launch_pass<double>(mode, nb, st, pc);
      else
        launch_pass<float>(mode, nb, st, pc);
      IZH_HIP(hipGetLastError());
    }
    if (scatter) {
      for (int r = 0; r < 6; r++) std::swap(b->r[r], b->r_alt[r]);
      std::swap(b->orig, b->orig_alt);
      std::swap(b->key, b->key_alt);
      g_cur = g_next;   // the slots now follow the scanned keys
    }
    fc.mode = last ? 2 : 1;
    fc.scan = keys ? 1 : 0;
    fc.inplace = scatter ? 0 : 1;
    fc.want_moments = s > 0;
    if (keys) {
      g_next = 1 - g_cur;
      fc.kg = g_next;
      fc.chunks = b->chunks + g_next * b->nchunks_cap;
      fc.nchunks = &b->d_scal->nchunks[g_next];
    }
    if (!dry) {
      ScopedTimer tm_(IZH_T_FINAL, st);
      hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, fc);
      IZH_HIP(hipGetLastError());
    }
    s++;
    if (last) return IZH_OK;
    const int64_t step = s - 1;
    if (!dry) {
      // charge, update3O and the E field of step s on the grouping the slots are in
      const Chunk *dch = b->chunks + g_cur * b->nchunks_cap;
      if (fork_green) {
        IZH_HIP(hipEventRecord(ev_dom, st));
        IZH_HIP(hipStreamWaitEvent(fstream, ev_dom, 0));
        IZH_CHECK(launch_green(f, fstream));
        IZH_HIP(hipEventRecord(ev_green, fstream));
      } else {
        IZH_CHECK(launch_green(f, st));
      }
      IZH_CHECK(launch_deposit(b, f, dch, &f->d_dom->kgrid[g_cur],
                               &b->d_scal->nchunks[g_cur], b->nchunks_cap));
      if (fork_green) IZH_HIP(hipStreamWaitEvent(st, ev_green, 0));
      IZH_CHECK(launch_convolve(f));
      IZH_CHECK(launch_efield(f));
      // the step's wakefield (src/Contrl/AccSimulator.f90:1272-1338) from the
      // same t positions, into E before the next pass kicks
      if (seg->wake && seg->wake->kind != 0) {
        WakeParams wp;
        IZH_CHECK(wake_params_of(b, f, seg->wake, &wp));
        IZH_CHECK(launch_wake(b, f, wp, b->d_scal, dch, &b->d_scal->nchunks[g_cur],
                              b->nchunks_cap));
        IZH_CHECK(launch_wake_add(f));
      }
    }
    // second half of the step: host bookkeeping now, applied by the next pass
    IZH_CHECK(segment_push(b, *seg, pb_prev, kb_prev, z));
    refpt_after[2 * step] = b->refpt[4];
    refpt_after[2 * step + 1] = b->refpt[5];
    tau2_prev = 2.0 * seg->tau1;
    return IZH_OK;
  }

  // Give the field its own stream back (after an error too).
  void release() {
    if (began) f->stream = fstream;
    began = false;
    if (ev_dom) (void)hipEventDestroy(ev_dom);
    if (ev_green) (void)hipEventDestroy(ev_green);
    ev_dom = ev_green = nullptr;
  }

  // After the last pass: wait for the stream and take the results.
  int end(izh_compdom *dom, double *moments) {
    release();
    IZH_HIP(hipMemcpyAsync(b->h_scal, b->d_scal, sizeof(DevScal), hipMemcpyDeviceToHost,
                           st));
    IZH_HIP(hipMemcpyAsync(f->h_dom, f->d_dom, sizeof(DevDom), hipMemcpyDeviceToHost, st));
    std::vector<double> recs(moments ? (size_t)total * kRunRecord : 0);
    if (moments && total > 0)
      IZH_HIP(hipMemcpyAsync(recs.data(), b->d_moments, total * kRunRecord * sizeof(double),
                             hipMemcpyDeviceToHost, st));
    IZH_HIP(hipStreamSynchronize(st));
    for (int64_t k = 0; moments && k < total; k++) {
      double *o = moments + k * IZH_RUN_RECORD;
      std::memcpy(o, recs.data() + k * kRunRecord, kRunRecord * sizeof(double));
      o[kRunRecord] = refpt_after[2 * k];
      o[kRunRecord + 1] = refpt_after[2 * k + 1];
    }
    const DevScal out = *b->h_scal;
    if (out.err != 0)
      return fail(IZH_ERR_STATE, "izh_run: %lld particles had a bad sort key or no sort slot",
                  (long long)out.err);
    if (out.rec != total)
      return fail(IZH_ERR_STATE, "izh_run: %lld step records for %lld steps",
                  (long long)out.rec, (long long)total);
    b->n = out.nslots;
    b->npt = (int64_t)out.npt;
    b->current = out.current;
    b->run_stats[0] = out.nchunks[g_next];
    b->run_stats[1] = out.overflow;
    b->run_stats[2] = out.nalive;
    b->run_stats[3] = ntiles;
    for (int k = 0; k < 6; k++) dom->range[k] = f->h_dom->range[k];
    for (int k = 0; k < 3; k++) dom->msize[k] = f->h_dom->msize[k];
    f->have_phi = 1;
    b->keys_valid = 1;
    b->keys_po = g_next;
    b->keys_field = f;
    b->keys_grid[0] = nx; b->keys_grid[1] = ny; b->keys_grid[2] = nz;
    return IZH_OK;
  }
};

int run_steps(izh_bunch *b, izh_field *f, izh_compdom *dom, double *z,
              const izh_segment *segs, int64_t nseg, double *moments) {
  Runner r;
  int rc = r.begin(b, f, nseg);
  for (int64_t k = 0; rc == IZH_OK && k <= nseg; k++)
    rc = r.pass(k < nseg ? &segs[k] : nullptr, z);
  if (rc != IZH_OK) {
    r.release();
    return rc;
  }
  return r.end(dom, moments);
}

// Steps of a lattice-scan job per captured graph: a whole number of lattice
// periods holding an even number of re-sorts, so the buffers and chunk
// lists a replay starts with are those the capture started with.
static int64_t job_block(int64_t nperiod, int resort) {
  int64_t b = nperiod;
  while (b % (2 * (int64_t)resort) != 0) b += nperiod;
  return b;
}

}  // namespace izh

// A lattice-scan job (no reference counterpart: the reference runs one scan
// point per process, utilities/Engscan.py:28-47): one bunch with its field
// and domain, stepped nsteps times through a periodic lattice.  Its passes
// are enqueued on the bunch's stream without waiting for the host; blocks of
// whole lattice periods go through one captured graph, replayed.
struct izh_job {
  izh_bunch *b = nullptr;
  izh_field *f = nullptr;
  izh_compdom *dom = nullptr;
  std::vector<izh_segment> period;
  int64_t nsteps = 0;
  double z = 0.0;
  izh::Runner r;
  bool use_graph = false;
  int64_t block = 0;             // steps per graph replay
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t replays = 0;
  hipEvent_t done_ev = nullptr;     // recorded after the work of each advance
  bool enqueued_all = false, finished = false;
};

namespace izh {
namespace {

struct HostState {
  double *r0, *alt0;
  uint32_t *orig, *key;
  int g_cur, g_next;
  bool operator==(const HostState &o) const {
    return r0 == o.r0 && alt0 == o.alt0 && orig == o.orig && key == o.key &&
           g_cur == o.g_cur && g_next == o.g_next;
  }
};

HostState host_state(const izh_job *j) {
  return {j->b->r[0], j->b->r_alt[0], j->b->orig, j->b->key, j->r.g_cur, j->r.g_next};
}

const izh_segment *job_seg(const izh_job *j, int64_t step) {
  return &j->period[(size_t)(step % (int64_t)j->period.size())];
}

// Passes s .. s+cnt-1 (steady passes: each finishes step s-1 and starts s).
int job_passes(izh_job *j, int64_t cnt, bool dry) {
  for (int64_t k = 0; k < cnt; k++) IZH_CHECK(j->r.pass(job_seg(j, j->r.s), &j->z, dry));
  return IZH_OK;
}

// One block of steady passes through the job's graph: captured on the first
// block (and launched), replayed after it with the host bookkeeping run dry.
int job_graph_block(izh_job *j) {
  hipStream_t st = j->b->stream;
  if (!j->exec) {
    const HostState before = host_state(j);
    IZH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    const int rc = job_passes(j, j->block, false);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &g);
    IZH_CHECK(rc);
    if (ec != hipSuccess)
      return fail(IZH_ERR_HIP, "izh_job_advance: graph capture: %s", hipGetErrorString(ec));
    j->graph = g;
    IZH_HIP(hipGraphInstantiate(&j->exec, g, nullptr, nullptr, 0));
    IZH_HIP(hipGraphLaunch(j->exec, st));
    j->replays = 1;
    // a replay must start from the buffers this capture started from
    if (!(host_state(j) == before)) j->use_graph = false;
    return IZH_OK;
  }
  IZH_CHECK(job_passes(j, j->block, true));
  IZH_HIP(hipGraphLaunch(j->exec, st));
  j->replays++;
  return IZH_OK;
}

}  // namespace
}  // namespace izh

using namespace izh;

extern "C" {

int izh_job_create(izh_bunch *b, izh_field *f, izh_compdom *dom, double z,
                   const izh_segment *period, int64_t nperiod, int64_t nsteps, double bcurr,
                   int use_graph, izh_job **out) {
  if (!b || !f || !dom || !period || !out || nperiod < 1 || nsteps < 1)
    return fail(IZH_ERR_ARG, "izh_job_create: bad argument");
  *out = nullptr;
  if (b->frame != kFrameZ)
    return fail(IZH_ERR_STATE, "izh_job_create: bunch is not in the z frame");
  if (dom->nx != f->nx || dom->ny != f->ny || dom->nz != f->nz)
    return fail(IZH_ERR_ARG, "izh_job_create: domain mesh differs from the field's");
  if (b->precision != f->precision)
    return fail(IZH_ERR_ARG, "izh_job_create: bunch and field precision differ");
  if (bcurr < (double)1.0e-10f || b->n == 0)
    return fail(IZH_ERR_ARG, "izh_job_create: a job runs the space-charge step (current "
                ">= 1e-10 A and particles; izh_step runs the maps alone)");
  for (int64_t k = 0; k < nperiod; k++) {
    const izh_segment &sg = period[k];
    const char *why = nullptr;
    const int kind = map1_kind(sg.e.bitype, sg.e.v, &why);
    if (kind < 0 || is_rf_cavity(sg.e.bitype))
      return fail(IZH_ERR_ARG, "izh_job_create: segment %lld: bitype %d: %s", (long long)k,
                  sg.e.bitype, kind < 0 ? why : "an RF cavity changes the reference energy, "
                  "so its steps do not repeat (izh_run runs it)");
    if (!(sg.tau1 >= 0) || (kind == kMap1Linear && sg.e.mapstp < 1))
      return fail(IZH_ERR_ARG, "izh_job_create: segment %lld: bad step or map steps",
                  (long long)k);
    if (sg.wake && sg.wake->kind != 0)
      return fail(IZH_ERR_ARG, "izh_job_create: segment %lld has a wake (izh_run runs it)",
                  (long long)k);
  }
  izh_job *j = new izh_job;
  j->b = b;
  j->f = f;
  j->dom = dom;
  j->period.assign(period, period + nperiod);
  for (auto &sg : j->period) sg.wake = nullptr;
  j->nsteps = nsteps;
  j->z = z;
  j->use_graph = use_graph != 0;
  j->r.fork_green = false;   // one stream per job (see Runner)
  int rc = j->r.begin(b, f, nsteps);
  j->block = job_block(nperiod, j->r.period);
  if (rc == IZH_OK) rc = j->r.pass(job_seg(j, 0), &j->z);   // pass 0 starts step 0
  if (rc == IZH_OK && (hipEventCreateWithFlags(&j->done_ev, hipEventDisableTiming) !=
                           hipSuccess || hipEventRecord(j->done_ev, b->stream) != hipSuccess))
    rc = fail(IZH_ERR_HIP, "izh_job_create: event");
  if (rc != IZH_OK) {
    j->r.release();
    delete j;
    return rc;
  }
  *out = j;
  return IZH_OK;
}

int izh_job_advance(izh_job *j, int64_t max_steps, int64_t *left) {
  if (!j || max_steps < 0) return fail(IZH_ERR_ARG, "izh_job_advance: bad argument");
  int64_t budget = max_steps;
  while (budget > 0 && !j->enqueued_all) {
    const int64_t s = j->r.s;               // next pass; it finishes step s-1
    if (s == j->nsteps) {                   // the last pass
      IZH_CHECK(j->r.pass(nullptr, &j->z));
      j->enqueued_all = true;
      budget--;
      break;
    }
    // a graph block covers steady passes s .. s+block-1 that start at a
    // block boundary and leave a steady pass after them
    if (j->use_graph && (s - 1) % j->block == 0 && s + j->block <= j->nsteps - 1 &&
        budget >= j->block) {
      IZH_CHECK(job_graph_block(j));
      budget -= j->block;
      continue;
    }
    IZH_CHECK(job_passes(j, 1, false));
    budget--;
  }
  if (budget < max_steps) IZH_HIP(hipEventRecord(j->done_ev, j->b->stream));
  if (left) *left = j->nsteps + 1 - j->r.s;
  return IZH_OK;
}

int izh_job_query(izh_job *j, int *state) {
  if (!j || !state) return fail(IZH_ERR_ARG, "izh_job_query: bad argument");
  const hipError_t e = hipEventQuery(j->done_ev);
  if (e != hipSuccess && e != hipErrorNotReady)
    return fail(IZH_ERR_HIP, "izh_job_query: %s", hipGetErrorString(e));
  const bool ran = e == hipSuccess;
  *state = !ran ? 0 : (j->enqueued_all ? 2 : 1);
  return IZH_OK;
}

int izh_job_finish(izh_job *j, double *z, double *moments) {
  if (!j) return fail(IZH_ERR_ARG, "izh_job_finish: bad argument");
  if (!j->enqueued_all) IZH_CHECK(izh_job_advance(j, INT64_MAX, nullptr));
  if (j->finished) return fail(IZH_ERR_STATE, "izh_job_finish: already finished");
  j->finished = true;
  IZH_CHECK(j->r.end(j->dom, moments));
  if (z) *z = j->z;
  return IZH_OK;
}

int izh_job_info(const izh_job *j, int64_t info[4]) {
  if (!j || !info) return fail(IZH_ERR_ARG, "izh_job_info: bad argument");
  info[0] = j->r.s;                 // passes enqueued
  info[1] = j->block;
  info[2] = j->use_graph ? 1 : 0;
  info[3] = j->replays;
  return IZH_OK;
}

int izh_job_destroy(izh_job *j) {
  if (!j) return IZH_OK;
  if (!j->finished) {
    // stop the run where it is: wait for what was enqueued
    (void)hipStreamSynchronize(j->b->stream);
    j->r.release();
  }
  if (j->exec) (void)hipGraphExecDestroy(j->exec);
  if (j->graph) (void)hipGraphDestroy(j->graph);
  if (j->done_ev) (void)hipEventDestroy(j->done_ev);
  delete j;
  return IZH_OK;
}

}  // extern "C"
