// izh_internal.hpp -- handles, error plumbing and launch helpers shared by the
// HIP translation units of libimpactz_hip.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

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

// Frame of the particle coordinates held by a bunch; the reference's call
// order moves the bunch through these (src/Contrl/AccSimulator.f90:1110-1429).
enum Frame : int {
  kFrameZ = 0,     // z as independent variable: x/Scxl, px, y/Scxl, py, phi, pt
  kFrameT1 = 1,    // after conv1st: 1st-order t frame, metres
  kFrameT0 = 2,    // after cvbkforth1st: 0th-order t frame, metres
};

}  // namespace izh

// Device-resident bunch in structure-of-arrays form.
struct izh_bunch {
  int device = 0;
  int precision = IZH_FP64;
  hipStream_t stream = nullptr;
  int64_t cap = 0;       // allocated rows
  int64_t n = 0;         // Nptlocal
  int64_t npt = 0;       // Npt (one process: == n after conv1st)
  double current = 0, mass = 0, charge = 0, freq = 0;
  double refpt[6] = {0, 0, 0, 0, 0, 0};
  izh_const c{};
  int frame = izh::kFrameZ;
  double wabs_sum = 0;   // sum |Pts1(8,:)|, bounds every cell of the grid charge
  // SoA state, Pts1 rows 1..9
  double *s[9] = {nullptr};
  double *alt[9] = {nullptr};   // compaction target, swapped with s
  // reduction scratch
  double *d_part = nullptr;     // kMaxRedBlocks * kMoments
  int64_t *d_cnt = nullptr;     // kMaxRedBlocks + 1 (kept counts, then offsets)
  double *d_out = nullptr;      // final reductions (kMoments + 8)
  double *h_out = nullptr;      // pinned mirror of d_out
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
  void *e = nullptr;            // T[3g]  Ex, Ey, Ez
  void *spec_a = nullptr;       // complex T[(nx+1)*ny*nz]
  void *spec_b = nullptr;       // complex T[(nx+1)*2ny*nz]
  void *green = nullptr;        // T[(nx+1)*(ny+1)*(nz+1)] transformed Green fn
  void *green_t = nullptr;      // T[(nx+1)*(ny+1)*(nz+1)] work
  double *corner = nullptr;     // double[(nx+2)*(ny+2)*(nz+2)] IGF antiderivative
  void *tw_x = nullptr, *tw_y = nullptr, *tw_z = nullptr;  // twiddles per length
  double rho_scale = 1.0;       // grid value = physical charge * rho_scale
  int have_rho = 0, have_phi = 0;
  double fixed_scale = 0;       // 2^e used by the fixed-point deposit
};

namespace izh {

// Launchers implemented in the kernel translation units.
int launch_drift(izh_bunch *b, double tau, double gam0, double beta0);
int launch_map_linear(izh_bunch *b, const double xm[36]);
int launch_conv1st(izh_bunch *b, double xrad, double yrad, double ptrange[6],
                   int64_t *nlost);
int launch_compact(izh_bunch *b, double xrad, double yrad, int64_t nkeep);
int launch_cvbk(izh_bunch *b);
int launch_moments(izh_bunch *b, double out[kMoments]);
int launch_deposit(izh_bunch *b, const izh_compdom *dom, izh_field *f);
int launch_solve(izh_field *f, const izh_compdom *dom);
int launch_efield(izh_field *f, const izh_compdom *dom);
int launch_gather_kick(izh_bunch *b, izh_field *f, const izh_compdom *dom,
                       double tau, double xycon, double tcon);

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

inline int64_t blocks_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  return g < 1 ? 1 : g;
}

}  // namespace izh
