// izh_api.hip -- the extern "C" entry points of libimpactz_hip.so.
//
// Host-side pieces of the step live here: the quadrupole map (RK6 on 18
// variables, computed once per half step on the host as the reference does),
// the domain update, the space-charge constants of scatter1 and the order of
// calls of one segment.  Everything per particle or per grid cell runs in
// the kernels of izh_particles.hip and izh_field.hip.
#include <cmath>
#include <cstdint>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "izh_internal.hpp"

namespace izh {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int field_alloc(izh_field *f);

// ------------------------------------------------------------- timers
namespace {
struct PendingTimer {
  int slot;
  hipEvent_t a, b;
};
struct TimerState {
  bool on = false;
  std::vector<PendingTimer> pending;
  PendingTimer open[IZH_T_COUNT];
  bool is_open[IZH_T_COUNT] = {false};
  double ms[IZH_T_COUNT] = {0};
  int64_t n[IZH_T_COUNT] = {0};
};
TimerState &timers() {
  static TimerState t;
  return t;
}
}  // namespace

bool timing_on() { return timers().on; }

void timer_begin(int slot, hipStream_t s) {
  TimerState &t = timers();
  if (!t.on || slot < 0 || slot >= IZH_T_COUNT || t.is_open[slot]) return;
  // launches captured into a graph run later, not between these events
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  PendingTimer p{slot, nullptr, nullptr};
  if (hipEventCreate(&p.a) != hipSuccess) return;
  if (hipEventCreate(&p.b) != hipSuccess) {
    (void)hipEventDestroy(p.a);
    return;
  }
  (void)hipEventRecord(p.a, s);
  t.open[slot] = p;
  t.is_open[slot] = true;
}

void timer_end(int slot, hipStream_t s) {
  TimerState &t = timers();
  if (slot < 0 || slot >= IZH_T_COUNT || !t.is_open[slot]) return;
  PendingTimer p = t.open[slot];
  (void)hipEventRecord(p.b, s);
  t.pending.push_back(p);
  t.is_open[slot] = false;
}

static void timers_collect() {
  TimerState &t = timers();
  for (PendingTimer &p : t.pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess &&
        hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      t.ms[p.slot] += ms;
      t.n[p.slot] += 1;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  t.pending.clear();
}

static void physconst(double freq, izh_const *c) {
  // src/DataStruct/PhysConst.f90:28-33
  c->clight = 299792458.0;
  c->pi = 2.0 * std::asin(1.0);
  c->scxl = c->clight / (2.0 * c->pi * freq);
  c->scfreq = freq;
}

template <typename P>
static int dmalloc(P **p, int64_t count) {
  if (hipMalloc((void **)p, (size_t)(count > 0 ? count : 1) * sizeof(P)) != hipSuccess)
    return fail(IZH_ERR_NOMEM, "device allocation of %lld x %zu bytes failed",
                (long long)count, sizeof(P));
  return IZH_OK;
}

template <typename P>
static void dfree(P *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// Moving rows (phase space + original index) for cap particles.
// The six rows share one allocation, row k at r[0] + k*cap, so a kernel can
// address them from one base pointer.
static int bunch_alloc_live(izh_bunch *b, int64_t cap) {
  IZH_CHECK(dmalloc(&b->r[0], 6 * cap));
  for (int r = 1; r < 6; r++) b->r[r] = b->r[0] + r * cap;
  IZH_CHECK(dmalloc(&b->orig, cap));
  b->cap = cap;
  return IZH_OK;
}

int bunch_alloc_alt(izh_bunch *b) {
  IZH_CHECK(dmalloc(&b->r_alt[0], 6 * b->cap));
  for (int r = 1; r < 6; r++) b->r_alt[r] = b->r_alt[0] + r * b->cap;
  IZH_CHECK(dmalloc(&b->orig_alt, b->cap));
  return IZH_OK;
}

static void bunch_free_all(izh_bunch *b) {
  dfree(b->r[0]);
  dfree(b->r_alt[0]);
  for (int r = 1; r < 6; r++) b->r[r] = b->r_alt[r] = nullptr;
  dfree(b->orig);
  dfree(b->orig_alt);
  dfree(b->key);
  dfree(b->key_alt);
  dfree(b->tpos);
  dfree(b->sink);
  dfree(b->qm0);
  dfree(b->w0);
  dfree(b->id0);
  dfree(b->d_scal);
  for (int k = 0; k < 2; k++) dfree(b->hist[k]);
  dfree(b->cursor);
  dfree(b->chunks);
  dfree(b->d_moments);
  dfree(b->d_chg);
  if (b->h_chg) (void)hipHostFree(b->h_chg);
  b->h_chg = nullptr;
  b->ntiles_cap = b->nchunks_cap = b->moments_cap = 0;
  b->cap = 0;
}

// Upload Pts1(9,n) (particle-major): rows 1-6 into the moving SoA rows,
// q/m, charge weight and id into the immutable rows, original index = i.
static int upload_pts(izh_bunch *b, const double *pts9, int64_t n) {
  touch_particles(b);
  if (n > 0xfffffffeLL)
    return fail(IZH_ERR_ARG, "bunch of %lld particles exceeds 2^32-2", (long long)n);
  if (n > b->cap) {
    IZH_HIP(hipStreamSynchronize(b->stream));
    bunch_free_all(b);
    IZH_CHECK(bunch_alloc_live(b, n));
  }
  dfree(b->qm0);
  dfree(b->w0);
  dfree(b->id0);
  IZH_CHECK(dmalloc(&b->qm0, n));
  IZH_CHECK(dmalloc(&b->w0, n));
  IZH_CHECK(dmalloc(&b->id0, n));
  std::vector<double> row((size_t)(n > 0 ? n : 1));
  b->wabs_sum = 0.0;
  b->wabs_max = 0.0;
  b->w_uniform = 1;
  b->w_value = n > 0 ? pts9[7] : 0.0;
  b->qm_uniform = 1;
  b->qm_value = n > 0 ? pts9[6] : 0.0;
  double *dst[9] = {b->r[0], b->r[1], b->r[2], b->r[3], b->r[4], b->r[5],
                    b->qm0, b->w0, b->id0};
  for (int r = 0; r < 9; r++) {
    for (int64_t i = 0; i < n; i++) row[i] = pts9[9 * i + r];
    if (r == 6)
      for (int64_t i = 0; i < n; i++)
        if (row[i] != b->qm_value) b->qm_uniform = 0;
    if (r == 7)
      for (int64_t i = 0; i < n; i++) {
        b->wabs_sum += std::fabs(row[i]);
        b->wabs_max = std::max(b->wabs_max, std::fabs(row[i]));
        if (row[i] != b->w_value) b->w_uniform = 0;
      }
    if (n > 0)
      IZH_HIP(hipMemcpyAsync(dst[r], row.data(), n * sizeof(double),
                             hipMemcpyHostToDevice, b->stream));
    IZH_HIP(hipStreamSynchronize(b->stream));
  }
  std::vector<uint32_t> idx((size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < n; i++) idx[i] = (uint32_t)i;
  if (n > 0)
    IZH_HIP(hipMemcpyAsync(b->orig, idx.data(), n * sizeof(uint32_t),
                           hipMemcpyHostToDevice, b->stream));
  IZH_HIP(hipStreamSynchronize(b->stream));
  b->n = n;
  b->n0 = n;
  return IZH_OK;
}

// REAL(4) padding of update_CompDom: eps = 1.0/(N-3.0) (src/Appl/CompDom.f90:547-549).
static double eps_real4(int n) { return (double)(1.0f / ((float)n - 3.0f)); }

// Space-charge constants of scatter1_BeamBunch (src/Appl/BeamBunch.f90:2065-2084),
// with fpei = Clight*Clight*1.0e-7 using the REAL(4) literal 1.0e-7.
static void sc_constants(const izh_bunch *b, double *xycon, double *tcon) {
  const izh_const &c = b->c;
  const double gam = -b->refpt[5];
  const double curr = b->current, mass = b->mass, chrg = b->charge;
  const double tmpscale = curr / c.scfreq;
  const double twopi = 2.0 * c.pi;
  const double fpei = c.clight * c.clight * (double)1.0e-7f;
  const double xk = 1.0 / c.scxl;
  const double bet = std::sqrt(gam * gam - 1.0) / gam;
  const double bbyk = bet / xk;
  const double gambet = gam * bet;
  const double brho = gambet / (c.clight * chrg) * mass;
  const double vz0 = bet * c.clight;
  const double perv0 = 2.0 * curr * fpei / (brho * vz0 * vz0 * gam * gam);
  *xycon = 0.5 * perv0 * gambet * bbyk * twopi / tmpscale;
  *tcon = bet * (*xycon) * gam * gam;
}

}  // namespace izh

using namespace izh;

extern "C" {

const char *izh_last_error(void) { return g_err.c_str(); }

int izh_timer_enable(int on) {
  timers_collect();
  timers().on = on != 0;
  return IZH_OK;
}

int izh_timer_read(double ms[IZH_T_COUNT], int64_t launches[IZH_T_COUNT],
                   int reset) {
  timers_collect();
  TimerState &t = timers();
  for (int k = 0; k < IZH_T_COUNT; k++) {
    if (ms) ms[k] = t.ms[k];
    if (launches) launches[k] = t.n[k];
    if (reset) {
      t.ms[k] = 0;
      t.n[k] = 0;
    }
  }
  return IZH_OK;
}

int izh_init(int device) {
  int count = 0;
  IZH_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count)
    return fail(IZH_ERR_ARG, "device %d not present (%d devices)", device, count);
  IZH_HIP(hipSetDevice(device));
  IZH_HIP(hipFree(nullptr));
  IZH_CHECK(scratch_setup(device));
  return IZH_OK;
}

int izh_physconst(double freq, izh_const *out) {
  if (!out || !(freq > 0)) return fail(IZH_ERR_ARG, "izh_physconst: bad argument");
  physconst(freq, out);
  return IZH_OK;
}

int izh_bunch_create(const double *pts9, int64_t n, double current,
                     double kinetic, double mass, double charge, double freq,
                     double phase0, int precision, izh_bunch **out) {
  if (!out || n < 0 || (n > 0 && !pts9) || !(mass > 0) || !(freq > 0) ||
      (precision != IZH_FP64 && precision != IZH_MIXED))
    return fail(IZH_ERR_ARG, "izh_bunch_create: bad argument");
  *out = nullptr;
  izh_bunch *b = new izh_bunch();
  IZH_HIP(hipGetDevice(&b->device));
  b->precision = precision;
  b->current = current;
  b->mass = mass;
  b->charge = charge;
  b->freq = freq;
  b->npt = n;
  // construct_BeamBunch: refptcl = phasini; refptcl(6) = -(W/m + 1)
  for (int k = 0; k < 6; k++) b->refpt[k] = phase0;
  b->refpt[5] = -(kinetic / mass + 1.0);
  physconst(freq, &b->c);
  int rc = IZH_OK;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(IZH_ERR_HIP, "izh_bunch_create: stream creation failed");
  }
  if ((rc = bunch_alloc_live(b, n)) != IZH_OK ||
      (rc = dmalloc(&b->d_part, (int64_t)kMaxRedBlocks * kPartStride)) != IZH_OK ||
      (rc = dmalloc(&b->d_out, (int64_t)kPartStride + 8)) != IZH_OK) {
    izh_bunch_destroy(b);
    return rc;
  }
  if (hipHostMalloc((void **)&b->h_out, (kPartStride + 8) * sizeof(double),
                    hipHostMallocDefault) != hipSuccess) {
    izh_bunch_destroy(b);
    return fail(IZH_ERR_NOMEM, "izh_bunch_create: pinned allocation failed");
  }
  if ((rc = upload_pts(b, pts9, n)) != IZH_OK) {
    izh_bunch_destroy(b);
    return rc;
  }
  b->frame = kFrameZ;
  *out = b;
  return IZH_OK;
}

int izh_bunch_create_from(const izh_bunch *src, int64_t offset, int64_t n, double current,
                          int precision, izh_bunch **out) {
  if (!src || !out || offset < 0 || n < 1 || offset + n > src->n ||
      (precision != IZH_FP64 && precision != IZH_MIXED))
    return fail(IZH_ERR_ARG, "izh_bunch_create_from: bad argument (slots %lld..%lld of %lld)",
                (long long)offset, (long long)(offset + n), (long long)src->n);
  if (src->frame != kFrameZ)
    return fail(IZH_ERR_STATE, "izh_bunch_create_from: source is not in the z frame");
  *out = nullptr;
  izh_bunch *b = new izh_bunch();
  IZH_HIP(hipGetDevice(&b->device));
  b->precision = precision;
  b->current = current;
  b->mass = src->mass;
  b->charge = src->charge;
  b->freq = src->freq;
  b->npt = n;
  for (int k = 0; k < 6; k++) b->refpt[k] = src->refpt[k];
  b->c = src->c;
  int rc = IZH_OK;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(IZH_ERR_HIP, "izh_bunch_create_from: stream creation failed");
  }
  if ((rc = bunch_alloc_live(b, n)) != IZH_OK ||
      (rc = dmalloc(&b->d_part, (int64_t)kMaxRedBlocks * kPartStride)) != IZH_OK ||
      (rc = dmalloc(&b->d_out, (int64_t)kPartStride + 8)) != IZH_OK ||
      (rc = dmalloc(&b->qm0, n)) != IZH_OK || (rc = dmalloc(&b->w0, n)) != IZH_OK ||
      (rc = dmalloc(&b->id0, n)) != IZH_OK) {
    izh_bunch_destroy(b);
    return rc;
  }
  if (hipHostMalloc((void **)&b->h_out, (kPartStride + 8) * sizeof(double),
                    hipHostMallocDefault) != hipSuccess) {
    izh_bunch_destroy(b);
    return fail(IZH_ERR_NOMEM, "izh_bunch_create_from: pinned allocation failed");
  }
  // Pts1(8) = I / (f N) sign(q) for the new particle count (distribution
  // readers, e.g. src/Appl/Distribution.f90:1660-1661)
  const double w = current / src->freq / (double)n * (src->charge < 0 ? -1.0 : 1.0);
  b->w_uniform = 1;
  b->w_value = w;
  b->wabs_sum = std::fabs(w) * (double)n;
  b->wabs_max = std::fabs(w);
  b->qm_uniform = src->qm_uniform;
  b->qm_value = src->qm_value;
  // the source's stream must have finished with its rows before the copy
  hipEvent_t ev;
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
    izh_bunch_destroy(b);
    return fail(IZH_ERR_HIP, "izh_bunch_create_from: event creation failed");
  }
  const bool joined = hipEventRecord(ev, src->stream) == hipSuccess &&
                      hipStreamWaitEvent(b->stream, ev, 0) == hipSuccess;
  (void)hipEventDestroy(ev);
  if (!joined) {
    izh_bunch_destroy(b);
    return fail(IZH_ERR_HIP, "izh_bunch_create_from: stream join failed");
  }
  if ((rc = launch_clone(src, b, offset, n, w)) != IZH_OK) {
    izh_bunch_destroy(b);
    return rc;
  }
  b->n = n;
  b->n0 = n;
  b->frame = kFrameZ;
  *out = b;
  return IZH_OK;
}

int izh_bunch_destroy(izh_bunch *b) {
  if (!b) return IZH_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  bunch_free_all(b);
  dfree(b->d_part);
  dfree(b->d_out);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->h_scal) (void)hipHostFree(b->h_scal);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return IZH_OK;
}

int izh_bunch_download(izh_bunch *b, double *pts9, int64_t cap, int64_t *n) {
  if (!b || !n || (b->n > 0 && !pts9))
    return fail(IZH_ERR_ARG, "izh_bunch_download: bad argument");
  if (cap < b->n)
    return fail(IZH_ERR_ARG, "izh_bunch_download: %lld rows needed, %lld given",
                (long long)b->n, (long long)cap);
  const int64_t nn = b->n, n0 = b->n0;
  std::vector<uint32_t> orig((size_t)(nn > 0 ? nn : 1));
  std::vector<double> rows((size_t)(6 * nn > 0 ? 6 * nn : 1));
  std::vector<double> fixed((size_t)(3 * n0 > 0 ? 3 * n0 : 1));
  for (int r = 0; r < 6 && nn > 0; r++)
    IZH_HIP(hipMemcpyAsync(rows.data() + r * nn, b->r[r], nn * sizeof(double),
                           hipMemcpyDeviceToHost, b->stream));
  if (nn > 0)
    IZH_HIP(hipMemcpyAsync(orig.data(), b->orig, nn * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, b->stream));
  const double *fx[3] = {b->qm0, b->w0, b->id0};
  for (int r = 0; r < 3 && n0 > 0; r++)
    IZH_HIP(hipMemcpyAsync(fixed.data() + r * n0, fx[r], n0 * sizeof(double),
                           hipMemcpyDeviceToHost, b->stream));
  IZH_HIP(hipStreamSynchronize(b->stream));
  // the reference keeps the surviving particles in their original order
  // (order-preserving compaction, src/Appl/BeamBunch.f90:1290,1345-1347):
  // place each row by its original index, then close the gaps
  std::vector<int64_t> at((size_t)(n0 > 0 ? n0 : 1), -1);
  for (int64_t i = 0; i < nn; i++) {
    if (orig[i] >= (uint64_t)n0 || at[orig[i]] >= 0)
      return fail(IZH_ERR_STATE, "izh_bunch_download: corrupt particle index");
    at[orig[i]] = i;
  }
  int64_t k = 0;
  for (int64_t o = 0; o < n0; o++) {
    const int64_t i = at[o];
    if (i < 0) continue;
    double *p = pts9 + 9 * k;
    for (int r = 0; r < 6; r++) p[r] = rows[r * nn + i];
    for (int r = 0; r < 3; r++) p[6 + r] = fixed[r * n0 + o];
    k++;
  }
  *n = nn;
  return IZH_OK;
}

int izh_bunch_upload(izh_bunch *b, const double *pts9, int64_t n) {
  if (!b || n < 0 || (n > 0 && !pts9))
    return fail(IZH_ERR_ARG, "izh_bunch_upload: bad argument");
  IZH_CHECK(upload_pts(b, pts9, n));
  b->frame = kFrameZ;
  return IZH_OK;
}

int izh_bunch_get_info(const izh_bunch *b, izh_bunch_info *info) {
  if (!b || !info) return fail(IZH_ERR_ARG, "izh_bunch_get_info: bad argument");
  info->current = b->current;
  info->mass = b->mass;
  info->charge = b->charge;
  info->npt = b->npt;
  info->nptlocal = b->n;
  for (int k = 0; k < 6; k++) info->refpt[k] = b->refpt[k];
  info->freq = b->freq;
  info->precision = b->precision;
  return IZH_OK;
}

int izh_bunch_set_info(izh_bunch *b, const izh_bunch_info *info) {
  if (!b || !info) return fail(IZH_ERR_ARG, "izh_bunch_set_info: bad argument");
  touch_particles(b);
  if (info->nptlocal != b->n)
    return fail(IZH_ERR_ARG, "izh_bunch_set_info: nptlocal is set by upload");
  b->current = info->current;
  b->mass = info->mass;
  b->charge = info->charge;
  b->npt = info->npt;
  for (int k = 0; k < 6; k++) b->refpt[k] = info->refpt[k];
  return IZH_OK;
}

// ------------------------------------------------------- element maps

int izh_rfdata_table(const double *rows4, int n, izh_rfdata **out) {
  if (!rows4 || n < 1 || !out) return fail(IZH_ERR_ARG, "izh_rfdata_table: bad argument");
  auto *d = new izh_rfdata();
  d->fourier = 0;
  d->z.resize(n); d->ez.resize(n); d->ezp.resize(n); d->ezpp.resize(n);
  // read1_Data shifts the positions so the table starts at 0
  // (src/DataStruct/Data.f90:195-199)
  const double z1 = rows4[0];
  for (int i = 0; i < n; i++) {
    d->z[i] = rows4[4 * i] - z1;
    d->ez[i] = rows4[4 * i + 1];
    d->ezp[i] = rows4[4 * i + 2];
    d->ezpp[i] = rows4[4 * i + 3];
  }
  *out = d;
  return IZH_OK;
}

int izh_rfdata_fourier(const double *coef, int n, izh_rfdata **out) {
  if (!coef || n < 1 || !out) return fail(IZH_ERR_ARG, "izh_rfdata_fourier: bad argument");
  auto *d = new izh_rfdata();
  d->fourier = 1;
  d->coef.assign(coef, coef + n);
  // read3_Data pads to an odd count (src/DataStruct/Data.f90:402-406)
  if (n % 2 == 0) d->coef.push_back(0.0);
  *out = d;
  return IZH_OK;
}

int izh_rfdata_destroy(izh_rfdata *d) {
  if (d && d->d_table) {
    (void)hipDeviceSynchronize();
    (void)hipFree(d->d_table);
  }
  delete d;
  return IZH_OK;
}

// The linear ODE the element maps share: the reference particle's phase
// (y[4]) and energy (y[5]), the cosine- and sine-like columns of three 2x2
// systems q' = p/Scxl, p' = -s q (y[6..17]) and, for the solenoids, the 2x2
// rotation of the Larmor frame (y[18..21]).  Spring constants:
//   quadrupole  s = (G/Brho, -G/Brho, 0)*Scxl  (src/Appl/Quadrupole.f90:283-291)
//   ConstFoc    s = (kx0^2, ky0^2, kz0^2)*Scxl (src/Appl/ConstFoc.f90:276-281)
//   solenoid    s = ((B/2Brho)^2, (B/2Brho)^2, 0)*Scxl, rotation rate
//               -B/(2 Brho)                    (src/Appl/Sol.f90:298-318)
//   RF cavity   the energy gain -qmcc Ez cos(phi) and the RF focusing of
//               intfunc1_DTL/_CCL/_SolRF (src/Appl/DTL.f90:286-347,
//               CCL.f90:269-328, SolRF.f90:287-353), plus the DTL's
//               quadrupoles and SolRF's solenoid field.
struct LinearElement {
  int type;      // 1, 2, 3 or 101-105
  double p[3];   // 1: gradient; 2: kx0^2, ky0^2, kz0^2; 3: Bz0
  // RF cavities: Param(1..4), the element length, the DTL's quadrupole
  // lengths and gradients (Param(7..10)), SolRF's Bz0 (Param(12))
  double zedge, escale, ww, theta0, length;
  double q[4];
  double bz0;
  const izh_rfdata *rf;
  int nvar() const { return (type == 3 || type == 105) ? 22 : 18; }
};

// getaxfldE_DTL/_CCDTL/_CCL/_SC (src/Appl/DTL.f90:350-393): linear
// interpolation in the table, bisection as the reference does it (and
// extrapolation outside it); getaxfldE_SolRF (src/Appl/SolRF.f90:356-466):
// the Fourier series inside the element and the field's range, else 0.
static void rf_axis_field(const izh_const *c, const LinearElement &e, double z, double &ez1,
                          double &ezp1) {
  const izh_rfdata &d = *e.rf;
  if (!d.fourier) {
    const double zz = z - e.zedge;
    int klo = 1, khi = (int)d.z.size();
    while (khi - klo > 1) {
      const int k = (khi + klo) / 2;
      if (d.z[k - 1] > zz) khi = k; else klo = k;
    }
    const double hstep = d.z[khi - 1] - d.z[klo - 1];
    double slope = (d.ez[khi - 1] - d.ez[klo - 1]) / hstep;
    ez1 = d.ez[klo - 1] + slope * (zz - d.z[klo - 1]);
    slope = (d.ezp[khi - 1] - d.ezp[klo - 1]) / hstep;
    ezp1 = d.ezp[klo - 1] + slope * (zz - d.z[klo - 1]);
    ez1 = ez1 * e.escale;
    ezp1 = ezp1 * e.escale;
    return;
  }
  const double epstol = (double)1.0e-10f;
  const double *F = d.coef.data() - 1;   // 1-based, as Fcoef
  ez1 = 0.0;
  ezp1 = 0.0;
  if ((z >= e.zedge && z <= e.zedge + e.length) || std::fabs(z - e.zedge) <= epstol ||
      std::fabs(z - e.zedge - e.length) <= epstol) {
    const double zlc = z - e.zedge;
    const int numpar1 = (int)(F[1] + 0.1);
    const double zstart1 = F[2], zend1 = F[3], zlength1 = F[4];
    if ((zlc >= zstart1 && zlc <= zend1) || std::fabs(zlc - zstart1) <= epstol ||
        std::fabs(zlc - zend1) <= epstol) {
      const double zmid = zlength1 / 2;
      const double zz = z - e.zedge - zstart1 - zmid;
      const double zlen = zlength1;
      const int ntmp = 4;
      ez1 = F[ntmp + 1] / 2;
      for (int i = 2; i <= (numpar1 - 1) / 2 + 1; i++) {
        const double arg = (i - 1) * 2 * c->pi * zz / zlen;
        const double cs = std::cos(arg), sn = std::sin(arg);
        ez1 = ez1 + F[2 * i - 2 + ntmp] * cs + F[2 * i - 1 + ntmp] * sn;
        ezp1 = ezp1 + (i - 1) * 2 * c->pi / zlen *
                          (-F[2 * i - 2 + ntmp] * sn + F[2 * i - 1 + ntmp] * cs);
      }
      ez1 = ez1 * e.escale;
      ezp1 = ezp1 * e.escale;
    }
  }
}

// getBgradfld_DTL (src/Appl/DTL.f90:395-418): the drift-tube quadrupoles'
// gradient at z.
static double dtl_gradient(const LinearElement &e, double z) {
  const double len1 = e.q[0], bgrad1 = e.q[1], len2 = e.length - e.q[2], bgrad2 = e.q[3];
  const double zz = z - e.zedge;
  if (zz < len1) return bgrad1;
  if (zz < len2) return 0.0;
  return bgrad2;
}

static void linear_rhs(const izh_const *c, const LinearElement &e, double qmcc, double t,
                       const double *y, double *f) {
  const double gamma0 = -y[5];
  const double beta0 = std::sqrt(gamma0 * gamma0 - 1.0) / gamma0;
  const double gbet = std::sqrt((gamma0 - 1.0) * (gamma0 + 1.0));
  for (int k = 0; k < 4; k++) f[k] = 0.0;
  f[4] = 1.0 / (beta0 * c->scxl);
  f[5] = 0.0;
  double s[3], rot = 0.0;
  if (e.type == 1) {
    const double brho = gbet / c->clight / qmcc;
    s[0] = e.p[0] / brho * c->scxl;
    s[1] = -e.p[0] / brho * c->scxl;
    s[2] = 0.0;
  } else if (e.type == 2) {
    for (int a = 0; a < 3; a++) s[a] = e.p[a] * c->scxl;
  } else if (e.type == 3) {
    const double brho = gbet / c->clight / qmcc;
    const double k = 0.5 * e.p[0] / brho;
    s[0] = s[1] = (k * k) * c->scxl;
    s[2] = 0.0;
    rot = -e.p[0] / brho / 2;
  } else {
    // the RF cavities (intfunc1_DTL, src/Appl/DTL.f90:286-347, and alike)
    double ez1, ezp1;
    rf_axis_field(c, e, t, ez1, ezp1);
    const double sinphi = std::sin(e.ww * y[4] + e.theta0);
    const double cosphi = std::cos(e.ww * y[4] + e.theta0);
    const double rfdsgn = ez1 * cosphi;
    f[5] = -qmcc * rfdsgn;
    const double b2 = beta0 * beta0, g2 = gamma0 * gamma0;
    const double acc = qmcc * rfdsgn / b2 / g2;
    const double s11tmp = 0.5 * (1.0 + 0.5 * g2) * (acc * acc) +
                          qmcc * 0.5 / (b2 * beta0) / (g2 * gamma0) * ez1 * sinphi * e.ww /
                              c->scxl;
    double kx = 0.0, ky = 0.0;
    if (e.type == 101) {
      const double brho = gbet / c->clight / qmcc;
      const double bgrad = dtl_gradient(e, t);
      kx = bgrad / brho;
      ky = -bgrad / brho;
    } else if (e.type == 105) {
      const double brho = gbet / c->clight / qmcc;
      const double kb = 0.5 * e.bz0 / brho;
      kx = ky = (kb * kb) / 2;
      rot = -e.bz0 / brho / 2;
    }
    s[0] = (s11tmp + kx) * c->scxl;
    s[1] = (s11tmp + ky) * c->scxl;
    const double s55 = -1.5 * qmcc / b2 / gamma0 * ezp1 * cosphi +
                       (b2 + 0.5) / (b2 * beta0) / gamma0 * qmcc * ez1 * sinphi * e.ww /
                           c->scxl +
                       1.5 * (1.0 - 0.5 * g2) * (acc * acc);
    s[2] = c->scxl * s55;
  }
  for (int blk = 0; blk < 3; blk++) {
    for (int col = 0; col < 2; col++) {
      const int o = 6 + 4 * blk + 2 * col;
      f[o] = y[o + 1] / c->scxl;
      f[o + 1] = -s[blk] * y[o];
    }
  }
  if (e.nvar() == 22) {
    f[18] = -rot * y[20];
    f[19] = -rot * y[21];
    f[20] = rot * y[18];
    f[21] = rot * y[19];
  }
}

// The entrance/exit terms of the RF maps (maplinear_DTL, src/Appl/DTL.f90:
// 189-220): dlt and thl from the on-axis field and its slope at z.
static void rf_edge(const izh_const *c, const LinearElement &e, double qmcc, double z,
                    const double refpt[6], double u, double &dlt, double &thl) {
  double ez, ez1;
  rf_axis_field(c, e, z, ez, ez1);
  const double g = -refpt[5];
  const double beta = std::sqrt(g * g - 1.0) / g;
  const double sinth = std::sin(e.ww * refpt[4] + e.theta0);
  const double costh = std::cos(e.ww * refpt[4] + e.theta0);
  const double uprim = qmcc * ez / beta * costh;
  const double qpw = 0.5 * qmcc * c->scxl / (u * e.ww);
  dlt = c->scxl * (0.5 * uprim / u - qpw * ez1 * sinth);
  thl = 1.5 * c->scxl * (uprim / u);
}

// maplinear_Quadrupole / _ConstFoc / _Sol / the RF cavities with their rk6i_*
// (src/Appl/Quadrupole.f90:120-252, ConstFoc.f90:116-243, Sol.f90:120-268,
// DTL.f90:136-284, SolRF.f90:126-285): an 8-stage 6th-order Runge-Kutta over
// mapstp steps of the ODE in z from t, then the 6x6 map with the energy
// scaling of each block (dlt = thl = 0 outside the RF cavities).
static void linear_map(const izh_const *c, const LinearElement &e, double t, double tau,
                       int mapstp, double refpt[6], double charge, double mass,
                       double xm[36]) {
  static const double kA[7][7] = {
      {1.0 / 9.0},
      {1.0 / 24.0, 3.0 / 24.0},
      {1.0 / 6.0, -3.0 / 6.0, 4.0 / 6.0},
      {278.0 / 544.0, -945.0 / 544.0, 840.0 / 544.0, 99.0 / 544.0},
      {-106.0 / 6.0, 273.0 / 6.0, -104.0 / 6.0, -107.0 / 6.0, 48.0 / 6.0},
      {110974.0 / 45648.0, -236799.0 / 45648.0, 68376.0 / 45648.0,
       103803.0 / 45648.0, -10240.0 / 45648.0, 1926.0 / 45648.0},
      {-101195.0 / 25994.0, 222534.0 / 25994.0, -71988.0 / 25994.0,
       -26109.0 / 25994.0, -20000.0 / 25994.0, -72.0 / 25994.0,
       22824.0 / 25994.0}};
  static const double kB[8] = {41.0 / 840.0, 0.0, 216.0 / 840.0, 27.0 / 840.0,
                               272.0 / 840.0, 27.0 / 840.0, 216.0 / 840.0,
                               41.0 / 840.0};
  constexpr int kMaxVar = 22;
  const bool rf = is_rf_cavity(e.type);
  const int nv = e.nvar();
  const double qmcc = charge / mass;
  double y[kMaxVar] = {0};
  y[4] = refpt[4];
  y[5] = refpt[5];
  y[6] = y[9] = y[10] = y[13] = y[14] = y[17] = 1.0;
  y[18] = y[21] = 1.0;
  const double h = tau / mapstp;
  const double gi = -refpt[5];
  const double ui = gi * (std::sqrt(gi * gi - 1.0) / gi);
  const double squi = std::sqrt(ui), sqi3 = squi * squi * squi;
  double dlti = 0.0, thli = 0.0;
  if (rf) rf_edge(c, e, qmcc, t, refpt, ui, dlti, thli);
  double k[8][kMaxVar] = {{0}}, yt[kMaxVar] = {0};
  // stage positions of rk6i_* (t + h/9, h/6, h/3, h/2, 2h/3, 5h/6, h)
  for (int step = 0; step < mapstp; step++) {
    const double ts = t + step * h;
    const double tt[8] = {ts, ts + h / 9.0, ts + h / 6.0, ts + h / 3.0, ts + 0.5 * h,
                          ts + 2.0 * h / 3.0, ts + 5.0 * h / 6.0, ts + h};
    for (int st = 0; st < 8; st++) {
      for (int v = 0; v < nv; v++) {
        double acc = 0.0;
        for (int q = 0; q < st; q++) acc += kA[st - 1][q] * k[q][v];
        yt[v] = y[v] + acc;
      }
      linear_rhs(c, e, qmcc, tt[st], st == 0 ? y : yt, k[st]);
      for (int v = 0; v < nv; v++) k[st][v] *= h;
    }
    for (int v = 0; v < nv; v++) {
      double acc = 0.0;
      for (int st = 0; st < 8; st++) acc += kB[st] * k[st][v];
      y[v] += acc;
    }
  }
  refpt[4] = y[4];
  refpt[5] = y[5];
  const double gf = -refpt[5];
  const double uf = gf * (std::sqrt(gf * gf - 1.0) / gf);
  const double squf = std::sqrt(uf), sqf3 = squf * squf * squf;
  double dltf = 0.0, thlf = 0.0;
  if (rf) rf_edge(c, e, qmcc, t + tau, refpt, uf, dltf, thlf);
  for (int q = 0; q < 36; q++) xm[q] = 0.0;
  // one 2x2 block from ODE columns o..o+3 with entrance/exit terms di, df,
  // times w (the solenoids' rotation entry; 1 otherwise), at rows r and
  // columns cc (src/Appl/DTL.f90:205-228)
  auto block = [&](int r, int cc, int o, double si, double sf, double di, double df,
                   double w) {
    xm[6 * r + cc] = w * (y[o] * si / sf + y[o + 2] * si / sf * di);
    xm[6 * (r + 1) + cc] =
        w * ((y[o + 1] - y[o] * df) * si * sf + (y[o + 3] - y[o + 2] * df) * si * sf * di);
    xm[6 * r + cc + 1] = w * y[o + 2] / (si * sf);
    xm[6 * (r + 1) + cc + 1] = w * (y[o + 3] - y[o + 2] * df) * sf / si;
  };
  if (e.nvar() == 22) {
    // src/Appl/Sol.f90:187-207, SolRF.f90:207-226: the x and y blocks times
    // the rotation
    block(0, 0, 6, squi, squf, dlti, dltf, y[18]);
    block(0, 2, 6, squi, squf, dlti, dltf, y[19]);
    block(2, 0, 10, squi, squf, dlti, dltf, y[20]);
    block(2, 2, 10, squi, squf, dlti, dltf, y[21]);
  } else {
    block(0, 0, 6, squi, squf, dlti, dltf, 1.0);
    block(2, 2, 10, squi, squf, dlti, dltf, 1.0);
  }
  // the longitudinal block scales with the cube (src/Appl/Quadrupole.f90:191-196)
  block(4, 4, 14, sqi3, sqf3, thli, thlf, 1.0);
}

// The linear element of an izh_element, or an error message.
static const char *linear_element(const izh_const *c, const izh_element &el,
                                  LinearElement &e) {
  e = LinearElement{};
  e.type = el.bitype;
  const double *v = el.v;
  if (el.bitype == 1) {
    // a field profile from a file (file ID > 1e-5) is not on this path
    if (v[1] > 1.0e-5) return "quadrupole fringe field from a file";
    e.p[0] = v[0];
  } else if (el.bitype == 2) {
    for (int a = 0; a < 3; a++) e.p[a] = v[a];
  } else if (el.bitype == 3) {
    e.p[0] = v[0];
  } else if (is_rf_cavity(el.bitype)) {
    if (!el.rf) return "an RF cavity needs its field (izh_rfdata)";
    if ((el.bitype == 105) != (el.rf->fourier != 0))
      return el.bitype == 105 ? "SolRF takes Fourier coefficients (izh_rfdata_fourier)"
                              : "this RF cavity takes a field table (izh_rfdata_table)";
    // getaxfldE_SolRF reads Fcoef(5 .. 4 + numpar1), numpar1 = Fcoef(1)
    // (src/Appl/SolRF.f90:385-417)
    if (el.bitype == 105 && (el.rf->coef.size() < 4 ||
                             (int)(el.rf->coef[0] + 0.1) < 1 ||
                             4 + (int)(el.rf->coef[0] + 0.1) > (int)el.rf->coef.size()))
      return "SolRF coefficients hold fewer values than their first value announces";
    if (!(v[1] > 0.0)) return "RF frequency must be positive";
    e.rf = el.rf;
    e.zedge = el.zedge;
    e.length = el.length;
    e.escale = v[0];
    e.ww = v[1] / c->scfreq;
    // Param(4)*asin(1.0)/90: a REAL(4) asin (src/Appl/DTL.f90:158)
    e.theta0 = v[2] * (double)std::asin(1.0f) / 90;
    for (int a = 0; a < 4; a++) e.q[a] = v[5 + a];   // DTL: Param(7..10)
    e.bz0 = v[10];                                  // SolRF: Param(12)
  } else {
    return "no linear map for this element type";
  }
  return nullptr;
}

int izh_element_maplinear(const izh_const *c, double z, double tau, const izh_element *e,
                          double refpt[6], double charge, double mass, double xm[36]) {
  if (!c || !e || !refpt || !xm || e->mapstp < 1 || !(mass > 0))
    return fail(IZH_ERR_ARG, "izh_element_maplinear: bad argument");
  LinearElement le;
  if (const char *why = linear_element(c, *e, le))
    return fail(IZH_ERR_ARG, "izh_element_maplinear: bitype %d: %s", e->bitype, why);
  linear_map(c, le, z, tau, e->mapstp, refpt, charge, mass, xm);
  return IZH_OK;
}

int izh_quad_maplinear(const izh_const *c, double tau, double gradient,
                       int mapstp, double refpt[6], double charge, double mass,
                       double xm[36]) {
  if (!c || !refpt || !xm || mapstp < 1 || !(mass > 0))
    return fail(IZH_ERR_ARG, "izh_quad_maplinear: bad argument");
  izh_element e{};
  e.bitype = 1;
  e.mapstp = mapstp;
  e.v[0] = gradient;
  return izh_element_maplinear(c, 0.0, tau, &e, refpt, charge, mass, xm);
}

// ------------------------------------------------------------- map push

int izh_map1(izh_bunch *b, double *z, double tau, const izh_element *e) {
  if (!b || !z || !e) return fail(IZH_ERR_ARG, "izh_map1: bad argument");
  if (b->frame != kFrameZ)
    return fail(IZH_ERR_STATE, "izh_map1: bunch is not in the z frame");
  const char *why = nullptr;
  const int kind = map1_kind(e->bitype, e->v, &why);
  if (kind < 0) return fail(IZH_ERR_ARG, "izh_map1: bitype %d: %s", e->bitype, why);
  if (kind == kMap1Drift) {
    // src/Appl/BeamBunch.f90:276-298
    const double gam0 = -b->refpt[5];
    const double beta0 = std::sqrt(gam0 * gam0 - 1.0) / gam0;
    IZH_CHECK(launch_drift(b, tau, gam0, beta0));
    b->refpt[4] = b->refpt[4] + tau / (b->c.scxl * beta0);
  } else if (kind == kMap1QuadK) {
    // transfmapK_Quadrupole (src/Appl/Quadrupole.f90:509-607)
    const double r6 = b->refpt[5];
    const double gambet0 = std::sqrt(r6 * r6 - 1.0);
    const double beta0 = std::sqrt(1.0 - 1.0 / (r6 * r6));
    IZH_CHECK(launch_quad_kmode(b, tau, e->v[0], b->charge / b->mass, gambet0, beta0));
    b->refpt[4] = b->refpt[4] + tau / beta0 / b->c.scxl;
  } else {
    double xm[36];
    IZH_CHECK(izh_element_maplinear(&b->c, *z, tau, e, b->refpt, b->charge, b->mass, xm));
    IZH_CHECK(launch_map_linear(b, xm));
  }
  *z = *z + tau;
  return IZH_OK;
}

int izh_drift1(izh_bunch *b, double *z, double tau, int bitype,
               double gradient, int mapstp) {
  if (bitype != 0 && bitype != 1)
    return fail(IZH_ERR_ARG, "izh_drift1: bitype %d not supported (0 drift, 1 quad; "
                "izh_map1 takes the other elements)", bitype);
  izh_element e{};
  e.bitype = bitype;
  e.mapstp = mapstp;
  e.v[0] = gradient;
  return izh_map1(b, z, tau, &e);
}

// ------------------------------------------------------------------ bends

static int dipole_check(izh_bunch *b, const izh_element *e, const char *who) {
  if (!b || !e) return fail(IZH_ERR_ARG, "%s: bad argument", who);
  if (e->bitype != 4) return fail(IZH_ERR_ARG, "%s: bitype %d is not a bend", who, e->bitype);
  if (!(e->length > 0.0) || e->v[0] == 0.0)
    return fail(IZH_ERR_ARG, "%s: a bend needs a length > 0 and an angle != 0", who);
  if (b->frame != kFrameZ) return fail(IZH_ERR_STATE, "%s: bunch is not in the z frame", who);
  return IZH_OK;
}

// The pole-face constants of run_AccSimulator (src/Contrl/AccSimulator.f90:
// 1058-1088) for the entrance (exit = false) or the exit, and the energy as
// Fpol_Dipole/Bpol_Dipole form it (src/Appl/Dipole.f90:478-485).
static DipoleMap dipole_face(const izh_bunch *b, const izh_element *e, bool exit) {
  DipoleMap m{};
  m.mode = exit ? kDipoleExit : kDipoleEntrance;
  m.scxl = b->c.scxl;
  m.gam0 = -b->refpt[5];
  m.beta0 = std::sqrt(1.0 - 1.0 / (m.gam0 * m.gam0));
  m.gambet = m.beta0 * m.gam0;
  m.qm0 = b->charge / b->mass;
  const double hgap = 2.0 * e->v[3];
  const double ang = exit ? e->v[5] : e->v[4];
  m.h0 = e->v[0] / e->length;
  m.h1 = exit ? e->v[7] : e->v[6];
  m.k1 = e->v[1];
  m.psi = m.h0 * hgap * e->v[8] * (1.0 + std::sin(ang) * std::sin(ang)) / std::cos(ang);
  m.tanphi = std::tan(ang);
  m.tanphib = std::tan(ang - m.psi);
  m.tanphi2 = m.tanphi * m.tanphi;
  m.secphi2 = 1.0 + m.tanphi2;
  m.secphi3 = m.secphi2 * (1.0 / std::cos(ang));
  m.secphib2 = 1.0 + m.tanphib * m.tanphib;
  return m;
}

int izh_fpol_dipole(izh_bunch *b, const izh_element *e) {
  IZH_CHECK(dipole_check(b, e, "izh_fpol_dipole"));
  const double gam0 = -b->refpt[5];
  const double beta0 = std::sqrt(1.0 - 1.0 / gam0 / gam0);
  b->refpt[4] = b->refpt[4] + e->length / (b->c.scxl * beta0);
  return launch_dipole(b, dipole_face(b, e, false));
}

int izh_bpol_dipole(izh_bunch *b, const izh_element *e) {
  IZH_CHECK(dipole_check(b, e, "izh_bpol_dipole"));
  return launch_dipole(b, dipole_face(b, e, true));
}

int izh_sector_dipole(izh_bunch *b, double *z, double tau, const izh_element *e) {
  IZH_CHECK(dipole_check(b, e, "izh_sector_dipole"));
  if (!z) return fail(IZH_ERR_ARG, "izh_sector_dipole: bad argument");
  // Sector_Dipole (src/Appl/Dipole.f90:712-754), beta0 from the run loop
  // (src/Contrl/AccSimulator.f90:1060-1061)
  const double g = -b->refpt[5];
  const double beta0 = std::sqrt(1.0 - 1.0 / g / g);
  DipoleMap m{};
  m.mode = kDipoleSector;
  m.scxl = b->c.scxl;
  m.beta0 = beta0;
  m.gambet = beta0 / std::sqrt(1.0 - beta0 * beta0);
  m.gam0 = std::sqrt(1.0 / (1.0 - beta0 * beta0));
  m.qm0 = b->charge / b->mass;
  const double h0 = e->v[0] / e->length;
  const double rho = 1.0 / h0, th = tau * h0;
  const double s = std::sin(th), c = std::cos(th), sh2 = std::sin(th / 2.0);
  m.len = tau;
  m.R11 = c;
  m.R12 = rho * s;
  m.R16 = rho * (1.0 - c);
  m.R21 = -s / rho;
  m.R22 = c;
  m.R26 = s;
  m.R51 = -m.R26;
  m.R52 = -m.R16;
  m.R56 = rho * (s - beta0 * beta0 * th);
  m.T111 = -s * s / (2.0 * rho);
  m.T112 = s * c;
  m.T116 = s * s;
  m.T122 = rho * c * sh2 * sh2;
  m.T126 = -rho * s * (-1.0 + c);
  m.T166 = -rho * s * s / 2.0;
  m.T144 = rho * (-1.0 + c) / 2.0;
  m.T216 = s / rho;
  m.T222 = -s / 2.0;
  m.T266 = -s;
  m.T244 = -s / 2.0;
  m.T314 = s;
  m.T324 = 2.0 * rho * sh2 * sh2;
  m.T346 = rho * (th - s);
  m.T522 = -0.5 * rho * s;
  m.T544 = m.T522;   // the reference keeps T544 = T522
  m.T526 = rho * (c - 1.0);
  IZH_CHECK(launch_dipole(b, m));
  *z = *z + tau;
  return IZH_OK;
}

int izh_isr_chicane(izh_bunch *b, double tau, const izh_element *e, const double *rd,
                    int64_t nrd) {
  IZH_CHECK(dipole_check(b, e, "izh_isr_chicane"));
  if (!rd && b->n > 0) return fail(IZH_ERR_ARG, "izh_isr_chicane: no deviates");
  // the kernel reads rd[orig], and orig runs over the particles of the upload
  if (b->n > 0 && nrd < b->n0)
    return fail(IZH_ERR_ARG, "izh_isr_chicane: %lld deviates for a bunch uploaded with "
                "%lld particles (one per uploaded particle, indexed by its upload row)",
                (long long)nrd, (long long)b->n0);
  // b0 = |B rho angle / L| (src/Contrl/AccSimulator.f90:1436-1439) and the
  // constants of ISRchicane (:1553-1570)
  const double clight = b->c.clight;
  const double gamma0 = -b->refpt[5];
  const double beta = std::sqrt((gamma0 + 1.0) * (gamma0 - 1.0)) / gamma0;
  const double brho = gamma0 * beta / clight * b->mass;
  const double b0 = std::fabs(brho * e->v[0] / e->length);
  const double re = 2.81794028946e-15, emass = 9.109382154e-31, hbar = 1.054571628e-34;
  const double qmass = 1.0 / 0.511e6;
  const double gambet0 = std::sqrt(gamma0 * gamma0 - 1.0);
  const double c1 = 2.0 * re * gambet0 * gambet0 / 3.0;
  const double c2 = 55.0 * re * hbar * gambet0 * gambet0 * gambet0 /
                    (24.0 * std::sqrt(3.0) * (emass * clight));
  const double bb = b0 * qmass * clight / gambet0;
  const double p1 = c1 * bb * bb * tau;
  const double sb = std::sqrt(bb);
  const double p2 = std::sqrt(c2) * (sb * sb * sb) * std::sqrt(tau);
  return launch_isr(b, gamma0, p1, p2, rd);
}

int izh_map_linear(izh_bunch *b, const double xm[36]) {
  if (!b || !xm) return fail(IZH_ERR_ARG, "izh_map_linear: bad argument");
  if (b->frame != kFrameZ)
    return fail(IZH_ERR_STATE, "izh_map_linear: bunch is not in the z frame");
  return launch_map_linear(b, xm);
}

// --------------------------------------------------------- conv1st

int izh_conv1st(izh_bunch *b, double tau, double ptrange[6], int64_t *nplc,
                int64_t *nptot, double xrad, double yrad, int flagbc) {
  (void)tau;
  if (!b) return fail(IZH_ERR_ARG, "izh_conv1st: bad argument");
  if (flagbc != 1)
    return fail(IZH_ERR_ARG, "izh_conv1st: only Flagbc=1 (3D open) is supported");
  if (b->frame != kFrameZ)
    return fail(IZH_ERR_STATE, "izh_conv1st: bunch is not in the z frame");
  double pr[6];
  int64_t nlost = 0;
  if (b->n > 0) {
    IZH_CHECK(launch_conv1st(b, xrad, yrad, pr, &nlost));
  } else {
    // empty bunch: the reference's initial +-1e20 (REAL(4)) range survives
    for (int k = 0; k < 3; k++) {
      pr[2 * k] = (double)1.0e20f;
      pr[2 * k + 1] = -(double)1.0e20f;
    }
  }
  b->frame = kFrameT1;
  if (nlost > 0) {
    if (!b->r_alt[0]) IZH_CHECK(bunch_alloc_alt(b));
    IZH_CHECK(launch_compact(b, xrad, yrad, b->n - nlost));
  }
  // one process: nptot = nplc; Current = Current*nptot/Npt (:1657-1663)
  const int64_t nnew = b->n;
  if (b->npt > 0) b->current = b->current * (double)nnew / (double)b->npt;
  b->npt = nnew;
  if (ptrange) std::memcpy(ptrange, pr, sizeof pr);
  if (nplc) *nplc = nnew;
  if (nptot) *nptot = nnew;
  return IZH_OK;
}

int izh_update_compdom(izh_compdom *dom, const double pr[6], int flagbc) {
  if (!dom || !pr) return fail(IZH_ERR_ARG, "izh_update_compdom: bad argument");
  if (flagbc != 1)
    return fail(IZH_ERR_ARG, "izh_update_compdom: only Flagbc=1 is supported");
  if (dom->nx < 4 || dom->ny < 4 || dom->nz < 4)
    return fail(IZH_ERR_ARG, "izh_update_compdom: mesh %dx%dx%d below 4",
                dom->nx, dom->ny, dom->nz);
  // src/Appl/CompDom.f90:547-557, 614-622
  const int nn[3] = {dom->nx, dom->ny, dom->nz};
  for (int a = 0; a < 3; a++) {
    const double eps = eps_real4(nn[a]);
    const double lo = pr[2 * a], hi = pr[2 * a + 1];
    dom->range[2 * a] = lo - eps * (hi - lo);
    dom->range[2 * a + 1] = hi + eps * (hi - lo);
    dom->msize[a] = (dom->range[2 * a + 1] - dom->range[2 * a]) / (nn[a] - 1);
  }
  return IZH_OK;
}

// ------------------------------------------------------------ fields

static bool pow2(int v) { return v >= 4 && (v & (v - 1)) == 0; }

int izh_field_create(int nx, int ny, int nz, int precision, izh_field **out) {
  if (!out) return fail(IZH_ERR_ARG, "izh_field_create: bad argument");
  *out = nullptr;
  if (!pow2(nx) || !pow2(ny) || !pow2(nz) || 2 * nx > 2048 || 2 * ny > 2048 ||
      2 * nz > 2048)
    return fail(IZH_ERR_ARG,
                "izh_field_create: mesh %dx%dx%d must be powers of two in [4,1024]",
                nx, ny, nz);
  if ((int64_t)(nx + 2) * (ny + 2) * (nz + 2) > INT32_MAX)
    return fail(IZH_ERR_ARG, "izh_field_create: mesh %dx%dx%d too large", nx, ny, nz);
  if (precision != IZH_FP64 && precision != IZH_MIXED)
    return fail(IZH_ERR_ARG, "izh_field_create: bad precision %d", precision);
  izh_field *f = new izh_field();
  IZH_HIP(hipGetDevice(&f->device));
  f->precision = precision;
  f->nx = nx;
  f->ny = ny;
  f->nz = nz;
  f->g = (int64_t)nx * ny * nz;
  if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess) {
    delete f;
    return fail(IZH_ERR_HIP, "izh_field_create: stream creation failed");
  }
  int rc = field_alloc(f);
  if (rc != IZH_OK) {
    izh_field_destroy(f);
    return rc;
  }
  *out = f;
  return IZH_OK;
}

int izh_field_destroy(izh_field *f) {
  if (!f) return IZH_OK;
  if (f->stream) (void)hipStreamSynchronize(f->stream);
  void *bufs[] = {f->rho_fx, f->rho,  f->phi,    f->e,    f->spec_a,
                  f->spec_b, f->green, f->green_t, f->corner, f->tw_x,
                  f->tw_y,   f->tw_z, f->d_dom,  f->chunks, f->d_wake};
  for (void *p : bufs)
    if (p) (void)hipFree(p);
  if (f->h_dom) (void)hipHostFree(f->h_dom);
  if (f->stream) (void)hipStreamDestroy(f->stream);
  delete f;
  return IZH_OK;
}

// Kernels of a bunch and a field run on their own streams while they share
// data: the consumer's stream waits for everything queued on the producer's.
static int join_streams(hipStream_t producer, hipStream_t consumer) {
  if (producer == consumer) return IZH_OK;
  hipEvent_t ev;
  IZH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  IZH_HIP(hipEventRecord(ev, producer));
  IZH_HIP(hipStreamWaitEvent(consumer, ev, 0));
  IZH_HIP(hipEventDestroy(ev));
  return IZH_OK;
}

static int check_dom(const izh_field *f, const izh_compdom *dom, const char *who) {
  if (!dom) return fail(IZH_ERR_ARG, "%s: no domain", who);
  if (dom->nx != f->nx || dom->ny != f->ny || dom->nz != f->nz)
    return fail(IZH_ERR_ARG, "%s: domain mesh %dx%dx%d != field %dx%dx%d", who,
                dom->nx, dom->ny, dom->nz, f->nx, f->ny, f->nz);
  for (int a = 0; a < 3; a++)
    if (!(dom->msize[a] > 0) || !std::isfinite(dom->msize[a]))
      return fail(IZH_ERR_ARG, "%s: mesh size %g not positive", who, dom->msize[a]);
  return IZH_OK;
}

int izh_charge(izh_bunch *b, const izh_compdom *dom, izh_field *f) {
  if (!b || !f) return fail(IZH_ERR_ARG, "izh_charge: bad argument");
  if (b->frame != kFrameT1)
    return fail(IZH_ERR_STATE, "izh_charge: call izh_conv1st first");
  IZH_CHECK(check_dom(f, dom, "izh_charge"));
  if (b->precision != f->precision)
    return fail(IZH_ERR_ARG, "izh_charge: bunch and field precision differ");
  IZH_CHECK(field_set_domain(f, dom));
  IZH_CHECK(join_streams(b->stream, f->stream));
  IZH_CHECK(launch_deposit_granular(b, f));
  IZH_CHECK(join_streams(f->stream, b->stream));
  return IZH_OK;
}

int izh_update3O(izh_field *f, const izh_compdom *dom) {
  if (!f) return fail(IZH_ERR_ARG, "izh_update3O: bad argument");
  IZH_CHECK(check_dom(f, dom, "izh_update3O"));
  if (!f->have_rho) return fail(IZH_ERR_STATE, "izh_update3O: no charge density");
  IZH_CHECK(field_set_domain(f, dom));
  IZH_CHECK(launch_solve(f));
  return IZH_OK;
}

int izh_cvbkforth1st(izh_bunch *b) {
  if (!b) return fail(IZH_ERR_ARG, "izh_cvbkforth1st: bad argument");
  if (b->frame != kFrameT1)
    return fail(IZH_ERR_STATE, "izh_cvbkforth1st: bunch is not in the t frame");
  IZH_CHECK(launch_cvbk(b));
  b->frame = kFrameT0;
  return IZH_OK;
}

int izh_kick1(izh_bunch *b, double tau, izh_field *f, const izh_compdom *dom,
              int flagbc) {
  if (!b || !f) return fail(IZH_ERR_ARG, "izh_kick1: bad argument");
  if (flagbc != 1) return fail(IZH_ERR_ARG, "izh_kick1: only Flagbc=1 is supported");
  if (b->frame != kFrameT0)
    return fail(IZH_ERR_STATE, "izh_kick1: call izh_cvbkforth1st first");
  // src/Appl/BeamBunch.f90:1731: no kick below 1e-20 A
  if (b->current < (double)1.0e-20f) {
    b->frame = kFrameZ;
    return IZH_OK;
  }
  IZH_CHECK(check_dom(f, dom, "izh_kick1"));
  if (!f->have_phi) return fail(IZH_ERR_STATE, "izh_kick1: no potential");
  if (b->precision != f->precision)
    return fail(IZH_ERR_ARG, "izh_kick1: bunch and field precision differ");
  IZH_CHECK(field_set_domain(f, dom));
  IZH_CHECK(launch_efield(f));
  IZH_CHECK(join_streams(f->stream, b->stream));
  double xycon, tcon;
  sc_constants(b, &xycon, &tcon);
  IZH_CHECK(launch_gather_kick(b, f, tau, xycon, tcon));
  IZH_CHECK(join_streams(b->stream, f->stream));
  b->frame = kFrameZ;
  return IZH_OK;
}

// ------------------------------------------------ Lorentz integrator

}  // extern "C"

namespace izh {
// The on-axis field of getfld_DTL and its first three derivatives at z
// (src/Appl/DTL.f90:811-833): read3_Data's series (Fcoef(1) = a0, then
// cosine and sine pairs over the element length, about its middle) with
// the reference's REAL(4) pi, times the field scale.
void rf_fourier_axis(const izh_rfdata &d, double escale, double length, double zedge,
                     double z, double ax[4]) {
  const double pi = (double)(2 * std::asin(1.0f));
  const double *F = d.coef.data() - 1;
  const int ndata = (int)d.coef.size();
  const double zz = z - (zedge + length / 2);
  double ez1 = F[1] / 2, ezp1 = 0.0, ezpp1 = 0.0, ezppp = 0.0;
  for (int i = 2; i <= (ndata - 1) / 2 + 1; i++) {
    const double k = (i - 1) * 2 * pi / length, arg = (i - 1) * 2 * pi * zz / length;
    const double cs = std::cos(arg), sn = std::sin(arg);
    ez1 = ez1 + F[2 * i - 2] * cs + F[2 * i - 1] * sn;
    ezp1 = ezp1 + k * (-F[2 * i - 2] * sn + F[2 * i - 1] * cs);
    ezpp1 = ezpp1 + (k * k) * (-F[2 * i - 2] * cs - F[2 * i - 1] * sn);
    ezppp = ezppp + (k * k * k) * (F[2 * i - 2] * sn - F[2 * i - 1] * cs);
  }
  ax[0] = ez1 * escale;
  ax[1] = ezp1 * escale;
  ax[2] = ezpp1 * escale;
  ax[3] = ezppp * escale;
}
// One Fourier section of getfld_SolRF (src/Appl/SolRF.f90:513-547,
// 577-611): count, start, end and period at F[o+1..o+4] (1-based),
// coefficients from F[o+5]; the series and its three derivatives at zz.
static void solrf_series(const double *F, int o, double zz, double pi, double d[4]) {
  const int np = (int)(F[o + 1] + 0.1), nt = o + 4;
  const double zlen = F[o + 4];
  d[0] = F[nt + 1] / 2;
  d[1] = d[2] = d[3] = 0.0;
  for (int i = 2; i <= (np - 1) / 2 + 1; i++) {
    const double arg = (i - 1) * 2 * pi * zz / zlen, k = (i - 1) * 2 * pi / zlen;
    const double cs = std::cos(arg), sn = std::sin(arg);
    d[0] = d[0] + F[2 * i - 2 + nt] * cs + F[2 * i - 1 + nt] * sn;
    d[1] = d[1] + k * (-F[2 * i - 2 + nt] * sn + F[2 * i - 1 + nt] * cs);
    d[2] = d[2] + (k * k) * (-F[2 * i - 2 + nt] * cs - F[2 * i - 1 + nt] * sn);
    d[3] = d[3] + (k * k * k) * (F[2 * i - 2 + nt] * sn - F[2 * i - 1 + nt] * cs);
  }
}

// getfld_SolRF's sections at the kick's z (every particle's pos(3)): which
// cover it, the RF section's field times the scale, the solenoid section's
// unscaled; ww (the real angular frequency) and ww/clite with the
// reference's REAL(4) pi and clite.  A message if the file is too short.
const char *solrf_axis(const izh_rfdata &d, double escale, double length, double zedge,
                       double z, ElemField &e) {
  const double pi = (double)(2 * std::asin(1.0f)), clite = (double)299792458.e0f;
  const double *F = d.coef.data() - 1;
  const int n = (int)d.coef.size();
  const int np1 = n >= 1 ? (int)(F[1] + 0.1) : 0;
  if (n < 4 || np1 < 1 || 8 + np1 > n || 8 + np1 + (int)(F[5 + np1] + 0.1) > n)
    return "coefficients hold fewer values than their two section headers announce";
  e.win = 0;
  e.ww = e.v[1] * 2 * pi;
  e.xl = e.ww / clite;
  if (!(z > zedge && z < zedge + length)) return nullptr;
  const double zlc = z - zedge;
  double dd[4];
  if (zlc >= F[2] && zlc <= F[3]) {
    solrf_series(F, 0, z - zedge - F[2] - F[4] / 2, pi, dd);
    for (int a = 0; a < 4; a++) e.ax[a] = dd[a] * escale;
    e.win |= 1;
  }
  const int o = 4 + np1;
  if (zlc >= F[o + 2] && zlc <= F[o + 3]) {
    solrf_series(F, o, z - zedge - F[o + 2] - F[o + 4] / 2, pi, e.bz);
    e.win |= 2;
  }
  return nullptr;
}
}  // namespace izh

extern "C" {

int izh_drift2(izh_bunch *b, double *z, double tau) {
  if (!b || !z) return fail(IZH_ERR_ARG, "izh_drift2: bad argument");
  if (b->frame != kFrameZ)
    return fail(IZH_ERR_STATE, "izh_drift2: bunch is not in the z frame");
  // src/Appl/BeamBunch.f90:481-504
  const double gamma0 = -b->refpt[5];
  const double beta0 = std::sqrt(1.0 - 1.0 / gamma0 / gamma0);
  IZH_CHECK(launch_drift2(b, tau, gamma0, beta0));
  b->refpt[4] = b->refpt[4] + 0.5 * tau / (b->c.scxl * beta0);
  *z = *z + 0.5 * tau;
  return IZH_OK;
}

static int kick2_impl(izh_bunch *b, double z, double tau, const izh_element *e,
                      izh_field *f, const izh_compdom *dom, int flagbc, bool wake) {
  if (!b || !e) return fail(IZH_ERR_ARG, "izh_kick2: bad argument");
  if (flagbc != 1) return fail(IZH_ERR_ARG, "izh_kick2: only Flagbc=1 is supported");
  if (const char *why = elem_field_unsupported(e->bitype, e->v))
    return fail(IZH_ERR_ARG, "izh_kick2: bitype %d: %s", e->bitype, why);
  // scatter2 after cvbkforth1st (0th-order t frame, current > 0); scatter20
  // on a z-frame bunch (src/Appl/BeamBunch.f90:2231,2487-2499)
  const bool sc = b->frame == kFrameT0 && b->current > 0.0;
  if (b->frame != kFrameZ && b->frame != kFrameT0)
    return fail(IZH_ERR_STATE, "izh_kick2: bunch is neither in the z frame nor after "
                "izh_cvbkforth1st");
  if (sc) {
    if (!f) return fail(IZH_ERR_ARG, "izh_kick2: space charge needs the field");
    IZH_CHECK(check_dom(f, dom, "izh_kick2"));
    if (!f->have_phi) return fail(IZH_ERR_STATE, "izh_kick2: no potential");
    if (b->precision != f->precision)
      return fail(IZH_ERR_ARG, "izh_kick2: bunch and field precision differ");
  }
  Kick2 k{};
  k.e.type = e->bitype;
  for (int a = 0; a < 9; a++) k.e.v[a] = e->v[a];
  k.e.length = e->length;
  k.e.zedge = e->zedge;
  if (e->bitype >= 101 && e->bitype <= 105) {
    if (!e->rf || !e->rf->fourier)
      return fail(IZH_ERR_ARG, "izh_kick2: bitype %d needs its Fourier coefficients "
                  "(izh_rfdata_fourier, as read3_Data reads them)", e->bitype);
    k.e.theta0 = e->v[2] * (double)std::asin(1.0f) / 90;
    if (e->bitype == 105) {
      if (const char *why = solrf_axis(*e->rf, e->v[0], e->length, e->zedge, z, k.e))
        return fail(IZH_ERR_ARG, "izh_kick2: SolRF: %s", why);
      k.e.bscale = e->v[10];
      k.e.tdiv = 2 * (double)(2 * std::asin(1.0f)) * b->c.scfreq;
    } else {
      rf_fourier_axis(*e->rf, e->v[0], e->length, e->zedge, z, k.e.ax);
      k.e.ww = e->v[1] / b->c.scfreq;
      k.e.xl = b->c.scxl / k.e.ww;
    }
  }
  k.gam = -b->refpt[5];
  k.clight = b->c.clight;
  k.xl = b->c.scxl;
  k.xk = 1.0 / b->c.scxl;
  k.beta0_c = std::sqrt(k.gam * k.gam - 1.0) / k.gam / b->c.clight;
  k.tg = b->refpt[4];
  k.z = z;
  k.tau = tau;
  k.qmcc = b->charge / b->mass;
  const double pos0[4] = {0.0, 0.0, z, k.tg};
  double f0[6];
  elem_field(k.e, k.xl, k.clight, pos0, f0);
  k.ez0 = f0[2];
  k.tmpgamma0 = k.gam + 0.5 * tau * k.qmcc * k.ez0;
  if (sc) {
    if (wake) {
      if (!f->have_wake)
        return fail(IZH_ERR_STATE, "izh_kick2wake: call izh_wakefield first");
      k.wake = f->d_wake;
      k.nzw = f->nz;
    }
    IZH_CHECK(field_set_domain(f, dom));
    IZH_CHECK(launch_efield(f));
    IZH_CHECK(join_streams(f->stream, b->stream));
    // Clight^2 * 1e-7 (REAL(4)) is 1/(4 pi eps0); the grid holds charge * rho_scale
    k.tmpscale = b->c.clight * b->c.clight * (double)1.0e-7f / f->rho_scale;
    // -beta0/Clight*tmpscale*eyn: the scaled factor enters through tmpscale
    IZH_CHECK(launch_kick2(b, f, k, true));
    IZH_CHECK(join_streams(b->stream, f->stream));
  } else {
    IZH_CHECK(launch_kick2(b, f, k, false));
  }
  b->refpt[5] = -(k.gam + tau * k.qmcc * k.ez0);
  b->frame = kFrameZ;
  return IZH_OK;
}

int izh_kick2(izh_bunch *b, double z, double tau, const izh_element *e, izh_field *f,
              const izh_compdom *dom, int flagbc) {
  return kick2_impl(b, z, tau, e, f, dom, flagbc, false);
}

int izh_kick2wake(izh_bunch *b, double z, double tau, const izh_element *e, izh_field *f,
                  const izh_compdom *dom, int flagbc) {
  return kick2_impl(b, z, tau, e, f, dom, flagbc, true);
}

int izh_step2(izh_bunch *b, izh_field *f, izh_compdom *dom, double *z, double tau1,
              const izh_element *e, double xrad, double yrad, double bcurr,
              const izh_wake *wake, double *moments) {
  if (!b || !f || !dom || !z || !e) return fail(IZH_ERR_ARG, "izh_step2: bad argument");
  if (const char *why = elem_field_unsupported(e->bitype, e->v))
    return fail(IZH_ERR_ARG, "izh_step2: bitype %d: %s", e->bitype, why);
  if ((e->bitype == 105 || e->bitype == 106) && e->v[13] > 0.0 && !wake)
    return fail(IZH_ERR_ARG, "izh_step2: bitype %d with a wake length (Param(15) > 0) "
                "needs its izh_wake", e->bitype);
  const bool wk = wake && wake->kind != 0;
  const double tau2 = 2.0 * tau1;
  IZH_CHECK(izh_drift2(b, z, tau2));                                        // :1112
  if (bcurr >= (double)1.0e-10f && b->n > 0) {
    double ptrange[6];
    int64_t nplc, nptot;
    IZH_CHECK(izh_conv1st(b, tau2, ptrange, &nplc, &nptot, xrad, yrad, 1));
    IZH_CHECK(izh_update_compdom(dom, ptrange, 1));
    IZH_CHECK(izh_charge(b, dom, f));
    IZH_CHECK(izh_update3O(f, dom));
    if (wk) IZH_CHECK(izh_wakefield(b, f, dom, wake));                      // :1272
    IZH_CHECK(izh_cvbkforth1st(b));
  }
  if (wk && bcurr >= (double)1.0e-10f && b->n > 0)
    IZH_CHECK(izh_kick2wake(b, *z, tau2, e, f, dom, 1));                    // :1448
  else
    IZH_CHECK(izh_kick2(b, *z, tau2, e, f, dom, 1));                        // :1452
  IZH_CHECK(izh_drift2(b, z, tau2));                                        // :1456
  if (moments) IZH_CHECK(izh_diagnostic1(b, moments, moments + 37, moments + 43));
  return IZH_OK;
}

int izh_diagnostic1(izh_bunch *b, double sums[37], double amax[6],
                    double *r2max) {
  if (!b) return fail(IZH_ERR_ARG, "izh_diagnostic1: bad argument");
  double m[kMoments];
  if (b->n > 0) {
    IZH_CHECK(launch_moments(b, m));
  } else {
    for (int k = 0; k < kMoments; k++) m[k] = 0.0;
  }
  if (sums) std::memcpy(sums, m, 37 * sizeof(double));
  if (amax) std::memcpy(amax, m + 37, 6 * sizeof(double));
  if (r2max) *r2max = m[43];
  return IZH_OK;
}

int izh_diagnostic1_rows(const double *rows, int64_t ld, int64_t n, double *out,
                         void *stream) {
  if (n < 0 || (n > 0 && !rows) || !out || ld < n)
    return fail(IZH_ERR_ARG, "izh_diagnostic1_rows: bad argument (rows %p, ld %lld, n %lld, "
                "out %p)", (const void *)rows, (long long)ld, (long long)n, (void *)out);
  if (ld > IZH_DIAG_ROWS_MAX)
    return fail(IZH_ERR_ARG, "izh_diagnostic1_rows: row length %lld beyond the maximum %lld",
                (long long)ld, (long long)IZH_DIAG_ROWS_MAX);
  return launch_moments_rows(rows, ld, n, out, (hipStream_t)stream);
}

int izh_chgupdate(izh_bunch *b, int nchge, const double *qmcclist, int64_t *counts) {
  if (!b || !qmcclist || !counts || nchge < 1 || nchge > kMaxChargeStates)
    return fail(IZH_ERR_ARG, "izh_chgupdate: bad argument (nchge %d, at most %d)", nchge,
                kMaxChargeStates);
  ChgList cl{};
  cl.n = nchge;
  cl.eps = (double)1.0e-4f;  // eps = 1.0e-4, a REAL(4) literal (BeamBunch.f90:700)
  for (int j = 0; j < nchge; j++) {
    if (qmcclist[j] == 0.0)
      return fail(IZH_ERR_ARG, "izh_chgupdate: charge state %d has q/mc^2 = 0", j + 1);
    cl.qmcc[j] = qmcclist[j];
  }
  if (b->qm_uniform || b->n == 0) {
    // every live particle has the same q/mc^2: the counts are n or 0
    for (int j = 0; j < nchge; j++)
      counts[j] = std::fabs((b->qm_value - cl.qmcc[j]) / cl.qmcc[j]) < cl.eps ? b->n : 0;
    return IZH_OK;
  }
  return launch_chg_count(b, cl, counts);
}

int izh_field_download(izh_field *f, int which, double *host) {
  if (!f || !host) return fail(IZH_ERR_ARG, "izh_field_download: bad argument");
  if (which < IZH_RHO || which > IZH_EZ)
    return fail(IZH_ERR_ARG, "izh_field_download: unknown buffer %d", which);
  IZH_HIP(hipStreamSynchronize(f->stream));
  // the grid holds physical values times rho_scale (1 in FP64); rho, phi and
  // E are linear in the charge, so one factor restores physical units
  const double un = 1.0 / f->rho_scale;
  const bool fp64 = f->precision == IZH_FP64;
  const size_t es = fp64 ? sizeof(double) : sizeof(float);
  const int stride = which >= IZH_EX ? 4 : 1;          // E is interleaved
  const int comp = which >= IZH_EX ? which - IZH_EX : 0;
  const void *src = which == IZH_RHO ? f->rho : (which == IZH_PHI ? f->phi : f->e);
  std::vector<char> tmp((size_t)f->g * stride * es);
  IZH_HIP(hipMemcpy(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < f->g; i++) {
    const int64_t k = i * stride + comp;
    const double v = fp64 ? ((const double *)tmp.data())[k]
                          : (double)((const float *)tmp.data())[k];
    host[i] = v * un;
  }
  return IZH_OK;
}

int izh_field_upload_rho(izh_field *f, const double *rho) {
  if (!f || !rho) return fail(IZH_ERR_ARG, "izh_field_upload_rho: bad argument");
  IZH_HIP(hipStreamSynchronize(f->stream));
  f->rho_scale = 1.0;
  if (f->precision == IZH_FP64) {
    IZH_HIP(hipMemcpy(f->rho, rho, f->g * sizeof(double), hipMemcpyHostToDevice));
  } else {
    double mx = 0.0;
    for (int64_t i = 0; i < f->g; i++) mx = std::fmax(mx, std::fabs(rho[i]));
    f->rho_scale = mx > 0 ? 1.0 / mx : 1.0;
    std::vector<float> tmp((size_t)f->g);
    for (int64_t i = 0; i < f->g; i++) tmp[i] = (float)(rho[i] * f->rho_scale);
    IZH_HIP(hipMemcpy(f->rho, tmp.data(), f->g * sizeof(float), hipMemcpyHostToDevice));
  }
  f->have_rho = 1;
  f->have_phi = 0;
  return IZH_OK;
}

// ---------------------------------------------------------- wakefields

}  // extern "C"

namespace izh {
// The kernel inputs of a wake for the field's mesh; the table moves to the
// device on first use.
int wake_params_of(const izh_bunch *b, const izh_field *f, const izh_wake *w,
                   WakeParams *p) {
  if (!w || w->kind < 1 || w->kind > 3)
    return fail(IZH_ERR_ARG, "wake: kind must be 1 (analytic), 2 (read-in table) or 3 (CSR)");
  if (f->nz > kWakeMaxNz)
    return fail(IZH_ERR_ARG, "wake: Nz = %d above %d", f->nz, kWakeMaxNz);
  *p = WakeParams{};
  p->kind = w->kind;
  p->nz = f->nz;
  p->flagbtw = w->flagbtw;
  p->scale = w->scale;
  if (w->kind == 1) {
    if (!(w->aa > 0) || !(w->gg > 0) || !(w->leng > 0))
      return fail(IZH_ERR_ARG, "wake: iris radius, gap and period must be positive");
    p->aa = w->aa;
    p->gg = w->gg;
    p->leng = w->leng;
  } else if (w->kind == 3) {
    if (!(w->r0 > 0) || !std::isfinite(w->r0) || !(w->blength > 0) ||
        !std::isfinite(w->zoff))
      return fail(IZH_ERR_ARG, "wake: a CSR wake needs a positive finite bend radius and "
                  "length and a finite offset");
    p->r0 = w->r0;
    p->zoff = w->zoff;
    p->blength = w->blength;
  } else {
    if (!w->table || w->table->fourier || w->table->z.size() < 2)
      return fail(IZH_ERR_ARG, "wake: a read-in wake needs a table of at least 2 rows "
                  "(izh_rfdata_table)");
    IZH_CHECK(wake_table_device(w->table, &p->tab));
    p->ntab = (int)w->table->z.size();
  }
  p->gam = -b->refpt[5];
  p->mass = b->mass;
  p->charge = b->charge;
  p->scfreq = b->c.scfreq;
  p->clight = b->c.clight;
  p->current = b->current;
  p->npt = (double)b->npt;
  sc_constants(b, &p->xycon, &p->tcon);
  if (!(b->current > 0)) p->xycon = p->tcon = 0.0;
  return IZH_OK;
}

// A SolRF/TWS element with a wake length switches the wake on; izh_step and
// izh_run take it from the segment (src/Contrl/AccSimulator.f90:859-870).
static bool needs_wake(const izh_element &e) {
  return (e.bitype == 105 || e.bitype == 106) && e.v[13] > 0.0;
}
}  // namespace izh

extern "C" {

int izh_wakefield(izh_bunch *b, izh_field *f, const izh_compdom *dom, const izh_wake *w) {
  if (!b || !f || !w) return fail(IZH_ERR_ARG, "izh_wakefield: bad argument");
  if (b->frame != kFrameT1)
    return fail(IZH_ERR_STATE, "izh_wakefield: call izh_conv1st first (t frame)");
  IZH_CHECK(check_dom(f, dom, "izh_wakefield"));
  WakeParams p;
  IZH_CHECK(wake_params_of(b, f, w, &p));
  IZH_CHECK(field_set_domain(f, dom));
  IZH_CHECK(join_streams(b->stream, f->stream));
  IZH_CHECK(launch_wake(b, f, p, nullptr, nullptr, nullptr, 0));
  IZH_CHECK(join_streams(f->stream, b->stream));
  return IZH_OK;
}

int izh_kick1wake(izh_bunch *b, double tau, izh_field *f, const izh_compdom *dom,
                  int flagbc) {
  if (!b || !f) return fail(IZH_ERR_ARG, "izh_kick1wake: bad argument");
  if (flagbc != 1) return fail(IZH_ERR_ARG, "izh_kick1wake: only Flagbc=1 is supported");
  if (b->frame != kFrameT0)
    return fail(IZH_ERR_STATE, "izh_kick1wake: call izh_cvbkforth1st first");
  if (b->current < (double)1.0e-20f) {   // src/Appl/BeamBunch.f90:4720
    b->frame = kFrameZ;
    return IZH_OK;
  }
  IZH_CHECK(check_dom(f, dom, "izh_kick1wake"));
  if (!f->have_phi) return fail(IZH_ERR_STATE, "izh_kick1wake: no potential");
  if (!f->have_wake) return fail(IZH_ERR_STATE, "izh_kick1wake: call izh_wakefield first");
  if (b->precision != f->precision)
    return fail(IZH_ERR_ARG, "izh_kick1wake: bunch and field precision differ");
  IZH_CHECK(field_set_domain(f, dom));
  IZH_CHECK(launch_efield(f));
  IZH_CHECK(launch_wake_add(f));
  IZH_CHECK(join_streams(f->stream, b->stream));
  double xycon, tcon;
  sc_constants(b, &xycon, &tcon);
  IZH_CHECK(launch_gather_kick(b, f, tau, xycon, tcon));
  IZH_CHECK(join_streams(b->stream, f->stream));
  b->frame = kFrameZ;
  return IZH_OK;
}

int izh_wake_download(izh_field *f, double *ex, double *ey, double *ez) {
  if (!f || !ex || !ey || !ez) return fail(IZH_ERR_ARG, "izh_wake_download: bad argument");
  if (!f->have_wake) return fail(IZH_ERR_STATE, "izh_wake_download: no wake computed");
  IZH_HIP(hipStreamSynchronize(f->stream));
  const size_t nz = (size_t)f->nz;
  IZH_HIP(hipMemcpy(ex, f->d_wake, nz * sizeof(double), hipMemcpyDeviceToHost));
  IZH_HIP(hipMemcpy(ey, f->d_wake + nz, nz * sizeof(double), hipMemcpyDeviceToHost));
  IZH_HIP(hipMemcpy(ez, f->d_wake + 2 * nz, nz * sizeof(double), hipMemcpyDeviceToHost));
  return IZH_OK;
}

// ---------------------------------------------------------------- step

int izh_step(izh_bunch *b, izh_field *f, izh_compdom *dom, double *z,
             const izh_segment *seg, double bcurr, double *moments) {
  if (!b || !f || !dom || !z || !seg) return fail(IZH_ERR_ARG, "izh_step: bad argument");
  if (needs_wake(seg->e) && !seg->wake)
    return fail(IZH_ERR_ARG, "izh_step: bitype %d with a wake length (Param(15) > 0) "
                "needs its izh_wake", seg->e.bitype);
  const bool wake = seg->wake && seg->wake->kind != 0;
  const double tau1 = seg->tau1, tau2 = 2.0 * tau1;
  // src/Contrl/AccSimulator.f90:1106-1110 first half push
  IZH_CHECK(izh_map1(b, z, tau1, &seg->e));
  // an empty bunch has nothing to deposit or kick (the reference would still
  // build a domain from its +-1e20 initial range; nothing reads it)
  if (bcurr >= (double)1.0e-10f && b->n > 0) {
    double ptrange[6];
    int64_t nplc, nptot;
    IZH_CHECK(izh_conv1st(b, tau2, ptrange, &nplc, &nptot, seg->xrad, seg->yrad, 1));
    IZH_CHECK(izh_update_compdom(dom, ptrange, 1));                         // :1156
    IZH_CHECK(izh_charge(b, dom, f));                                       // :1191
    IZH_CHECK(izh_update3O(f, dom));                                        // :1263
    if (wake) IZH_CHECK(izh_wakefield(b, f, dom, seg->wake));               // :1272
    IZH_CHECK(izh_cvbkforth1st(b));                                         // :1399
    if (wake)
      IZH_CHECK(izh_kick1wake(b, tau2, f, dom, 1));                         // :1417
    else
      IZH_CHECK(izh_kick1(b, tau2, f, dom, 1));                             // :1425
  }
  IZH_CHECK(izh_map1(b, z, tau1, &seg->e));                                // :1429
  if (moments) IZH_CHECK(izh_diagnostic1(b, moments, moments + 37, moments + 43));
  return IZH_OK;
}

int izh_run(izh_bunch *b, izh_field *f, izh_compdom *dom, double *z,
            const izh_segment *segs, int64_t nseg, double bcurr,
            double *moments) {
  if (!b || !f || !dom || !z || (nseg > 0 && !segs) || nseg < 0)
    return fail(IZH_ERR_ARG, "izh_run: bad argument");
  if (b->frame != kFrameZ)
    return fail(IZH_ERR_STATE, "izh_run: bunch is not in the z frame");
  if (dom->nx != f->nx || dom->ny != f->ny || dom->nz != f->nz)
    return fail(IZH_ERR_ARG, "izh_run: domain mesh differs from the field's");
  if (b->precision != f->precision)
    return fail(IZH_ERR_ARG, "izh_run: bunch and field precision differ");
  for (int64_t s = 0; s < nseg; s++) {
    const izh_segment &sg = segs[s];
    const char *why = nullptr;
    const int kind = map1_kind(sg.e.bitype, sg.e.v, &why);
    if (kind < 0)
      return fail(IZH_ERR_ARG, "izh_run: segment %lld: bitype %d: %s", (long long)s,
                  sg.e.bitype, why);
    if (!(sg.tau1 >= 0) || (kind == kMap1Linear && sg.e.mapstp < 1))
      return fail(IZH_ERR_ARG, "izh_run: segment %lld: bad step or map steps", (long long)s);
    if (needs_wake(sg.e) && !sg.wake)
      return fail(IZH_ERR_ARG, "izh_run: segment %lld: bitype %d with a wake length "
                  "(Param(15) > 0) needs its izh_wake", (long long)s, sg.e.bitype);
    if (sg.wake && sg.wake->kind != 0) {
      WakeParams p;
      IZH_CHECK(wake_params_of(b, f, sg.wake, &p));
    }
  }
  if (nseg == 0) return IZH_OK;
  // no space charge or nothing to carry it: the reference loop is maps only
  if (bcurr < (double)1.0e-10f || b->n == 0) {
    for (int64_t s = 0; s < nseg; s++) {
      double *rec = moments ? moments + s * IZH_RUN_RECORD : nullptr;
      IZH_CHECK(izh_step(b, f, dom, z, &segs[s], bcurr, rec));
      if (rec) {
        rec[kMoments] = (double)b->npt;
        rec[kMoments + 1] = b->refpt[4];
        rec[kMoments + 2] = b->refpt[5];
      }
    }
    return IZH_OK;
  }
  return run_steps(b, f, dom, z, segs, nseg, moments);
}

int izh_bunch_set_tuning(izh_bunch *b, const izh_tuning *t) {
  if (!b || !t) return fail(IZH_ERR_ARG, "izh_bunch_set_tuning: bad argument");
  if (t->resort_period < 0 || t->resort_period > 1 << 20)
    return fail(IZH_ERR_ARG, "izh_bunch_set_tuning: resort_period %d outside 0..2^20",
                t->resort_period);
  b->resort_period = t->resort_period;
  return IZH_OK;
}

int izh_bunch_get_tuning(const izh_bunch *b, izh_tuning *t) {
  if (!b || !t) return fail(IZH_ERR_ARG, "izh_bunch_get_tuning: bad argument");
  *t = izh_tuning{};
  t->resort_period = b->resort_period > 0 ? b->resort_period : kDefaultResortPeriod;
  return IZH_OK;
}

int izh_run_stats(const izh_bunch *b, int64_t stats[4]) {
  if (!b || !stats) return fail(IZH_ERR_ARG, "izh_run_stats: bad argument");
  for (int k = 0; k < 4; k++) stats[k] = b->run_stats[k];
  return IZH_OK;
}

int izh_probe_copy(int64_t nbytes, double *ms) { return launch_probe_copy(nbytes, ms); }

}  // extern "C"
