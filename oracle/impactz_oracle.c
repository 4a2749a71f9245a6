/*
 * impactz_oracle.c -- CPU restatement of the IMPACT-Z space-charge step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Each function
 * cites the reference routine it restates; operation order and the REAL(4)
 * literals of the reference are kept where they change FP64 results:
 *   - update_CompDom pads by eps = 1.0/(N-3.0) evaluated in REAL(4)
 *     (src/Appl/CompDom.f90:547-549);
 *   - conv1st wraps the phase with pilc = 2.0*asin(1.0) in REAL(4)
 *     (src/Appl/BeamBunch.f90:1268);
 *   - scatter1 uses fpei = Clight*Clight*1.0e-7 with a REAL(4) 1.0e-7
 *     (src/Appl/BeamBunch.f90:2071);
 *   - diagnostic1 takes cube roots with the REAL(4) exponent 1.0/3.0
 *     (src/Contrl/Output.f90:281).
 * The only intentional deviation: with a deck current below 1e-10 A the
 * reference still calls kick1 with the previous step's field
 * (src/Contrl/AccSimulator.f90:1124,1417-1425); or_step skips that kick.
 */
#include "impactz_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Pts1(c, i) with 1-based component c, 0-based particle i. */
#define P(c, i) pts[9 * (int64_t)(i) + (c) - 1]

void or_physconst(double freq, or_const *c) {
  /* src/DataStruct/PhysConst.f90:28-33 */
  c->clight = 299792458.0;
  c->pi = 2.0 * asin(1.0);
  c->scxl = c->clight / (2.0 * c->pi * freq);
  c->scfreq = freq;
}

/* ---------------------------------------------------------------- maps */

/* getaxfldE_DTL / _CCDTL / _CCL / _SC (src/Appl/DTL.f90:350-393): linear
 * interpolation in read1_Data's table found by bisection on zdat(1..Ndata)
 * (extrapolating outside it), times the field scale. */
static void axfld_table(const or_element *e, double z, double *ez1, double *ezp1) {
  const or_rfdata *d = e->rf;
  double zz = z - e->zedge, hstep, slope;
  int klo = 1, khi = d->n, k;
  while (khi - klo > 1) {
    k = (khi + klo) / 2;
    if (d->zdat[k - 1] > zz) khi = k;
    else klo = k;
  }
  hstep = d->zdat[khi - 1] - d->zdat[klo - 1];
  slope = (d->edat[khi - 1] - d->edat[klo - 1]) / hstep;
  *ez1 = d->edat[klo - 1] + slope * (zz - d->zdat[klo - 1]);
  slope = (d->epdat[khi - 1] - d->epdat[klo - 1]) / hstep;
  *ezp1 = d->epdat[klo - 1] + slope * (zz - d->zdat[klo - 1]);
  *ez1 = *ez1 * e->v[0];
  *ezp1 = *ezp1 * e->v[0];
}

/* getaxfldE_SolRF (src/Appl/SolRF.f90:356-466): the Fourier series of
 * read3_Data's coefficients, inside the element and the field's own range
 * (with a REAL(4) 1e-10 tolerance at the ends), else 0. */
static void axfld_fourier(const or_const *c, const or_element *e, double z, double *ez1,
                          double *ezp1) {
  const double *F = e->rf->fcoef - 1; /* Fcoef(1..) */
  const double epstol = (double)1.0e-10f;
  double zedge = e->zedge, len = e->length;
  *ez1 = 0.0;
  *ezp1 = 0.0;
  if ((z >= zedge && z <= zedge + len) || fabs(z - zedge) <= epstol ||
      fabs(z - zedge - len) <= epstol) {
    double zlc = z - zedge;
    int numpar1 = (int)(F[1] + 0.1), i;
    double zstart1 = F[2], zend1 = F[3], zlength1 = F[4];
    if ((zlc >= zstart1 && zlc <= zend1) || fabs(zlc - zstart1) <= epstol ||
        fabs(zlc - zend1) <= epstol) {
      double zmid = zlength1 / 2, zz = z - zedge - zstart1 - zmid, zlen = zlength1;
      *ez1 = F[5] / 2;
      for (i = 2; i <= (numpar1 - 1) / 2 + 1; i++) {
        double kk = (i - 1) * 2 * c->pi / zlen;
        double arg = (i - 1) * 2 * c->pi * zz / zlen;
        *ez1 = *ez1 + F[2 * i + 2] * cos(arg) + F[2 * i + 3] * sin(arg);
        *ezp1 = *ezp1 + kk * (-F[2 * i + 2] * sin(arg) + F[2 * i + 3] * cos(arg));
      }
      *ez1 = *ez1 * e->v[0];
      *ezp1 = *ezp1 * e->v[0];
    }
  }
}

static void axfld(const or_const *c, const or_element *e, double z, double *ez1,
                  double *ezp1) {
  if (e->rf->fourier) axfld_fourier(c, e, z, ez1, ezp1);
  else axfld_table(e, z, ez1, ezp1);
}

/* Right-hand side of the linear map ODE at position t: the reference
 * particle's phase and energy (y5, y6), the cosine- and sine-like
 * solutions of the three planes (y7..y18) and, for the solenoids, the
 * rotation pair (y19..y22).
 *   quadrupole: intfunc1_Quadrupole, constant gradient
 *               (src/Appl/Quadrupole.f90:254-305);
 *   ConstFoc:   intfunc1_ConstFoc (src/Appl/ConstFoc.f90:245-293);
 *   solenoid:   intfunc1_Sol (src/Appl/Sol.f90:271-321);
 *   RF cavity:  intfunc1_DTL (src/Appl/DTL.f90:286-347, with the
 *               drift-tube quadrupoles of getBgradfld_DTL, :395-418),
 *               intfunc1_CCDTL / _CCL / _SC (CCL.f90:269-328) and
 *               intfunc1_SolRF (SolRF.f90:287-353, Bz0 = Param(12)). */
static void lin_rhs(const or_const *c, const or_element *e, double qmcc, double t,
                    const double *y, double *f) {
  int type = e->bitype;
  const double *pv = e->v;
  double gamma0 = -y[5];
  double beta0 = sqrt(gamma0 * gamma0 - 1.0) / gamma0;
  double gbet = sqrt((gamma0 - 1.0) * (gamma0 + 1.0));
  double brho, s11, s33, s55, c2alpha = 0.0;
  f[0] = f[1] = f[2] = f[3] = 0.0;
  f[4] = 1.0 / (beta0 * c->scxl);
  f[5] = 0.0;
  if (type == 1) {
    brho = gbet / c->clight / qmcc;
    s11 = pv[0] / brho * c->scxl;
    s33 = -pv[0] / brho * c->scxl;
    s55 = 0.0;
  } else if (type == 2) {
    s11 = pv[0] * c->scxl;
    s33 = pv[1] * c->scxl;
    s55 = pv[2] * c->scxl;
  } else if (type == 3) {
    double k;
    brho = gbet / c->clight / qmcc;
    k = 0.5 * pv[0] / brho;
    s11 = (k * k) * c->scxl;
    s33 = (k * k) * c->scxl;
    s55 = 0.0;
    c2alpha = -pv[0] / brho / 2;
  } else {
    double ez1, ezp1, ww, theta0, sinphi, cosphi, rfdsgn, s11tmp, r;
    axfld(c, e, t, &ez1, &ezp1);
    ww = pv[1] / c->scfreq;
    theta0 = pv[2] * (double)asinf(1.0f) / 90;
    sinphi = sin(ww * y[4] + theta0);
    cosphi = cos(ww * y[4] + theta0);
    rfdsgn = ez1 * cosphi;
    f[5] = -qmcc * rfdsgn;
    brho = gbet / c->clight / qmcc;
    r = qmcc * rfdsgn / (beta0 * beta0) / (gamma0 * gamma0);
    s11tmp = 0.5 * (1.0 + 0.5 * gamma0 * gamma0) * (r * r) +
             qmcc * 0.5 / (beta0 * beta0 * beta0) / (gamma0 * gamma0 * gamma0) * ez1 *
                 sinphi * ww / c->scxl;
    if (type == 101) {
      /* getBgradfld_DTL: Param(7..10) = v[5..8] */
      double zz = t - e->zedge, bgrad;
      if (zz < pv[5]) bgrad = pv[6];
      else if (zz < e->length - pv[7]) bgrad = 0.0;
      else bgrad = pv[8];
      s11 = (s11tmp + bgrad / brho) * c->scxl;
      s33 = (s11tmp - bgrad / brho) * c->scxl;
    } else if (type == 105) {
      double b0 = pv[10], kb = 0.5 * b0 / brho;
      s11 = (s11tmp + (kb * kb) / 2) * c->scxl;
      s33 = (s11tmp + (kb * kb) / 2) * c->scxl;
      c2alpha = -b0 / brho / 2;
    } else {
      s11 = s11tmp * c->scxl;
      s33 = s11tmp * c->scxl;
    }
    s55 = -1.5 * qmcc / (beta0 * beta0) / gamma0 * ezp1 * cosphi +
          (beta0 * beta0 + 0.5) / (beta0 * beta0 * beta0) / gamma0 * qmcc * ez1 * sinphi *
              ww / c->scxl +
          1.5 * (1.0 - 0.5 * gamma0 * gamma0) * (r * r);
    s55 = c->scxl * s55;
  }
  f[6] = y[7] / c->scxl;
  f[7] = -s11 * y[6];
  f[8] = y[9] / c->scxl;
  f[9] = -s11 * y[8];
  f[10] = y[11] / c->scxl;
  f[11] = -s33 * y[10];
  f[12] = y[13] / c->scxl;
  f[13] = -s33 * y[12];
  f[14] = y[15] / c->scxl;
  f[15] = -s55 * y[14];
  f[16] = y[17] / c->scxl;
  f[17] = -s55 * y[16];
  if (type == 3 || type == 105) {
    f[18] = -c2alpha * y[20];
    f[19] = -c2alpha * y[21];
    f[20] = c2alpha * y[18];
    f[21] = c2alpha * y[19];
  }
}

/* dlt and thl of maplinear_DTL at the entrance or exit (src/Appl/DTL.f90:
 * 189-220): from the on-axis field and slope at z and the reference phase
 * and energy in refpt. */
static void rf_edge_terms(const or_const *c, const or_element *e, double qmcc, double z,
                          const double refpt[6], double *dlt, double *thl) {
  double ez, ez1, g = -refpt[5], beta = sqrt(g * g - 1.0) / g, u = g * beta;
  double ww = e->v[1] / c->scfreq, theta0 = e->v[2] * (double)asinf(1.0f) / 90;
  double sinth, costh, uprim, qpw;
  axfld(c, e, z, &ez, &ez1);
  sinth = sin(ww * refpt[4] + theta0);
  costh = cos(ww * refpt[4] + theta0);
  uprim = qmcc * ez / beta * costh;
  qpw = 0.5 * qmcc * c->scxl / (u * ww);
  *dlt = c->scxl * (0.5 * uprim / u - qpw * ez1 * sinth);
  *thl = 1.5 * c->scxl * (uprim / u);
}

int or_element_map(const or_const *c, double t, double tau, const or_element *el,
                   double refpt[6], double charge, double mass, double xm[36]) {
  /* maplinear_Quadrupole / _ConstFoc / _Sol / _DTL / _SolRF with rk6i_*
   * (src/Appl/Quadrupole.f90:120-252, ConstFoc.f90:116-243, Sol.f90:120-268,
   * DTL.f90:136-284, SolRF.f90:126-285) */
  enum { NV = 22 };
  double y[NV], yt[NV], f[NV], a[NV], b[NV], cc[NV], d[NV], e[NV], g[NV], o[NV],
      p[NV];
  double h, gi, betai, ui, squi, sqi3, gf, betaf, uf, squf, sqf3, tt;
  double dlti = 0.0, thli = 0.0, dltf = 0.0, thlf = 0.0;
  double qmcc = charge / mass;
  int i, k, nv, rf, bitype = el->bitype, mapstp = el->mapstp;
  const double *v = el->v;
  rf = bitype >= 101 && bitype <= 105;
  if (!((bitype >= 1 && bitype <= 3) || rf) || mapstp < 1) return -1;
  /* a quadrupole's field from a file (file ID > 1e-5) is not restated */
  if (bitype == 1 && v[1] > 1.0e-5) return -1;
  if (rf && (!el->rf || (bitype == 105) != (el->rf->fourier != 0))) return -1;
  nv = (bitype == 3 || bitype == 105) ? 22 : 18;
  memset(y, 0, sizeof y);
  memset(f, 0, sizeof f);
  y[4] = refpt[4];
  y[5] = refpt[5];
  y[6] = 1.0;
  y[9] = 1.0;
  y[10] = 1.0;
  y[13] = 1.0;
  y[14] = 1.0;
  y[17] = 1.0;
  y[18] = 1.0;
  y[21] = 1.0;
  h = tau / mapstp;
  gi = -refpt[5];
  betai = sqrt(gi * gi - 1.0) / gi;
  ui = gi * betai;
  squi = sqrt(ui);
  sqi3 = squi * squi * squi;
  if (rf) rf_edge_terms(c, el, qmcc, t, refpt, &dlti, &thli);
  for (i = 0; i < mapstp; i++) {
    tt = t + i * h;  /* rk6i_*: t = tint + i*h after each step */
    lin_rhs(c, el, qmcc, tt, y, f);
    for (k = 0; k < nv; k++) { a[k] = h * f[k]; yt[k] = y[k] + a[k] / 9.0; }
    lin_rhs(c, el, qmcc, tt + h / 9.0, yt, f);
    for (k = 0; k < nv; k++) {
      b[k] = h * f[k];
      yt[k] = y[k] + (a[k] + 3.0 * b[k]) / 24.0;
    }
    lin_rhs(c, el, qmcc, tt + h / 6.0, yt, f);
    for (k = 0; k < nv; k++) {
      cc[k] = h * f[k];
      yt[k] = y[k] + (a[k] - 3.0 * b[k] + 4.0 * cc[k]) / 6.0;
    }
    lin_rhs(c, el, qmcc, tt + h / 3.0, yt, f);
    for (k = 0; k < nv; k++) {
      d[k] = h * f[k];
      yt[k] = y[k] + (278.0 * a[k] - 945.0 * b[k] + 840.0 * cc[k] + 99.0 * d[k]) / 544.0;
    }
    lin_rhs(c, el, qmcc, tt + 0.5 * h, yt, f);
    for (k = 0; k < nv; k++) {
      e[k] = h * f[k];
      yt[k] = y[k] + (-106.0 * a[k] + 273.0 * b[k] - 104.0 * cc[k] - 107.0 * d[k] +
                      48.0 * e[k]) / 6.0;
    }
    lin_rhs(c, el, qmcc, tt + 2.0 * h / 3.0, yt, f);
    for (k = 0; k < nv; k++) {
      g[k] = h * f[k];
      yt[k] = y[k] + (110974.0 * a[k] - 236799.0 * b[k] + 68376.0 * cc[k] +
                      103803.0 * d[k] - 10240.0 * e[k] + 1926.0 * g[k]) / 45648.0;
    }
    lin_rhs(c, el, qmcc, tt + 5.0 * h / 6.0, yt, f);
    for (k = 0; k < nv; k++) {
      o[k] = h * f[k];
      yt[k] = y[k] + (-101195.0 * a[k] + 222534.0 * b[k] - 71988.0 * cc[k] -
                      26109.0 * d[k] - 2.0e4 * e[k] - 72.0 * g[k] + 22824.0 * o[k]) /
                         25994.0;
    }
    lin_rhs(c, el, qmcc, tt + h, yt, f);
    for (k = 0; k < nv; k++) {
      p[k] = h * f[k];
      y[k] = y[k] + (41.0 * a[k] + 216.0 * cc[k] + 27.0 * d[k] + 272.0 * e[k] +
                     27.0 * g[k] + 216.0 * o[k] + 41.0 * p[k]) / 840.0;
    }
  }
  refpt[4] = y[4];
  refpt[5] = y[5];
  gf = -refpt[5];
  betaf = sqrt(gf * gf - 1.0) / gf;
  uf = gf * betaf;
  squf = sqrt(uf);
  sqf3 = squf * squf * squf;
  if (rf) rf_edge_terms(c, el, qmcc, t + tau, refpt, &dltf, &thlf);
  /* xm row-major: xm[6*(j-1)+(k-1)] = xm(j,k); dlt = thl = 0 outside the
   * RF cavities (src/Appl/DTL.f90:205-220, SolRF.f90:207-231) */
  for (k = 0; k < 36; k++) xm[k] = 0.0;
  if (nv == 22) {
    /* each transverse block times the rotation (src/Appl/Sol.f90:187-207) */
    int blk;
    for (blk = 0; blk < 4; blk++) {
      int r = blk < 2 ? 0 : 2, cl = (blk % 2) * 2, oo = r == 0 ? 6 : 10;
      double w = y[18 + blk];
      xm[6 * r + cl] = w * (y[oo] * squi / squf + y[oo + 2] * squi / squf * dlti);
      xm[6 * (r + 1) + cl] = w * ((y[oo + 1] - y[oo] * dltf) * squi * squf +
                                  (y[oo + 3] - y[oo + 2] * dltf) * squi * squf * dlti);
      xm[6 * r + cl + 1] = w * y[oo + 2] / (squi * squf);
      xm[6 * (r + 1) + cl + 1] = w * (y[oo + 3] - y[oo + 2] * dltf) * squf / squi;
    }
  } else {
    xm[0] = y[6] * squi / squf + y[8] * squi / squf * dlti;
    xm[6] = (y[7] - y[6] * dltf) * squi * squf + (y[9] - y[8] * dltf) * squi * squf * dlti;
    xm[1] = y[8] / (squi * squf);
    xm[7] = (y[9] - y[8] * dltf) * squf / squi;
    xm[14] = y[10] * squi / squf + y[12] * squi / squf * dlti;
    xm[20] = (y[11] - y[10] * dltf) * squi * squf +
             (y[13] - y[12] * dltf) * squi * squf * dlti;
    xm[15] = y[12] / (squi * squf);
    xm[21] = (y[13] - y[12] * dltf) * squf / squi;
  }
  xm[28] = y[14] * sqi3 / sqf3 + y[16] * sqi3 / sqf3 * thli;
  xm[34] = (y[15] - y[14] * thlf) * sqi3 * sqf3 + (y[17] - y[16] * thlf) * sqi3 * sqf3 * thli;
  xm[29] = y[16] / (sqi3 * sqf3);
  xm[35] = (y[17] - y[16] * thlf) * sqf3 / sqi3;
  return 0;
}

void or_quad_map(const or_const *c, double tau, double grad, int mapstp,
                 double refpt[6], double charge, double mass, double xm[36]) {
  or_element e;
  memset(&e, 0, sizeof e);
  e.bitype = 1;
  e.mapstp = mapstp;
  e.v[0] = grad;
  or_element_map(c, 0.0, tau, &e, refpt, charge, mass, xm);
}

/* transfmapK_Quadrupole (src/Appl/Quadrupole.f90:509-607): the chromatic
 * map of a quadrupole given by K = G/Brho, per particle. */
static void quad_kmode(const or_const *c, double *pts, int64_t n, double refpt[6],
                       double tau, double kq, double qmass) {
  double gambet0 = sqrt(refpt[5] * refpt[5] - 1.0);
  double beta0 = sqrt(1.0 - 1.0 / (refpt[5] * refpt[5]));
  int64_t i;
  for (i = 0; i < n; i++) {
    double gam = -refpt[5] - P(6, i);
    double gambetz = sqrt(gam * gam - 1.0 - P(2, i) * P(2, i) - P(4, i) * P(4, i));
    double betaz = gambetz / gam;
    double kstr, rtkstr, rtkstrzz, cs, ss, t;
    double xm11, xm12, xm21, xm22, xm33, xm34, xm43, xm44, tmp[4];
    gambetz = sqrt(gam * gam - 1.0);
    kstr = P(7, i) / qmass * gambet0 / gambetz * kq;
    rtkstr = sqrt(fabs(kstr));
    rtkstrzz = rtkstr * tau;
    if (kstr > 0.0) {
      cs = cos(rtkstrzz);
      ss = sin(rtkstrzz);
      xm11 = cs;
      xm12 = ss / rtkstr;
      xm21 = -rtkstr * ss;
      xm22 = cs;
      t = exp(rtkstrzz);
      cs = (t * t + 1.0) / (t + t);
      ss = (t * t - 1.0) / (t + t);
      xm33 = cs;
      xm34 = ss / rtkstr;
      xm43 = ss * rtkstr;
      xm44 = cs;
    } else if (kstr < 0.0) {
      t = exp(rtkstrzz);
      cs = (t * t + 1.0) / (t + t);
      ss = (t * t - 1.0) / (t + t);
      xm11 = cs;
      xm12 = ss / rtkstr;
      xm21 = rtkstr * ss;
      xm22 = cs;
      cs = cos(rtkstrzz);
      ss = sin(rtkstrzz);
      xm33 = cs;
      xm34 = ss / rtkstr;
      xm43 = -ss * rtkstr;
      xm44 = cs;
    } else {
      xm11 = 1.0;
      xm12 = tau;
      xm21 = 0.0;
      xm22 = 1.0;
      xm33 = 1.0;
      xm34 = tau;
      xm43 = 0.0;
      xm44 = 1.0;
    }
    tmp[0] = xm11 * P(1, i) + xm12 * P(2, i) / gambetz / c->scxl;
    tmp[1] = gambetz * c->scxl * xm21 * P(1, i) + xm22 * P(2, i);
    tmp[2] = xm33 * P(3, i) + xm34 * P(4, i) / gambetz / c->scxl;
    tmp[3] = gambetz * c->scxl * xm43 * P(3, i) + xm44 * P(4, i);
    P(5, i) = P(5, i) + (1.0 / betaz - 1.0 / beta0) * tau / c->scxl;
    P(1, i) = tmp[0];
    P(2, i) = tmp[1];
    P(3, i) = tmp[2];
    P(4, i) = tmp[3];
  }
  refpt[4] = refpt[4] + tau / beta0 / c->scxl;
}

int or_map1(const or_const *c, double *pts, int64_t n, double refpt[6],
            double *z, double tau, const or_element *e, double charge, double mass) {
  int64_t i;
  int bitype = e->bitype;
  const double *v = e->v;
  if (bitype == 0) {
    /* exact drift, src/Appl/BeamBunch.f90:276-298 */
    double gam0 = -refpt[5];
    double gambet0 = sqrt(gam0 * gam0 - 1.0);
    double beta0 = gambet0 / gam0;
    for (i = 0; i < n; i++) {
      double tpx = P(2, i), tpy = P(4, i), tpt = P(6, i);
      double d = tpt - gam0;
      double tmph = sqrt(d * d - 1.0 - tpx * tpx - tpy * tpy);
      P(1, i) = P(1, i) + tpx * tau / tmph / c->scxl;
      P(3, i) = P(3, i) + tpy * tau / tmph / c->scxl;
      P(5, i) = P(5, i) - (1.0 / beta0 + (tpt - gam0) / tmph) * tau / c->scxl;
    }
    refpt[4] = refpt[4] + tau / (c->scxl * beta0);
  } else if (bitype == 1 && v[1] < -10.0) {
    return -1;  /* undulator map (src/Appl/BeamBunch.f90:302-304): not restated */
  } else if (bitype == 1 && v[1] < -1.0e-6) {
    quad_kmode(c, pts, n, refpt, tau, v[0], charge / mass);  /* :305-307 */
  } else {
    /* 6x6 map, src/Appl/BeamBunch.f90:308-320, 436-449 */
    double xm[36];
    if (bitype >= 2 && v[3] < 0.0) return -1;  /* ideal RF cavity model (:325-329) */
    if (or_element_map(c, *z, tau, e, refpt, charge, mass, xm) != 0) return -1;
    for (i = 0; i < n; i++) {
      double t[6];
      int j, k;
      for (j = 0; j < 6; j++) {
        t[j] = 0.0;
        for (k = 0; k < 6; k++) t[j] = t[j] + P(k + 1, i) * xm[6 * j + k];
      }
      for (j = 0; j < 6; j++) P(j + 1, i) = t[j];
    }
  }
  *z = *z + tau;
  return 0;
}

/* ------------------------------------------------------- frame changes */

int64_t or_conv1st(const or_const *c, double *pts, int64_t n,
                   const double refpt[6], double xrad, double yrad,
                   double ptrange[6], int64_t *npt_total, double *current) {
  const double pilc = (double)(2.0f * asinf(1.0f));
  const double xl = c->scxl;
  const double gam = -refpt[5];
  const double rad = (xrad + yrad) / 2.0;
  int64_t i0, i, ilost = 0;
  int k;
  for (k = 0; k < 3; k++) {
    ptrange[2 * k] = (double)1.0e20f;
    ptrange[2 * k + 1] = -(double)1.0e20f;
  }
  for (i0 = 0; i0 < n; i0++) {
    int ntmp5;
    double tmp5, rcpgammai, tmpbet, betai, tmpx, tmpy, tmp0, radtest;
    i = i0 - ilost;
    ntmp5 = (int)(P(5, i0) / pilc);
    tmp5 = P(5, i0) - ntmp5 * pilc;
    P(5, i0) = tmp5 - (ntmp5 % 2) * pilc;
    rcpgammai = 1.0 / (-P(6, i0) + gam);
    tmpbet = 1.0 - rcpgammai * rcpgammai * (1.0 + P(2, i0) * P(2, i0) + P(4, i0) * P(4, i0));
    betai = sqrt(fabs(tmpbet));
    tmpx = P(1, i0) * xl;
    P(1, i) = (P(1, i0) - P(2, i0) * P(5, i0) * rcpgammai) * xl;
    tmp0 = fmin(tmpx, P(1, i));
    if (ptrange[0] > tmp0) ptrange[0] = tmp0;
    tmp0 = fmax(tmpx, P(1, i));
    if (ptrange[1] < tmp0) ptrange[1] = tmp0;
    P(2, i) = P(2, i0);
    tmpy = P(3, i0) * xl;
    P(3, i) = (P(3, i0) - P(4, i0) * P(5, i0) * rcpgammai) * xl;
    tmp0 = fmin(tmpy, P(3, i));
    if (ptrange[2] > tmp0) ptrange[2] = tmp0;
    tmp0 = fmax(tmpy, P(3, i));
    if (ptrange[3] < tmp0) ptrange[3] = tmp0;
    P(4, i) = P(4, i0);
    P(5, i) = -gam * betai * P(5, i0) * xl;
    if (ptrange[4] > P(5, i)) ptrange[4] = P(5, i);
    if (ptrange[5] < P(5, i)) ptrange[5] = P(5, i);
    P(6, i) = P(6, i0);
    radtest = sqrt(P(1, i) * P(1, i) + P(3, i) * P(3, i));
    if (radtest >= rad || tmpbet <= 0.0) ilost++;
    P(7, i) = P(7, i0);
    P(8, i) = P(8, i0);
    P(9, i) = P(9, i0);
  }
  n -= ilost;
  /* one process: nptot = nplc; Current = Current*nptot/Npt (:1657-1663) */
  *current = *current * (double)n / (double)(*npt_total);
  *npt_total = n;
  return n;
}

void or_cvbkforth1st(const or_const *c, double *pts, int64_t n,
                     const double refpt[6]) {
  const double xl = c->scxl;
  const double xk = 1.0 / xl;
  const double gamma0 = -refpt[5];
  int64_t i;
  for (i = 0; i < n; i++) {
    double rcpgammai = 1.0 / (-P(6, i) + gamma0);
    double betai = sqrt(1.0 - rcpgammai * rcpgammai *
                                  (1.0 + P(2, i) * P(2, i) + P(4, i) * P(4, i)));
    P(5, i) = P(5, i) * xk / (-gamma0 * betai);
    P(1, i) = P(1, i) * xk + P(2, i) * P(5, i) * rcpgammai;
    P(3, i) = P(3, i) * xk + P(4, i) * P(5, i) * rcpgammai;
  }
  for (i = 0; i < n; i++) {
    double rcpgammai = 1.0 / (-P(6, i) + gamma0);
    double betai = sqrt(1.0 - rcpgammai * rcpgammai *
                                  (1.0 + P(2, i) * P(2, i) + P(4, i) * P(4, i)));
    P(1, i) = P(1, i) * xl;
    P(3, i) = P(3, i) * xl;
    P(5, i) = -gamma0 * betai * P(5, i) * xl;
  }
}

/* ------------------------------------------------------------ domain */

void or_update_compdom(const double pr[6], int nx, int ny, int nz,
                       double range[6], double msize[3]) {
  const double epsx = (double)(1.0f / ((float)nx - 3.0f));
  const double epsy = (double)(1.0f / ((float)ny - 3.0f));
  const double epsz = (double)(1.0f / ((float)nz - 3.0f));
  range[0] = pr[0] - epsx * (pr[1] - pr[0]);
  range[1] = pr[1] + epsx * (pr[1] - pr[0]);
  range[2] = pr[2] - epsy * (pr[3] - pr[2]);
  range[3] = pr[3] + epsy * (pr[3] - pr[2]);
  range[4] = pr[4] - epsz * (pr[5] - pr[4]);
  range[5] = pr[5] + epsz * (pr[5] - pr[4]);
  msize[0] = (range[1] - range[0]) / (nx - 1);
  msize[1] = (range[3] - range[2]) / (ny - 1);
  msize[2] = (range[5] - range[4]) / (nz - 1);
}

/* ----------------------------------------------------------- deposit */

void or_deposit(const double *pts, int64_t n, int nx, int ny, int nz,
                const double range[6], const double msize[3], double *rho) {
  const double hxi = 1.0 / msize[0], hyi = 1.0 / msize[1], hzi = 1.0 / msize[2];
  const double hx = msize[0], hy = msize[1], hz = msize[2];
  const double xmin = range[0], ymin = range[2], zmin = range[4];
  const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
  int64_t i;
  memset(rho, 0, sizeof(double) * (size_t)nx * ny * nz);
  for (i = 0; i < n; i++) {
    int ix = (int)((P(1, i) - xmin) * hxi + 1.0);
    double ab = ((xmin - P(1, i)) + ix * hx) * hxi;
    int jx = (int)((P(3, i) - ymin) * hyi + 1.0);
    double cd = ((ymin - P(3, i)) + jx * hy) * hyi;
    int kx = (int)((P(5, i) - zmin) * hzi + 1.0);
    double ef = ((zmin - P(5, i)) + kx * hz) * hzi;
    double w = P(8, i);
    int64_t o;
    if (ix < 1 || ix + 1 > nx || jx < 1 || jx + 1 > ny || kx < 1 || kx + 1 > nz)
      continue;
    o = (ix - 1) * sx + (jx - 1) * sy + (kx - 1) * sz;
    rho[o] += ab * cd * ef * w;
    rho[o + sy] += ab * (1.0 - cd) * ef * w;
    rho[o + sy + sz] += ab * (1.0 - cd) * (1.0 - ef) * w;
    rho[o + sz] += ab * cd * (1.0 - ef) * w;
    rho[o + sx + sz] += (1.0 - ab) * cd * (1.0 - ef) * w;
    rho[o + sx + sy + sz] += (1.0 - ab) * (1.0 - cd) * (1.0 - ef) * w;
    rho[o + sx + sy] += (1.0 - ab) * (1.0 - cd) * ef * w;
    rho[o + sx] += (1.0 - ab) * cd * ef * w;
  }
}

/* ------------------------------------------------------ Green function */

static double igf_antideriv(double x, double y, double z) {
  /* src/Appl/Field.f90:401-415 */
  double rr = sqrt(x * x + y * y + z * z);
  double aa = -z * z * atan(x * y / (z * rr)) / 2.0;
  double bb = -y * y * atan(x * z / (y * rr)) / 2.0;
  double cc = -x * x * atan(y * z / (x * rr)) / 2.0;
  double dd = y * z * log(x + rr);
  double ee = x * z * log(y + rr);
  double ff = x * y * log(z + rr);
  return aa + bb + cc + dd + ee + ff;
}

void or_green_igf(int nx, int ny, int nz, double hx, double hy, double hz,
                  double *g) {
  const int tx = nx + 2, ty = ny + 2, tz = nz + 2;
  double *t = (double *)malloc(sizeof(double) * (size_t)tx * ty * tz);
  const double vol = hx * hy * hz;
  int i, j, k;
#define T(a, b, c) t[(a) + (int64_t)tx * ((b) + (int64_t)ty * (c))]
  for (k = 0; k < tz; k++)
    for (j = 0; j < ty; j++)
      for (i = 0; i < tx; i++)
        T(i, j, k) = igf_antideriv(i * hx - hx / 2.0, j * hy - hy / 2.0,
                                   k * hz - hz / 2.0) / vol;
  for (k = 0; k <= nz; k++)
    for (j = 0; j <= ny; j++)
      for (i = 0; i <= nx; i++)
        g[i + (int64_t)(nx + 1) * (j + (int64_t)(ny + 1) * k)] =
            T(i + 1, j + 1, k + 1) - T(i, j + 1, k + 1) - T(i + 1, j, k + 1) +
            T(i, j, k + 1) - T(i + 1, j + 1, k) + T(i, j + 1, k) +
            T(i + 1, j, k) - T(i, j, k);
#undef T
  free(t);
}

/* ------------------------------------------------------------- solver */

/* In-place radix-2 complex FFT of length m (power of 2) on interleaved data,
 * sign -1 forward, +1 backward, unnormalized. */
static void fft1(double *re, double *im, int m, int sign) {
  int i, j, k, len;
  for (i = 1, j = 0; i < m; i++) {
    int bit = m >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double tr = re[i], ti = im[i];
      re[i] = re[j]; im[i] = im[j];
      re[j] = tr; im[j] = ti;
    }
  }
  for (len = 2; len <= m; len <<= 1) {
    int half = len >> 1;
    for (k = 0; k < half; k++) {
      double ang = sign * 2.0 * 3.14159265358979323846 * k / len;
      double wr = cos(ang), wi = sin(ang);
      for (i = k; i < m; i += len) {
        double ur = re[i], ui = im[i];
        double vr = re[i + half] * wr - im[i + half] * wi;
        double vi = re[i + half] * wi + im[i + half] * wr;
        re[i] = ur + vr; im[i] = ui + vi;
        re[i + half] = ur - vr; im[i + half] = ui - vi;
      }
    }
  }
}

static void fft3(double *re, double *im, int mx, int my, int mz, int sign) {
  int mmax = mx > my ? mx : my;
  double *lr, *li;
  int64_t a, b;
  mmax = mmax > mz ? mmax : mz;
  lr = (double *)malloc(sizeof(double) * mmax);
  li = (double *)malloc(sizeof(double) * mmax);
  for (b = 0; b < (int64_t)my * mz; b++) fft1(re + b * mx, im + b * mx, mx, sign);
  for (b = 0; b < mz; b++)
    for (a = 0; a < mx; a++) {
      int64_t base = a + (int64_t)mx * my * b;
      int t;
      for (t = 0; t < my; t++) { lr[t] = re[base + (int64_t)mx * t]; li[t] = im[base + (int64_t)mx * t]; }
      fft1(lr, li, my, sign);
      for (t = 0; t < my; t++) { re[base + (int64_t)mx * t] = lr[t]; im[base + (int64_t)mx * t] = li[t]; }
    }
  for (b = 0; b < (int64_t)mx * my; b++) {
    int64_t s = (int64_t)mx * my;
    int t;
    for (t = 0; t < mz; t++) { lr[t] = re[b + s * t]; li[t] = im[b + s * t]; }
    fft1(lr, li, mz, sign);
    for (t = 0; t < mz; t++) { re[b + s * t] = lr[t]; im[b + s * t] = li[t]; }
  }
  free(lr);
  free(li);
}

static int is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

int or_solve_openbc(const double *rho, int nx, int ny, int nz, double hx,
                    double hy, double hz, double *phi) {
  const int mx = 2 * nx, my = 2 * ny, mz = 2 * nz;
  const int64_t m = (int64_t)mx * my * mz;
  double *g, *ar, *ai, *gr, *gi, scale;
  int i, j, k;
  if (!is_pow2(nx) || !is_pow2(ny) || !is_pow2(nz)) return -1;
  g = (double *)malloc(sizeof(double) * (size_t)(nx + 1) * (ny + 1) * (nz + 1));
  ar = (double *)calloc((size_t)m, sizeof(double));
  ai = (double *)calloc((size_t)m, sizeof(double));
  gr = (double *)malloc(sizeof(double) * (size_t)m);
  gi = (double *)calloc((size_t)m, sizeof(double));
  if (!g || !ar || !ai || !gr || !gi) return -2;
  or_green_igf(nx, ny, nz, hx, hy, hz, g);
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++)
        ar[i + (int64_t)mx * (j + (int64_t)my * k)] =
            rho[i + (int64_t)nx * (j + (int64_t)ny * k)];
  /* even extension: G(m) = g(min(m, M-m)) (src/Appl/Field.f90:446-503) */
  for (k = 0; k < mz; k++)
    for (j = 0; j < my; j++)
      for (i = 0; i < mx; i++) {
        int a = i <= nx ? i : mx - i, b = j <= ny ? j : my - j, c = k <= nz ? k : mz - k;
        gr[i + (int64_t)mx * (j + (int64_t)my * k)] =
            g[a + (int64_t)(nx + 1) * (b + (int64_t)(ny + 1) * c)];
      }
  fft3(ar, ai, mx, my, mz, -1);
  fft3(gr, gi, mx, my, mz, -1);
  for (int64_t t = 0; t < m; t++) {
    double r = ar[t] * gr[t] - ai[t] * gi[t];
    double s = ar[t] * gi[t] + ai[t] * gr[t];
    ar[t] = r;
    ai[t] = s;
  }
  fft3(ar, ai, mx, my, mz, +1);
  scale = 1.0 / (double)m;
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++)
        phi[i + (int64_t)nx * (j + (int64_t)ny * k)] =
            ar[i + (int64_t)mx * (j + (int64_t)my * k)] * scale;
  free(g); free(ar); free(ai); free(gr); free(gi);
  return 0;
}

void or_solve_direct(const double *rho, int nx, int ny, int nz, double hx,
                     double hy, double hz, double *phi) {
  double *g = (double *)malloc(sizeof(double) * (size_t)(nx + 1) * (ny + 1) * (nz + 1));
  int i, j, k, a, b, c;
  or_green_igf(nx, ny, nz, hx, hy, hz, g);
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++) {
        double s = 0.0;
        for (c = 0; c < nz; c++)
          for (b = 0; b < ny; b++)
            for (a = 0; a < nx; a++)
              s += rho[a + (int64_t)nx * (b + (int64_t)ny * c)] *
                   g[abs(i - a) + (int64_t)(nx + 1) * (abs(j - b) + (int64_t)(ny + 1) * abs(k - c))];
        phi[i + (int64_t)nx * (j + (int64_t)ny * k)] = s;
      }
  free(g);
}

/* ------------------------------------------------------- field + kick */

void or_efield(const double *phi, int nx, int ny, int nz, const double msize[3],
               double *ex, double *ey, double *ez) {
  const double hxi = 1.0 / msize[0], hyi = 1.0 / msize[1], hzi = 1.0 / msize[2];
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  int i, j, k;
#define F(a, b, c) phi[(a) + sy * (b) + sz * (c)]
#define IDX(a, b, c) ((a) + sy * (b) + sz * (c))
  for (k = 0; k < nz; k++)
    for (j = 0; j < ny; j++) {
      ex[IDX(0, j, k)] = hxi * (F(0, j, k) - F(1, j, k));
      for (i = 1; i < nx - 1; i++)
        ex[IDX(i, j, k)] = 0.5 * hxi * (F(i - 1, j, k) - F(i + 1, j, k));
      ex[IDX(nx - 1, j, k)] = hxi * (F(nx - 2, j, k) - F(nx - 1, j, k));
    }
  for (k = 0; k < nz; k++) {
    for (i = 0; i < nx; i++) ey[IDX(i, 0, k)] = hyi * (F(i, 0, k) - F(i, 1, k));
    for (j = 1; j < ny - 1; j++)
      for (i = 0; i < nx; i++)
        ey[IDX(i, j, k)] = 0.5 * hyi * (F(i, j - 1, k) - F(i, j + 1, k));
    for (i = 0; i < nx; i++)
      ey[IDX(i, ny - 1, k)] = hyi * (F(i, ny - 2, k) - F(i, ny - 1, k));
  }
  for (j = 0; j < ny; j++)
    for (i = 0; i < nx; i++) ez[IDX(i, j, 0)] = hzi * (F(i, j, 0) - F(i, j, 1));
  for (k = 1; k < nz - 1; k++)
    for (j = 0; j < ny; j++)
      for (i = 0; i < nx; i++)
        ez[IDX(i, j, k)] = 0.5 * hzi * (F(i, j, k - 1) - F(i, j, k + 1));
  for (j = 0; j < ny; j++)
    for (i = 0; i < nx; i++)
      ez[IDX(i, j, nz - 1)] = hzi * (F(i, j, nz - 2) - F(i, j, nz - 1));
#undef F
#undef IDX
}

static double cic(const double *e, int64_t o, int64_t sy, int64_t sz, double ab,
                  double cd, double ef) {
  /* term order of src/Appl/BeamBunch.f90:2130-2137 */
  return e[o] * ab * cd * ef + e[o + sy] * ab * (1.0 - cd) * ef +
         e[o + sy + sz] * ab * (1.0 - cd) * (1.0 - ef) +
         e[o + sz] * ab * cd * (1.0 - ef) +
         e[o + 1 + sz] * (1.0 - ab) * cd * (1.0 - ef) +
         e[o + 1 + sy + sz] * (1.0 - ab) * (1.0 - cd) * (1.0 - ef) +
         e[o + 1 + sy] * (1.0 - ab) * (1.0 - cd) * ef +
         e[o + 1] * (1.0 - ab) * cd * ef;
}

void or_scatter1(const or_const *c, double *pts, int64_t n, int nx, int ny,
                 int nz, const double *ex, const double *ey, const double *ez,
                 const double range[6], const double msize[3], double gam,
                 double curr, double tau, double mass, double chrg) {
  const double tmpscale = curr / c->scfreq;
  const double twopi = 2.0 * c->pi;
  const double fpei = c->clight * c->clight * (double)1.0e-7f;
  const double xk = 1.0 / c->scxl;
  const double bet = sqrt(gam * gam - 1.0) / gam;
  const double bbyk = bet / xk;
  const double gambet = gam * bet;
  const double brho = gambet / (c->clight * chrg) * mass;
  const double vz0 = bet * c->clight;
  const double perv0 = 2.0 * curr * fpei / (brho * vz0 * vz0 * gam * gam);
  const double xycon = 0.5 * perv0 * gambet * bbyk * twopi / tmpscale;
  const double tcon = bet * xycon * gam * gam;
  const double hxi = 1.0 / msize[0], hyi = 1.0 / msize[1], hzi = 1.0 / msize[2];
  const double hx = msize[0], hy = msize[1], hz = msize[2];
  const double xmin = range[0], ymin = range[2], zmin = range[4];
  const int64_t sy = nx, sz = (int64_t)nx * ny;
  int64_t i;
  (void)nz;
  for (i = 0; i < n; i++) {
    int ix = (int)((P(1, i) - xmin) * hxi + 1.0);
    double ab = ((xmin - P(1, i)) + ix * hx) * hxi;
    int jx = (int)((P(3, i) - ymin) * hyi + 1.0);
    double cd = ((ymin - P(3, i)) + jx * hy) * hyi;
    int kx = (int)((P(5, i) - zmin) * hzi + 1.0);
    double ef = ((zmin - P(5, i)) + kx * hz) * hzi;
    int64_t o = (ix - 1) + sy * (jx - 1) + sz * (kx - 1);
    double exn = cic(ex, o, sy, sz, ab, cd, ef) * gam;
    double eyn = cic(ey, o, sy, sz, ab, cd, ef) * gam;
    double ezn = cic(ez, o, sy, sz, ab, cd, ef);
    double rcpgammai = 1.0 / (-P(6, i) + gam);
    double tmpsq = 1.0 - rcpgammai * rcpgammai * (1.0 + P(2, i) * P(2, i) + P(4, i) * P(4, i));
    double betai = sqrt(tmpsq);
    P(1, i) = P(1, i) * xk;
    P(2, i) = P(2, i) + tau * xycon * exn;
    P(3, i) = P(3, i) * xk;
    P(4, i) = P(4, i) + tau * xycon * eyn;
    P(5, i) = P(5, i) * xk / (-gam * betai);
    P(6, i) = P(6, i) - tau * tcon * ezn;
  }
}

void or_kick1(const or_const *c, double *pts, int64_t n, int nx, int ny, int nz,
              const double *phi, const double range[6], const double msize[3],
              const double refpt[6], double curr, double tau, double mass,
              double chrg) {
  size_t g = (size_t)nx * ny * nz;
  double *e;
  if (curr < (double)1.0e-20f) return;
  e = (double *)malloc(3 * g * sizeof(double));
  or_efield(phi, nx, ny, nz, msize, e, e + g, e + 2 * g);
  or_scatter1(c, pts, n, nx, ny, nz, e, e + g, e + 2 * g, range, msize,
              -refpt[5], curr, tau, mass, chrg);
  free(e);
}

/* ------------------------------------------------------------- wakefields */

/* The line charge density and centroids of the run loop
 * (src/Contrl/AccSimulator.f90:1276-1321, one process): each particle's
 * t-frame z binned linearly on the Nz mesh lines (unit weight), x and y
 * weighted for the centroids; rho in C/m of the lab frame (the bin width
 * divided by gamma). */
static void wake_line(const double *pts, int64_t n, const double range[6], int nz,
                      double current, int64_t np, double scfreq, double gam,
                      double charge, double *hzwake, double *rho, double *xb, double *yb) {
  double hz = (range[5] - range[4]) / (nz - 1);
  double *dens = (double *)calloc(nz + 1, sizeof(double));
  double *xw = (double *)calloc(nz + 1, sizeof(double));
  double *yw = (double *)calloc(nz + 1, sizeof(double));
  int64_t i;
  int k;
  for (i = 0; i < n; i++) {
    int iz = (int)((P(5, i) - range[4]) / hz + 1);
    double ab = ((range[4] - P(5, i)) + iz * hz) / hz;
    /* a particle on the last mesh line (iz = Nz, ab = 1) puts its zero
     * upper weight in the extra row nz, as the reference's loop writes it
     * one past densz(Nz) */
    if (iz < 1) iz = 1;
    if (iz > nz) iz = nz;
    xw[iz - 1] += P(1, i) * ab;
    xw[iz] += P(1, i) * (1.0 - ab);
    yw[iz - 1] += P(3, i) * ab;
    yw[iz] += P(3, i) * (1.0 - ab);
    dens[iz - 1] += ab;
    dens[iz] += 1.0 - ab;
  }
  hz = hz / gam;
  for (k = 0; k < nz; k++) {
    rho[k] = dens[k] * current / scfreq / (double)np / hz * (charge / fabs(charge));
    xb[k] = dens[k] != 0.0 ? xw[k] / dens[k] : 0.0;
    yb[k] = dens[k] != 0.0 ? yw[k] / dens[k] : 0.0;
  }
  *hzwake = hz;
  free(dens);
  free(xw);
  free(yw);
}

/* A wake function at distance zz behind its source: the analytic
 * short-range wakes of wakefield_FieldQuant (K. Bane's fits with the
 * reference's constants, src/Appl/Field.f90:621-642,692-706; plane 0
 * longitudinal, 1 and 2 transverse), or the read-in table of
 * wakefieldread_FieldQuant with its interpolation weights as written there
 * (ab = iwk*hwk - zwk, not divided by hwk; :820-830). */
static double wake_fn(const or_const *c, const or_wake *w, int plane, double zz) {
  if (w->kind == 1) {
    const double pilc = 2 * asin(1.0), z0 = 120 * pilc;
    const double aa = w->aa, gg = w->gg, leng = w->leng;
    if (plane == 0) {
      double zz00 = (double)0.41f * pow(aa / leng, (double)0.8f) *
                    pow(gg / leng, (double)1.6f) * aa;
      return z0 * c->clight * exp(-sqrt(zz / zz00)) / (pilc * aa * aa);
    } else {
      double zz00 = 1.0 * (double)0.169f * aa * pow(aa / leng, (double)1.17f) *
                    pow(gg / aa, (double)0.38f);
      double coef1 = 4 * z0 * c->clight * zz00 / (pilc * aa * aa * aa * aa);
      return coef1 * (1.0 - (1. + sqrt(zz / zz00)) * exp(-sqrt(zz / zz00)));
    }
  } else {
    const double *col = plane == 0 ? w->edat : (plane == 1 ? w->epdat : w->eppdat);
    double hwk = w->zdat[1] - w->zdat[0], zwk = zz - w->zdat[0];
    int iwk = (int)(zwk / hwk + 1), iwk1 = iwk + 1;
    double ab = iwk * hwk - zwk;
    /* rows past the table read the Data module's zero-filled arrays */
    double e1 = iwk <= w->n ? col[iwk - 1] : 0.0, e2 = iwk1 <= w->n ? col[iwk1 - 1] : 0.0;
    return e1 * ab + e2 * (1.0 - ab);
  }
}

int or_wakefield(const or_const *c, const double *pts, int64_t n, const double range[6],
                 int nz, double current, int64_t np, double gam, double charge,
                 const or_wake *w, double *ex, double *ey, double *ez) {
  /* src/Contrl/AccSimulator.f90:1272-1338 with wakefield_FieldQuant /
   * wakefieldread_FieldQuant (src/Appl/Field.f90:590-1001).  The cyclic
   * convolution those routines take by FFT on 2*Nz points is, with their
   * Green function (half weight at distance 0 for the longitudinal wake,
   * none for the transverse ones) is, for bin k, the sum over bin k itself
   * (longitudinal, half weight) and the bins j > k at distance (j-k)*hz,
   * computed here directly. */
  double *rho = (double *)malloc(nz * sizeof(double));
  double *xb = (double *)malloc(nz * sizeof(double));
  double *yb = (double *)malloc(nz * sizeof(double));
  double hz;
  int k, j, lon, tra;
  if (w->kind < 1 || w->kind > 3) return -1;
  wake_line(pts, n, range, nz, current, np, c->scfreq, gam, charge, &hz, rho, xb, yb);
  if (w->kind == 3) {
    /* the CSR branch (src/Contrl/AccSimulator.f90:1344-1396): zwkmin =
     * range(5)/gamma + zoff, hzwake/gamma, exwake = eywake = 0 */
    or_csr_wake(nz, w->r0, range[4] / gam + w->zoff, hz, w->blength, rho, gam, ez);
    for (k = 0; k < nz; k++) ex[k] = ey[k] = 0.0;
    free(rho);
    free(xb);
    free(yb);
    return 0;
  }
  lon = !(w->kind == 2 && w->flagbtw == 4);
  tra = !(w->kind == 2 && (w->flagbtw == 2 || w->flagbtw == 3));
  for (k = 0; k < nz; k++) {
    double sz = 0.0, sx = 0.0, sy = 0.0;
    if (lon) sz = 0.5 * wake_fn(c, w, 0, 0.0) * rho[k];
    for (j = k + 1; j < nz; j++) {
      double zz = (j - k) * hz;
      if (lon) sz += wake_fn(c, w, 0, zz) * rho[j];
      if (tra) {
        sx += wake_fn(c, w, 1, zz) * rho[j] * xb[j];
        sy += wake_fn(c, w, w->kind == 1 ? 1 : 2, zz) * rho[j] * yb[j];
      }
    }
    ez[k] = w->scale * (-sz * hz);
    ex[k] = w->scale * (sx * hz);
    ey[k] = w->scale * (sy * hz);
  }
  free(rho);
  free(xb);
  free(yb);
  return 0;
}

/* ------------------------------------------------------------- CSR wake */

/* The integrated CSR kernels of the four regimes of Saldin et al.
 * (src/Appl/Field.f90:1253-1301): IcsrCaseA (entrance transient),
 * IcsrCaseB (steady state), IcsrCaseC (exit transient, source inside the
 * bend), IcsrCaseD (exit transient, source after the bend: the retarded
 * angle psi is the root of funcd, found by or_csr_root).  Integer powers
 * are products, as the compiled reference evaluates them. */
static double or_csr_a(double phih, double ssh, double xk2) {
  const double bb = 2 * phih + phih * phih * phih / 3 - 2 * ssh;
  const double cc = phih * phih + phih * phih * phih * phih / 12 - 2 * ssh * phih;
  const double yh = (-bb + sqrt(bb * bb - 4 * cc)) / 2;
  const double phpy = phih + yh;
  return xk2 * (-(2 * phpy + phih * phih * phih) /
                (phpy * phpy + phih * phih * phih * phih / 4) + 1.0 / ssh);
}

static double or_csr_b(double ssh, double xk2) {
  const double aa = sqrt(64.0 + 144.0 * (ssh * ssh));
  const double uuh = pow(aa + 12 * ssh, 1.0 / 3.0) - pow(aa - 12 * ssh, 1.0 / 3.0);
  return xk2 * (-4 * uuh * (uuh * uuh + 8) / ((uuh * uuh + 4.0) * (uuh * uuh + 12.0)));
}

static double or_csr_c(double phimh, double xxbarh, double ssh, double xk2) {
  const double bb = 2 * (phimh + xxbarh) - 2 * ssh + phimh * phimh * phimh / 3 +
                    phimh * phimh * xxbarh;
  const double cc = (phimh + xxbarh) * (phimh + xxbarh) +
                    phimh * phimh * (phimh * phimh + 4 * phimh * xxbarh) / 12 -
                    2 * ssh * (phimh + xxbarh);
  const double yh = (-bb + sqrt(bb * bb - 4 * cc)) / 2;
  const double phpxya = phimh + xxbarh + yh;
  const double phpxyb = phimh * xxbarh + phimh * phimh / 2.0;
  return xk2 * (-2.0 * (phpxya + phimh * phpxyb) / (phpxya * phpxya + phpxyb * phpxyb) +
                1.0 / ssh);
}

/* funcd (src/Appl/Field.f90:1373-1384): f(psi) and f'(psi). */
static void or_csr_funcd(double psi, double xxh, double deltas, double *f, double *df) {
  *f = psi * psi * psi * psi / 12 + xxh * psi * psi * psi / 3 + psi * psi +
       (2 * xxh - 2 * deltas) * psi - 2 * deltas * xxh + xxh * xxh;
  *df = psi * psi * psi / 3 + xxh * psi * psi + 2 * psi + 2 * xxh - 2 * deltas;
}

/* root (src/Appl/Field.f90:1305-1371): Newton-Raphson with bisection
 * safeguard in [x1, x2].  The reference PAUSEs when the root is not
 * bracketed or the iterations run out; here the iteration goes on as it
 * would after the pause. */
static double or_csr_root(double x1, double x2, double xacc, int maxit, double zeta,
                          double xxh) {
  double fl, fh, df, f, xl, xh;
  or_csr_funcd(x1, xxh, zeta, &fl, &df);
  or_csr_funcd(x2, xxh, zeta, &fh, &df);
  if (fabs(fl) < xacc) return x1;
  if (fabs(fh) < xacc) return x2;
  if (fl < 0.0) { xl = x1; xh = x2; } else { xh = x1; xl = x2; }
  double rts = 0.5 * (x1 + x2);
  double dxold = fabs(x2 - x1), dx = dxold;
  or_csr_funcd(rts, xxh, zeta, &f, &df);
  for (int j = 1; j <= maxit; j++) {
    if (((rts - xh) * df - f) * ((rts - xl) * df - f) > 0.0 || fabs(2.0 * f) > fabs(dxold * df)) {
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
    or_csr_funcd(rts, xxh, zeta, &f, &df);
    if (f < 0.0) xl = rts; else xh = rts;
  }
  return rts;
}

static double or_csr_d(double xxbarh, double psimax, double ssh, double xk2, double epstol,
                       int nmax) {
  const double psi = or_csr_root(-epstol, psimax * (1.0 + epstol), epstol, nmax, ssh, xxbarh);
  const double psipxa = psi + xxbarh;
  const double psipxb = psi * (xxbarh + psi / 2.0);
  return xk2 * (-2.0 * (psipxa + psi * psipxb) / (psipxa * psipxa + psipxb * psipxb) +
                1.0 / ssh);
}

void or_csr_wake(int nx, double r0, double ptmin, double hx, double blength,
                 const double *rhonew, double gam, double *ezwake) {
  /* csrwakeTrIGF_FieldQuant (src/Appl/Field.f90:1007-1251): the 1D CSR
   * wake on the Nx mesh lines, each a sum over the lines behind it of the
   * line density times the kernel integrated over the line's cell (the IGF
   * differences), regime by regime: inside the bend (xx <= blength) the
   * entrance transient A up to the retarded slip point and the steady state
   * B after it; past the bend the exit transients C and D. */
  const double epstol = 2.0e-9;
  const int nmax = 100;
  const double pilc = 2 * asin(1.0);
  const double phim = blength / r0;
  const double xconst = 1. / (4 * pilc * 8.854187817e-12);
  const double xk2 = gam / r0;
  const double g3 = gam * gam * gam;
#define RHO(j) rhonew[(j) - 1]
#define SSH(xp) ((xx - (xp)) * g3 / r0)
  ezwake[0] = 0.0;
  for (int i = 2; i <= nx; i++) {
    const double xx = ptmin + (i - 1) * hx;
    const double xr = xx / r0;
    const double xxl = xr * xr * xr * r0 / 24;
    const int ni = i;
    double e = 0.0;
    if (xx <= 0) {
    } else if (xx <= blength) {
      const double xslp = (xx / r0) * r0 / 2 / gam / gam + xxl;
      const int islp0 = (int)((xx - xslp - ptmin) / hx);
      const int islp1 = islp0 + 1;
      const double phih = xx / r0 * gam;
      /* case A */
      for (int j = 2; j <= islp1 - 1; j++) {
        const double t2 = or_csr_a(phih, SSH(ptmin + (j - 1) * hx + hx / 2), xk2);
        const double t1 = or_csr_a(phih, SSH(ptmin + (j - 1) * hx - hx / 2), xk2);
        e = e + RHO(j) * (t1 - t2);
      }
      if (islp1 >= 2) {
        double t2 = or_csr_a(phih, SSH(xx - xslp), xk2);
        double t1 = or_csr_a(phih, SSH(ptmin + (islp1 - 1) * hx - hx / 2), xk2);
        e = e + RHO(islp1) * (t1 - t2);
        t2 = or_csr_a(phih, SSH(ptmin + hx / 2), xk2);
        t1 = or_csr_a(phih, SSH(ptmin), xk2);
        e = e + RHO(1) * (t1 - t2);
      } else if (islp1 == 1) {
        const double t2 = or_csr_a(phih, SSH(xx - xslp), xk2);
        const double t1 = or_csr_a(phih, SSH(ptmin + (islp1 - 1) * hx), xk2);
        e = e + RHO(islp1) * (t1 - t2);
      }
      /* case B */
      const int jstart = islp1 > 1 ? islp1 : 1;
      for (int j = jstart + 2; j <= ni - 1; j++) {
        const double t2 = or_csr_b(SSH(ptmin + (j - 1) * hx + hx / 2), xk2);
        const double t1 = or_csr_b(SSH(ptmin + (j - 1) * hx - hx / 2), xk2);
        e = e + RHO(j) * (t1 - t2);
      }
      if (jstart <= ni - 2) {
        double t2 = or_csr_b(SSH(ptmin + (ni - 1) * hx), xk2);
        double t1 = or_csr_b(SSH(ptmin + (ni - 1) * hx - hx / 2), xk2);
        e = e + RHO(ni) * (t1 - t2);
        if (islp1 < 1) {
          t2 = or_csr_b(SSH(ptmin + hx / 2), xk2);
          t1 = or_csr_b(SSH(ptmin), xk2);
          e = e + RHO(1) * (t1 - t2);
        } else {
          const int j = islp1 + 1;
          t2 = or_csr_b(SSH(ptmin + (j - 1) * hx + hx / 2), xk2);
          t1 = or_csr_b(SSH(xx - xslp), xk2);
          e = e + RHO(j) * (t1 - t2);
        }
      } else if (jstart == ni - 1) {
        const double t2 = or_csr_b(SSH(ptmin + (ni - 1) * hx), xk2);
        const double t1 = or_csr_b(SSH(xx - xslp), xk2);
        e = e + RHO(ni) * (t1 - t2);
      }
    } else {
      const double xxbar = xx - blength;
      const double xslp = (r0 * phim + xxbar) / 2 / gam / gam +
                          r0 * phim * phim * phim / 24 * (r0 * phim + 4 * xxbar) /
                              (r0 * phim + xxbar);
      const int islp0 = (int)((xx - xslp - ptmin) / hx);
      const int islp1 = islp0 + 1;
      const double xslpn = xxbar / 2.0 / gam / gam;
      const int islpn = (int)((xx - xslpn - ptmin) / hx);
      const int islpn1 = islpn + 1;
      /* case C */
      const double xxbarh = xxbar * gam / r0;
      const double phimh = phim * gam;
      for (int j = 2; j <= islp1 - 1; j++) {
        const double t2 = or_csr_c(phimh, xxbarh, SSH(ptmin + (j - 1) * hx + hx / 2), xk2);
        const double t1 = or_csr_c(phimh, xxbarh, SSH(ptmin + (j - 1) * hx - hx / 2), xk2);
        e = e + RHO(j) * (t1 - t2);
      }
      if (islp1 >= 2) {
        double t2 = or_csr_c(phimh, xxbarh, SSH(xx - xslp), xk2);
        double t1 = or_csr_c(phimh, xxbarh, SSH(ptmin + (islp1 - 1) * hx - hx / 2), xk2);
        e = e + RHO(islp1) * (t1 - t2);
        t2 = or_csr_c(phimh, xxbarh, SSH(ptmin + hx / 2), xk2);
        t1 = or_csr_c(phimh, xxbarh, SSH(ptmin), xk2);
        e = e + RHO(1) * (t1 - t2);
      } else if (islp1 == 1) {
        const double t2 = or_csr_c(phimh, xxbarh, SSH(xx - xslp), xk2);
        const double t1 = or_csr_c(phimh, xxbarh, SSH(ptmin), xk2);
        e = e + RHO(islp1) * (t1 - t2);
      }
      /* case D */
      const double psimax = phim * gam;
      const int jstart = islp1 > 0 ? islp1 : 0;
      for (int j = jstart + 2; j <= islpn1 - 1; j++) {
        const double t2 = or_csr_d(xxbarh, psimax, SSH(ptmin + (j - 1) * hx + hx / 2), xk2,
                                   epstol, nmax);
        const double t1 = or_csr_d(xxbarh, psimax, SSH(ptmin + (j - 1) * hx - hx / 2), xk2,
                                   epstol, nmax);
        e = e + RHO(j) * (t1 - t2);
      }
      if (islpn1 >= jstart + 2) {
        double t2 = or_csr_d(xxbarh, psimax, xslpn * g3 / r0, xk2, epstol, nmax);
        double t1 = or_csr_d(xxbarh, psimax, SSH(ptmin + (islpn1 - 1) * hx - hx / 2), xk2,
                             epstol, nmax);
        e = e + RHO(islpn1) * (t1 - t2);
        t2 = or_csr_d(xxbarh, psimax, SSH(ptmin + jstart * hx + hx / 2), xk2, epstol, nmax);
        t1 = or_csr_d(xxbarh, psimax, xslp * g3 / r0, xk2, epstol, nmax);
        e = e + RHO(jstart + 1) * (t1 - t2);
      } else if (islpn1 == jstart + 1) {
        const double t2 = or_csr_d(xxbarh, psimax, xslpn * g3 / r0, xk2, epstol, nmax);
        const double t1 = or_csr_d(xxbarh, psimax, xslp * g3 / r0, xk2, epstol, nmax);
        e = e + RHO(jstart + 1) * (t1 - t2);
      }
    }
    ezwake[i - 1] = e * xconst;
  }
#undef RHO
#undef SSH
}

void or_kick1wake(const or_const *c, double *pts, int64_t n, int nx, int ny, int nz,
                  const double *phi, const double range[6], const double msize[3],
                  const double refpt[6], double curr, double tau, double mass,
                  double chrg, const double *ex, const double *ey, const double *ez) {
  /* kick1wake_BeamBunch (src/Appl/BeamBunch.f90:4694-5046): the wake of each
   * z plane added to the grid field, divided by the constants scatter1
   * multiplies it with, then scatter1 */
  size_t g = (size_t)nx * ny * nz, pl = (size_t)nx * ny, q;
  double *e, gam, bet, gambet, brho, vz0, perv0, xycon, tcon, tmpscale, fpei, twopi, xk,
      bbyk;
  int k;
  if (curr < (double)1.0e-20f) return;
  e = (double *)malloc(3 * g * sizeof(double));
  or_efield(phi, nx, ny, nz, msize, e, e + g, e + 2 * g);
  gam = -refpt[5];
  twopi = 2.0 * c->pi;
  tmpscale = curr / c->scfreq;
  fpei = c->clight * c->clight * (double)1.0e-7f;
  xk = 1 / c->scxl;
  bet = sqrt(gam * gam - 1.0) / gam;
  bbyk = bet / xk;
  gambet = gam * bet;
  brho = gambet / (c->clight * chrg) * mass;
  vz0 = bet * c->clight;
  perv0 = 2.0 * curr * fpei / (brho * vz0 * vz0 * gam * gam);
  xycon = 0.5 * perv0 * gambet * bbyk * twopi / tmpscale;
  tcon = bet * xycon * gam * gam;
  for (k = 0; k < nz; k++)
    for (q = 0; q < pl; q++) {
      size_t o = (size_t)k * pl + q;
      e[o] = e[o] + chrg * ex[k] / xycon / mass / bet / gam;
      e[g + o] = e[g + o] + chrg * ey[k] / xycon / mass / bet / gam;
      e[2 * g + o] = e[2 * g + o] + chrg * ez[k] / tcon / mass;
    }
  or_scatter1(c, pts, n, nx, ny, nz, e, e + g, e + 2 * g, range, msize, gam, curr, tau,
              mass, chrg);
  free(e);
}

/* ----------------------------------------- nonlinear integrator (Flagmap=2) */

/* ---------------------------------------------------------------- bends */

/* Transport coordinates of particle i (src/Appl/Dipole.f90:489-497,
 * 758-764): t[0..5] = x, x', y, y', z, delta with the particle's own
 * gamma beta_z = bz. */
static void to_transport(const or_const *c, const double *pts, int64_t i, double bz,
                         double beta0, double gambet, double qm0, double t[6]) {
  t[0] = P(1, i) * c->scxl;
  t[1] = P(2, i) / bz;
  t[2] = P(3, i) * c->scxl;
  t[3] = P(4, i) / bz;
  t[4] = -P(5, i) * beta0 * c->scxl;
  t[5] = -P(6, i) / beta0 / gambet - (P(7, i) - qm0) / qm0;
}

/* and back (src/Appl/Dipole.f90:516-523, 779-786); gb = |gamma beta| */
static void from_transport(const or_const *c, double *pts, int64_t i, double gb,
                           double beta0, double gambet, double qm0, const double t[6]) {
  double bz = gb / sqrt(t[1] * t[1] + t[3] * t[3] + 1.0);
  P(1, i) = t[0] / c->scxl;
  P(2, i) = t[1] * bz;
  P(3, i) = t[2] / c->scxl;
  P(4, i) = t[3] * bz;
  P(5, i) = -t[4] / c->scxl / beta0;
  P(6, i) = -beta0 * gambet * (t[5] + (P(7, i) - qm0) / qm0);
}

void or_dipole_face(const or_const *c, double *pts, int64_t n, double refpt[6],
                    const or_element *e, int exit, double charge, double mass) {
  /* the run loop's constants (src/Contrl/AccSimulator.f90:1058-1088) */
  const double *v = e->v;
  double gam0 = -refpt[5];
  double beta0 = sqrt(1.0 - 1.0 / gam0 / gam0);
  double gambet = beta0 * gam0;
  double qm0 = charge / mass;
  double hgap = 2 * v[3];
  double h0 = v[0] / e->length, k1 = v[1];
  double ang = exit ? v[5] : v[4];
  double h1 = exit ? v[7] : v[6];
  double psi = h0 * hgap * v[8] * (1.0 + sin(ang) * sin(ang)) / cos(ang);
  double tanphi = tan(ang), tanphib = tan(ang - psi);
  double tanphi2 = tanphi * tanphi, secphi2 = 1.0 + tanphi2;
  double secphi3 = secphi2 * (1.0 / cos(ang)), secphib2 = 1.0 + tanphib * tanphib;
  int64_t i;
  if (!exit) refpt[4] = refpt[4] + e->length / (c->scxl * beta0);
  for (i = 0; i < n; i++) {
    double gamn = gam0 - P(6, i);
    double bz = sqrt(gamn * gamn - 1.0 - P(2, i) * P(2, i) - P(4, i) * P(4, i));
    double t[6], o[6];
    double x, xp, y, yp, d;
    to_transport(c, pts, i, bz, beta0, gambet, qm0, t);
    x = t[0]; xp = t[1]; y = t[2]; yp = t[3]; d = t[5];
    if (!exit) {  /* Fpol_Dipole (src/Appl/Dipole.f90:499-512) */
      o[0] = x - 0.5 * h0 * tanphi2 * x * x + 0.5 * h0 * secphi2 * y * y;
      o[1] = h0 * tanphi * x + xp + h0 * tanphi2 * x * xp +
             (0.5 * h0 * h1 * secphi3 + k1 * tanphi) * x * x -
             (0.5 * h0 * h1 * secphi3 + k1 * tanphi - h0 * h0 * tanphi * tanphi2 -
              0.5 * h0 * h0 * tanphi) * y * y -
             h0 * tanphi2 * y * yp - h0 * tanphi * x * d;
      o[2] = y + h0 * tanphi2 * x * y;
      o[3] = -h0 * tanphib * y + yp - h0 * tanphi2 * x * yp - h0 * secphi2 * xp * y -
             (h0 * h1 * secphi3 + 2 * k1 * tanphi) * x * y +
             (h0 * tanphi - h0 * psi * secphib2) * y * d;
    } else {      /* Bpol_Dipole (src/Appl/Dipole.f90:562-576) */
      o[0] = x + 0.5 * h0 * tanphi2 * x * x - 0.5 * h0 * secphi2 * y * y;
      o[1] = h0 * tanphi * x + xp - h0 * tanphi2 * x * xp +
             (0.5 * h0 * h1 * secphi3 + k1 * tanphi - h0 * h0 * tanphi2 * tanphi / 2) * x * x -
             (0.5 * h0 * h1 * secphi3 + k1 * tanphi + h0 * h0 * tanphi * tanphi2 / 2) * y * y +
             h0 * tanphi2 * y * yp - h0 * tanphi * x * d;
      o[2] = y - h0 * tanphi2 * x * y;
      o[3] = -h0 * tanphib * y + yp + h0 * tanphi2 * x * yp + h0 * secphi2 * xp * y -
             (h0 * h1 * secphi3 + 2 * k1 * tanphi - h0 * h0 * secphi2 * tanphi) * x * y +
             (h0 * tanphi - h0 * psi * secphib2) * y * d;
    }
    o[4] = t[4];
    o[5] = t[5];
    from_transport(c, pts, i, sqrt(gamn * gamn - 1.0), beta0, gambet, qm0, o);
  }
}

void or_sector_dipole(const or_const *c, double *pts, int64_t n, const double refpt[6],
                      double *z, double tau, const or_element *e, double charge,
                      double mass) {
  double g = -refpt[5];
  double beta0 = sqrt(1.0 - 1.0 / g / g);   /* src/Contrl/AccSimulator.f90:1060-1061 */
  double qm0 = charge / mass;
  double h0 = e->v[0] / e->length;
  double rho = 1.0 / h0, th = tau * h0;
  double s = sin(th), co = cos(th), sh = sin(th / 2.0);
  /* first- and second-order transport elements (src/Appl/Dipole.f90:712-743) */
  double R11 = co, R12 = rho * s, R16 = rho * (1.0 - co), R21 = -s / rho, R22 = co;
  double R26 = s, R51 = -R26, R52 = -R16, R56 = rho * (s - beta0 * beta0 * th);
  double T111 = -s * s / (2.0 * rho), T112 = s * co, T116 = s * s;
  double T122 = rho * co * sh * sh, T126 = -rho * s * (-1.0 + co);
  double T166 = -rho * s * s / 2.0, T144 = rho * (-1.0 + co) / 2.0;
  double T216 = s / rho, T222 = -s / 2.0, T266 = -s, T244 = -s / 2.0;
  double T314 = s, T324 = 2.0 * rho * sh * sh, T346 = rho * (th - s);
  double T522 = -0.5 * rho * s, T544 = T522, T526 = rho * (co - 1.0);
  double gambet = beta0 / sqrt(1.0 - beta0 * beta0);
  double gam0 = sqrt(1.0 / (1.0 - beta0 * beta0));
  int64_t i;
  for (i = 0; i < n; i++) {
    double gami = -P(6, i) + gam0;
    double gbi = sqrt(gami * gami - 1.0);
    double bz = sqrt(gbi * gbi - P(2, i) * P(2, i) - P(4, i) * P(4, i));
    double t[6], o[6];
    double x0, xp0, y0, yp0, z0, eta;
    to_transport(c, pts, i, bz, beta0, gambet, qm0, t);
    x0 = t[0]; xp0 = t[1]; y0 = t[2]; yp0 = t[3]; z0 = t[4]; eta = t[5];
    o[0] = R11 * x0 + R12 * xp0 + R16 * eta + T111 * x0 * x0 + T112 * x0 * xp0 +
           T116 * x0 * eta + T122 * xp0 * xp0 + T126 * xp0 * eta + T166 * eta * eta +
           T144 * yp0 * yp0;
    o[1] = R21 * x0 + R22 * xp0 + R26 * eta + T216 * x0 * eta + T222 * xp0 * xp0 +
           T266 * eta * eta + T244 * yp0 * yp0;
    o[2] = y0 + tau * yp0 + T314 * x0 * yp0 + T324 * xp0 * yp0 + T346 * yp0 * eta;
    o[3] = yp0;
    o[4] = R51 * x0 + R52 * xp0 + z0 + R56 * eta + T522 * xp0 * xp0 + T544 * yp0 * yp0 +
           T526 * xp0 * eta;
    o[5] = eta;
    from_transport(c, pts, i, gbi, beta0, gambet, qm0, o);
  }
  *z = *z + tau;
}

void or_isr_chicane(const or_const *c, double *pts, int64_t n, const double refpt[6],
                    double tau, const or_element *e, double mass, const double *rd) {
  /* b0 of the run loop (src/Contrl/AccSimulator.f90:1436-1439) */
  double gamma0 = -refpt[5];
  double beta = sqrt((gamma0 + 1.0) * (gamma0 - 1.0)) / gamma0;
  double brho = gamma0 * beta / c->clight * mass;
  double b0 = fabs(brho * e->v[0] / e->length);
  /* ISRchicane (src/Contrl/AccSimulator.f90:1553-1582) */
  double re = 2.81794028946e-15, emass = 9.109382154e-31, hbar = 1.054571628e-34;
  double qmass = 1.0 / 0.511e6;
  double gambet0 = sqrt(gamma0 * gamma0 - 1.0);
  double c1 = 2 * re * gambet0 * gambet0 / 3;
  double c2 = 55.0 * re * hbar * pow(gambet0, 3) / (24.0 * sqrt(3.0) * (emass * c->clight));
  double bb = b0 * qmass * c->clight / gambet0;
  double p1 = c1 * bb * bb * tau;
  double p2 = sqrt(c2) * pow(sqrt(bb), 3) * sqrt(tau);
  int64_t i;
  for (i = 0; i < n; i++) {
    double gam = gamma0 - P(6, i);
    double gb = sqrt(gam * gam - 1.0);
    double ptot = gb * gb * (p1 + p2 * rd[i]);
    double gb2 = gb - ptot;
    P(2, i) = P(2, i) - ptot * P(2, i) / gb;
    P(4, i) = P(4, i) - ptot * P(4, i) / gb;
    P(6, i) = gamma0 - sqrt(1.0 + gb2 * gb2);
  }
}

void or_drift2(const or_const *c, double *pts, int64_t n, double refpt[6], double *z,
               double tau) {
  /* drift2_BeamBunch (src/Appl/BeamBunch.f90:467-509) */
  const double xl = c->scxl;
  const double gamma0 = -refpt[5];
  const double beta0 = sqrt(1.0 - 1.0 / gamma0 / gamma0);
  int64_t i;
  for (i = 0; i < n; i++) {
    double pz = sqrt((gamma0 - P(6, i)) * (gamma0 - P(6, i)) - 1.0 - P(2, i) * P(2, i) -
                     P(4, i) * P(4, i));
    P(1, i) = P(1, i) + 0.5 * tau * P(2, i) / (xl * pz);
    P(3, i) = P(3, i) + 0.5 * tau * P(4, i) / (xl * pz);
    P(5, i) = P(5, i) + 0.5 * tau * ((gamma0 - P(6, i)) / (xl * pz) - 1.0 / (beta0 * xl));
  }
  refpt[4] = refpt[4] + 0.5 * tau / (xl * beta0);
  *z = *z + 0.5 * tau;
}

/* getfld_BeamLineElem for the analytic fields (src/Appl/BeamLineElem.f90:
 * 827-870): E (V/m) and B (T) at pos = (x, y, z, t).  Returns -1 for the
 * elements not restated. */
/* getfld_DTL / _CCDTL / _CCL / _SC (src/Appl/DTL.f90:769-855, CCL.f90,
 * CCDTL.f90, SC.f90): the DTL's drift-tube quadrupoles, then the on-axis
 * field of read3_Data's Fourier series (Fcoef(1) = a0, then cosine and
 * sine pairs over the element length about its middle) and its paraxial
 * expansion, with the REAL(4) constants of the reference (clite, pi,
 * asin(1.0)). */
static int rf_field(const or_const *c, int bitype, const double *v, double length,
                    double zedge, const or_rfdata *rf, const double pos[4], double f[6]) {
  const double clite = (double)299792458.e0f, pi = (double)(2 * asinf(1.0f));
  const double escale = v[0], ww = v[1] / c->scfreq, xl = c->scxl / ww;
  const double theta0 = v[2] * (double)asinf(1.0f) / 90;
  const double zmid = zedge + length / 2;
  const double *F = rf->fcoef - 1;
  double zz, ez1, ezp1 = 0.0, ezpp1 = 0.0, ezppp = 0.0, tt, tmpcos, tmpsin, f1, f1p, r2;
  int i;
  if (!rf || !rf->fourier) return -1;
  if (bitype == 101) {
    double bgrad;
    zz = pos[2] - zedge;
    if (zz < v[5]) bgrad = v[6];
    else if (zz < length - v[7]) bgrad = 0.0;
    else bgrad = v[8];
    f[3] = bgrad * pos[1];
    f[4] = bgrad * pos[0];
  }
  zz = pos[2] - zmid;
  ez1 = F[1] / 2;
  for (i = 2; i <= (rf->n - 1) / 2 + 1; i++) {
    double k = (i - 1) * 2 * pi / length, arg = (i - 1) * 2 * pi * zz / length;
    double cs = cos(arg), sn = sin(arg);
    ez1 = ez1 + F[2 * i - 2] * cs + F[2 * i - 1] * sn;
    ezp1 = ezp1 + k * (-F[2 * i - 2] * sn + F[2 * i - 1] * cs);
    ezpp1 = ezpp1 + (k * k) * (-F[2 * i - 2] * cs - F[2 * i - 1] * sn);
    ezppp = ezppp + (k * k * k) * (F[2 * i - 2] * sn - F[2 * i - 1] * cs);
  }
  ez1 = ez1 * escale;
  ezp1 = ezp1 * escale;
  ezpp1 = ezpp1 * escale;
  ezppp = ezppp * escale;
  tt = pos[3];
  tmpcos = cos(ww * tt + theta0);
  tmpsin = sin(ww * tt + theta0);
  f1 = -(ezpp1 + ez1 / xl / xl) / 4;
  f1p = -(ezppp + ezp1 / xl / xl) / 4;
  r2 = pos[0] * pos[0] + pos[1] * pos[1];
  f[0] += -pos[0] * (ezp1 / 2 + f1p * r2 / 4) * tmpcos;
  f[1] += -pos[1] * (ezp1 / 2 + f1p * r2 / 4) * tmpcos;
  f[2] += (ez1 + f1 * r2) * tmpcos;
  f[3] += pos[1] / (xl * clite) * (ez1 / 2 + f1 * r2 / 4) * tmpsin;
  f[4] += -pos[0] / (xl * clite) * (ez1 / 2 + f1 * r2 / 4) * tmpsin;
  return 0;
}

/* One Fourier section of getfld_SolRF (src/Appl/SolRF.f90:513-547,577-611):
 * the series whose count, start, end and period sit at Fcoef(o+1..o+4),
 * coefficients from Fcoef(o+5), at zz from the section's middle; returns
 * the field and its first three derivatives (unscaled). */
static void solrf_series(const double *F, int o, double zz, double pi, double d[4]) {
  int np = (int)(F[o + 1] + 0.1), i, nt = o + 4;
  double zlen = F[o + 4];
  d[0] = F[nt + 1] / 2;
  d[1] = d[2] = d[3] = 0.0;
  for (i = 2; i <= (np - 1) / 2 + 1; i++) {
    double arg = (i - 1) * 2 * pi * zz / zlen, k = (i - 1) * 2 * pi / zlen;
    double cs = cos(arg), sn = sin(arg);
    d[0] = d[0] + F[2 * i - 2 + nt] * cs + F[2 * i - 1 + nt] * sn;
    d[1] = d[1] + k * (-F[2 * i - 2 + nt] * sn + F[2 * i - 1 + nt] * cs);
    d[2] = d[2] + (k * k) * (-F[2 * i - 2 + nt] * cs - F[2 * i - 1 + nt] * sn);
    d[3] = d[3] + (k * k * k) * (F[2 * i - 2 + nt] * sn - F[2 * i - 1 + nt] * cs);
  }
}

/* getfld_SolRF (src/Appl/SolRF.f90:482-620): inside the element, the RF
 * field of the first Fourier section (its paraxial expansion at the real
 * angular frequency) and the solenoid field of the second, scaled by
 * Param(12), each within its own z range. */
static int solrf_field(const or_const *c, const double *v, double length, double zedge,
                       const or_rfdata *rf, const double pos[4], double f[6]) {
  const double clite = (double)299792458.e0f, pi = (double)(2 * asinf(1.0f));
  const double *F;
  int np1;
  double zlc, d[4];
  if (!rf || !rf->fourier) return -1;
  F = rf->fcoef - 1;
  if (!(pos[2] > zedge && pos[2] < zedge + length)) return 0;
  zlc = pos[2] - zedge;
  np1 = (int)(F[1] + 0.1);
  if (zlc >= F[2] && zlc <= F[3]) {
    double ww = v[1] * 2 * pi, xlrep = ww / clite;
    double theta0 = v[2] * (double)asinf(1.0f) / 90;
    double ez1, ezp1, ezpp1, ezppp, tt, tmpcos, tmpsin, f1, f1p, r2;
    solrf_series(F, 0, pos[2] - zedge - F[2] - F[4] / 2, pi, d);
    ez1 = d[0] * v[0];
    ezp1 = d[1] * v[0];
    ezpp1 = d[2] * v[0];
    ezppp = d[3] * v[0];
    tt = pos[3] / (2 * pi * c->scfreq);
    tmpcos = cos(ww * tt + theta0);
    tmpsin = sin(ww * tt + theta0);
    f1 = -(ezpp1 + ez1 * xlrep * xlrep) / 4;
    f1p = -(ezppp + ezp1 * xlrep * xlrep) / 4;
    r2 = pos[0] * pos[0] + pos[1] * pos[1];
    f[0] += -pos[0] * (ezp1 / 2 + f1p * r2 / 4) * tmpcos;
    f[1] += -pos[1] * (ezp1 / 2 + f1p * r2 / 4) * tmpcos;
    f[2] += (ez1 + f1 * r2) * tmpcos;
    f[3] += pos[1] * xlrep / clite * (ez1 / 2 + f1 * r2 / 4) * tmpsin;
    f[4] += -pos[0] * xlrep / clite * (ez1 / 2 + f1 * r2 / 4) * tmpsin;
  }
  {
    int o = 4 + np1;
    double bscale = v[10], r2;
    if (zlc >= F[o + 2] && zlc <= F[o + 3]) {
      solrf_series(F, o, pos[2] - zedge - F[o + 2] - F[o + 4] / 2, pi, d);
      r2 = pos[0] * pos[0] + pos[1] * pos[1];
      f[3] = f[3] - bscale * d[1] / 2 * pos[0] + bscale * d[3] * pos[0] * r2 / 16;
      f[4] = f[4] - bscale * d[1] / 2 * pos[1] + bscale * d[3] * pos[1] * r2 / 16;
      f[5] = f[5] + bscale * d[0] - bscale * d[2] * r2 / 4;
    }
  }
  return 0;
}

int or_elem_field(const or_const *c, int bitype, const double *v, double length,
                  double zedge, const or_rfdata *rf, const double pos[4], double f[6]) {
  int k;
  for (k = 0; k < 6; k++) f[k] = 0.0;
  if (bitype >= 101 && bitype <= 104) return rf_field(c, bitype, v, length, zedge, rf, pos, f);
  if (bitype == 105) return solrf_field(c, v, length, zedge, rf, pos, f);
  if (bitype == 0) return 0;  /* getfld_DriftTube: no field */
  if (bitype == 1) {
    /* getfld_Quadrupole (src/Appl/Quadrupole.f90:369-402), file ID <= 0 */
    if (v[1] > 0.0) return -1;
    f[3] = v[0] * pos[1];
    f[4] = v[0] * pos[0];
    return 0;
  }
  if (bitype == 2) {
    /* getfld_ConstFoc (src/Appl/ConstFoc.f90:297-322): qmcc = gambetz = 1 */
    const double qmcc = 1.0, pz = 1.0;
    f[2] = v[2] * pos[3] * pz * pz * pz * c->scxl / qmcc;
    f[3] = -v[0] * pos[1] * pz / qmcc / c->clight;
    f[4] = v[1] * pos[0] * pz / qmcc / c->clight;
    return 0;
  }
  if (bitype == 3) {
    /* getfld_Sol (src/Appl/Sol.f90:388-435): linear ramps over 2*Param(4) */
    const double x02 = 2 * v[2];
    const double zz = pos[2] - zedge;
    const double b0 = v[0];
    if (zz >= 0.0 && zz < x02) {
      const double bb = b0 / x02;
      f[3] = -bb / 2 * pos[0];
      f[4] = -bb / 2 * pos[1];
      f[5] = bb * zz;
    } else if (zz >= x02 && zz <= length - x02) {
      f[5] = b0;
    } else if (zz > length - x02 && zz <= length) {
      const double bb = -b0 / x02;
      const double zshift = length - x02;
      f[3] = -bb / 2 * pos[0];
      f[4] = -bb / 2 * pos[1];
      f[5] = b0 + bb * (zz - zshift);
    }
    return 0;
  }
  return -1;
}

/* The implicit Lorentz push of scatter2/scatter20 (src/Appl/BeamBunch.f90:
 * 3026-3066): two iterations with the longitudinal momentum averaged. */
static void lorentz_push(double *p, double ex, double ey, double ez, double bx, double by,
                         double bz, double tau, double clight, double tmpgamma0,
                         double gam, double qmcc, double ez0) {
  const double q = p[6];
  const double tmp12 = -tau * clight * bz * q / 2;
  const double tmp13 = tau * ex * q / 2;
  const double tmp23 = tau * ey * q / 2;
  const double tmp1 = tmpgamma0 * ex * tau * q;
  const double tmp2 = clight * by * tau * q;
  const double tmp3 = tmpgamma0 * ey * tau * q;
  const double tmp4 = clight * bx * tau * q;
  const double tmp5 = (ez0 * qmcc - ez * q) * tau;
  double pz = sqrt((gam - p[5]) * (gam - p[5]) - 1.0 - p[1] * p[1] - p[3] * p[3]);
  const double p1 = p[1], p2 = p[3], p3 = p[5];
  int ii;
  for (ii = 0; ii < 2; ii++) {
    const double a12 = tmp12 / pz, a13 = tmp13 / pz, a23 = tmp23 / pz;
    const double b1 = p1 - a12 * p2 - a13 * p3 + tmp1 / pz - tmp2;
    const double b2 = a12 * p1 + p2 - a23 * p3 + tmp3 / pz + tmp4;
    const double b3 = -a13 * p1 - a23 * p2 + p3 + tmp5;
    const double det = 1 + a12 * a12 - a13 * a13 - a23 * a23;
    const double s11 = 1 - a23 * a23, s12 = -a12 + a13 * a23, s13 = a12 * a23 - a13;
    const double s21 = a12 + a13 * a23, s22 = 1 - a13 * a13, s23 = -a23 - a13 * a12;
    const double s31 = -a12 * a23 - a13, s32 = -a23 + a12 * a13, s33 = 1 + a12 * a12;
    double tmpgamma00, tmpsq2;
    p[1] = (s11 * b1 + s12 * b2 + s13 * b3) / det;
    p[3] = (s21 * b1 + s22 * b2 + s23 * b3) / det;
    p[5] = (s31 * b1 + s32 * b2 + s33 * b3) / det;
    tmpgamma00 = gam + tau * qmcc * ez0;
    tmpsq2 = (tmpgamma00 - p[5]) * (tmpgamma00 - p[5]) - 1.0 - p[1] * p[1] - p[3] * p[3];
    if (tmpsq2 > 0.0) pz = 0.5 * pz + 0.5 * sqrt(tmpsq2);
  }
}

int or_kick2(const or_const *c, double *pts, int64_t n, int nx, int ny, int nz,
             const double *phi, const double range[6], const double msize[3],
             double refpt[6], double curr, double tau, double mass, double chrg,
             int bitype, const double *v, double length, double zedge,
             const or_rfdata *rf, double z, const double *wake) {
  /* kick2_BeamBunch (src/Appl/BeamBunch.f90:2187-2508) with scatter2
   * (:2510-3071) or, without current, scatter20 (:3075-3531), for the
   * elements with analytic fields */
  const double xk = 1.0 / c->scxl, xl = c->scxl;
  const double tmpscale = c->clight * c->clight * (double)1.0e-7f;
  const double tg = refpt[4];
  double gam = -refpt[5];
  const double beta0 = sqrt(gam * gam - 1.0) / gam;
  const double qmcc = chrg / mass;
  double pos[4] = {0.0, 0.0, z, tg}, fld[6], wk[3] = {0.0, 0.0, 0.0};
  double ez0, tmpgamma0;
  int64_t i;
  if (or_elem_field(c, bitype, v, length, zedge, rf, pos, fld) != 0) return -1;
  ez0 = fld[2];
  tmpgamma0 = gam + 0.5 * tau * qmcc * ez0;
  if (curr > 0.0) {
    size_t g = (size_t)nx * ny * nz;
    double *e = (double *)malloc(3 * g * sizeof(double));
    const double *ex = e, *ey = e + g, *ezg = e + 2 * g;
    const double hxi = 1.0 / msize[0], hyi = 1.0 / msize[1], hzi = 1.0 / msize[2];
    const double hx = msize[0], hy = msize[1], hz = msize[2];
    const double xmin = range[0], ymin = range[2], zmin = range[4];
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    or_efield(phi, nx, ny, nz, msize, e, e + g, e + 2 * g);
    for (i = 0; i < n; i++) {
      int ix = (int)((P(1, i) - xmin) * hxi + 1.0);
      double ab = ((xmin - P(1, i)) + ix * hx) * hxi;
      int jx = (int)((P(3, i) - ymin) * hyi + 1.0);
      double cd = ((ymin - P(3, i)) + jx * hy) * hyi;
      int kx = (int)((P(5, i) - zmin) * hzi + 1.0);
      double ef = ((zmin - P(5, i)) + kx * hz) * hzi;
      int64_t o = (ix - 1) + sy * (jx - 1) + sz * (kx - 1);
      double exn = cic(ex, o, sy, sz, ab, cd, ef) * gam;
      double eyn = cic(ey, o, sy, sz, ab, cd, ef) * gam;
      double ezn = cic(ezg, o, sy, sz, ab, cd, ef);
      double rcpgammai = 1.0 / (-P(6, i) + gam);
      double betai = sqrt(1.0 - rcpgammai * rcpgammai *
                                    (1 + P(2, i) * P(2, i) + P(4, i) * P(4, i)));
      P(5, i) = P(5, i) * xk / (-gam * betai);
      P(1, i) = P(1, i) * xk;
      P(3, i) = P(3, i) * xk;
      pos[0] = P(1, i) * xl;
      pos[1] = P(3, i) * xl;
      pos[2] = z;
      pos[3] = P(5, i) + tg;
      or_elem_field(c, bitype, v, length, zedge, rf, pos, fld);
      if (wake) {
        /* kick2wake with scatter2wake (src/Appl/BeamBunch.f90:5048-5938): the
         * wake of the particle's z, linear in its CIC weight, added to E */
        int k1 = kx < 1 ? 1 : (kx > nz - 1 ? nz - 1 : kx);
        wk[0] = wake[k1 - 1] * ef + wake[k1] * (1.0 - ef);
        wk[1] = wake[nz + k1 - 1] * ef + wake[nz + k1] * (1.0 - ef);
        wk[2] = wake[2 * nz + k1 - 1] * ef + wake[2 * nz + k1] * (1.0 - ef);
      }
      lorentz_push(&P(1, i), tmpscale * exn + fld[0] + wk[0], tmpscale * eyn + fld[1] + wk[1],
                   tmpscale * ezn + fld[2] + wk[2], -beta0 / c->clight * tmpscale * eyn + fld[3],
                   beta0 / c->clight * tmpscale * exn + fld[4], fld[5], tau, c->clight,
                   tmpgamma0, gam, qmcc, ez0);
    }
    free(e);
  } else {
    for (i = 0; i < n; i++) {
      pos[0] = P(1, i) * xl;
      pos[1] = P(3, i) * xl;
      pos[2] = z;
      pos[3] = P(5, i) + tg;
      or_elem_field(c, bitype, v, length, zedge, rf, pos, fld);
      lorentz_push(&P(1, i), fld[0], fld[1], fld[2], fld[3], fld[4], fld[5], tau,
                   c->clight, tmpgamma0, gam, qmcc, ez0);
    }
  }
  gam = gam + tau * qmcc * ez0;
  refpt[5] = -gam;
  return 0;
}

int64_t or_step2(const or_const *c, double *pts, int64_t n, or_bunch_state *st,
                 int nx, int ny, int nz, double tau1, const or_element *e,
                 double pipe_radius, const or_wake *wake, double *rho_out,
                 double *phi_out) {
  /* one segment of the reference loop with Flagmap=2
   * (src/Contrl/AccSimulator.f90:1106-1113,1446-1458) */
  size_t g = (size_t)nx * ny * nz;
  double *rho = rho_out ? rho_out : (double *)calloc(g, sizeof(double));
  double *phi = phi_out ? phi_out : (double *)calloc(g, sizeof(double));
  double ptrange[6], range[6] = {0, 0, 0, 0, 0, 0}, msize[3] = {1, 1, 1};
  const double tau2 = 2.0 * tau1;
  int rc = 0;
  double curr = 0.0, *wk = NULL;
  or_drift2(c, pts, n, st->refpt, &st->z, tau2);
  if (st->bcurr >= (double)1.0e-10f) {
    n = or_conv1st(c, pts, n, st->refpt, pipe_radius, pipe_radius, ptrange,
                   &st->npt, &st->current);
    or_update_compdom(ptrange, nx, ny, nz, range, msize);
    or_deposit(pts, n, nx, ny, nz, range, msize, rho);
    rc = or_solve_openbc(rho, nx, ny, nz, msize[0], msize[1], msize[2], phi);
    if (rc == 0 && wake && wake->kind) {
      /* src/Contrl/AccSimulator.f90:1272-1338, then kick2wake (:1446-1450) */
      wk = (double *)malloc(3 * (size_t)nz * sizeof(double));
      rc = or_wakefield(c, pts, n, range, nz, st->current, st->npt, -st->refpt[5],
                        st->charge, wake, wk, wk + nz, wk + 2 * nz);
    }
    or_cvbkforth1st(c, pts, n, st->refpt);
    curr = st->current;
  }
  if (rc == 0)
    rc = or_kick2(c, pts, n, nx, ny, nz, phi, range, msize, st->refpt, curr, tau2,
                  st->mass, st->charge, e->bitype, e->v, e->length, e->zedge, e->rf,
                  st->z, curr > 0.0 ? wk : NULL);
  free(wk);
  if (rc == 0) or_drift2(c, pts, n, st->refpt, &st->z, tau2);
  if (!rho_out) free(rho);
  if (!phi_out) free(phi);
  return rc == 0 ? n : -1;
}

/* -------------------------------------------------------- diagnostics */

void or_diag_sums(const double *pts, int64_t n, double s[37], double amax[6],
                  double *r2max) {
  int64_t i;
  int j;
  for (j = 0; j < 37; j++) s[j] = 0.0;
  if (n > 0) {
    for (j = 0; j < 6; j++) amax[j] = fabs(P(j + 1, 0));
    *r2max = P(1, 0) * P(1, 0) + P(3, 0) * P(3, 0);
  } else {
    for (j = 0; j < 6; j++) amax[j] = 0.0;
    *r2max = 0.0;
  }
  for (i = 0; i < n; i++) {
    double x = P(1, i), px = P(2, i), y = P(3, i), py = P(4, i), z = P(5, i),
           pz = P(6, i), r2;
    s[0] += x;  s[6] += x * x;  s[15] += x * x * x;  s[16] += x * x * x * x;
    s[12] += x * px;  s[27] += x * px * px;  s[28] += x * px * px * px;
    s[29] += x * x * px;  s[30] += x * x * px * px;  s[31] += x * x * x * px;
    s[1] += px;  s[7] += px * px;  s[17] += px * px * px;  s[18] += px * px * px * px;
    s[2] += y;  s[8] += y * y;  s[19] += y * y * y;  s[20] += y * y * y * y;
    s[13] += y * py;  s[32] += y * py * py;  s[33] += y * py * py * py;
    s[34] += y * y * py;  s[35] += y * y * py * py;  s[36] += y * y * y * py;
    s[3] += py;  s[9] += py * py;  s[21] += py * py * py;  s[22] += py * py * py * py;
    s[4] += z;  s[10] += z * z;  s[23] += fabs(z * z * z);  s[24] += z * z * z * z;
    s[14] += z * pz;
    s[5] += pz;  s[11] += pz * pz;  s[25] += pz * pz * pz;  s[26] += pz * pz * pz * pz;
    for (j = 0; j < 6; j++)
      if (amax[j] < fabs(P(j + 1, i))) amax[j] = fabs(P(j + 1, i));
    r2 = x * x + y * y;
    if (*r2max < r2) *r2max = r2;
  }
}

void or_diag_row(const or_const *c, const double t[37], const double gm[6],
                 double r2max, int64_t npt, const double refpt[6], double mass,
                 double out[41]) {
  const double qmc = mass / 1.0e6, xl = c->scxl, xt = 180.0 / c->pi;
  const double third = (double)(1.0f / 3.0f);
  const double den1 = 1.0 / (double)npt;
  const double gam = -refpt[5];
  const double gambet = sqrt(gam * gam - 1.0);
  const double bet = sqrt(gam * gam - 1.0) / gam;
  const double energy = (gam - 1.0) * qmc;
  double x0 = t[0] * den1, px0 = t[1] * den1, y0 = t[2] * den1, py0 = t[3] * den1;
  double z0 = t[4] * den1, pz0 = t[5] * den1;
  double sqx = t[6] * den1, sqpx = t[7] * den1, sqy = t[8] * den1;
  double sqpy = t[9] * den1, sqz = t[10] * den1, sqpz = t[11] * den1;
  double s1 = sqx - x0 * x0, s2 = sqpx - px0 * px0, s3 = sqy - y0 * y0;
  double s4 = sqpy - py0 * py0, s5 = sqz - z0 * z0, s6 = sqpz - pz0 * pz0;
  double xpx = t[12] * den1 - x0 * px0, ypy = t[13] * den1 - y0 * py0;
  double zpz = t[14] * den1 - z0 * pz0;
  double cubx = t[15] * den1, fthx = t[16] * den1, cubpx = t[17] * den1;
  double fthpx = t[18] * den1, cuby = t[19] * den1, fthy = t[20] * den1;
  double cubpy = t[21] * den1, fthpy = t[22] * den1, cubz = t[23] * den1;
  double fthz = t[24] * den1, cubpz = t[25] * den1, fthpz = t[26] * den1;
#define M3(c3, sq, m0) pow(fabs((c3) - 3 * (sq) * (m0) + 2 * (m0) * (m0) * (m0)), third)
#define M4(f4, c3, sq, m0) \
  sqrt(sqrt(fabs((f4) - 4 * (c3) * (m0) + 6 * (sq) * (m0) * (m0) - 3 * (m0) * (m0) * (m0) * (m0))))
  double x03 = M3(cubx, sqx, x0), x04 = M4(fthx, cubx, sqx, x0);
  double px03 = M3(cubpx, sqpx, px0), px04 = M4(fthpx, cubpx, sqpx, px0);
  double y03 = M3(cuby, sqy, y0), y04 = M4(fthy, cuby, sqy, y0);
  double py03 = M3(cubpy, sqpy, py0), py04 = M4(fthpy, cubpy, sqpy, py0);
  double z03 = M3(cubz, sqz, z0), z04 = M4(fthz, cubz, sqz, z0);
  double pz03 = M3(cubpz, sqpz, pz0), pz04 = M4(fthpz, cubpz, sqpz, pz0);
#undef M3
#undef M4
  double epx = sqrt(fmax(s1 * s2 - xpx * xpx, 0.0));
  double epy = sqrt(fmax(s3 * s4 - ypy * ypy, 0.0));
  double epz = sqrt(fmax(s5 * s6 - zpz * zpz, 0.0));
  double xrms = sqrt(s1), pxrms = sqrt(s2), yrms = sqrt(s3), pyrms = sqrt(s4);
  double zrms = sqrt(s5), pzrms = sqrt(s6);
  double v[41] = {
      refpt[4], gam, energy, bet, sqrt(r2max) * xl,
      x0 * xl, xrms * xl, px0 / gambet, pxrms / gambet, -xpx / epx, epx * xl,
      y0 * xl, yrms * xl, py0 / gambet, pyrms / gambet, -ypy / epy, epy * xl,
      z0 * xt, zrms * xt, pz0 * qmc, pzrms * qmc, -zpz / epz, epz * qmc * xt,
      gm[0] * xl, gm[1] / gambet, gm[2] * xl, gm[3] / gambet, gm[4] * xt, gm[5] * qmc,
      x03 * xl, px03 / gambet, y03 * xl, py03 / gambet, z03 * xt, pz03 * qmc,
      x04 * xl, px04 / gambet, y04 * xl, py04 / gambet, z04 * xt, pz04 * qmc};
  memcpy(out, v, sizeof v);
}

/* chgupdate_BeamBunch (src/Appl/BeamBunch.f90:688-728): particles per charge
 * state, a particle counting for state j when |(Pts1(7)-qmcc(j))/qmcc(j)| is
 * below the REAL(4) literal 1.0e-4 (:700). */
void or_chgupdate(const double *pts, int64_t n, int nchge, const double *qmcc,
                  int64_t *counts) {
  const double eps = (double)1.0e-4f;
  for (int j = 0; j < nchge; j++) counts[j] = 0;
  for (int64_t i = 0; i < n; i++)
    for (int j = 0; j < nchge; j++) {
      const double dd = fabs((pts[9 * i + 6] - qmcc[j]) / qmcc[j]);
      if (dd < eps) counts[j]++;
    }
}

/* ------------------------------------------------------------ one step */

int64_t or_step(const or_const *c, double *pts, int64_t n, or_bunch_state *st,
                int nx, int ny, int nz, double tau1, const or_element *e,
                double pipe_radius, const or_wake *wake, double *rho_out,
                double *phi_out) {
  size_t g = (size_t)nx * ny * nz;
  double *rho = rho_out ? rho_out : (double *)malloc(g * sizeof(double));
  double *phi = phi_out ? phi_out : (double *)malloc(g * sizeof(double));
  double ptrange[6], range[6], msize[3];
  const double tau2 = 2.0 * tau1;
  int rc = or_map1(c, pts, n, st->refpt, &st->z, tau1, e, st->charge,
                   st->mass);
  if (rc == 0 && st->bcurr >= (double)1.0e-10f) {
    n = or_conv1st(c, pts, n, st->refpt, pipe_radius, pipe_radius, ptrange,
                   &st->npt, &st->current);
    or_update_compdom(ptrange, nx, ny, nz, range, msize);
    or_deposit(pts, n, nx, ny, nz, range, msize, rho);
    rc = or_solve_openbc(rho, nx, ny, nz, msize[0], msize[1], msize[2], phi);
    if (rc == 0 && wake && wake->kind) {
      /* src/Contrl/AccSimulator.f90:1272-1338, then kick1wake (:1417-1421) */
      double *wk = (double *)malloc(3 * (size_t)nz * sizeof(double));
      rc = or_wakefield(c, pts, n, range, nz, st->current, st->npt, -st->refpt[5],
                        st->charge, wake, wk, wk + nz, wk + 2 * nz);
      or_cvbkforth1st(c, pts, n, st->refpt);
      if (rc == 0)
        or_kick1wake(c, pts, n, nx, ny, nz, phi, range, msize, st->refpt, st->current,
                     tau2, st->mass, st->charge, wk, wk + nz, wk + 2 * nz);
      free(wk);
    } else {
      or_cvbkforth1st(c, pts, n, st->refpt);
      if (rc == 0)
        or_kick1(c, pts, n, nx, ny, nz, phi, range, msize, st->refpt, st->current,
                 tau2, st->mass, st->charge);
    }
  }
  if (rc == 0)
    rc = or_map1(c, pts, n, st->refpt, &st->z, tau1, e, st->charge,
                 st->mass);
  if (!rho_out) free(rho);
  if (!phi_out) free(phi);
  return rc == 0 ? n : -1;
}
