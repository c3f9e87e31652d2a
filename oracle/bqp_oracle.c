/* CPU FP64 oracle — TEST INFRASTRUCTURE ONLY (see bqp_oracle.h).
 *
 * Sources restated (formulas from the published algorithms; structure follows
 * the reference functions cited beside each block):
 *   Montenbruck & Gill, "Satellite Orbits" (2000) Sec. 3.2.4-3.2.5: Cunningham
 *     V/W recursion (eqs. 3.29-3.30) and the acceleration sums (eq. 3.33).
 *   Holmes & Featherstone (2002): full normalisation N_nm and the normalised
 *     column recursion constants.
 *   Vallado (2013) Table 8-4: piecewise exponential atmosphere.
 *   Gustafsson (1991): PI step-size control, with the reference's constants.
 */
#include "bqp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(n, m) ((n) * ((n) + 1) / 2 + (m))

struct orc_field {
  int nmax, mmax, form;
  double mu, radius, gm_r2;
  double* cbar; /* triangular, n = 0..nmax */
  double* sbar;
  /* recursion constants for rows n = 0..nmax+1 (triangular) */
  double* ra;   /* a_nm (normalised) or pt1 = (2n-1)/(n-m) (compat) */
  double* rb;   /* b_nm (normalised) or pt21 = (n+m-1)/(n-m) (compat) */
  double* sect; /* sectoral factor per m */
  /* summation factors per (n, m), n = 0..nmax */
  double* f1;
  double* f2;
  double* f3;
  double* cu; /* compat: de-normalised C, S */
  double* su;
};

/* ---------------------------------------------------------------- field */

/* Full normalisation factor N_nm = sqrt((2-d_m0)(2n+1)(n-m)!/(n+m)!), by a
 * direct product in long double (the reference uses ln-factorials,
 * LY/src/config.rs:142-152, which loses ~1e-15 relative). */
static double norm_factor(int n, int m) {
  long double p = (long double)((m == 0) ? 1 : 2) * (long double)(2 * n + 1);
  for (int k = n - m + 1; k <= n + m; ++k) p /= (long double)k;
  return (double)sqrtl(p);
}

orc_field* orc_field_create(int nmax, int mmax, const double* cbar, const double* sbar,
                            double mu, double radius, int form) {
  if (nmax < 0 || nmax > ORC_NMAX || mmax < 0 || mmax > nmax || !(mu > 0) || !(radius > 0))
    return NULL;
  if (form == ORC_GRAVITY_COMPAT && nmax > 100) return NULL;
  orc_field* f = (orc_field*)calloc(1, sizeof(orc_field));
  f->nmax = nmax;
  f->mmax = mmax;
  f->form = form;
  f->mu = mu;
  f->radius = radius;
  f->gm_r2 = mu / (radius * radius);
  size_t nt = (size_t)IDX(nmax + 1, 0);   /* terms for n <= nmax */
  size_t nr = (size_t)IDX(nmax + 2, 0);   /* rows n <= nmax+1 */
  f->cbar = (double*)calloc(nt, sizeof(double));
  f->sbar = (double*)calloc(nt, sizeof(double));
  f->cu = (double*)calloc(nt, sizeof(double));
  f->su = (double*)calloc(nt, sizeof(double));
  f->f1 = (double*)calloc(nt, sizeof(double));
  f->f2 = (double*)calloc(nt, sizeof(double));
  f->f3 = (double*)calloc(nt, sizeof(double));
  f->ra = (double*)calloc(nr, sizeof(double));
  f->rb = (double*)calloc(nr, sizeof(double));
  f->sect = (double*)calloc((size_t)nmax + 2, sizeof(double));
  for (int n = 0; n <= nmax; ++n)
    for (int m = 0; m <= n; ++m) {
      if (m > mmax) continue;
      f->cbar[IDX(n, m)] = cbar[IDX(n, m)];
      f->sbar[IDX(n, m)] = (m == 0) ? 0.0 : sbar[IDX(n, m)];
    }
  if (form == ORC_GRAVITY_NORMALISED) {
    for (int m = 1; m <= nmax + 1; ++m)
      f->sect[m] = (m == 1) ? sqrt(3.0) : sqrt((2.0 * m + 1.0) / (2.0 * m));
    for (int n = 1; n <= nmax + 1; ++n)
      for (int m = 0; m < n; ++m) {
        double dn = n, dm = m;
        f->ra[IDX(n, m)] = sqrt((2 * dn + 1) * (2 * dn - 1) / ((dn - dm) * (dn + dm)));
        f->rb[IDX(n, m)] = (n >= m + 2)
            ? sqrt((2 * dn + 1) * (dn + dm - 1) * (dn - dm - 1) /
                   ((dn - dm) * (dn + dm) * (2 * dn - 3)))
            : 0.0;
      }
    for (int n = 0; n <= nmax; ++n)
      for (int m = 0; m <= n; ++m) {
        double dn = n, dm = m, k = (2 * dn + 1) / (2 * dn + 3);
        if (m == 0) {
          f->f1[IDX(n, 0)] = sqrt(k * (dn + 1) * (dn + 2) / 2.0);
        } else {
          f->f1[IDX(n, m)] = 0.5 * sqrt(k * (dn + dm + 1) * (dn + dm + 2));
          f->f2[IDX(n, m)] = 0.5 * sqrt(k * (dn - dm + 1) * (dn - dm + 2)) * (m == 1 ? sqrt(2.0) : 1.0);
        }
        f->f3[IDX(n, m)] = sqrt(k * (dn + dm + 1) * (dn - dm + 1));
      }
  } else {
    /* reference semantics: de-normalised coefficients (LY/src/config.rs:186-213)
     * and the un-normalised pt1/pt21 constants (SC/src/gravity.rs:534-600). */
    for (int m = 1; m <= nmax + 1; ++m) f->sect[m] = 2.0 * m - 1.0;
    for (int n = 1; n <= nmax + 1; ++n)
      for (int m = 0; m < n; ++m) {
        f->ra[IDX(n, m)] = (2.0 * n - 1.0) / (double)(n - m);
        f->rb[IDX(n, m)] = (double)(n + m - 1) / (double)(n - m);
      }
    for (int n = 0; n <= nmax; ++n)
      for (int m = 0; m <= n; ++m) {
        double nf = norm_factor(n, m);
        f->cu[IDX(n, m)] = f->cbar[IDX(n, m)] * nf;
        f->su[IDX(n, m)] = f->sbar[IDX(n, m)] * nf;
      }
  }
  return f;
}

void orc_field_destroy(orc_field* f) {
  if (!f) return;
  free(f->cbar); free(f->sbar); free(f->cu); free(f->su);
  free(f->f1); free(f->f2); free(f->f3);
  free(f->ra); free(f->rb); free(f->sect);
  free(f);
}

/* ---------------------------------------------------------------- gravity */

static __thread double tls_v[(ORC_NMAX + 3) * (ORC_NMAX + 4) / 2];
static __thread double tls_w[(ORC_NMAX + 3) * (ORC_NMAX + 4) / 2];

/* V/W fill then term-by-term M-G summation, the two-phase structure of
 * SC/src/gravity.rs:2638-2675 (legendre_vw_dispatch :758 + summation :1719).
 * The fill is m-outer: sectoral seed, then the column recursion. */
int orc_gravity_bf(const orc_field* f, const double pos[3], double acc[3]) {
  const double x = pos[0], y = pos[1], z = pos[2];
  const double r2 = fma(x, x, fma(y, y, z * z));
  if (!(r2 > 0.0) || !isfinite(r2)) {
    acc[0] = acc[1] = acc[2] = NAN;
    return ORC_INVALID_INPUT;
  }
  const int N = f->nmax, M = f->mmax;
  const double R = f->radius;
  const double rr = 1.0 / r2;
  const double xt = x * R * rr, yt = y * R * rr, zt = z * R * rr, rho = R * R * rr;
  double* V = tls_v;
  double* W = tls_w;
  const int top = N + 1; /* rows 0..N+1 are needed */
  const int mtop = (M + 1 < top) ? M + 1 : top;
  /* column m = 0 .. mtop */
  double vs = R / sqrt(r2), ws = 0.0; /* V00, W00 */
  for (int m = 0; m <= mtop; ++m) {
    if (m > 0) {
      const double s = f->sect[m];
      const double nv = s * fma(xt, vs, -(yt * ws));
      const double nw = s * fma(xt, ws, yt * vs);
      vs = nv;
      ws = nw;
    }
    V[IDX(m, m)] = vs;
    W[IDX(m, m)] = ws;
    double v1 = vs, w1 = ws, v2 = 0.0, w2 = 0.0;
    for (int n = m + 1; n <= top; ++n) {
      const double A = f->ra[IDX(n, m)] * zt;
      const double B = f->rb[IDX(n, m)] * rho;
      const double v = fma(A, v1, -(B * v2));
      const double w = fma(A, w1, -(B * w2));
      V[IDX(n, m)] = v;
      W[IDX(n, m)] = w;
      v2 = v1; v1 = v;
      w2 = w1; w1 = w;
    }
  }
  /* columns beyond mtop are never read (terms need m' + 1 <= M + 1) */
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (f->form == ORC_GRAVITY_NORMALISED) {
    for (int n = 0; n <= N; ++n) {
      const int mm = (n < M) ? n : M;
      const int r = n + 1;
      for (int m = 0; m <= mm; ++m) {
        const double C = f->cbar[IDX(n, m)], S = f->sbar[IDX(n, m)];
        if (C == 0.0 && S == 0.0) continue;
        const double f1 = f->f1[IDX(n, m)], f3 = f->f3[IDX(n, m)];
        if (m == 0) {
          ax = fma(-C * f1, V[IDX(r, 1)], ax);
          ay = fma(-C * f1, W[IDX(r, 1)], ay);
          az = fma(-C * f3, V[IDX(r, 0)], az);
        } else {
          const double f2 = f->f2[IDX(n, m)];
          const double vp = V[IDX(r, m + 1)], wp = W[IDX(r, m + 1)];
          const double vm = V[IDX(r, m - 1)], wm = W[IDX(r, m - 1)];
          const double v0 = V[IDX(r, m)], w0 = W[IDX(r, m)];
          ax += f1 * (-fma(C, vp, S * wp)) + f2 * fma(C, vm, S * wm);
          ay += f1 * fma(S, vp, -(C * wp)) + f2 * fma(S, vm, -(C * wm));
          az += f3 * (-fma(C, v0, S * w0));
        }
      }
    }
  } else {
    /* un-normalised M-G sums with cf2 = (d+2)(d+1), dm1 = d+1, d = n-m
     * (SC/src/gravity.rs:1719-1856 packed terms, :2075 raw form) */
    for (int n = 0; n <= N; ++n) {
      const int mm = (n < M) ? n : M;
      const int r = n + 1;
      for (int m = 0; m <= mm; ++m) {
        const double C = f->cu[IDX(n, m)], S = f->su[IDX(n, m)];
        if (C == 0.0 && S == 0.0) continue;
        const double d = (double)(n - m);
        if (m == 0) {
          ax = fma(-C, V[IDX(r, 1)], ax);
          ay = fma(-C, W[IDX(r, 1)], ay);
          az = fma((d + 1.0) * (-C), V[IDX(r, 0)], az);
        } else {
          const double cf2 = (d + 2.0) * (d + 1.0);
          const double vp = V[IDX(r, m + 1)], wp = W[IDX(r, m + 1)];
          const double vm = V[IDX(r, m - 1)], wm = W[IDX(r, m - 1)];
          const double v0 = V[IDX(r, m)], w0 = W[IDX(r, m)];
          ax += 0.5 * (-fma(C, vp, S * wp)) + 0.5 * cf2 * fma(C, vm, S * wm);
          ay += 0.5 * fma(S, vp, -(C * wp)) + 0.5 * cf2 * fma(S, vm, -(C * wm));
          az += (d + 1.0) * (-fma(C, v0, S * w0));
        }
      }
    }
  }
  acc[0] = ax * f->gm_r2;
  acc[1] = ay * f->gm_r2;
  acc[2] = az * f->gm_r2;
  return ORC_OK;
}

int orc_gravity_bf_batch(const orc_field* f, const double* pos, int64_t n, double* acc,
                         int threads) {
  int bad = 0;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads) reduction(| : bad)
#endif
  for (int64_t i = 0; i < n; ++i) bad |= (orc_gravity_bf(f, pos + 3 * i, acc + 3 * i) != 0);
  (void)threads;
  return bad ? ORC_INVALID_INPUT : ORC_OK;
}

/* ------------------------------------------------------------- atmosphere */

/* Vallado (2013) Table 8-4: base altitude h0 (km), nominal density rho0
 * (kg/m^3) and scale height H (km) for each band [h0, next h0). */
static const double vallado_h0[28] = {0, 25, 30, 40, 50, 60, 70, 80, 90, 100, 110, 120, 130, 140,
                                      150, 180, 200, 250, 300, 350, 400, 450, 500, 600, 700, 800,
                                      900, 1000};
static const double vallado_rho0[28] = {
    1.225, 3.899e-2, 1.774e-2, 3.972e-3, 1.057e-3, 3.206e-4, 8.770e-5, 1.905e-5, 3.396e-6,
    5.297e-7, 9.661e-8, 2.438e-8, 8.484e-9, 3.845e-9, 2.070e-9, 5.464e-10, 2.789e-10,
    7.248e-11, 2.418e-11, 9.518e-12, 3.725e-12, 1.585e-12, 6.967e-13, 1.454e-13, 3.614e-14,
    1.170e-14, 5.245e-15, 3.019e-15};
static const double vallado_H[28] = {7.249, 6.349, 6.682, 7.554, 8.382, 7.714, 6.549, 5.799,
                                     5.382, 5.877, 7.263, 9.473, 12.636, 16.149, 22.523, 29.740,
                                     37.105, 45.546, 53.628, 53.298, 58.515, 60.828, 63.822,
                                     71.835, 88.667, 124.64, 181.05, 268.00};

double orc_density(int model, double h_m) {
  if (model == ORC_DRAG_VALLADO) {
    const double hk = h_m * 1e-3;
    if (hk < 0.0) return vallado_rho0[0];
    int b = 0;
    for (int i = 1; i < 28; ++i)
      if (hk >= vallado_h0[i]) b = i;
    return vallado_rho0[b] * exp(-(hk - vallado_h0[b]) / vallado_H[b]);
  }
  if (model == ORC_DRAG_COMPAT) {
    /* LY/src/rhs.rs:1728-1788, AtmModel::Exponential */
    const double hk = h_m * 1e-3;
    if (hk <= 0.0) return 1.225;
    if (hk >= 1000.0) return 0.0;
    return 1.225 * exp(-hk / 8.5);
  }
  return 0.0;
}

/* ------------------------------------------------------------------- RHS */

/* Cowell RHS in the inertial frame: rotate to the body frame by a uniform
 * z-rotation, evaluate gravity, rotate back, add drag (LY/src/rhs.rs:1119-1163
 * compute_drag; a = -1/2 * (Cd A/m) * rho * |v_rel| * v_rel, v_rel = v - w x r). */
void orc_rhs(const orc_model* md, double bc, double t, const double y[6], double dy[6]) {
  const double th = fma(md->omega, t, md->theta0);
  const double c = cos(th), s = sin(th);
  double pb[3], ab[3];
  pb[0] = fma(c, y[0], s * y[1]);
  pb[1] = fma(c, y[1], -(s * y[0]));
  pb[2] = y[2];
  orc_gravity_bf(md->field, pb, ab);
  double ax = fma(c, ab[0], -(s * ab[1]));
  double ay = fma(s, ab[0], c * ab[1]);
  double az = ab[2];
  if (md->drag != ORC_DRAG_NONE && bc != 0.0) {
    const double r = sqrt(fma(y[0], y[0], fma(y[1], y[1], y[2] * y[2])));
    const double rho = orc_density(md->drag, r - md->r_atm);
    if (rho > 0.0) {
      const double vrx = fma(md->omega, y[1], y[3]);
      const double vry = fma(-md->omega, y[0], y[4]);
      const double vrz = y[5];
      const double vr = sqrt(fma(vrx, vrx, fma(vry, vry, vrz * vrz)));
      const double k = -0.5 * bc * rho * vr;
      ax = fma(k, vrx, ax);
      ay = fma(k, vry, ay);
      az = fma(k, vrz, az);
    }
  }
  dy[0] = y[3];
  dy[1] = y[4];
  dy[2] = y[5];
  dy[3] = ax;
  dy[4] = ay;
  dy[5] = az;
}

/* ------------------------------------------------------- Encke pieces */

static double orc_field_mu(const orc_field* f) { return f->mu; }

/* Two-body propagation by Lagrange coefficients in the eccentric-anomaly
 * change x (Battin 1999 §4.5; the reference propagates its baseline the same
 * way from equinoctial elements, SC/src/lib.rs:2191-2572, 3448):
 *   n dt = x + s1 (1 - cos x) - c1 sin x,  s1 = (r0.v0)/sqrt(mu a),  c1 = 1 - r0/a
 *   r = a + (r0 - a) cos x + s1 a sin x
 *   f = 1 - (a/r0)(1 - cos x),      g = dt - (x - sin x)/n
 *   fd = -sqrt(mu a) sin x/(r r0),  gd = 1 - (a/r)(1 - cos x)
 * solved by Halley's method (the reference's Halley-type longitude solve,
 * tolerance 1e-12, is restated at 1e-15 rad). */
int orc_two_body(double mu, const double s0[6], double dt, double out[6], double* xio) {
  const double r0 = sqrt(fma(s0[0], s0[0], fma(s0[1], s0[1], s0[2] * s0[2])));
  const double v2 = fma(s0[3], s0[3], fma(s0[4], s0[4], s0[5] * s0[5]));
  const double rv = fma(s0[0], s0[3], fma(s0[1], s0[4], s0[2] * s0[5]));
  const double ainv = 2.0 / r0 - v2 / mu;
  if (!(ainv > 0.0) || !isfinite(ainv)) return ORC_INVALID_INPUT;
  const double a = 1.0 / ainv;
  const double sqa = sqrt(a);
  const double n = sqrt(mu * ainv * ainv * ainv);
  const double s1 = rv / (sqrt(mu) * sqa);
  const double c1 = 1.0 - r0 * ainv;
  const double M = n * dt;
  double x = (xio && isfinite(*xio)) ? *xio : M;
  int ok = 0;
  for (int it = 0; it < 50; ++it) {
    const double sx = sin(x), cx = cos(x);
    const double F = x + s1 * (1.0 - cx) - c1 * sx - M;
    const double F1 = 1.0 + s1 * sx - c1 * cx;
    const double F2 = s1 * cx + c1 * sx;
    const double d = -2.0 * F * F1 / (2.0 * F1 * F1 - F * F2);
    x += d;
    if (fabs(d) <= 1e-15 * fmax(1.0, fabs(x))) {
      ok = 1;
      break;
    }
  }
  if (!ok || !isfinite(x)) return ORC_INVALID_INPUT;
  const double sx = sin(x), cx = cos(x);
  const double r = a + (r0 - a) * cx + s1 * a * sx;
  const double f = 1.0 - (a / r0) * (1.0 - cx);
  const double g = dt - (x - sx) / n;
  const double fd = -sqrt(mu * a) * sx / (r * r0);
  const double gd = 1.0 - (a / r) * (1.0 - cx);
  for (int i = 0; i < 3; ++i) {
    out[i] = fma(f, s0[i], g * s0[3 + i]);
    out[3 + i] = fma(fd, s0[i], gd * s0[3 + i]);
  }
  if (xio) *xio = x;
  return ORC_OK;
}

/* Battin's Encke correction (LY/src/rhs.rs:1071-1104, Battin 1999 §9.3):
 * mu (rb/|rb|^3 - r/|r|^3) = mu/|r|^3 (f(q) rb - dr),
 * q = dr.(dr + 2 rb)/|rb|^2,  f(q) = q (3 + 3q + q^2)/(1 + (1+q)^{3/2}). */
void orc_battin(double mu, const double rb[3], const double dr[3], double acc[3]) {
  const double rb2 = fma(rb[0], rb[0], fma(rb[1], rb[1], rb[2] * rb[2]));
  double r[3];
  for (int i = 0; i < 3; ++i) r[i] = rb[i] + dr[i];
  const double r2 = fma(r[0], r[0], fma(r[1], r[1], r[2] * r[2]));
  if (!(rb2 > 0.0 && r2 > 0.0)) return;
  const double drb = fma(dr[0], rb[0], fma(dr[1], rb[1], dr[2] * rb[2]));
  const double dd = fma(dr[0], dr[0], fma(dr[1], dr[1], dr[2] * dr[2]));
  const double q = (dd + 2.0 * drb) / rb2;
  const double q1 = 1.0 + q;
  const double fq = q * (3.0 + q * (3.0 + q)) / (1.0 + q1 * sqrt(q1));
  const double k = mu / (r2 * sqrt(r2));
  for (int i = 0; i < 3; ++i) acc[i] += k * fma(fq, rb[i], -dr[i]);
}

/* Right-hand side of either formulation.  Encke: the state is the deviation
 * d from the two-body baseline started at base (time t0); the total state is
 * baseline + d; the acceleration is the full field plus mu r/|r|^3 (the
 * baseline carries the central term), drag on the total state, plus Battin's
 * correction (LY/src/rhs.rs:3885-4118). */
typedef struct rhs_ctx {
  const orc_model* md;
  double bc;
  double base[6];
  double t0;
  double x;  /* warm start of the baseline's Kepler solve */
  int bad;   /* baseline failed */
} rhs_ctx;

/* Per-attempt trace of the calling thread's integrations (test hook): one
 * record of ORC_TRACE_FIELDS doubles per step attempt. */
static __thread double* tls_trace = NULL;
static __thread int64_t tls_trace_cap = 0, tls_trace_n = 0;

void orc_trace_begin(double* buf, int64_t cap_records) {
  tls_trace = buf;
  tls_trace_cap = cap_records;
  tls_trace_n = 0;
}

int64_t orc_trace_end(void) {
  const int64_t n = tls_trace_n;
  tls_trace = NULL;
  tls_trace_cap = tls_trace_n = 0;
  return n;
}

static void trace_attempt(double t, double hs, double en, double thresh, int accepted, double h_next) {
  if (!tls_trace) return;
  if (tls_trace_n < tls_trace_cap) {
    double* r = tls_trace + ORC_TRACE_FIELDS * tls_trace_n;
    r[0] = t;
    r[1] = hs;
    r[2] = en;
    r[3] = thresh;
    r[4] = accepted ? 1.0 : 0.0;
    r[5] = h_next;
  }
  ++tls_trace_n;
}

static void eval_rhs(rhs_ctx* c, double t, const double y[6], double dy[6]) {
  if (c->md->formulation == ORC_TEST_EXP) { /* y' = y, solution y0 exp(t - t0) */
    (void)t;
    for (int i = 0; i < 6; ++i) dy[i] = y[i];
    return;
  }
  if (c->md->formulation != ORC_ENCKE) {
    orc_rhs(c->md, c->bc, t, y, dy);
    return;
  }
  double b[6], tot[6];
  if (orc_two_body(orc_field_mu(c->md->field), c->base, t - c->t0, b, &c->x) != ORC_OK) {
    c->bad = 1;
    for (int i = 0; i < 6; ++i) dy[i] = NAN;
    return;
  }
  for (int i = 0; i < 6; ++i) tot[i] = b[i] + y[i];
  double a[6];
  orc_rhs(c->md, c->bc, t, tot, a);
  const double mu = orc_field_mu(c->md->field);
  const double r2 = fma(tot[0], tot[0], fma(tot[1], tot[1], tot[2] * tot[2]));
  const double k = mu / (r2 * sqrt(r2));
  double acc[3];
  for (int i = 0; i < 3; ++i) acc[i] = fma(k, tot[i], a[3 + i]);
  orc_battin(mu, b, y, acc);
  dy[0] = y[3];
  dy[1] = y[4];
  dy[2] = y[5];
  dy[3] = acc[0];
  dy[4] = acc[1];
  dy[5] = acc[2];
}

/* Total (inertial) state of the context at t. */
static int total_state(rhs_ctx* c, double t, const double y[6], double out[6]) {
  if (c->md->formulation != ORC_ENCKE) {
    memcpy(out, y, 6 * sizeof(double));
    return ORC_OK;
  }
  double b[6];
  const int rc = orc_two_body(orc_field_mu(c->md->field), c->base, t - c->t0, b, &c->x);
  for (int i = 0; i < 6; ++i) out[i] = b[i] + y[i];
  return rc;
}

/* -------------------------------------------------------------- tableaus */

static const bqp_tableau T_DOPRI5 = BQP_DOPRI5_TABLEAU;
static const bqp_tableau T_DOP853 = BQP_DOP853_TABLEAU;
static const bqp_tableau T_RK87 = BQP_RK87_TABLEAU;

static const bqp_tableau* tableau_of(int method) {
  switch (method) {
    case BQP_METHOD_DOPRI5: return &T_DOPRI5;
    case BQP_METHOD_DOP853: return &T_DOP853;
    case BQP_METHOD_RK87: return &T_RK87;
    default: return NULL;
  }
}

int orc_tableau(int method, bqp_tableau* out) {
  const bqp_tableau* t = tableau_of(method);
  if (!t) return ORC_INVALID_INPUT;
  memcpy(out, t, sizeof(*t));
  return ORC_OK;
}

/* ------------------------------------------------------------ integrator */

static int finite6(const double* v) {
  for (int i = 0; i < 6; ++i)
    if (!isfinite(v[i])) return 0;
  return 1;
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* One explicit RK step (LY/src/odesolve/solver.rs:1965-2173 rk_step):
 * stage sums by fma-axpy with h*a_ij (zero coefficients skipped), solution by
 * fma-axpy with h*b_j, error sums with Kahan compensation. */
static void rk_step(rhs_ctx* rc, const bqp_tableau* T, double t,
                    const double y[6], double h, int reuse_k1, double k[BQP_MAX_STAGES][6],
                    double yn[6], double err[6], double err3[6], int64_t* evals) {
  double yt[6];
  if (!reuse_k1) {
    eval_rhs(rc, t, y, k[0]);
    ++*evals;
  }
  for (int s = 1; s < T->stages; ++s) {
    memcpy(yt, y, sizeof(yt));
    for (int j = 0; j < s; ++j) {
      const double sc = h * T->a[s][j];
      if (sc == 0.0) continue;
      for (int i = 0; i < 6; ++i) yt[i] = fma(k[j][i], sc, yt[i]);
    }
    eval_rhs(rc, t + T->c[s] * h, yt, k[s]);
    ++*evals;
  }
  double comp[6] = {0}, comp3[6] = {0};
  memcpy(yn, y, 6 * sizeof(double));
  memset(err, 0, 6 * sizeof(double));
  memset(err3, 0, 6 * sizeof(double));
  for (int j = 0; j < T->stages; ++j) {
    const double hb = h * T->b[j], he = h * T->e[j];
    if (hb != 0.0)
      for (int i = 0; i < 6; ++i) yn[i] = fma(k[j][i], hb, yn[i]);
    if (he != 0.0)
      for (int i = 0; i < 6; ++i) {
        const double term = he * k[j][i];
        const double yk = term - comp[i];
        const double acc = err[i] + yk;
        comp[i] = (acc - err[i]) - yk;
        err[i] = acc;
      }
    if (T->has_e3 && T->e3[j] != 0.0) {
      const double he3 = h * T->e3[j];
      for (int i = 0; i < 6; ++i) {
        const double term = he3 * k[j][i];
        const double yk = term - comp3[i];
        const double acc = err3[i] + yk;
        comp3[i] = (acc - err3[i]) - yk;
        err3[i] = acc;
      }
    }
  }
}

static double radius_of(const double* y) { return sqrt(fma(y[0], y[0], fma(y[1], y[1], y[2] * y[2]))); }

/* The minimum-radius test (r_min > 0).  With a flattening f > 0 it is the
 * reference's ground event, LY/src/events.rs:87-145 check_ground: the surface
 * radius of the ellipsoid (equatorial r_atm, polar r_atm (1 - f)) along the
 * position's direction is R = a b |r| / sqrt(a^2 z^2 + b^2 (x^2 + y^2)), and
 * the row fails once |r| - R drops below r_min - r_atm (the reference's
 * GROUND_ALTITUDE, 100 km). */
static int above_ground(const orc_model* md, const double* y) {
  const double d = radius_of(y);
  if (md->ground_flattening == 0.0) return d >= md->r_min;
  const double a = md->r_atm, b = md->r_atm * (1.0 - md->ground_flattening);
  const double den = sqrt(fma(a * a, y[2] * y[2], (b * b) * fma(y[0], y[0], y[1] * y[1])));
  const double surface = (den > 0.0) ? a * b * d / den : b;
  return d - surface >= md->r_min - md->r_atm;
}

/* Sample times an integration lands on without restarting its controller
 * (forced steps onto sample times, LY/src/odesolve/solver.rs:1356-1368):
 * each step is shortened onto the next time in (ta, tb], the inertial state
 * there is written to out[k], and the time sum's compensation restarts. */
typedef struct orc_stops {
  const double* t;  /* increasing sample times */
  int64_t m;
  int64_t next;     /* first sample not yet written */
  double* out;      /* m x 6 */
} orc_stops;

/* One adaptive integration of y over [ta, tb] with a fresh controller:
 * LY/src/odesolve/solver.rs:1099-1925 integrate_internal with the PI
 * controller macros :445-529 (ORC_CONTROL_SCALED / _ABSOLUTE), or
 * LY/src/odesolve/lightyear_compat.rs:100-600 (ORC_CONTROL_COMPAT).  The step
 * budget is per call, as the reference's per-segment solver call.
 * h_open > 0 replaces the solver's default opening step (clamped to
 * [h_min, h_max]; the solver path's `h0`, LY/src/integrator.rs:490).  On
 * return *h_exit holds the step size the controller asked for after its last
 * accepted step (solver.rs:497-498 final_controller_h), 0 if none. */
static int integrate_segment(rhs_ctx* rc, const orc_integ* cfg, const bqp_tableau* T, double y[6],
                             double ta, double tb, orc_stats* loc, double h_open, double* h_exit,
                             orc_stops* stops) {
  const double span = tb - ta;
  if (h_exit) *h_exit = 0.0;
  if (span == 0.0) return ORC_OK;
  int64_t seg_steps = 0, seg_rejects = 0;
  const int compat = (cfg->control == ORC_CONTROL_COMPAT);
  const double q = compat ? 5.0 : (double)T->order_err + 1.0;
  const double inv_q = 1.0 / q;
  double h_min, h_max, h, eps = cfg->eps;
  if (compat) {
    eps = fmax(eps, 1e-12);
    h_min = 1e-10;
    h_max = (cfg->h_max > 0.0 && isfinite(cfg->h_max)) ? fmax(cfg->h_max, h_min) : span;
    h = clampd(fabs(span / 100.0), 1e-3, fmax(h_max, 1e-3));
  } else {
    h_max = (cfg->h_max > 0.0 && isfinite(cfg->h_max)) ? cfg->h_max : span;
    h_min = (cfg->h_min > 0.0) ? cfg->h_min : 1e-12;
    if (h_max < h_min) h_max = h_min;
    const double h0 = (h_open > 0.0) ? h_open
                      : (cfg->h0 > 0.0) ? cfg->h0 : ((span <= 60.0) ? span / 2.0 : span / 100.0);
    h = clampd(h0, h_min, h_max);
  }
  const int64_t max_steps = cfg->max_steps > 0 ? cfg->max_steps : 500000;
  const int max_rejects = cfg->max_rejects > 0 ? cfg->max_rejects : 50;
  double k[BQP_MAX_STAGES][6], yn[6], err[6], err3[6], k1c[6];
  int have_k1 = 0;
  if (compat) { /* the compat loop evaluates dy once up front and reuses it */
    eval_rhs(rc, ta, y, k1c);
    loc->evals++;
    have_k1 = 1;
  }
  double t = ta, t_comp = 0.0, err_prev = 0.0;
  int have_prev = 0, just_rej = 0, rejects = 0;
  const int blend3 = T->has_e3 && (cfg->control == ORC_CONTROL_SCALED ||
                                   (cfg->control == ORC_CONTROL_ABSOLUTE && eps >= 1e-6));
  while (t != tb && tb - t > 0.0) {
    if ((compat ? seg_steps + seg_rejects : seg_steps) >= max_steps) {
      loc->status = ORC_MAX_STEPS;
      return loc->status;
    }
    const double tgt = (stops && stops->next < stops->m && stops->t[stops->next] < tb) ? stops->t[stops->next] : tb;
    const double t_att = t;
    const double rem = tgt - t;
    double hs = (h < rem) ? h : rem;
    const int lands = (hs == rem);
    if (!compat && hs < h_min && !lands) {
      loc->status = ORC_STEP_UNDERFLOW;
      return loc->status;
    }
    if (have_k1) memcpy(k[0], k1c, sizeof(k1c));
    rk_step(rc, T, t, y, hs, have_k1, k, yn, err, err3, &loc->evals);
    if (!finite6(yn) || !finite6(err)) {
      ++rejects;
      ++loc->rejects;
      ++seg_rejects;
      if (!compat) have_k1 = 0;
      if (rejects > max_rejects) {
        loc->status = ORC_NON_FINITE;
        return loc->status;
      }
      h = clampd(hs * 0.1, h_min, h_max);
      continue;
    }
    double en, thresh;
    if (cfg->control == ORC_CONTROL_SCALED) {
      double s5 = 0.0, s3 = 0.0;
      for (int i = 0; i < 6; ++i) {
        const double atol = (i < 3) ? cfg->atol_pos : cfg->atol_vel;
        const double sc = atol + cfg->rtol * fmax(fabs(y[i]), fabs(yn[i]));
        const double den = (sc > 0.0) ? sc : atol;
        const double e5 = err[i] / den;
        s5 = fma(e5, e5, s5);
        if (blend3) {
          const double e3 = err3[i] / den;
          s3 = fma(e3, e3, s3);
        }
      }
      const double rms5 = sqrt(s5 / 6.0);
      if (blend3) {
        const double rms3 = sqrt(s3 / 6.0);
        en = sqrt(fma(rms5, rms5, 0.01 * rms3 * rms3));
      } else {
        en = rms5;
      }
      thresh = 1.0;
    } else {
      double m5 = 0.0, m3 = 0.0;
      for (int i = 0; i < 6; ++i) {
        m5 = fmax(m5, fabs(err[i]));
        if (blend3 && !compat) m3 = fmax(m3, fabs(err3[i]));
      }
      en = (blend3 && !compat) ? sqrt(fma(m5, m5, 0.01 * m3 * m3)) : m5;
      thresh = eps;
    }
    if (!isfinite(en)) {
      ++rejects;
      ++loc->rejects;
      ++seg_rejects;
      if (!compat) have_k1 = 0;
      if (rejects > max_rejects) {
        loc->status = ORC_NON_FINITE;
        return loc->status;
      }
      h = clampd(hs * 0.1, h_min, h_max);
      continue;
    }
    if (en <= thresh || hs <= h_min) {
      double tn;
      if (lands) {
        tn = tgt;
      } else {
        const double ky = hs - t_comp;
        tn = t + ky;
        t_comp = (tn - t) - ky;
      }
      ++loc->steps;
      ++seg_steps;
      rejects = 0;
      if (T->fsal) {
        memcpy(k1c, k[T->stages - 1], sizeof(k1c));
        have_k1 = 1;
      } else {
        have_k1 = 0;
      }
      t = tn;
      memcpy(y, yn, 6 * sizeof(double));
      if (rc->md->r_min > 0.0) {
        double tot[6];
        if (total_state(rc, t, y, tot) != ORC_OK || !above_ground(rc->md, tot)) {
          loc->status = ORC_BELOW_MIN_RADIUS;
          return loc->status;
        }
      }
      const double eff = (en > 0.0) ? en : thresh * 0.1;
      double fac, maxg;
      if (compat) {
        fac = (have_prev && err_prev > 0.0)
                  ? 0.9 * pow(eps / eff, 0.7 / 5.0) * pow(err_prev / eps, 0.4 / 5.0)
                  : 0.9 * pow(eps / eff, 0.2);
        maxg = just_rej ? 2.0 : 5.0;
      } else {
        if (cfg->control == ORC_CONTROL_SCALED)
          fac = (have_prev && err_prev > 0.0)
                    ? 0.9 * pow(1.0 / eff, 0.7 / q) * pow(err_prev, 0.4 / q)
                    : 0.9 * pow(eff, -inv_q);
        else
          fac = (have_prev && err_prev > 0.0)
                    ? 0.9 * pow(eps / eff, 0.7 / q) * pow(err_prev / eps, 0.4 / q)
                    : 0.9 * pow(eps / eff, inv_q);
        const double base = fmin(1.0 + 4.0 * (5.0 / q), 5.0);
        maxg = just_rej ? fmin(base, 2.0) : base;
      }
      fac = clampd(fac, 0.2, maxg);
      h = clampd(hs * fac, h_min, h_max);
      trace_attempt(t_att, hs, en, thresh, 1, h);
      if (h_exit) *h_exit = h;
      err_prev = en;
      have_prev = 1;
      just_rej = 0;
      if (stops && stops->next < stops->m && t == stops->t[stops->next]) {
        double tot[6];
        if (total_state(rc, t, y, tot) != ORC_OK) {
          loc->status = ORC_NON_FINITE;
          return loc->status;
        }
        memcpy(stops->out + 6 * stops->next, tot, sizeof(tot));
        ++stops->next;
        t_comp = 0.0;
      }
    } else {
      ++rejects;
      ++loc->rejects;
      ++seg_rejects;
      if (!compat) have_k1 = 0;
      if (rejects > max_rejects) {
        loc->status = ORC_MAX_REJECTS;
        return loc->status;
      }
      double fac;
      if (cfg->control == ORC_CONTROL_SCALED)
        fac = 0.9 * pow(en, -inv_q);
      else
        fac = 0.9 * pow(eps / en, compat ? 0.2 : inv_q);
      fac = clampd(fac, 0.1, 0.5);
      h = clampd(hs * fac, h_min, h_max);
      if (compat) {
        err_prev = en;
        have_prev = 1;
      }
      just_rej = 1;
      trace_attempt(t_att, hs, en, thresh, 0, h);
    }
  }
  (void)seg_rejects;
  return ORC_OK;
}

/* Propagation of one row.  Cowell: one segment.  Encke (ORC_ENCKE): the
 * segment driver of LY/src/integrator.rs:3418-3563 -- segments of at most
 * ORC_ENCKE_SEGMENT_S (:236 MAX_RECTIFICATION_SEGMENT_S, bounded_segment_end
 * :953-961), each integrating the deviation from 0 on a two-body baseline
 * started from the current total state, then rebased.  (The reference's
 * 10 km deviation test only decides how the last segment's result is
 * expressed, :941-951, so the total state is the same.) */
static int propagate_impl(const orc_model* md, const orc_integ* cfg, double bc, const double y0[6],
                          double t0, double tf, double yf[6], orc_stats* st, orc_stops* stops) {
  orc_stats loc = {0, 0, 0, ORC_OK};
  const bqp_tableau* T = tableau_of(cfg->method);
  double y[6];
  memcpy(y, y0, sizeof(y));
  if (!T || !finite6(y0) || !isfinite(t0) || !isfinite(tf) || !(tf >= t0)) {
    loc.status = ORC_INVALID_INPUT;
    goto fail;
  }
  if (md->r_min > 0.0 && !above_ground(md, y)) {
    loc.status = ORC_BELOW_MIN_RADIUS;
    goto fail;
  }
  if (stops) /* samples at the start time are the initial state */
    while (stops->next < stops->m && stops->t[stops->next] == t0) {
      memcpy(stops->out + 6 * stops->next, y, sizeof(y));
      ++stops->next;
    }
  rhs_ctx rc;
  memset(&rc, 0, sizeof(rc));
  rc.md = md;
  rc.bc = bc;
  if (tf == t0) {
    /* nothing to integrate */
  } else if (md->formulation != ORC_ENCKE) {
    if (integrate_segment(&rc, cfg, T, y, t0, tf, &loc, 0.0, NULL, stops) != ORC_OK) goto fail;
  } else {
    /* Step-size carry (LY/src/integrator.rs:341-402 hcarry_take/_store): a
     * rebased segment opens with the step size the controller asked for at
     * the end of the previous segment of the same row, on the adaptive
     * solver path (not the DOPRI5 compat loop, :420-486) and only when the
     * step cap is above 60 s (HCARRY_UNCLAMPED_DT_MAX_S, :370; h_max <= 0
     * here means no cap).  The carry starts empty for every row. */
    const int carry_ok = !cfg->no_hcarry && cfg->control != ORC_CONTROL_COMPAT &&
                         !(cfg->h_max > 0.0 && cfg->h_max <= 60.0);
    double carried = 0.0;
    double ts = t0;
    while (1) {
      const double te = (tf - ts > ORC_ENCKE_SEGMENT_S) ? ts + ORC_ENCKE_SEGMENT_S : tf;
      memcpy(rc.base, y, sizeof(y));
      rc.t0 = ts;
      rc.x = NAN;
      double d[6] = {0, 0, 0, 0, 0, 0};
      double probe[6];
      if (orc_two_body(md->field->mu, rc.base, 0.0, probe, NULL) != ORC_OK) {
        loc.status = ORC_INVALID_INPUT;  /* baseline is not an ellipse */
        goto fail;
      }
      double h_exit = 0.0;
      if (integrate_segment(&rc, cfg, T, d, ts, te, &loc, carry_ok ? carried : 0.0, &h_exit, stops) != ORC_OK)
        goto fail;
      carried = h_exit;
      if (total_state(&rc, te, d, y) != ORC_OK) {
        loc.status = ORC_NON_FINITE;
        goto fail;
      }
      ts = te;
      if (te == tf) break;
    }
  }
  memcpy(yf, y, sizeof(y));
  if (st) *st = loc;
  return ORC_OK;
fail:
  for (int i = 0; i < 6; ++i) yf[i] = NAN;
  if (stops) /* a failed row's unreached samples are NaN */
    for (int64_t k = stops->next; k < stops->m; ++k)
      for (int i = 0; i < 6; ++i) stops->out[6 * k + i] = NAN;
  if (st) *st = loc;
  return loc.status;
}

int orc_propagate(const orc_model* md, const orc_integ* cfg, double bc, const double y0[6],
                  double t0, double tf, double yf[6], orc_stats* st) {
  return propagate_impl(md, cfg, bc, y0, t0, tf, yf, st, NULL);
}

/* Sampled propagation (LY/src/batch.rs:247-390 integrate_batch_native_into
 * with t_eval; DT/src/lib.rs:237-353 propagate(times_s=...)): one integration
 * from t0 to the last sample time whose steps are shortened onto every
 * sample time, with the controller carried through the samples (Cowell: one
 * controller for the whole row; Encke: one per 5,400 s segment, samples
 * inside a segment do not restart it).  out holds m x 6 inertial states. */
int orc_propagate_sampled(const orc_model* md, const orc_integ* cfg, double bc, const double y0[6], double t0,
                          const double* t_eval, int64_t m, double* out, orc_stats* st) {
  double yf[6];
  orc_stops stops = {t_eval, m, 0, out};
  int bad = (m <= 0 || !isfinite(t0));
  for (int64_t k = 0; k < m && !bad; ++k)
    if (!isfinite(t_eval[k]) || t_eval[k] < t0 || (k > 0 && !(t_eval[k] > t_eval[k - 1]))) bad = 1;
  if (bad) {
    for (int64_t k = 0; k < m * 6; ++k) out[k] = NAN;
    if (st) {
      orc_stats z = {0, 0, 0, ORC_INVALID_INPUT};
      *st = z;
    }
    return ORC_INVALID_INPUT;
  }
  return propagate_impl(md, cfg, bc, y0, t0, t_eval[m - 1], yf, st, &stops);
}

int orc_propagate_sampled_batch(const orc_model* md, const orc_integ* cfg, const double* y0, double t0,
                                const double* t_eval, int64_t m, const double* bc, int64_t n, double* out,
                                int32_t* status, int64_t* steps, int64_t* evals, int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    orc_stats st;
    orc_propagate_sampled(md, cfg, bc ? bc[i] : 0.0, y0 + 6 * i, t0, t_eval, m, out + 6 * m * i, &st);
    if (status) status[i] = st.status;
    if (steps) steps[i] = st.steps;
    if (evals) evals[i] = st.evals;
  }
  (void)threads;
  for (int64_t i = 0; i < n && status; ++i)
    if (status[i] != ORC_OK) return status[i];
  return ORC_OK;
}

/* Ragged row-parallel batch (LY/src/session.rs:2325-2436): rows are
 * independent, results keyed by index, failed rows NaN-filled.  Rows are
 * handed to threads one at a time (a row is milliseconds to seconds of
 * work, so small samples still keep every thread busy). */
int orc_propagate_batch(const orc_model* md, const orc_integ* cfg, const double* y0,
                        const double* t0, const double* tf, const double* bc, int64_t n,
                        double* yf, int32_t* status, int64_t* steps, int64_t* evals,
                        int threads) {
  int first_bad = 0;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int64_t i = 0; i < n; ++i) {
    orc_stats st;
    orc_propagate(md, cfg, bc ? bc[i] : 0.0, y0 + 6 * i, t0[i], tf[i], yf + 6 * i, &st);
    if (status) status[i] = st.status;
    if (steps) steps[i] = st.steps;
    if (evals) evals[i] = st.evals;
  }
  (void)threads;
  for (int64_t i = 0; i < n && status; ++i)
    if (status[i] != ORC_OK) {
      first_bad = status[i];
      break;
    }
  return first_bad;
}

/* ---------------------------------------------------------------- elements */

/* Keplerian elements [a (m), e, i, raan, argp, mean anomaly] (rad) to an
 * inertial state; Kepler's equation by Newton iteration. */
void orc_kep2eci(double mu, const double kep[6], double eci[6]) {
  const double a = kep[0], e = kep[1], inc = kep[2], raan = kep[3], argp = kep[4];
  const double M = kep[5];
  double E = (e < 0.8) ? M : M_PI;
  for (int it = 0; it < 60; ++it) {
    const double d = (E - e * sin(E) - M) / (1.0 - e * cos(E));
    E -= d;
    if (fabs(d) < 1e-15) break;
  }
  const double cE = cos(E), sE = sin(E), sq = sqrt(1.0 - e * e);
  const double xp = a * (cE - e), yp = a * sq * sE;
  const double nmean = sqrt(mu / (a * a * a));
  const double den = 1.0 - e * cE;
  const double vxp = -a * nmean * sE / den, vyp = a * nmean * sq * cE / den;
  const double cO = cos(raan), sO = sin(raan), ci = cos(inc), si = sin(inc), cw = cos(argp),
               sw = sin(argp);
  const double p1 = cO * cw - sO * sw * ci, p2 = sO * cw + cO * sw * ci, p3 = sw * si;
  const double q1 = -cO * sw - sO * cw * ci, q2 = -sO * sw + cO * cw * ci, q3 = cw * si;
  eci[0] = p1 * xp + q1 * yp;
  eci[1] = p2 * xp + q2 * yp;
  eci[2] = p3 * xp + q3 * yp;
  eci[3] = p1 * vxp + q1 * vyp;
  eci[4] = p2 * vxp + q2 * vyp;
  eci[5] = p3 * vxp + q3 * vyp;
}
