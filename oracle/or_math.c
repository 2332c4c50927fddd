/*
 * or_math.c -- oracle (test infrastructure only): Boys function, Rys roots and
 * weights, Cartesian normalisation and the Cartesian->spherical matrices.
 * CPU restatement of crates/integral/src/math/{boys.rs,rys.rs,norm.rs,
 * solid_harmonics.rs} and of the table recipe in xtask/src/gen_rys_tables.rs.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "ig_oracle.h"

/* ------------------------------------------------------------------------ */
/* Boys function                                                             */
/* ------------------------------------------------------------------------ */

/* boys.rs:63-65 */
static double asymptotic_boundary(int m) { return 33.0 + 2.5 * (double)m; }

#define BOYS_K 8
#define BOYS_STRIDE (OR_BOYS_MAX_M + BOYS_K + 1) /* boys.rs:85 */
#define BOYS_ROW (BOYS_STRIDE + 1)               /* boys.rs:91 */
#define BOYS_H 0.1                               /* boys.rs:95 */
#define BOYS_TMAX 93.0                           /* boys.rs:123 */
#define BOYS_NGRID 930                           /* boys.rs:126 */

/* non-alternating series for the top order, boys.rs:399-418 */
static double boys_series(int m, double t, double et) {
  double two_t = 2.0 * t;
  double term = 1.0 / (double)(2 * m + 1);
  double sum = term;
  for (int k = 1; k < 1000; ++k) {
    term *= two_t / (double)(2 * m + 2 * k + 1);
    sum += term;
    if (term <= sum * 1e-17) break;
  }
  return et * sum;
}

/* large-T closed form, boys.rs:423-430 */
static double boys_asymptotic(int m, double t) {
  double val = 0.5 * sqrt(M_PI / t);
  double two_t = 2.0 * t;
  for (int k = 1; k <= m; ++k) val *= (double)(2 * k - 1) / two_t;
  return val;
}

/* boys.rs:366-384 */
void or_boys_reference(int m_max, double t, double *out) {
  double et = exp(-t);
  out[m_max] = (t < asymptotic_boundary(m_max)) ? boys_series(m_max, t, et)
                                                  : boys_asymptotic(m_max, t);
  for (int m = m_max; m > 0; --m) out[m - 1] = (2.0 * t * out[m] + et) / (double)(2 * m - 1);
}

static double *g_boys_table = NULL;
static double g_inv_fact[BOYS_K + 1];
static double g_inv_odd[OR_BOYS_MAX_M + 1];
static pthread_once_t g_boys_once = PTHREAD_ONCE_INIT;

/* boys.rs:131-136 (table of F_0..F_{STRIDE-1} and e^{-T0} per grid point) */
static void build_boys_table(void) {
  g_boys_table = (double *)malloc(sizeof(double) * (BOYS_NGRID + 1) * BOYS_ROW);
  for (int i = 0; i <= BOYS_NGRID; ++i) {
    double t0 = (double)i * BOYS_H;
    double *row = g_boys_table + (size_t)i * BOYS_ROW;
    or_boys_reference(BOYS_STRIDE - 1, t0, row);
    row[BOYS_STRIDE] = exp(-t0);
  }
  g_inv_fact[0] = 1.0;
  for (int k = 1; k <= BOYS_K; ++k) g_inv_fact[k] = g_inv_fact[k - 1] / (double)k;
  g_inv_odd[0] = 1.0;
  for (int m = 1; m <= OR_BOYS_MAX_M; ++m) g_inv_odd[m] = 1.0 / (double)(2 * m - 1);
}

/* boys.rs:146-193: Horner on the nearest grid row for the top order, e^{-T}
 * from the tabulated e^{-T0} times a Horner exp, then downward recurrence. */
void or_boys(int m_max, double t, double *out) {
  pthread_once(&g_boys_once, build_boys_table);
  if (t < BOYS_TMAX && m_max + BOYS_K < BOYS_STRIDE) {
    long i = lround(t * (1.0 / BOYS_H));
    double neg_d = (double)i * BOYS_H - t;
    const double *row = g_boys_table + (size_t)i * BOYS_ROW;
    const double *base = row + m_max;
    double top = base[BOYS_K] * g_inv_fact[BOYS_K];
    for (int k = BOYS_K - 1; k >= 0; --k) top = top * neg_d + base[k] * g_inv_fact[k];
    out[m_max] = top;
    if (m_max > 0) {
      double ep = g_inv_fact[BOYS_K];
      for (int k = BOYS_K - 1; k >= 0; --k) ep = ep * neg_d + g_inv_fact[k];
      double et = row[BOYS_STRIDE] * ep;
      double two_t = 2.0 * t;
      for (int m = m_max; m > 0; --m) out[m - 1] = (two_t * out[m] + et) * g_inv_odd[m];
    }
  } else {
    or_boys_reference(m_max, t, out);
  }
}

/* ------------------------------------------------------------------------ */
/* Rys quadrature                                                            */
/* ------------------------------------------------------------------------ */

#define GL_POINTS 128 /* rys.rs:100 */

/* Implicit-shift QL on a symmetric tridiagonal matrix tracking only the first
 * eigenvector row (Golub-Welsch shortcut).  rys.rs:409-480 */
static void tridiagonal_ql(int n, double *d, double *e, double *z0) {
  if (n <= 1) return;
  const int max_iter = 60;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      while (m < n - 1) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
        ++m;
      }
      if (m == l) break;
      ++iter;
      if (iter > max_iter) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int underflow = 0;
      int i = m;
      while (i > l) {
        --i;
        double f = s * e[i];
        double b = c * e[i];
        double r2 = hypot(f, g);
        e[i + 1] = r2;
        if (r2 == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          underflow = 1;
          break;
        }
        s = f / r2;
        c = g / r2;
        double dd = d[i + 1] - p;
        double r3 = (d[i] - dd) * s + 2.0 * c * b;
        p = s * r3;
        d[i + 1] = dd + p;
        g = c * r3 - b;
        f = z0[i + 1];
        z0[i + 1] = s * z0[i] + c * f;
        z0[i] = c * z0[i] - s * f;
      }
      if (underflow) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
}

/* sort (node, weight) ascending by node; rys.rs:389-399 */
static void sort_pairs(int n, const double *d, const double *z0, double mu0, double *roots,
                       double *weights) {
  int idx[GL_POINTS];
  for (int i = 0; i < n; ++i) idx[i] = i;
  for (int i = 1; i < n; ++i) { /* insertion sort, stable */
    int v = idx[i];
    int j = i - 1;
    while (j >= 0 && d[idx[j]] > d[v]) {
      idx[j + 1] = idx[j];
      --j;
    }
    idx[j + 1] = v;
  }
  for (int o = 0; o < n; ++o) {
    roots[o] = d[idx[o]];
    weights[o] = mu0 * z0[idx[o]] * z0[idx[o]];
  }
}

/* rys.rs:358-400 */
static void golub_welsch(int n, const double *alpha, const double *beta, double *roots,
                         double *weights) {
  double mu0 = beta[0];
  if (n == 1) {
    roots[0] = alpha[0];
    weights[0] = mu0;
    return;
  }
  double d[OR_MAX_RYS], e[OR_MAX_RYS], z0[OR_MAX_RYS];
  for (int i = 0; i < n; ++i) {
    d[i] = alpha[i];
    z0[i] = 0.0;
  }
  z0[0] = 1.0;
  for (int i = 1; i < n; ++i) e[i - 1] = sqrt(beta[i] > 0.0 ? beta[i] : 0.0);
  e[n - 1] = 0.0;
  tridiagonal_ql(n, d, e, z0);
  sort_pairs(n, d, z0, mu0, roots, weights);
}

static double g_gl_nodes[GL_POINTS], g_gl_wts[GL_POINTS];
static double g_lag_nodes[OR_MAX_RYS][OR_MAX_RYS], g_lag_wts[OR_MAX_RYS][OR_MAX_RYS];

/* Gauss-Legendre on [0,1] via Golub-Welsch, rys.rs:326-352; generalized
 * Gauss-Laguerre(-1/2) nodes, rys.rs:541-569 */
static void build_quadratures(void) {
  double d[GL_POINTS], e[GL_POINTS], z0[GL_POINTS];
  for (int i = 0; i < GL_POINTS; ++i) {
    d[i] = 0.5;
    e[i] = 0.0;
    z0[i] = 0.0;
  }
  z0[0] = 1.0;
  for (int i = 1; i < GL_POINTS; ++i) {
    double l = (double)i;
    e[i - 1] = sqrt(0.25 * l * l / (4.0 * l * l - 1.0));
  }
  e[GL_POINTS - 1] = 0.0;
  tridiagonal_ql(GL_POINTS, d, e, z0);
  sort_pairs(GL_POINTS, d, z0, 1.0, g_gl_nodes, g_gl_wts);

  for (int m = 1; m <= OR_MAX_RYS; ++m) {
    double a[OR_MAX_RYS], b[OR_MAX_RYS];
    for (int k = 0; k < m; ++k) {
      double kf = (double)k;
      a[k] = 2.0 * kf - 0.5 + 1.0;
      b[k] = kf * (kf - 0.5);
    }
    b[0] = sqrt(M_PI);
    golub_welsch(m, a, b, g_lag_nodes[m - 1], g_lag_wts[m - 1]);
  }
}

/* discretized Stieltjes, rys.rs:283-322 */
static void stieltjes(int n, const double *x, const double *w, double *alpha, double *beta) {
  double p_prev[GL_POINTS], p_curr[GL_POINTS], p_next[GL_POINTS];
  double norm = 0.0, xnorm = 0.0;
  for (int m = 0; m < GL_POINTS; ++m) {
    p_prev[m] = 0.0;
    p_curr[m] = 1.0;
    double wp = w[m];
    norm += wp;
    xnorm += x[m] * wp;
  }
  beta[0] = norm;
  alpha[0] = xnorm / norm;
  for (int k = 1; k < n; ++k) {
    double a = alpha[k - 1], bb = beta[k - 1];
    double nn = 0.0, xn = 0.0;
    for (int m = 0; m < GL_POINTS; ++m) {
      p_next[m] = (x[m] - a) * p_curr[m] - bb * p_prev[m];
      double wp = w[m] * p_next[m] * p_next[m];
      nn += wp;
      xn += x[m] * wp;
    }
    alpha[k] = xn / nn;
    beta[k] = nn / norm;
    norm = nn;
    memcpy(p_prev, p_curr, sizeof(p_curr));
    memcpy(p_curr, p_next, sizeof(p_next));
  }
}

static pthread_once_t g_quad_once = PTHREAD_ONCE_INIT;

/* finite branch: rys.rs:265-281 */
static void finite_branch(int n, double t, double *roots, double *weights) {
  pthread_once(&g_quad_once, build_quadratures);
  double x[GL_POINTS], w[GL_POINTS];
  for (int m = 0; m < GL_POINTS; ++m) {
    double u = g_gl_nodes[m];
    x[m] = u * u;
    w[m] = g_gl_wts[m] * exp(-t * x[m]);
  }
  double alpha[OR_MAX_RYS], beta[OR_MAX_RYS];
  stieltjes(n, x, w, alpha, beta);
  golub_welsch(n, alpha, beta, roots, weights);
}

/* rys.rs:519-528 */
static void laguerre_branch(int n, double t, double *roots, double *weights) {
  pthread_once(&g_quad_once, build_quadratures);
  double inv_t = 1.0 / t, scale = 0.5 / sqrt(t);
  for (int i = 0; i < n; ++i) {
    roots[i] = g_lag_nodes[n - 1][i] * inv_t;
    weights[i] = g_lag_wts[n - 1][i] * scale;
  }
}

/* rys.rs:197-211 */
void or_rys_reference(int n, double t, double *roots, double *weights) {
  if (t >= asymptotic_boundary(2 * n - 1))
    laguerre_branch(n, t, roots, weights);
  else
    finite_branch(n, t, roots, weights);
}

/* ---- Chebyshev-in-T tables, recipe of xtask/src/gen_rys_tables.rs:1-190 ---- */
#define CHEB_NC 17
#define FIT_TOL 2e-14
#define CHECK_PER_CELL 33
#define MAX_SUB 256

static double *g_root_coef[OR_MAX_RYS]; /* [sub][i][NC] */
static double *g_wt_coef[OR_MAX_RYS];
static int g_nsub[OR_MAX_RYS];
static double g_worst[OR_MAX_RYS];
static pthread_once_t g_table_once = PTHREAD_ONCE_INIT;

/* half-c0 Clenshaw, rys.rs:259-269 */
static double chebev(const double *c, double xi) {
  double two_xi = 2.0 * xi, d = 0.0, dd = 0.0;
  for (int k = CHEB_NC - 1; k >= 1; --k) {
    double sv = d;
    d = two_xi * d - dd + c[k];
    dd = sv;
  }
  return xi * d - dd + 0.5 * c[0];
}

/* gen_rys_tables.rs:151-180: sample at the Chebyshev nodes, DCT */
static void fit_cell(int n, double a, double b, double *rc, double *wc) {
  double bma = 0.5 * (b - a), bpa = 0.5 * (b + a);
  double nr[CHEB_NC][OR_MAX_RYS], nw[CHEB_NC][OR_MAX_RYS];
  for (int j = 0; j < CHEB_NC; ++j) {
    double y = cos(M_PI * ((double)j + 0.5) / (double)CHEB_NC);
    finite_branch(n, y * bma + bpa, nr[j], nw[j]);
  }
  memset(rc, 0, sizeof(double) * n * CHEB_NC);
  memset(wc, 0, sizeof(double) * n * CHEB_NC);
  double fac = 2.0 / (double)CHEB_NC;
  for (int k = 0; k < CHEB_NC; ++k)
    for (int j = 0; j < CHEB_NC; ++j) {
      double ckj = cos(M_PI * (double)k * ((double)j + 0.5) / (double)CHEB_NC);
      for (int i = 0; i < n; ++i) {
        rc[i * CHEB_NC + k] += fac * nr[j][i] * ckj;
        wc[i * CHEB_NC + k] += fac * nw[j][i] * ckj;
      }
    }
}

/* gen_rys_tables.rs:118-150 */
static double max_error(int n, double t_hi, int n_sub, double *rc_all, double *wc_all) {
  double width = t_hi / (double)n_sub, worst = 0.0;
  double rr[OR_MAX_RYS], ww[OR_MAX_RYS];
  for (int s = 0; s < n_sub; ++s) {
    double a = (double)s * width, b = a + width;
    double *rc = rc_all + (size_t)s * n * CHEB_NC, *wc = wc_all + (size_t)s * n * CHEB_NC;
    fit_cell(n, a, b, rc, wc);
    for (int j = 0; j < CHECK_PER_CELL; ++j) {
      double t = a + ((double)j / (double)(CHECK_PER_CELL - 1)) * (b - a);
      finite_branch(n, t, rr, ww);
      double xi = (2.0 * t - (a + b)) / (b - a);
      for (int i = 0; i < n; ++i) {
        double er = fabs(chebev(rc + i * CHEB_NC, xi) - rr[i]);
        double ew = fabs(chebev(wc + i * CHEB_NC, xi) - ww[i]);
        if (er > worst) worst = er;
        if (ew > worst) worst = ew;
      }
    }
  }
  return worst;
}

static void build_rys_tables(void) {
  for (int n = 1; n <= OR_MAX_RYS; ++n) {
    double t_hi = asymptotic_boundary(2 * n - 1);
    double *rc = (double *)malloc(sizeof(double) * MAX_SUB * n * CHEB_NC);
    double *wc = (double *)malloc(sizeof(double) * MAX_SUB * n * CHEB_NC);
    int n_sub = 1;
    double worst;
    for (;;) {
      worst = max_error(n, t_hi, n_sub, rc, wc);
      if (worst <= FIT_TOL || n_sub >= MAX_SUB) break;
      ++n_sub;
    }
    g_nsub[n - 1] = n_sub;
    g_worst[n - 1] = worst;
    g_root_coef[n - 1] = rc;
    g_wt_coef[n - 1] = wc;
  }
}

int or_rys_table_nsub(int n) {
  pthread_once(&g_table_once, build_rys_tables);
  return (n >= 1 && n <= OR_MAX_RYS) ? g_nsub[n - 1] : 0;
}
double or_rys_table_worst(int n) {
  pthread_once(&g_table_once, build_rys_tables);
  return (n >= 1 && n <= OR_MAX_RYS) ? g_worst[n - 1] : 0.0;
}

/* production path rys.rs:174-185, finite branch by table: rys.rs:230-252 */
void or_rys(int n, double t, double *roots, double *weights) {
  double t_hi = asymptotic_boundary(2 * n - 1);
  if (t >= t_hi) {
    laguerre_branch(n, t, roots, weights);
    return;
  }
  pthread_once(&g_table_once, build_rys_tables);
  int n_sub = g_nsub[n - 1];
  double width = t_hi / (double)n_sub;
  int s = (int)(t / width);
  if (s >= n_sub) s = n_sub - 1;
  double a = (double)s * width, b = a + width;
  double xi = (2.0 * t - (a + b)) / (b - a);
  const double *rc = g_root_coef[n - 1] + (size_t)s * n * CHEB_NC;
  const double *wc = g_wt_coef[n - 1] + (size_t)s * n * CHEB_NC;
  for (int i = 0; i < n; ++i) {
    roots[i] = chebev(rc + i * CHEB_NC, xi);
    weights[i] = chebev(wc + i * CHEB_NC, xi);
  }
}

/* ------------------------------------------------------------------------ */
/* normalisation and c2s                                                     */
/* ------------------------------------------------------------------------ */

static double double_factorial(int n) { /* norm.rs:27-35 */
  double acc = 1.0;
  for (int k = n; k > 1; k -= 2) acc *= (double)k;
  return acc;
}

/* norm.rs:42-49 */
double or_cart_norm(double alpha, int lx, int ly, int lz) {
  int l = lx + ly + lz;
  double df = double_factorial(2 * lx - 1) * double_factorial(2 * ly - 1) *
              double_factorial(2 * lz - 1);
  return pow(2.0 * alpha / M_PI, 0.75) * sqrt(pow(4.0 * alpha, (double)l)) / sqrt(df);
}

static double factorial(int n) {
  double r = 1.0;
  for (int k = 2; k <= n; ++k) r *= (double)k;
  return r;
}
static double binom(int n, int k) {
  if (k < 0 || k > n || n < 0) return 0.0;
  return factorial(n) / (factorial(k) * factorial(n - k));
}
static double cos_half_pi(int n) {
  int r = ((n % 4) + 4) % 4;
  return r == 0 ? 1.0 : (r == 2 ? -1.0 : 0.0);
}
static double sin_half_pi(int n) {
  int r = ((n % 4) + 4) % 4;
  return r == 1 ? 1.0 : (r == 3 ? -1.0 : 0.0);
}

/* Racah-normalised real solid harmonic coefficient, solid_harmonics.rs:87-143 */
static double racah_coeff(int l, int m, int lx, int ly, int lz) {
  int bm = m < 0 ? -m : m;
  if (bm > l) return 0.0;
  int cosine = m >= 0;
  double norm = bm == 0 ? sqrt(factorial(l - bm) / factorial(l + bm))
                        : sqrt(2.0 * factorial(l - bm) / factorial(l + bm));
  double acc = 0.0;
  int kmax = (l - bm) / 2;
  for (int k = 0; k <= kmax; ++k) {
    int zpow = l - 2 * k - bm;
    double gamma = ((k % 2 == 0) ? 1.0 : -1.0) * pow(2.0, -(double)l) * binom(l, k) *
                   binom(2 * l - 2 * k, l) * factorial(l - 2 * k) / factorial(l - 2 * k - bm);
    if (gamma == 0.0) continue;
    for (int a = 0; a <= k; ++a)
      for (int b = 0; b <= k - a; ++b) {
        int c = k - a - b;
        for (int p = 0; p <= bm; ++p) {
          int q = bm - p;
          if (2 * a + p != lx || 2 * b + q != ly || 2 * c + zpow != lz) continue;
          double trig = cosine ? cos_half_pi(q) : sin_half_pi(q);
          if (trig == 0.0) continue;
          double multinomial = factorial(k) / (factorial(a) * factorial(b) * factorial(c));
          acc += gamma * multinomial * binom(bm, p) * trig;
        }
      }
  }
  return norm * acc;
}

static double gaussian_moment_1d(int n, double beta) { /* solid_harmonics.rs:175-182 */
  if (n % 2 == 1) return 0.0;
  return double_factorial(n - 1) / pow(2.0 * beta, (double)(n / 2)) * sqrt(M_PI / beta);
}

static double monomial_to_raw(int l) { /* solid_harmonics.rs:166-172 */
  return l < 2 ? 1.0 : sqrt(4.0 * M_PI / (double)(2 * l + 1));
}

/* solid_harmonics.rs:218-250: rows in m_order (p: x,y,z = m +1,-1,0) */
void or_c2s_matrix(int l, double *out) {
  int nc = (l + 1) * (l + 2) / 2;
  int comps[28][3];
  int q = 0;
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      comps[q][0] = lx;
      comps[q][1] = ly;
      comps[q][2] = l - lx - ly;
      ++q;
    }
  int ms[13];
  if (l == 1) {
    ms[0] = 1;
    ms[1] = -1;
    ms[2] = 0;
  } else {
    for (int i = 0; i < 2 * l + 1; ++i) ms[i] = -l + i;
  }
  double n_mono = or_cart_norm(1.0, l, 0, 0);
  double raw_shell = monomial_to_raw(l) * monomial_to_raw(l);
  for (int row = 0; row < 2 * l + 1; ++row) {
    double coeffs[28];
    for (int i = 0; i < nc; ++i)
      coeffs[i] = racah_coeff(l, ms[row], comps[i][0], comps[i][1], comps[i][2]);
    double qq = 0.0;
    for (int i = 0; i < nc; ++i) {
      if (coeffs[i] == 0.0) continue;
      for (int j = 0; j < nc; ++j) {
        if (coeffs[j] == 0.0) continue;
        double bare = gaussian_moment_1d(comps[i][0] + comps[j][0], 2.0) *
                      gaussian_moment_1d(comps[i][1] + comps[j][1], 2.0) *
                      gaussian_moment_1d(comps[i][2] + comps[j][2], 2.0);
        qq += coeffs[i] * coeffs[j] * raw_shell * n_mono * n_mono * bare;
      }
    }
    double kappa = 1.0 / sqrt(qq);
    for (int i = 0; i < nc; ++i) out[row * nc + i] = kappa * coeffs[i];
  }
}
