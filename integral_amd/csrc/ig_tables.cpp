// ig_tables.cpp -- host-side constant tables for the device kernels:
//   * Rys roots/weights as piecewise degree-12 polynomials in T on
//     equal-width cells of [0, 33 + 2.5(2n-1)), fit to a discretized-Stieltjes
//     Gauss rule (128-point Gauss-Legendre in u = sqrt(x)) and Golub-Welsch.
//     Each cell is fit as a Chebyshev series, then stored in the power basis of
//     the cell variable xi in [-1, 1] (converted in long double) so the device
//     evaluates it by Horner: one FMA per term instead of Clenshaw's two.  The
//     number of cells per n is the smallest whose stored polynomials meet 3e-15
//     absolute on 33 samples per cell.  Fit method and seam follow
//     crates/integral/src/math/rys.rs (:104-185, :230-269, :275-569) and
//     xtask/src/gen_rys_tables.rs:1-190.
//   * generalized Gauss-Laguerre(-1/2) nodes for the large-T branch (rys.rs:519-569).
//   * Cartesian->spherical matrices scaled by sqrt(4pi/(2l+1)) for l >= 2
//     (math/solid_harmonics.rs:87-250, spherical.rs:183-194).
//   * primitive normalisation N(alpha; l,0,0) (math/norm.rs:42-49).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "ig_internal.h"

namespace ig {
namespace {

constexpr int kGL = 128;
constexpr int kNC = kRysTerms;  // 13 coefficients (degree 12)

double seam(int n) { return 33.0 + 2.5 * double(2 * n - 1); }

// Symmetric tridiagonal eigen-solve by implicit-shift QL, carrying only the
// first row of the eigenvector matrix.  d: diagonal, e[i] couples i,i+1.
void ql_first_row(int n, double *d, double *e, double *z) {
  for (int l = 0; l < n; ++l) {
    for (int iter = 0; iter < 60; ++iter) {
      int m = l;
      for (; m < n - 1; ++m)
        if (std::fabs(e[m]) <= 2.220446049250313e-16 * (std::fabs(d[m]) + std::fabs(d[m + 1])))
          break;
      if (m == l) break;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        double rr = std::hypot(f, g);
        e[i + 1] = rr;
        if (rr == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / rr;
        c = g / rr;
        double dd = d[i + 1] - p;
        double t = (d[i] - dd) * s + 2.0 * c * b;
        p = s * t;
        d[i + 1] = dd + p;
        g = c * t - b;
        double zi1 = z[i + 1];
        z[i + 1] = s * z[i] + c * zi1;
        z[i] = c * z[i] - s * zi1;
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
}

// Gauss rule from Jacobi coefficients (alpha, beta), nodes ascending.
void gauss_from_jacobi(int n, const double *alpha, const double *beta, double *x, double *w) {
  if (n == 1) {
    x[0] = alpha[0];
    w[0] = beta[0];
    return;
  }
  std::vector<double> d(alpha, alpha + n), e(n, 0.0), z(n, 0.0);
  z[0] = 1.0;
  for (int i = 1; i < n; ++i) e[i - 1] = std::sqrt(std::max(beta[i], 0.0));
  ql_first_row(n, d.data(), e.data(), z.data());
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
  for (int k = 0; k < n; ++k) {
    x[k] = d[order[k]];
    w[k] = beta[0] * z[order[k]] * z[order[k]];
  }
}

struct Quadratures {
  double u[kGL], uw[kGL];
  double lag_x[kMaxRys + 1][kMaxRys], lag_w[kMaxRys + 1][kMaxRys];
  Quadratures() {
    // Gauss-Legendre on [0,1]: monic shifted Legendre, a_k = 1/2, b_k = k^2/(4(4k^2-1)).
    double a[kGL], b[kGL];
    for (int k = 0; k < kGL; ++k) {
      a[k] = 0.5;
      b[k] = k == 0 ? 1.0 : 0.25 * double(k) * double(k) / (4.0 * double(k) * double(k) - 1.0);
    }
    gauss_from_jacobi(kGL, a, b, u, uw);
    // generalized Laguerre(-1/2): a_k = 2k + 1/2, b_k = k(k - 1/2), b_0 = Gamma(1/2).
    for (int n = 1; n <= kMaxRys; ++n) {
      double aa[kMaxRys], bb[kMaxRys];
      for (int k = 0; k < n; ++k) {
        aa[k] = 2.0 * k + 0.5;
        bb[k] = double(k) * (double(k) - 0.5);
      }
      bb[0] = std::sqrt(M_PI);
      gauss_from_jacobi(n, aa, bb, lag_x[n], lag_w[n]);
    }
  }
};

const Quadratures &quad() {
  static Quadratures q;
  return q;
}

// n-point Gauss rule of dmu = 1/2 x^{-1/2} e^{-Tx} dx on [0,1] by the
// discretized Stieltjes procedure on X_m = u_m^2, W_m = w_m e^{-T u_m^2}.
void stieltjes_rule(int n, double T, double *x, double *w) {
  const Quadratures &Q = quad();
  double X[kGL], W[kGL], pm[kGL], pc[kGL];
  double alpha[kMaxRys], beta[kMaxRys];
  double norm = 0.0, xnorm = 0.0;
  for (int m = 0; m < kGL; ++m) {
    X[m] = Q.u[m] * Q.u[m];
    W[m] = Q.uw[m] * std::exp(-T * X[m]);
    pm[m] = 0.0;
    pc[m] = 1.0;
    norm += W[m];
    xnorm += X[m] * W[m];
  }
  alpha[0] = xnorm / norm;
  beta[0] = norm;
  for (int k = 1; k < n; ++k) {
    double nn = 0.0, xn = 0.0;
    for (int m = 0; m < kGL; ++m) {
      double pn = (X[m] - alpha[k - 1]) * pc[m] - beta[k - 1] * pm[m];
      pm[m] = pc[m];
      pc[m] = pn;
      double t = W[m] * pn * pn;
      nn += t;
      xn += X[m] * t;
    }
    alpha[k] = xn / nn;
    beta[k] = nn / norm;
    norm = nn;
  }
  gauss_from_jacobi(n, alpha, beta, x, w);
}

// Chebyshev coefficients of every root and weight of the n-point rule on [a,b].
void fit(int n, double a, double b, double *rc, double *wc) {
  double xs[kNC][kMaxRys], ws[kNC][kMaxRys];
  for (int j = 0; j < kNC; ++j) {
    double y = std::cos(M_PI * (j + 0.5) / kNC);
    stieltjes_rule(n, 0.5 * (b - a) * y + 0.5 * (b + a), xs[j], ws[j]);
  }
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < kNC; ++k) {
      double sr = 0.0, sw = 0.0;
      for (int j = 0; j < kNC; ++j) {
        double ck = std::cos(M_PI * k * (j + 0.5) / kNC);
        sr += xs[j][i] * ck;
        sw += ws[j][i] * ck;
      }
      rc[i * kNC + k] = 2.0 / kNC * sr;
      wc[i * kNC + k] = 2.0 / kNC * sw;
    }
}

// Half-c0 Chebyshev series -> power basis in xi, in place (long double).
void cheb_to_power(double *c) {
  long double Tkm1[kNC] = {}, Tk[kNC] = {}, out[kNC] = {};
  Tkm1[0] = 1.0L;  // T_0
  Tk[1] = 1.0L;    // T_1
  out[0] = 0.5L * c[0];
  for (int j = 0; j < kNC; ++j) out[j] += (long double)c[1] * Tk[j];
  for (int k = 2; k < kNC; ++k) {
    long double Tn[kNC] = {};
    for (int j = 0; j < kNC; ++j) {  // T_k = 2 xi T_{k-1} - T_{k-2}
      if (j >= 1) Tn[j] += 2.0L * Tk[j - 1];
      Tn[j] -= Tkm1[j];
    }
    for (int j = 0; j < kNC; ++j) {
      Tkm1[j] = Tk[j];
      Tk[j] = Tn[j];
      out[j] += (long double)c[k] * Tn[j];
    }
  }
  for (int j = 0; j < kNC; ++j) c[j] = double(out[j]);
}

double horner(const double *m, double xi) {
  double r = m[kNC - 1];
  for (int k = kNC - 2; k >= 0; --k) r = std::fma(r, xi, m[k]);
  return r;
}

double cells_error(int n, int cells, std::vector<double> &rc, std::vector<double> &wc) {
  double width = seam(n) / cells, worst = 0.0;
  rc.assign(size_t(cells) * n * kNC, 0.0);
  wc.assign(size_t(cells) * n * kNC, 0.0);
  for (int s = 0; s < cells; ++s) {
    double a = s * width, b = a + width;
    double *r = rc.data() + size_t(s) * n * kNC, *w = wc.data() + size_t(s) * n * kNC;
    fit(n, a, b, r, w);
    for (int i = 0; i < n; ++i) {
      cheb_to_power(r + i * kNC);
      cheb_to_power(w + i * kNC);
    }
    for (int j = 0; j < 33; ++j) {
      double T = a + (b - a) * j / 32.0;
      double xr[kMaxRys], wr[kMaxRys];
      stieltjes_rule(n, T, xr, wr);
      double xi = (2.0 * T - (a + b)) / (b - a);
      for (int i = 0; i < n; ++i) {
        worst = std::max(worst, std::fabs(horner(r + i * kNC, xi) - xr[i]));
        worst = std::max(worst, std::fabs(horner(w + i * kNC, xi) - wr[i]));
      }
    }
  }
  return worst;
}

double factorial(int n) {
  double r = 1.0;
  for (int k = 2; k <= n; ++k) r *= k;
  return r;
}
double choose(int n, int k) {
  return (k < 0 || k > n) ? 0.0 : factorial(n) / (factorial(k) * factorial(n - k));
}
double dfact(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

}  // namespace

const RysHostTables &rys_host_tables() {
  static RysHostTables t;
  static std::once_flag once;
  std::call_once(once, [] {
    std::vector<double> rc, wc;
    t.data.clear();
    for (int n = 1; n <= kMaxRys; ++n) {
      // counts grow slowly with n: start the search just below the last one
      int cells = n == 1 ? 1 : std::max(1, t.nsub[n - 1] - 3);
      double err;
      for (;; ++cells) {
        err = cells_error(n, cells, rc, wc);
        if (err <= 3e-15 || cells >= 256) break;
      }
      t.nsub[n] = cells;
      t.worst[n] = err;
      t.root_off[n] = int(t.data.size());
      t.data.insert(t.data.end(), rc.begin(), rc.end());
      t.wt_off[n] = int(t.data.size());
      t.data.insert(t.data.end(), wc.begin(), wc.end());
      t.lag_off[n] = int(t.data.size());
      const Quadratures &Q = quad();
      t.data.insert(t.data.end(), Q.lag_x[n], Q.lag_x[n] + n);
      t.data.insert(t.data.end(), Q.lag_w[n], Q.lag_w[n] + n);
    }
  });
  return t;
}

// Boys function table F_m(T0), T0 = 0, 0.1, ..., 93.0 (rows), m = 0..33, and
// exp(-T0) as the last column (math/boys.rs:63-136 layout and grid).  Each row
// comes from the series F_m(T) = e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
// at m = 33 in long double and the downward recurrence
// F_{m-1} = (2T F_m + e^-T) / (2m-1), which is stable.
const std::vector<double> &boys_host_table() {
  static std::vector<double> t;
  static std::once_flag once;
  std::call_once(once, [] {
    t.assign(size_t(kBoysRows) * kBoysCols, 0.0);
    for (int i = 0; i < kBoysRows; ++i) {
      const long double T = 0.1L * i, eT = std::exp(-T);
      long double term = 1.0L / (2 * kBoysMaxM + 1), sum = term;
      for (int k = 1; k < 2000; ++k) {
        term *= 2.0L * T / (2 * kBoysMaxM + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-22L) break;
      }
      long double f = eT * sum;
      double *row = t.data() + size_t(i) * kBoysCols;
      row[kBoysMaxM] = double(f);
      for (int m = kBoysMaxM; m > 0; --m) {
        f = (2.0L * T * f + eT) / (2 * m - 1);
        row[m - 1] = double(f);
      }
      row[kBoysMaxM + 1] = double(eT);
    }
  });
  return t;
}

void rys_host_eval(int n, double T, double *x, double *w) {
  const RysHostTables &t = rys_host_tables();
  double hi = seam(n);
  if (T >= hi) {
    const double *L = t.data.data() + t.lag_off[n];
    double inv = 1.0 / T, sc = 0.5 / std::sqrt(T);
    for (int i = 0; i < n; ++i) {
      x[i] = L[i] * inv;
      w[i] = L[n + i] * sc;
    }
    return;
  }
  const double v = T * (double(t.nsub[n]) / hi);  // cell coordinate, as on the device
  const int s = std::min(int(v), t.nsub[n] - 1);
  const double xi = 2.0 * (v - double(s)) - 1.0;
  for (int i = 0; i < n; ++i) {
    x[i] = horner(t.data.data() + t.root_off[n] + (size_t(s) * n + i) * kNC, xi);
    w[i] = horner(t.data.data() + t.wt_off[n] + (size_t(s) * n + i) * kNC, xi);
  }
}

double shell_norm(double alpha, int l) {
  return std::pow(2.0 * alpha / M_PI, 0.75) * std::sqrt(std::pow(4.0 * alpha, double(l))) /
         std::sqrt(dfact(2 * l - 1));
}

// Racah real solid harmonic coefficient of x^lx y^ly z^lz for order m (m >= 0
// cosine type, m < 0 sine type).
static double solid_harmonic(int l, int m, int lx, int ly, int lz) {
  int M = std::abs(m);
  double norm = std::sqrt((M == 0 ? 1.0 : 2.0) * factorial(l - M) / factorial(l + M));
  double acc = 0.0;
  for (int k = 0; k <= (l - M) / 2; ++k) {
    double gam = ((k & 1) ? -1.0 : 1.0) * std::ldexp(1.0, -l) * choose(l, k) *
                 choose(2 * l - 2 * k, l) * factorial(l - 2 * k) / factorial(l - 2 * k - M);
    int zp = l - 2 * k - M;
    for (int a = 0; a <= k; ++a)
      for (int b = 0; a + b <= k; ++b) {
        int c = k - a - b;
        for (int p = 0; p <= M; ++p) {
          int q = M - p;
          if (2 * a + p != lx || 2 * b + q != ly || 2 * c + zp != lz) continue;
          int r = q & 3;
          double trig = m >= 0 ? (r == 0 ? 1.0 : (r == 2 ? -1.0 : 0.0))
                               : (r == 1 ? 1.0 : (r == 3 ? -1.0 : 0.0));
          acc += gam * factorial(k) / (factorial(a) * factorial(b) * factorial(c)) * choose(M, p) *
                 trig;
        }
      }
  }
  return norm * acc;
}

std::vector<double> c2s_scaled(int l) {
  int nc = ncart(l), nf = 2 * l + 1;
  std::vector<int> ms(nf);
  if (l == 1)
    ms = {1, -1, 0};
  else
    for (int i = 0; i < nf; ++i) ms[i] = i - l;
  std::vector<double> out(size_t(nf) * nc);
  double raw = l < 2 ? 1.0 : std::sqrt(4.0 * M_PI / (2 * l + 1));
  double nmono = shell_norm(1.0, l);
  auto moment = [](int n) {  // integral of x^n exp(-2x^2) dx
    return (n & 1) ? 0.0 : dfact(n - 1) / std::pow(4.0, n / 2) * std::sqrt(M_PI / 2.0);
  };
  for (int r = 0; r < nf; ++r) {
    std::vector<double> c(nc);
    for (int i = 0; i < nc; ++i) {
      int lx, ly, lz;
      cart_powers(l, i, lx, ly, lz);
      c[i] = solid_harmonic(l, ms[r], lx, ly, lz);
    }
    double q = 0.0;
    for (int i = 0; i < nc; ++i)
      for (int j = 0; j < nc; ++j) {
        if (c[i] == 0.0 || c[j] == 0.0) continue;
        int ix, iy, iz, jx, jy, jz;
        cart_powers(l, i, ix, iy, iz);
        cart_powers(l, j, jx, jy, jz);
        q += c[i] * c[j] * raw * raw * nmono * nmono * moment(ix + jx) * moment(iy + jy) *
             moment(iz + jz);
      }
    for (int i = 0; i < nc; ++i) out[size_t(r) * nc + i] = raw * c[i] / std::sqrt(q);
  }
  return out;
}

}  // namespace ig
