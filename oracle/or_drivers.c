/*
 * or_drivers.c -- oracle (test infrastructure only): basis data model, dense
 * ERI tensor, Schwarz bounds, integral-direct J/K and the density-contracted
 * ERI gradient.  CPU restatement of crates/integral/src/{shell.rs,
 * integrals.rs, spherical.rs, direct.rs, direct_grad.rs, grad.rs}.
 * OpenMP parallelism is over canonical bra shell pairs with per-thread
 * accumulators, the call-site parallelism the reference documents
 * (README.md:68-80, direct.rs:19-20, direct_grad.rs:140-142).
 */
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "or_internal.h"

static __thread char g_err[256];
void or_set_error(const char *msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }
const char *or_last_error(void) { return g_err; }

/* Cartesian triples per degree, lx descending then ly descending (am.rs:23-31) */
#define MAX_DEG 14
static int g_comps[MAX_DEG][120][3];
static int g_comps_ready = 0;
static void build_comps(void) {
  for (int l = 0; l < MAX_DEG; ++l) {
    int q = 0;
    for (int lx = l; lx >= 0; --lx)
      for (int ly = l - lx; ly >= 0; --ly) {
        g_comps[l][q][0] = lx;
        g_comps[l][q][1] = ly;
        g_comps[l][q][2] = l - lx - ly;
        ++q;
      }
  }
  g_comps_ready = 1;
}
const int (*or_cart_comps(int l))[3] { return (const int(*)[3])g_comps[l]; }

/* ------------------------------------------------------------------------ */
/* basis: shell.rs:64-118 (validation), :182-187 (primitive_coeff),           */
/* :226-275 (nao, offsets, atoms by bitwise-equal centres)                   */
/* ------------------------------------------------------------------------ */
int or_basis_create(int nsh, const int *l, const int *kind, const double *centers,
                    const int *nprim, const double *exps, const double *coefs,
                    or_basis **out) {
  if (!g_comps_ready) build_comps();
  *out = NULL;
  if (nsh < 0) {
    or_set_error("negative shell count");
    return OR_ERR_INVALID_ARGUMENT;
  }
  size_t ntot = 0;
  for (int s = 0; s < nsh; ++s) {
    if (nprim[s] <= 0) {
      or_set_error("empty contraction");
      return OR_ERR_EMPTY_CONTRACTION;
    }
    if (l[s] < 0 || l[s] > OR_MAX_L) {
      or_set_error("angular momentum too high");
      return OR_ERR_L_TOO_HIGH;
    }
    ntot += (size_t)nprim[s];
  }
  or_basis *b = (or_basis *)calloc(1, sizeof(or_basis));
  b->nsh = nsh;
  b->sh = (or_shell *)calloc((size_t)nsh, sizeof(or_shell));
  b->store = (double *)malloc(sizeof(double) * 4 * (ntot ? ntot : 1));
  b->atoms = (double *)malloc(sizeof(double) * 3 * (nsh ? nsh : 1));
  double *e = b->store, *c = e + ntot, *ef = c + ntot, *e2 = ef + ntot;
  size_t p = 0;
  int off = 0;
  for (int s = 0; s < nsh; ++s) {
    or_shell *sh = &b->sh[s];
    sh->l = l[s];
    sh->kind = kind[s];
    sh->nprim = nprim[s];
    memcpy(sh->center, centers + 3 * s, 3 * sizeof(double));
    sh->exps = e + p;
    sh->coefs = c + p;
    sh->eff = ef + p;
    sh->eff2a = e2 + p;
    for (int i = 0; i < nprim[s]; ++i) {
      e[p + i] = exps[p + i];
      c[p + i] = coefs[p + i];
      double a = exps[p + i];
      ef[p + i] = (a == 0.0) ? coefs[p + i] : coefs[p + i] * or_cart_norm(a, l[s], 0, 0);
      e2[p + i] = 2.0 * a * ef[p + i]; /* direct_grad.rs:189-199 */
    }
    p += (size_t)nprim[s];
    sh->ncart = or_ncart(l[s]);
    sh->nfunc = kind[s] ? 2 * l[s] + 1 : sh->ncart;
    sh->off = off;
    off += sh->nfunc;
    if (kind[s]) { /* spherical.rs:183-194 */
      sh->c2s = (double *)malloc(sizeof(double) * sh->nfunc * sh->ncart);
      or_c2s_matrix(l[s], sh->c2s);
      double ratio = l[s] < 2 ? 1.0 : sqrt(4.0 * M_PI / (double)(2 * l[s] + 1));
      for (int q = 0; q < sh->nfunc * sh->ncart; ++q) sh->c2s[q] *= ratio;
    }
    int found = -1;
    for (int a = 0; a < b->natom; ++a)
      if (memcmp(b->atoms + 3 * a, sh->center, 3 * sizeof(double)) == 0) {
        found = a;
        break;
      }
    if (found < 0) {
      memcpy(b->atoms + 3 * b->natom, sh->center, 3 * sizeof(double));
      found = b->natom++;
    }
    sh->atom = found;
  }
  b->nao = off;
  b->nao_cart = 0;
  for (int s = 0; s < nsh; ++s) b->nao_cart += b->sh[s].ncart;
  *out = b;
  return OR_OK;
}

void or_basis_destroy(or_basis *b) {
  if (!b) return;
  for (int s = 0; s < b->nsh; ++s) free(b->sh[s].c2s);
  free(b->sh);
  free(b->store);
  free(b->atoms);
  free(b);
}
int or_basis_nao(const or_basis *b) { return b->nao; }
int or_basis_nao_cart(const or_basis *b) { return b->nao_cart; }
int or_basis_natom(const or_basis *b) { return b->natom; }
int or_basis_nsh(const or_basis *b) { return b->nsh; }

static or_shref shref(const or_shell *s, int l, const double *coefs) {
  or_shref r;
  r.l = l;
  r.center = s->center;
  r.nprim = s->nprim;
  r.exps = s->exps;
  r.coefs = coefs;
  return r;
}

/* one axis of a row-major 4-tensor by an n_out x n_in matrix, spherical.rs:208-240 */
static void apply_axis(const double *in, const int *dims, int axis, const double *mat, int n_out,
                       double *out) {
  int n_in = dims[axis];
  int outer = 1, inner = 1;
  for (int x = 0; x < axis; ++x) outer *= dims[x];
  for (int x = axis + 1; x < 4; ++x) inner *= dims[x];
  memset(out, 0, sizeof(double) * (size_t)outer * n_out * inner);
  for (int o = 0; o < outer; ++o)
    for (int q = 0; q < n_out; ++q) {
      double *dst = out + ((size_t)o * n_out + q) * inner;
      for (int j = 0; j < n_in; ++j) {
        double m = mat[q * n_in + j];
        if (m == 0.0) continue;
        const double *src = in + ((size_t)o * n_in + j) * inner;
        for (int s = 0; s < inner; ++s) dst[s] += m * src[s];
      }
    }
}

/* spherical.rs:307-325: in-place via a temporary of the same capacity.
 * mats[k] is n_out[k] x dims[k] or NULL; returns the new length. */
static size_t transform4(double *block, double *tmp, const int *dims_in, const double *const *mats,
                         const int *n_out) {
  int dims[4] = {dims_in[0], dims_in[1], dims_in[2], dims_in[3]};
  for (int ax = 0; ax < 4; ++ax) {
    if (!mats[ax]) continue;
    apply_axis(block, dims, ax, mats[ax], n_out[ax], tmp);
    dims[ax] = n_out[ax];
    memcpy(block, tmp, sizeof(double) * (size_t)dims[0] * dims[1] * dims[2] * dims[3]);
  }
  return (size_t)dims[0] * dims[1] * dims[2] * dims[3];
}

#define MAXBLK (28 * 28 * 28 * 28)

/* function-space block for shells (i,j,k,l) into out (zeroed here) */
static void func_block(const or_basis *b, int engine, int i, int j, int k, int l, double *out,
                       double *tmp, int64_t *nprimq) {
  const or_shell *s[4] = {&b->sh[i], &b->sh[j], &b->sh[k], &b->sh[l]};
  or_shref r[4];
  int dims[4], nout[4];
  const double *mats[4];
  size_t n = 1;
  for (int q = 0; q < 4; ++q) {
    r[q] = shref(s[q], s[q]->l, s[q]->eff);
    dims[q] = s[q]->ncart;
    nout[q] = s[q]->nfunc;
    mats[q] = s[q]->c2s;
    n *= (size_t)dims[q];
  }
  memset(out, 0, sizeof(double) * n);
  or_cart_quartet(engine, r, out, nprimq);
  transform4(out, tmp, dims, mats, nout);
}

int or_eri_block(const or_basis *b, int engine, int i, int j, int k, int l, double *out) {
  double *blk = (double *)malloc(sizeof(double) * MAXBLK);
  double *tmp = (double *)malloc(sizeof(double) * MAXBLK);
  func_block(b, engine, i, j, k, l, blk, tmp, NULL);
  size_t n = (size_t)b->sh[i].nfunc * b->sh[j].nfunc * b->sh[k].nfunc * b->sh[l].nfunc;
  memcpy(out, blk, sizeof(double) * n);
  free(blk);
  free(tmp);
  return OR_OK;
}

int or_eri_block_cart(const or_basis *b, int engine, int i, int j, int k, int l, double *out) {
  const or_shell *s[4] = {&b->sh[i], &b->sh[j], &b->sh[k], &b->sh[l]};
  or_shref r[4];
  size_t n = 1;
  for (int q = 0; q < 4; ++q) {
    r[q] = shref(s[q], s[q]->l, s[q]->eff);
    n *= (size_t)s[q]->ncart;
  }
  memset(out, 0, sizeof(double) * n);
  or_cart_quartet(engine, r, out, NULL);
  return OR_OK;
}

/* integrals.rs:1227-1236 */
static const int PERMS8[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                 {2, 3, 0, 1}, {2, 3, 1, 0}, {3, 2, 0, 1}, {3, 2, 1, 0}};

/* integrals.rs:1271-1312 */
static void scatter_s8(double *out, size_t nao, const int *sidx, const int *offs, const int *n,
                       const double *block) {
  int seen[8][4];
  int n_seen = 0;
  for (int p = 0; p < 8; ++p) {
    int tup[4];
    for (int q = 0; q < 4; ++q) tup[q] = sidx[PERMS8[p][q]];
    int dup = 0;
    for (int s = 0; s < n_seen; ++s)
      if (memcmp(seen[s], tup, sizeof(tup)) == 0) dup = 1;
    if (dup) continue;
    memcpy(seen[n_seen++], tup, sizeof(tup));
    size_t stride[4] = {0, 0, 0, 0}, base = 0;
    for (int q = 0; q < 4; ++q) {
      int src_axis = PERMS8[p][q];
      size_t st = 1;
      for (int e = 0; e < 3 - q; ++e) st *= nao;
      stride[src_axis] = st;
      base += (size_t)offs[src_axis] * st;
    }
    size_t src = 0;
    for (int a = 0; a < n[0]; ++a)
      for (int bb = 0; bb < n[1]; ++bb)
        for (int c = 0; c < n[2]; ++c)
          for (int d = 0; d < n[3]; ++d)
            out[base + a * stride[0] + bb * stride[1] + c * stride[2] + d * stride[3]] =
                block[src++];
  }
}

static int npairs(const or_basis *b) { return b->nsh * (b->nsh + 1) / 2; }
static void pair_of(int64_t p, int *i, int *j) {
  int ii = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) / 2.0);
  while ((int64_t)ii * (ii + 1) / 2 > p) --ii;
  while ((int64_t)(ii + 1) * (ii + 2) / 2 <= p) ++ii;
  *i = ii;
  *j = (int)(p - (int64_t)ii * (ii + 1) / 2);
}

static int eri_dense_impl(const or_basis *b, int engine, const double *q, double tau,
                          double *out, int64_t *stats, int nthreads) {
  size_t nao = (size_t)b->nao;
  memset(out, 0, sizeof(double) * nao * nao * nao * nao);
  int np = npairs(b);
  int64_t total = 0, skipped = 0;
  int nsh = b->nsh;
#pragma omp parallel num_threads(nthreads) reduction(+ : total, skipped)
  {
    double *blk = (double *)malloc(sizeof(double) * MAXBLK);
    double *tmp = (double *)malloc(sizeof(double) * MAXBLK);
#pragma omp for schedule(dynamic, 1)
    for (int p = 0; p < np; ++p) {
      int si, sj;
      pair_of(p, &si, &sj);
      for (int sk = 0; sk <= si; ++sk) {
        int l_top = (sk == si) ? sj : sk;
        for (int sl = 0; sl <= l_top; ++sl) {
          ++total;
          if (q && q[si * nsh + sj] * q[sk * nsh + sl] < tau) {
            ++skipped;
            continue;
          }
          func_block(b, engine, si, sj, sk, sl, blk, tmp, NULL);
          int sidx[4] = {si, sj, sk, sl};
          int offs[4], n[4];
          for (int x = 0; x < 4; ++x) {
            offs[x] = b->sh[sidx[x]].off;
            n[x] = b->sh[sidx[x]].nfunc;
          }
          scatter_s8(out, nao, sidx, offs, n, blk);
        }
      }
    }
    free(blk);
    free(tmp);
  }
  if (stats) {
    stats[0] = total;
    stats[1] = skipped;
  }
  return OR_OK;
}

int or_eri_dense(const or_basis *b, int engine, double *out, int nthreads) {
  return eri_dense_impl(b, engine, NULL, 0.0, out, NULL, nthreads);
}

/* integrals.rs:842-862 */
int or_schwarz(const or_basis *b, int engine, double *q, int nthreads) {
  int nsh = b->nsh;
#pragma omp parallel num_threads(nthreads)
  {
    double *blk = (double *)malloc(sizeof(double) * MAXBLK);
    double *tmp = (double *)malloc(sizeof(double) * MAXBLK);
#pragma omp for schedule(dynamic, 1)
    for (int ij = 0; ij < nsh * nsh; ++ij) {
      int i = ij / nsh, j = ij % nsh;
      int ni = b->sh[i].nfunc, nj = b->sh[j].nfunc;
      func_block(b, engine, i, j, i, j, blk, tmp, NULL);
      double mx = 0.0;
      for (int mu = 0; mu < ni; ++mu)
        for (int nu = 0; nu < nj; ++nu) {
          double v = fabs(blk[((mu * nj + nu) * ni + mu) * nj + nu]);
          if (v > mx) mx = v;
        }
      q[ij] = sqrt(mx);
    }
    free(blk);
    free(tmp);
  }
  return OR_OK;
}

int or_eri_screened(const or_basis *b, int engine, double tau, double *out, int64_t *stats,
                    int nthreads) {
  double *q = (double *)malloc(sizeof(double) * (size_t)b->nsh * b->nsh);
  or_schwarz(b, engine, q, nthreads);
  int rc = eri_dense_impl(b, engine, q, tau, out, stats, nthreads);
  free(q);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* direct J/K                                                                */
/* ------------------------------------------------------------------------ */

/* direct.rs:103-125 */
int or_density_shell_bounds(const or_basis *b, int ndens, const double *const *dens,
                            double *bounds) {
  int nsh = b->nsh, nao = b->nao;
  for (int d = 0; d < ndens; ++d)
    for (int i = 0; i < nsh; ++i)
      for (int j = 0; j < nsh; ++j) {
        double mx = 0.0;
        for (int a = 0; a < b->sh[i].nfunc; ++a) {
          const double *row = dens[d] + (size_t)(b->sh[i].off + a) * nao + b->sh[j].off;
          for (int c = 0; c < b->sh[j].nfunc; ++c)
            if (fabs(row[c]) > mx) mx = fabs(row[c]);
        }
        bounds[((size_t)d * nsh + i) * nsh + j] = mx;
      }
  return OR_OK;
}

/* direct.rs:387-404 */
static inline void add_contraction(double *J, double *K, const double *D, size_t n, double g,
                                   int a, int b, int c, int d) {
  J[a * n + b] += g * D[c * n + d];
  if (K) K[a * n + c] += g * D[b * n + d];
}

/* direct.rs:353-383 */
static void scatter_unique(double *J, double *K, const double *D, size_t n, double g, int a,
                           int b, int c, int d) {
  add_contraction(J, K, D, n, g, a, b, c, d);
  if (a != b) add_contraction(J, K, D, n, g, b, a, c, d);
  if (c != d) {
    add_contraction(J, K, D, n, g, a, b, d, c);
    if (a != b) add_contraction(J, K, D, n, g, b, a, d, c);
  }
  if (a != c || b != d) {
    add_contraction(J, K, D, n, g, c, d, a, b);
    if (c != d) add_contraction(J, K, D, n, g, d, c, a, b);
    if (a != b) {
      add_contraction(J, K, D, n, g, c, d, b, a);
      if (c != d) add_contraction(J, K, D, n, g, d, c, b, a);
    }
  }
}

int or_direct_jk(const or_basis *b, int engine, const double *schwarz, int ndens,
                 const double *const *dens, const double *bounds, double threshold,
                 int want_exchange, double *const *J, double *const *K, int64_t bra_begin,
                 int64_t bra_end, int nthreads, int64_t *stats) {
  if (ndens < 1 || ndens > 64) {
    or_set_error("direct contraction needs 1..=64 densities");
    return OR_ERR_INVALID_ARGUMENT;
  }
  int nsh = b->nsh;
  size_t nao = (size_t)b->nao, n2 = nao * nao;
  double *q = (double *)schwarz;
  if (!q) {
    q = (double *)malloc(sizeof(double) * nsh * nsh);
    or_schwarz(b, engine, q, nthreads);
  }
  int64_t np = npairs(b);
  if (bra_end > np || bra_end < 0) bra_end = np;
  int64_t nq = 0, npq = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : nq, npq)
  {
    double *blk = (double *)malloc(sizeof(double) * MAXBLK);
    double *tmp = (double *)malloc(sizeof(double) * MAXBLK);
    double *acc = (double *)calloc((size_t)ndens * 2 * n2, sizeof(double));
#pragma omp for schedule(dynamic, 1)
    for (int64_t p = bra_begin; p < bra_end; ++p) {
      int i, j;
      pair_of(p, &i, &j);
      double qab = q[i * nsh + j];
      if (qab == 0.0) continue; /* direct.rs:165-168 */
      for (int k = 0; k <= i; ++k) {
        int l_top = (k == i) ? j : k;
        for (int l = 0; l <= l_top; ++l) {
          double pair_bound = qab * q[k * nsh + l];
          if (pair_bound < threshold) continue;
          uint64_t mask = 0;
          for (int d = 0; d < ndens; ++d) {
            const double *bd = bounds + (size_t)d * nsh * nsh;
            double db = bd[i * nsh + j];
            double v;
            if ((v = bd[k * nsh + l]) > db) db = v;
            if ((v = bd[i * nsh + k]) > db) db = v;
            if ((v = bd[i * nsh + l]) > db) db = v;
            if ((v = bd[j * nsh + k]) > db) db = v;
            if ((v = bd[j * nsh + l]) > db) db = v;
            if (pair_bound * db >= threshold) mask |= (uint64_t)1 << d;
          }
          if (!mask) continue;
          func_block(b, engine, i, j, k, l, blk, tmp, &npq);
          ++nq;
          /* direct.rs:284-326 */
          const or_shell *si = &b->sh[i], *sj = &b->sh[j], *sk = &b->sh[k], *sl = &b->sh[l];
          int na = si->nfunc, nb = sj->nfunc, nc = sk->nfunc, nd = sl->nfunc;
          for (int a = 0; a < na; ++a) {
            int mu = si->off + a;
            int b_hi = (i == j) ? a + 1 : nb;
            for (int bb = 0; bb < b_hi; ++bb) {
              int nu = sj->off + bb;
              for (int c = 0; c < nc; ++c) {
                int lam = sk->off + c;
                int d_hi = (k == l) ? c + 1 : nd;
                size_t base = (((size_t)a * nb + bb) * nc + c) * nd;
                for (int d = 0; d < d_hi; ++d) {
                  int sig = sl->off + d;
                  if (i == k && j == l && (mu < lam || (mu == lam && nu < sig))) continue;
                  double g = blk[base + d];
                  for (int di = 0; di < ndens; ++di)
                    if (mask & ((uint64_t)1 << di))
                      scatter_unique(acc + (size_t)di * 2 * n2,
                                     want_exchange ? acc + (size_t)di * 2 * n2 + n2 : NULL,
                                     dens[di], nao, g, mu, nu, lam, sig);
                }
              }
            }
          }
        }
      }
    }
#pragma omp critical
    {
      for (int d = 0; d < ndens; ++d) {
        for (size_t x = 0; x < n2; ++x) J[d][x] += acc[(size_t)d * 2 * n2 + x];
        if (want_exchange)
          for (size_t x = 0; x < n2; ++x) K[d][x] += acc[(size_t)d * 2 * n2 + n2 + x];
      }
    }
    free(acc);
    free(blk);
    free(tmp);
  }
  if (!schwarz) free(q);
  if (stats) {
    stats[0] = nq;
    stats[1] = npq;
  }
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* gradient                                                                  */
/* ------------------------------------------------------------------------ */

/* symmetrised shell-block maxima, direct_grad.rs:292-312 */
static void shell_block_max(const or_basis *b, const double *m, double *out) {
  int nsh = b->nsh, nao = b->nao;
  memset(out, 0, sizeof(double) * nsh * nsh);
  for (int i = 0; i < nsh; ++i)
    for (int j = 0; j < nsh; ++j) {
      double mx = 0.0;
      for (int a = 0; a < b->sh[i].nfunc; ++a) {
        const double *row = m + (size_t)(b->sh[i].off + a) * nao + b->sh[j].off;
        for (int c = 0; c < b->sh[j].nfunc; ++c)
          if (fabs(row[c]) > mx) mx = fabs(row[c]);
      }
      int lo = i < j ? i : j, hi = i < j ? j : i;
      double sym = out[hi * nsh + lo] > mx ? out[hi * nsh + lo] : mx;
      out[hi * nsh + lo] = sym;
      out[lo * nsh + hi] = sym;
    }
}

typedef struct {
  int ncoul, nexch;
  const double *cfac, *xfac;
  const double *const *cl, *const *cr, *const *xl, *const *xr;
  double *clmax, *crmax, *xlmax, *xrmax; /* per term nsh*nsh */
} grad_density;

/* direct_grad.rs:626-644 */
static double block_bound(const grad_density *g, int nsh, int i, int j, int k, int l) {
#define AT(m, a, bb) (m)[(a)*nsh + (bb)]
  double bound = 0.0;
  for (int t = 0; t < g->ncoul; ++t) {
    const double *L = g->clmax + (size_t)t * nsh * nsh, *R = g->crmax + (size_t)t * nsh * nsh;
    bound += 0.5 * fabs(g->cfac[t]) * (AT(L, i, j) * AT(R, k, l) + AT(L, k, l) * AT(R, i, j));
  }
  for (int t = 0; t < g->nexch; ++t) {
    const double *L = g->xlmax + (size_t)t * nsh * nsh, *R = g->xrmax + (size_t)t * nsh * nsh;
    bound += 0.25 * fabs(g->xfac[t]) *
             (AT(L, i, k) * AT(R, j, l) + AT(L, j, k) * AT(R, i, l) + AT(L, i, l) * AT(R, j, k) +
              AT(L, j, l) * AT(R, i, k));
  }
#undef AT
  return bound;
}

/* weight block (function space, then pulled back to Cartesian), direct_grad.rs:490-547 */
static void weight_block(const or_basis *b, const grad_density *g, const int *idx,
                         double degeneracy, double *w, double *tmp) {
  size_t n = (size_t)b->nao;
  const or_shell *s[4] = {&b->sh[idx[0]], &b->sh[idx[1]], &b->sh[idx[2]], &b->sh[idx[3]]};
  int oi = s[0]->off, oj = s[1]->off, ok = s[2]->off, ol = s[3]->off;
  int ni = s[0]->nfunc, nj = s[1]->nfunc, nk = s[2]->nfunc, nl = s[3]->nfunc;
  memset(w, 0, sizeof(double) * (size_t)ni * nj * nk * nl);
  for (int t = 0; t < g->ncoul; ++t) {
    const double *left = g->cl[t], *right = g->cr[t];
    double f = 0.5 * degeneracy * g->cfac[t];
    for (int a = 0; a < ni; ++a)
      for (int bb = 0; bb < nj; ++bb) {
        size_t mn = (size_t)(oi + a) * n + oj + bb;
        double l_mn = f * left[mn], r_mn = f * right[mn];
        double *row = w + ((size_t)a * nj + bb) * nk * nl;
        for (int c = 0; c < nk; ++c) {
          size_t ls = (size_t)(ok + c) * n + ol;
          for (int d = 0; d < nl; ++d) row[c * nl + d] += l_mn * right[ls + d] + r_mn * left[ls + d];
        }
      }
  }
  for (int t = 0; t < g->nexch; ++t) {
    const double *left = g->xl[t], *right = g->xr[t];
    double f = 0.25 * degeneracy * g->xfac[t];
    for (int a = 0; a < ni; ++a) {
      size_t mu = (size_t)(oi + a) * n;
      for (int bb = 0; bb < nj; ++bb) {
        size_t nu = (size_t)(oj + bb) * n;
        double *row = w + ((size_t)a * nj + bb) * nk * nl;
        for (int c = 0; c < nk; ++c) {
          size_t lam = (size_t)(ok + c);
          double l_ml = f * left[mu + lam], r_ml = f * right[mu + lam];
          double l_nl = f * left[nu + lam], r_nl = f * right[nu + lam];
          for (int d = 0; d < nl; ++d) {
            size_t sig = (size_t)(ol + d);
            row[c * nl + d] += l_ml * right[nu + sig] + l_nl * right[mu + sig] +
                               left[mu + sig] * r_nl + left[nu + sig] * r_ml;
          }
        }
      }
    }
  }
  /* pull back with s2c = c2s^T (n_cart x n_func) */
  int dims[4] = {ni, nj, nk, nl}, nout[4];
  const double *mats[4];
  double *s2c[4] = {NULL, NULL, NULL, NULL};
  for (int q = 0; q < 4; ++q) {
    nout[q] = s[q]->ncart;
    mats[q] = NULL;
    if (s[q]->c2s) {
      int nf = s[q]->nfunc, nc = s[q]->ncart;
      s2c[q] = (double *)malloc(sizeof(double) * nf * nc);
      for (int f = 0; f < nf; ++f)
        for (int c = 0; c < nc; ++c) s2c[q][c * nf + f] = s[q]->c2s[f * nc + c];
      mats[q] = s2c[q];
    }
  }
  transform4(w, tmp, dims, mats, nout);
  for (int q = 0; q < 4; ++q) free(s2c[q]);
}

/* raised/lowered block of position pos (direct_grad.rs:554-620) */
static void shifted_block(const or_basis *b, int engine, const int *idx, int pos, int lp,
                          int weighted, double *out, int64_t *npq) {
  or_shref r[4];
  size_t n = 1;
  for (int q = 0; q < 4; ++q) {
    const or_shell *s = &b->sh[idx[q]];
    int l = (q == pos) ? lp : s->l;
    r[q] = shref(s, l, (q == pos && weighted) ? s->eff2a : s->eff);
    n *= (size_t)or_ncart(l);
  }
  memset(out, 0, sizeof(double) * n);
  or_cart_quartet(engine, r, out, npq);
}

/* centre-derivative dot, direct_grad.rs:425-484 */
static void grad_quartet(const or_basis *b, int engine_plan, const int *idx, double degeneracy,
                         const grad_density *g, double *out, double *w, double *tmp,
                         double *raised, double *lowered, int64_t *npq) {
  weight_block(b, g, idx, degeneracy, w, tmp);
  int ls[4], dims[4];
  int deg = 1;
  for (int q = 0; q < 4; ++q) {
    ls[q] = b->sh[idx[q]].l;
    dims[q] = or_ncart(ls[q]);
    deg *= b->sh[idx[q]].nprim;
  }
  int engine = engine_plan;
  if (engine == OR_ENGINE_AUTO) engine = or_select_engine(ls[0] + ls[1], ls[2] + ls[3], deg);
  double fourth[3] = {0, 0, 0};
  for (int pos = 0; pos < 3; ++pos) {
    int l = ls[pos];
    shifted_block(b, engine, idx, pos, l + 1, 1, raised, npq);
    if (l > 0) shifted_block(b, engine, idx, pos, l - 1, 0, lowered, npq);
    int outer = 1, inner = 1;
    for (int x = 0; x < pos; ++x) outer *= dims[x];
    for (int x = pos + 1; x < 4; ++x) inner *= dims[x];
    int nn = dims[pos], n_up = or_ncart(l + 1), n_down = or_ncart(l - 1);
    const int(*comps)[3] = or_cart_comps(l);
    double gg[3] = {0, 0, 0};
    for (int o = 0; o < outer; ++o)
      for (int a = 0; a < nn; ++a) {
        const double *wr = w + ((size_t)o * nn + a) * inner;
        for (int axis = 0; axis < 3; ++axis) {
          int up[3] = {comps[a][0], comps[a][1], comps[a][2]};
          up[axis] += 1;
          const double *rr = raised + ((size_t)o * n_up + or_cart_index(up)) * inner;
          double dot = 0.0;
          for (int s = 0; s < inner; ++s) dot += wr[s] * rr[s];
          gg[axis] += dot;
          if (comps[a][axis] > 0) {
            int dn[3] = {comps[a][0], comps[a][1], comps[a][2]};
            dn[axis] -= 1;
            const double *lr = lowered + ((size_t)o * n_down + or_cart_index(dn)) * inner;
            double dot2 = 0.0;
            for (int s = 0; s < inner; ++s) dot2 += wr[s] * lr[s];
            gg[axis] -= (double)comps[a][axis] * dot2;
          }
        }
      }
    int atom = b->sh[idx[pos]].atom;
    for (int axis = 0; axis < 3; ++axis) {
      out[atom * 3 + axis] += gg[axis];
      fourth[axis] -= gg[axis];
    }
  }
  int atom = b->sh[idx[3]].atom;
  for (int axis = 0; axis < 3; ++axis) out[atom * 3 + axis] += fourth[axis];
}

int or_direct_grad(const or_basis *b, int engine, const double *schwarz, int ncoul,
                   const double *cfac, const double *const *cleft,
                   const double *const *cright, int nexch, const double *xfac,
                   const double *const *xleft, const double *const *xright,
                   double threshold, double *out, int64_t bra_begin, int64_t bra_end,
                   int nthreads, int64_t *stats) {
  for (int s = 0; s < b->nsh; ++s)
    if (b->sh[s].l > OR_MAX_GRAD_L) {
      or_set_error("angular momentum too high for gradient");
      return OR_ERR_L_TOO_HIGH_FOR_GRADIENT;
    }
  int nsh = b->nsh;
  size_t nn = (size_t)nsh * nsh;
  grad_density g;
  g.ncoul = ncoul;
  g.nexch = nexch;
  g.cfac = cfac;
  g.xfac = xfac;
  g.cl = cleft;
  g.cr = cright;
  g.xl = xleft;
  g.xr = xright;
  g.clmax = (double *)malloc(sizeof(double) * nn * (ncoul ? ncoul : 1));
  g.crmax = (double *)malloc(sizeof(double) * nn * (ncoul ? ncoul : 1));
  g.xlmax = (double *)malloc(sizeof(double) * nn * (nexch ? nexch : 1));
  g.xrmax = (double *)malloc(sizeof(double) * nn * (nexch ? nexch : 1));
  for (int t = 0; t < ncoul; ++t) {
    shell_block_max(b, cleft[t], g.clmax + t * nn);
    shell_block_max(b, cright[t], g.crmax + t * nn);
  }
  for (int t = 0; t < nexch; ++t) {
    shell_block_max(b, xleft[t], g.xlmax + t * nn);
    shell_block_max(b, xright[t], g.xrmax + t * nn);
  }
  double *q = (double *)schwarz;
  if (!q) {
    q = (double *)malloc(sizeof(double) * nn);
    or_schwarz(b, engine, q, nthreads);
  }
  int64_t np = npairs(b);
  if (bra_end > np || bra_end < 0) bra_end = np;
  int natom = b->natom;
  int64_t nq = 0, npq = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : nq, npq)
  {
    double *w = (double *)malloc(sizeof(double) * MAXBLK);
    double *tmp = (double *)malloc(sizeof(double) * MAXBLK);
    double *raised = (double *)malloc(sizeof(double) * MAXBLK);
    double *lowered = (double *)malloc(sizeof(double) * MAXBLK);
    double *acc = (double *)calloc((size_t)natom * 3, sizeof(double));
#pragma omp for schedule(dynamic, 1)
    for (int64_t p = bra_begin; p < bra_end; ++p) {
      int i, j;
      pair_of(p, &i, &j);
      double qab = q[i * nsh + j];
      if (qab == 0.0 && threshold > 0.0) continue; /* direct_grad.rs:349-352 */
      for (int k = 0; k <= i; ++k) {
        int l_top = (k == i) ? j : k;
        for (int l = 0; l <= l_top; ++l) {
          int at = b->sh[i].atom;
          if (b->sh[j].atom == at && b->sh[k].atom == at && b->sh[l].atom == at) continue;
          double pair_bound = qab * q[k * nsh + l];
          if (pair_bound < threshold) continue;
          double degeneracy = (i == j ? 1.0 : 2.0) * (k == l ? 1.0 : 2.0) *
                              ((i == k && j == l) ? 1.0 : 2.0);
          if (pair_bound * degeneracy * block_bound(&g, nsh, i, j, k, l) < threshold) continue;
          int idx[4] = {i, j, k, l};
          grad_quartet(b, engine, idx, degeneracy, &g, acc, w, tmp, raised, lowered, &npq);
          ++nq;
        }
      }
    }
#pragma omp critical
    for (int x = 0; x < natom * 3; ++x) out[x] += acc[x];
    free(acc);
    free(w);
    free(tmp);
    free(raised);
    free(lowered);
  }
  if (!schwarz) free(q);
  free(g.clmax);
  free(g.crmax);
  free(g.xlmax);
  free(g.xrmax);
  if (stats) {
    stats[0] = nq;
    stats[1] = npq;
  }
  return OR_OK;
}

/* dense reference over all ordered quartets, all four centres differentiated
 * (grad.rs:426-466, combine_center_derivative grad.rs:838-868, deriv.rs:90-147) */
int or_eri_grad_contract(const or_basis *b, int engine, const double *gamma, double *out,
                         int nthreads) {
  for (int s = 0; s < b->nsh; ++s)
    if (b->sh[s].l > OR_MAX_GRAD_L) {
      or_set_error("angular momentum too high for gradient");
      return OR_ERR_L_TOO_HIGH_FOR_GRADIENT;
    }
  int nsh = b->nsh, natom = b->natom;
  size_t nao = (size_t)b->nao;
  memset(out, 0, sizeof(double) * natom * 3);
#pragma omp parallel num_threads(nthreads)
  {
    double *raised = (double *)malloc(sizeof(double) * MAXBLK);
    double *lowered = (double *)malloc(sizeof(double) * MAXBLK);
    double *der = (double *)malloc(sizeof(double) * MAXBLK);
    double *tmp = (double *)malloc(sizeof(double) * MAXBLK);
    double *acc = (double *)calloc((size_t)natom * 3, sizeof(double));
#pragma omp for schedule(dynamic, 1)
    for (int si = 0; si < nsh; ++si)
      for (int sj = 0; sj < nsh; ++sj)
        for (int sk = 0; sk < nsh; ++sk)
          for (int sl = 0; sl < nsh; ++sl) {
            int idx[4] = {si, sj, sk, sl};
            int ls[4], dims[4], nf[4], offs[4];
            const double *mats[4];
            int deg = 1;
            for (int q = 0; q < 4; ++q) {
              const or_shell *s = &b->sh[idx[q]];
              ls[q] = s->l;
              dims[q] = s->ncart;
              nf[q] = s->nfunc;
              offs[q] = s->off;
              mats[q] = s->c2s;
              deg *= s->nprim;
            }
            int eng = engine == OR_ENGINE_AUTO
                          ? or_select_engine(ls[0] + ls[1], ls[2] + ls[3], deg)
                          : engine;
            size_t nblk = (size_t)dims[0] * dims[1] * dims[2] * dims[3];
            for (int pos = 0; pos < 4; ++pos) {
              int l = ls[pos];
              shifted_block(b, eng, idx, pos, l + 1, 1, raised, NULL);
              if (l > 0) shifted_block(b, eng, idx, pos, l - 1, 0, lowered, NULL);
              int outer = 1, inner = 1;
              for (int x = 0; x < pos; ++x) outer *= dims[x];
              for (int x = pos + 1; x < 4; ++x) inner *= dims[x];
              int n_up = or_ncart(l + 1), n_down = or_ncart(l - 1);
              const int(*comps)[3] = or_cart_comps(l);
              for (int axis = 0; axis < 3; ++axis) {
                for (int o = 0; o < outer; ++o)
                  for (int a = 0; a < dims[pos]; ++a) {
                    int up[3] = {comps[a][0], comps[a][1], comps[a][2]};
                    up[axis] += 1;
                    const double *rr = raised + ((size_t)o * n_up + or_cart_index(up)) * inner;
                    double *dd = der + ((size_t)o * dims[pos] + a) * inner;
                    for (int s = 0; s < inner; ++s) dd[s] = rr[s];
                    if (comps[a][axis] > 0) {
                      int dn[3] = {comps[a][0], comps[a][1], comps[a][2]};
                      dn[axis] -= 1;
                      const double *lr =
                          lowered + ((size_t)o * n_down + or_cart_index(dn)) * inner;
                      double cnt = (double)comps[a][axis];
                      for (int s = 0; s < inner; ++s) dd[s] -= cnt * lr[s];
                    }
                  }
                (void)nblk;
                transform4(der, tmp, dims, mats, nf);
                double dot = 0.0;
                size_t src = 0;
                for (int a = 0; a < nf[0]; ++a)
                  for (int bb = 0; bb < nf[1]; ++bb)
                    for (int c = 0; c < nf[2]; ++c)
                      for (int d = 0; d < nf[3]; ++d)
                        dot += gamma[(((offs[0] + a) * nao + offs[1] + bb) * nao + offs[2] + c) *
                                         nao +
                                     offs[3] + d] *
                               der[src++];
                acc[b->sh[idx[pos]].atom * 3 + axis] += dot;
              }
            }
          }
#pragma omp critical
    for (int x = 0; x < natom * 3; ++x) out[x] += acc[x];
    free(acc);
    free(raised);
    free(lowered);
    free(der);
    free(tmp);
  }
  return OR_OK;
}
