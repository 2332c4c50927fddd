/*
 * or_eri.c -- oracle (test infrastructure only): contracted Cartesian shell-
 * quartet ERI blocks by the two reference engines.
 *   Rys:    crates/integral/src/engine/rys.rs:121-308, driver integrals.rs:1086-1114
 *   OS/HGP: crates/integral/src/engine/os_eri.rs:188-288 (pair data),
 *           :455-697 (contracted VRR), :3040-3143 (HRR in contracted space)
 * Block layout: row-major (a,b,c,d), d fastest, Cartesian order of am.rs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "or_internal.h"

static const double TWO_PI_2_5 = 34.98683665524972; /* 2 pi^{5/2} */

/* ------------------------------------------------------------------------ */
/* Rys engine                                                                */
/* ------------------------------------------------------------------------ */

#define L1 (OR_MAX_L + 1)
#define LT1 (2 * OR_MAX_L + 1)

/* engine/rys.rs:234-308 */
static void build_axis(int la, int lb, int lc, int ld, double c00, double cp0, double b10,
                       double b01, double b00, double ab, double cd, double *out4 /* L1^4 */) {
  int nbra = la + lb, nket = lc + ld;
  double g[LT1][LT1];
  g[0][0] = 1.0;
  if (nbra >= 1) g[1][0] = c00;
  for (int n = 1; n < nbra; ++n) g[n + 1][0] = c00 * g[n][0] + (double)n * b10 * g[n - 1][0];
  for (int m = 0; m < nket; ++m)
    for (int n = 0; n <= nbra; ++n) {
      double term = cp0 * g[n][m];
      if (m >= 1) term += (double)m * b01 * g[n][m - 1];
      if (n >= 1) term += (double)n * b00 * g[n - 1][m];
      g[n][m + 1] = term;
    }
  double h[LT1][L1][LT1];
  for (int i = 0; i <= nbra; ++i)
    for (int m = 0; m <= nket; ++m) h[i][0][m] = g[i][m];
  for (int j = 1; j <= lb; ++j)
    for (int i = 0; i <= nbra - j; ++i)
      for (int m = 0; m <= nket; ++m) h[i][j][m] = h[i + 1][j - 1][m] + ab * h[i][j - 1][m];
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      double t[LT1][L1];
      for (int k = 0; k <= nket; ++k) t[k][0] = h[i][j][k];
      for (int l = 1; l <= ld; ++l)
        for (int k = 0; k <= nket - l; ++k) t[k][l] = t[k + 1][l - 1] + cd * t[k][l - 1];
      for (int k = 0; k <= lc; ++k)
        for (int l = 0; l <= ld; ++l) out4[((i * L1 + j) * L1 + k) * L1 + l] = t[k][l];
    }
}

/* one primitive quartet, engine/rys.rs:121-226 */
static void rys_primitive(const or_shref *s, int pa, int pb, int pc, int pd, double scale,
                          double *out) {
  int la = s[0].l, lb = s[1].l, lc = s[2].l, ld = s[3].l;
  double ea = s[0].exps[pa], eb = s[1].exps[pb], ec = s[2].exps[pc], ed = s[3].exps[pd];
  const double *A = s[0].center, *B = s[1].center, *C = s[2].center, *D = s[3].center;
  double p = ea + eb, q = ec + ed;
  double P[3], Q[3], pq_vec[3], pa_[3], qc_[3], ab_[3], cd_[3];
  double dab = 0.0, dcd = 0.0, rpq = 0.0;
  for (int x = 0; x < 3; ++x) {
    P[x] = (ea * A[x] + eb * B[x]) / p;
    Q[x] = (ec * C[x] + ed * D[x]) / q;
    dab += (A[x] - B[x]) * (A[x] - B[x]);
    dcd += (C[x] - D[x]) * (C[x] - D[x]);
  }
  double kab = exp(-(ea * eb / p) * dab);
  double kcd = exp(-(ec * ed / q) * dcd);
  double pqs = p + q;
  double rho = p * q / pqs;
  for (int x = 0; x < 3; ++x) {
    pq_vec[x] = P[x] - Q[x];
    rpq += pq_vec[x] * pq_vec[x];
    pa_[x] = P[x] - A[x];
    qc_[x] = Q[x] - C[x];
    ab_[x] = A[x] - B[x];
    cd_[x] = C[x] - D[x];
  }
  double t = rho * rpq;
  double pref = 2.0 * M_PI * M_PI * sqrt(M_PI) / (p * q * sqrt(pqs)) * kab * kcd;
  int nroots = (la + lb + lc + ld) / 2 + 1;
  double roots[OR_MAX_RYS], wts[OR_MAX_RYS];
  or_rys(nroots, t, roots, wts);

  int na = or_ncart(la), nb = or_ncart(lb), nc = or_ncart(lc), nd = or_ncart(ld);
  static __thread double ix[L1 * L1 * L1 * L1], iy[L1 * L1 * L1 * L1], iz[L1 * L1 * L1 * L1];
  for (int r = 0; r < nroots; ++r) {
    double x = roots[r];
    double b00 = x / (2.0 * pqs);
    double b10 = (1.0 - x * q / pqs) / (2.0 * p);
    double b01 = (1.0 - x * p / pqs) / (2.0 * q);
    double cq = x * q / pqs, cp = x * p / pqs;
    double *tabs[3] = {ix, iy, iz};
    for (int ax = 0; ax < 3; ++ax) {
      double c00 = pa_[ax] - cq * pq_vec[ax];
      double cp0 = qc_[ax] + cp * pq_vec[ax];
      build_axis(la, lb, lc, ld, c00, cp0, b10, b01, b00, ab_[ax], cd_[ax], tabs[ax]);
    }
    double pw = scale * pref * wts[r];
    const int(*ca)[3] = or_cart_comps(la);
    const int(*cb)[3] = or_cart_comps(lb);
    const int(*cc)[3] = or_cart_comps(lc);
    const int(*cdd)[3] = or_cart_comps(ld);
    for (int ia = 0; ia < na; ++ia)
      for (int ib = 0; ib < nb; ++ib)
        for (int ic = 0; ic < nc; ++ic)
          for (int id = 0; id < nd; ++id) {
#define IDX(t, a, b, c, d) t[(((a)*L1 + (b)) * L1 + (c)) * L1 + (d)]
            double v = IDX(ix, ca[ia][0], cb[ib][0], cc[ic][0], cdd[id][0]) *
                       IDX(iy, ca[ia][1], cb[ib][1], cc[ic][1], cdd[id][1]) *
                       IDX(iz, ca[ia][2], cb[ib][2], cc[ic][2], cdd[id][2]);
#undef IDX
            out[((ia * nb + ib) * nc + ic) * nd + id] += pw * v;
          }
  }
}

/* integrals.rs:1086-1114 */
static void contract_rys(const or_shref *s, double *out, int64_t *nprimq) {
  for (int a = 0; a < s[0].nprim; ++a)
    for (int b = 0; b < s[1].nprim; ++b)
      for (int c = 0; c < s[2].nprim; ++c)
        for (int d = 0; d < s[3].nprim; ++d) {
          double scale = s[0].coefs[a] * s[1].coefs[b] * s[2].coefs[c] * s[3].coefs[d];
          rys_primitive(s, a, b, c, d, scale, out);
        }
  if (nprimq) *nprimq += (int64_t)s[0].nprim * s[1].nprim * s[2].nprim * s[3].nprim;
}

/* ------------------------------------------------------------------------ */
/* OS/HGP engine                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
  double zeta, center[3], kappa, c1, c2, inv_2zeta, r1[3];
} prim_pair;

#define PAIR_NEGLIGIBLE 1e-32 /* os_eri.rs:211 */

/* os_eri.rs:221-249 */
static int build_pairs(const or_shref *s1, const or_shref *s2, prim_pair *out) {
  double d2 = 0.0;
  for (int x = 0; x < 3; ++x) d2 += (s1->center[x] - s2->center[x]) * (s1->center[x] - s2->center[x]);
  int n = 0;
  for (int i = 0; i < s1->nprim; ++i)
    for (int j = 0; j < s2->nprim; ++j) {
      double e1 = s1->exps[i], e2 = s2->exps[j];
      double c1 = s1->coefs[i], c2 = s2->coefs[j];
      double zeta = e1 + e2;
      double kappa = exp(-(e1 * e2 / zeta) * d2);
      if (kappa * fabs(c1 * c2) < PAIR_NEGLIGIBLE) continue;
      prim_pair *pp = &out[n++];
      pp->zeta = zeta;
      for (int x = 0; x < 3; ++x) {
        pp->center[x] = (e1 * s1->center[x] + e2 * s2->center[x]) / zeta;
        pp->r1[x] = pp->center[x] - s1->center[x];
      }
      pp->kappa = kappa;
      pp->c1 = c1;
      pp->c2 = c2;
      pp->inv_2zeta = 0.5 / zeta;
    }
  return n;
}

static int first_axis(const int *t) { return t[0] > 0 ? 0 : (t[1] > 0 ? 1 : 2); }

/* Contracted VRR [e0|f0]^(0) for all deg(e)<=ne, deg(f)<=nf (os_eri.rs:26-41
 * relations), scaled by ((ca cb) cc) cd and summed over primitive quartets. */
static void os_contract(const or_shref *s, const prim_pair *bra, int nbra, const prim_pair *ket,
                        int nket, double *cef /* naddr(ne) x naddr(nf) */, double *work) {
  int ne = s[0].l + s[1].l, nf = s[2].l + s[3].l, L = ne + nf;
  int n_e = or_naddr(ne), n_f = or_naddr(nf);
  int M = L + 1;
  double *V = work; /* [ae][af][m] */
  double F[4 * OR_MAX_L + 2];
  for (int ib = 0; ib < nbra; ++ib) {
    const prim_pair *bp = &bra[ib];
    double p = bp->zeta;
    double bc = bp->c1 * bp->c2;
    for (int ik = 0; ik < nket; ++ik) {
      const prim_pair *kp = &ket[ik];
      double q = kp->zeta;
      double pq = p + q;
      double rpq = 0.0, W[3], WP[3], WQ[3];
      for (int x = 0; x < 3; ++x) {
        double d = bp->center[x] - kp->center[x];
        rpq += d * d;
        W[x] = (p * bp->center[x] + q * kp->center[x]) / pq;
        WP[x] = W[x] - bp->center[x];
        WQ[x] = W[x] - kp->center[x];
      }
      double t = (p * q / pq) * rpq;
      double pref = TWO_PI_2_5 / (p * q * sqrt(pq)) * bp->kappa * kp->kappa;
      or_boys(L, t, F);
      double scale = (bc * kp->c1) * kp->c2;
      double i2p = bp->inv_2zeta, i2q = kp->inv_2zeta, i2pq = 0.5 / pq;
      double q_pq = q / pq, p_pq = p / pq;
      /* [00|00]^m */
      for (int m = 0; m <= L; ++m) V[m] = pref * F[m];
      /* bra raise with f = 0 */
      for (int de = 1; de <= ne; ++de) {
        const int(*ce)[3] = or_cart_comps(de);
        for (int ie = 0; ie < or_ncart(de); ++ie) {
          int e[3] = {ce[ie][0], ce[ie][1], ce[ie][2]};
          int i = first_axis(e);
          int e1[3] = {e[0], e[1], e[2]};
          e1[i] -= 1;
          int a1 = or_addr(e1);
          int cnt = e1[i];
          int a2 = -1;
          if (cnt > 0) {
            int e2[3] = {e1[0], e1[1], e1[2]};
            e2[i] -= 1;
            a2 = or_addr(e2);
          }
          double *dst = V + (size_t)or_addr(e) * n_f * M;
          const double *s1 = V + (size_t)a1 * n_f * M;
          const double *s2 = a2 >= 0 ? V + (size_t)a2 * n_f * M : NULL;
          for (int m = 0; m <= L - de; ++m) {
            double v = bp->r1[i] * s1[m] + WP[i] * s1[m + 1];
            if (cnt > 0) v += (double)cnt * i2p * (s2[m] - q_pq * s2[m + 1]);
            dst[m] = v;
          }
        }
      }
      /* ket raise */
      for (int df = 1; df <= nf; ++df) {
        const int(*cf)[3] = or_cart_comps(df);
        for (int jf = 0; jf < or_ncart(df); ++jf) {
          int f[3] = {cf[jf][0], cf[jf][1], cf[jf][2]};
          int j = first_axis(f);
          int f1[3] = {f[0], f[1], f[2]};
          f1[j] -= 1;
          int af = or_addr(f), af1 = or_addr(f1);
          int cntf = f1[j];
          int af2 = -1;
          if (cntf > 0) {
            int f2[3] = {f1[0], f1[1], f1[2]};
            f2[j] -= 1;
            af2 = or_addr(f2);
          }
          for (int de = 0; de <= ne; ++de) {
            const int(*ce)[3] = or_cart_comps(de);
            for (int ie = 0; ie < or_ncart(de); ++ie) {
              int e[3] = {ce[ie][0], ce[ie][1], ce[ie][2]};
              int ae = or_addr(e);
              int cnte = e[j];
              int aem = -1;
              if (cnte > 0) {
                int em[3] = {e[0], e[1], e[2]};
                em[j] -= 1;
                aem = or_addr(em);
              }
              double *dst = V + ((size_t)ae * n_f + af) * M;
              const double *s1 = V + ((size_t)ae * n_f + af1) * M;
              const double *s2 = af2 >= 0 ? V + ((size_t)ae * n_f + af2) * M : NULL;
              const double *s3 = aem >= 0 ? V + ((size_t)aem * n_f + af1) * M : NULL;
              for (int m = 0; m <= L - de - df; ++m) {
                double v = kp->r1[j] * s1[m] + WQ[j] * s1[m + 1];
                if (cntf > 0) v += (double)cntf * i2q * (s2[m] - p_pq * s2[m + 1]);
                if (cnte > 0) v += (double)cnte * i2pq * s3[m + 1];
                dst[m] = v;
              }
            }
          }
        }
      }
      /* contract [e0|f0]^(0) */
      for (int ae = or_naddr(s[0].l - 1); ae < n_e; ++ae)
        for (int af = or_naddr(s[2].l - 1); af < n_f; ++af)
          cef[(size_t)ae * n_f + af] += scale * V[((size_t)ae * n_f + af) * M];
    }
  }
}

/* HRR A->B then C->D in contracted space (os_eri.rs:3040-3143 relations). */
static void os_hrr(const or_shref *s, const double *cef, double *out, double *work) {
  int la = s[0].l, lb = s[1].l, lc = s[2].l, ld = s[3].l;
  int ne = la + lb, nf = lc + ld;
  int n_e = or_naddr(ne), n_f = or_naddr(nf);
  double AB[3], CD[3];
  for (int x = 0; x < 3; ++x) {
    AB[x] = s[0].center[x] - s[1].center[x];
    CD[x] = s[2].center[x] - s[3].center[x];
  }
  /* bra: X_j[a(addr)][b(cart idx deg j)][f]  */
  size_t layer = (size_t)n_e * 28 * n_f;
  double *X0 = work, *X1 = work + layer;
  for (int ae = 0; ae < n_e; ++ae)
    for (int af = 0; af < n_f; ++af) X0[((size_t)ae * 28 + 0) * n_f + af] = cef[(size_t)ae * n_f + af];
  for (int j = 1; j <= lb; ++j) {
    const int(*cb)[3] = or_cart_comps(j);
    for (int da = la; da <= ne - j; ++da) {
      const int(*ca)[3] = or_cart_comps(da);
      for (int ia = 0; ia < or_ncart(da); ++ia) {
        int a[3] = {ca[ia][0], ca[ia][1], ca[ia][2]};
        int aa = or_addr(a);
        for (int ib = 0; ib < or_ncart(j); ++ib) {
          int b[3] = {cb[ib][0], cb[ib][1], cb[ib][2]};
          int i = first_axis(b);
          int bm[3] = {b[0], b[1], b[2]};
          bm[i] -= 1;
          int ibm = or_cart_index(bm);
          int ap[3] = {a[0], a[1], a[2]};
          ap[i] += 1;
          int aap = or_addr(ap);
          double *dst = X1 + ((size_t)aa * 28 + ib) * n_f;
          const double *s1 = X0 + ((size_t)aap * 28 + ibm) * n_f;
          const double *s2 = X0 + ((size_t)aa * 28 + ibm) * n_f;
          for (int af = 0; af < n_f; ++af) dst[af] = s1[af] + AB[i] * s2[af];
        }
      }
    }
    double *tmp = X0;
    X0 = X1;
    X1 = tmp;
  }
  /* X0 now holds (a b | f) for deg a = la, deg b = lb */
  int na = or_ncart(la), nb = or_ncart(lb), nc = or_ncart(lc), nd = or_ncart(ld);
  /* ket: Z_k[c(addr)][d(cart idx deg k)] per (a,b) */
  double *Ybuf0 = (double *)malloc(sizeof(double) * (size_t)n_f * 28);
  double *Ybuf1 = (double *)malloc(sizeof(double) * (size_t)n_f * 28);
  const int(*ca)[3] = or_cart_comps(la);
  for (int ia = 0; ia < na; ++ia) {
    int aa = or_addr(ca[ia]);
    for (int ib = 0; ib < nb; ++ib) {
      const double *src = X0 + ((size_t)aa * 28 + ib) * n_f;
      double *Z0 = Ybuf0, *Z1 = Ybuf1;
      for (int af = 0; af < n_f; ++af) Z0[(size_t)af * 28] = src[af];
      for (int k = 1; k <= ld; ++k) {
        const int(*cdk)[3] = or_cart_comps(k);
        for (int dc = lc; dc <= nf - k; ++dc) {
          const int(*cc)[3] = or_cart_comps(dc);
          for (int ic = 0; ic < or_ncart(dc); ++ic) {
            int c[3] = {cc[ic][0], cc[ic][1], cc[ic][2]};
            int ac = or_addr(c);
            for (int id = 0; id < or_ncart(k); ++id) {
              int d[3] = {cdk[id][0], cdk[id][1], cdk[id][2]};
              int i = first_axis(d);
              int dm[3] = {d[0], d[1], d[2]};
              dm[i] -= 1;
              int idm = or_cart_index(dm);
              int cp[3] = {c[0], c[1], c[2]};
              cp[i] += 1;
              int acp = or_addr(cp);
              Z1[(size_t)ac * 28 + id] = Z0[(size_t)acp * 28 + idm] + CD[i] * Z0[(size_t)ac * 28 + idm];
            }
          }
        }
        double *tmp = Z0;
        Z0 = Z1;
        Z1 = tmp;
      }
      const int(*cc)[3] = or_cart_comps(lc);
      for (int ic = 0; ic < nc; ++ic) {
        int ac = or_addr(cc[ic]);
        for (int id = 0; id < nd; ++id)
          out[((ia * nb + ib) * nc + ic) * nd + id] += Z0[(size_t)ac * 28 + id];
      }
    }
  }
  free(Ybuf0);
  free(Ybuf1);
}

static void contract_os(const or_shref *s, double *out, int64_t *nprimq) {
  int ne = s[0].l + s[1].l, nf = s[2].l + s[3].l, L = ne + nf;
  int n_e = or_naddr(ne), n_f = or_naddr(nf);
  prim_pair *bra = (prim_pair *)malloc(sizeof(prim_pair) * (size_t)s[0].nprim * s[1].nprim);
  prim_pair *ket = (prim_pair *)malloc(sizeof(prim_pair) * (size_t)s[2].nprim * s[3].nprim);
  int nb = build_pairs(&s[0], &s[1], bra);
  int nk = build_pairs(&s[2], &s[3], ket);
  double *cef = (double *)calloc((size_t)n_e * n_f, sizeof(double));
  size_t vsz = (size_t)n_e * n_f * (L + 1);
  size_t hsz = 2 * (size_t)n_e * 28 * n_f;
  double *work = (double *)malloc(sizeof(double) * (vsz > hsz ? vsz : hsz));
  os_contract(s, bra, nb, ket, nk, cef, work);
  os_hrr(s, cef, out, work);
  if (nprimq) *nprimq += (int64_t)nb * nk;
  free(work);
  free(cef);
  free(bra);
  free(ket);
}

/* integrals.rs:374-388 */
int or_select_engine(int ne, int nf, int degree) {
  int l_total = ne + nf;
  int fast = (ne > nf ? ne : nf) <= 6;
  int threshold;
  if (l_total <= 5)
    threshold = 1;
  else if (l_total <= 16)
    threshold = fast ? 3 : 16;
  else
    return OR_ENGINE_RYS;
  return degree >= threshold ? OR_ENGINE_OS : OR_ENGINE_RYS;
}

/* contracted Cartesian block, accumulated (+=) into out */
void or_cart_quartet(int engine, const or_shref *s, double *out, int64_t *nprimq) {
  if (engine == OR_ENGINE_AUTO)
    engine = or_select_engine(s[0].l + s[1].l, s[2].l + s[3].l,
                              s[0].nprim * s[1].nprim * s[2].nprim * s[3].nprim);
  if (engine == OR_ENGINE_OS)
    contract_os(s, out, nprimq);
  else
    contract_rys(s, out, nprimq);
}
