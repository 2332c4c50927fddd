// ig_cls.h -- compile-time quartet classes shared by the kernel translation
// units: class constants, element offsets, geometry loads, the team quartet
// block and the list of compiled classes.  Included inside namespace ig.
#pragma once

#include "ig_device.h"

namespace ig {

__host__ __device__ constexpr int pick_team(int n4) {
  return n4 <= 32 ? 1 : n4 <= 64 ? 2 : n4 <= 128 ? 4 : n4 <= 256 ? 8 : n4 <= 512 ? 16
                                                                   : n4 <= 1024 ? 32 : 64;
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// classes with at most this many Cartesian block elements run lane-per-quartet
// with the block in VGPRs
constexpr int kRegBlockMax = 81;

template <int LA_, int LB_, int LC_, int LD_>
struct Cls {
  static constexpr int LA = LA_, LB = LB_, LC = LC_, LD = LD_;
  static constexpr int na = ncart(LA), nb = ncart(LB), nc = ncart(LC), nd = ncart(LD);
  static constexpr int N4 = na * nb * nc * nd;
  static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
  static constexpr int T = (LA + 1) * (LB + 1) * (LC + 1) * (LD + 1);
  static constexpr int G = pick_team(N4);
  static constexpr int E = (N4 + G - 1) / G;
  static constexpr int TEAMS = 64 / G;
  // tables of one primitive quartet + the records of a chunk (quartet_block)
  static constexpr int WORK = 3 * NR * T + (G < 16 ? G : 16) * (14 + 2 * NR);
};

// Packed per-element table offsets (x | y<<10 | z<<20) of element e.
template <class C>
__device__ __forceinline__ int elem_offsets(int e) {
  int id = e % C::nd;
  int r = e / C::nd;
  int ic = r % C::nc;
  r /= C::nc;
  int ib = r % C::nb;
  int ia = r / C::nb;
  int p[4][3];
  cart_powers(C::LA, ia, p[0][0], p[0][1], p[0][2]);
  cart_powers(C::LB, ib, p[1][0], p[1][1], p[1][2]);
  cart_powers(C::LC, ic, p[2][0], p[2][1], p[2][2]);
  cart_powers(C::LD, id, p[3][0], p[3][1], p[3][2]);
  int packed = 0;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    int o = ((p[0][ax] * (C::LB + 1) + p[1][ax]) * (C::LC + 1) + p[2][ax]) * (C::LD + 1) + p[3][ax];
    packed |= o << (10 * ax);
  }
  return packed;
}

struct QuartetGeom {
  int sA, sB, sC, sD;
  double A[3], B[3], C[3], D[3];
};

__device__ __forceinline__ void load_geom(const BasisDev &b, int sA, int sB, int sC, int sD,
                                          QuartetGeom &g) {
  g.sA = sA;
  g.sB = sB;
  g.sC = sC;
  g.sD = sD;
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    g.A[x] = b.xyz[3 * sA + x];
    g.B[x] = b.xyz[3 * sB + x];
    g.C[x] = b.xyz[3 * sC + x];
    g.D[x] = b.xyz[3 * sD + x];
  }
}

// Per-primitive-quartet records of a team in LDS, filled for a chunk of up
// to PQ primitive quartets at once so the global loads (pair data, Rys
// coefficients) of the whole chunk are in flight together:
//   0 zeta_P, 1 zeta_Q, 2 1/(zeta_P+zeta_Q), 3 1/(2 zeta_P), 4 1/(2 zeta_Q),
//   5-7 P, 8-10 Q, 11 prefactor, 12 T, 13 attenuation s, then NR roots and
//   NR weights.
constexpr int kRecHead = 14;
__host__ __device__ constexpr int rec_size(int nr) { return kRecHead + 2 * nr; }
__host__ __device__ constexpr int rec_chunk(int g) { return g < 16 ? g : 16; }

template <int NR, int G>
__device__ __forceinline__ void prim_records(const PairsDev &p, const RysDev &R, int ob, int ok,
                                             int nkp, int c0, int nc, double omega, int tl,
                                             double *__restrict__ rec) {
  constexpr int RS = rec_size(NR);
  for (int q = tl; q < nc; q += G) {
    const int f = c0 + q, ib = f / nkp, ik = f - ib * nkp;
    const double pz = p.zeta[ob + ib], qz = p.zeta[ok + ik];
    const double P[3] = {p.px[ob + ib], p.py[ob + ib], p.pz[ob + ib]};
    const double Q[3] = {p.px[ok + ik], p.py[ok + ik], p.pz[ok + ik]};
    const double rs = rsqrt(pz + qz), ipq = rs * rs;
    const double rho = pz * qz * ipq;
    double r2 = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x) r2 += (P[x] - Q[x]) * (P[x] - Q[x]);
    double Tq = rho * r2;
    double pref = kTwoPi52 * p.kz[ob + ib] * p.kz[ok + ik] * rs;
    double s_att = 1.0;
    if (omega > 0.0) {  // erf(w r)/r: roots/weights at sT, x -> s x, w -> sqrt(s) w
      s_att = omega * omega / (rho + omega * omega);
      Tq *= s_att;
      pref *= sqrt(s_att);
    }
    double *r = rec + q * RS;
    r[0] = pz;
    r[1] = qz;
    r[2] = ipq;
    r[3] = p.i2z[ob + ib];
    r[4] = p.i2z[ok + ik];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      r[5 + x] = P[x];
      r[8 + x] = Q[x];
    }
    r[11] = pref;
    r[12] = Tq;
    r[13] = s_att;
  }
  team_sync();
  for (int t = tl; t < nc * 2 * NR; t += G) {
    const int q = t / (2 * NR), idx = t - q * (2 * NR);
    double *r = rec + q * RS;
    const double v = rys_series(R, NR, r[12], idx);
    r[kRecHead + idx] = idx < NR ? v * r[13] : v;
  }
  team_sync();
}

// Coefficients of the (root, axis) 2D table of record r: the vertical
// recurrence terms and the transfer distances (ab, cd), selected per axis.
struct AxisTerms {
  double c00, cp0, b10, b01, b00, g00;
};
template <int NR>
__device__ __forceinline__ AxisTerms axis_terms(const double *r, int root, int ax,
                                                const QuartetGeom &g) {
  const double x = r[kRecHead + root];
  const double pz = r[0], qz = r[1], ipq = r[2];
  const double fq = x * qz * ipq, fp = x * pz * ipq;
  const double Pa = r[5 + ax], Qa = r[8 + ax];
  const double Aa = ax == 0 ? g.A[0] : (ax == 1 ? g.A[1] : g.A[2]);
  const double Ca = ax == 0 ? g.C[0] : (ax == 1 ? g.C[1] : g.C[2]);
  AxisTerms t;
  t.c00 = (Pa - Aa) - fq * (Pa - Qa);
  t.cp0 = (Qa - Ca) + fp * (Pa - Qa);
  t.b10 = (1.0 - fq) * r[3];
  t.b01 = (1.0 - fp) * r[4];
  t.b00 = 0.5 * x * ipq;
  // the z table carries the weight and prefactor
  t.g00 = ax == 2 ? r[11] * r[kRecHead + NR + root] : 1.0;
  return t;
}

// Contracted Cartesian block of quartet (pb | pk) accumulated into acc[]
// (element tl + G*k).  `work` is this team's LDS area (C::WORK doubles):
// the 3*NR tables of one primitive quartet, then the records of a chunk.
template <class C>
__device__ __forceinline__ void quartet_block(const PairsDev &p, const RysDev &R, int pb, int pk,
                                              const QuartetGeom &g, double omega, int tl,
                                              const int *offs, double *acc,
                                              double *__restrict__ work) {
  constexpr int G = C::G, NR = C::NR, T = C::T, PQ = rec_chunk(C::G), RS = rec_size(C::NR);
  double *tab = work;
  double *rec = work + 3 * NR * T;
  double AB[3], CD[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    AB[x] = g.A[x] - g.B[x];
    CD[x] = g.C[x] - g.D[x];
  }
  const int ob = p.pp_off[pb], nbp = p.pp_n[pb];
  const int ok = p.pp_off[pk], nkp = p.pp_n[pk];
  const int npq = nbp * nkp;
  for (int c0 = 0; c0 < npq; c0 += PQ) {
    const int nc = npq - c0 < PQ ? npq - c0 : PQ;
    prim_records<NR, G>(p, R, ob, ok, nkp, c0, nc, omega, tl, rec);
    for (int q = 0; q < nc; ++q) {
      const double *r = rec + q * RS;
      // per (root, axis) 2D tables
      for (int task = tl; task < 3 * NR; task += G) {
        const int root = task / 3, ax = task - 3 * root;
        const AxisTerms a = axis_terms<NR>(r, root, ax, g);
        const double ab = ax == 0 ? AB[0] : (ax == 1 ? AB[1] : AB[2]);
        const double cd = ax == 0 ? CD[0] : (ax == 1 ? CD[1] : CD[2]);
        rys_axis_table<C::LA, C::LB, C::LC, C::LD>(a.c00, a.cp0, a.b10, a.b01, a.b00, ab, cd,
                                                   tab + task * T, a.g00);
      }
      team_sync();
      // assembly, one root at a time so only 3E table reads are live
#pragma unroll 1
      for (int root = 0; root < NR; ++root) {
        const double *tx = tab + (3 * root) * T, *ty = tx + T, *tz = ty + T;
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
          if (tl + G * k < C::N4) {
            const int o = offs[k];
            acc[k] += tx[o & 1023] * ty[(o >> 10) & 1023] * tz[o >> 20];
          }
        }
      }
      team_sync();
    }
  }
}

// J/K digestion of one quartet's Cartesian block (LDS, `blk`, element
// ((a*nb+b)*nc+c)*nd+d) by the G lanes of a team:
// Jp[ab] += 2f g D[cd], Jp[cd] += 2f g D[ab], Kp[ac] += f g D[bd],
// Kp[bc] += f g D[ad], Kp[ad] += f g D[bc], Kp[bd] += f g D[ac], with
// f = 1/|stabiliser| of the shell quartet, for each effective density whose
// screening bit is set in `mask`.
template <class C, int G>
__device__ __forceinline__ void team_digest(const BasisDev &b, const DensDev &dd,
                                            const QuartetGeom &g, bool same_pair,
                                            unsigned long long mask, const double *blk, int tl) {
  const int ncart2 = b.nao_cart * b.nao_cart;
  const double f = (g.sA == g.sB ? 0.5 : 1.0) * (g.sC == g.sD ? 0.5 : 1.0) * (same_pair ? 0.5 : 1.0);
  const int a0 = b.off_cart[g.sA], b0 = b.off_cart[g.sB], c0 = b.off_cart[g.sC],
            d0 = b.off_cart[g.sD];
  const int ncrt = b.nao_cart;
  constexpr int na = C::na, nb = C::nb, nc = C::nc, nd = C::nd;
  constexpr int NJ = na * nb + nc * nd;
  const int NO = NJ + (dd.want_k ? na * nc + nb * nc + na * nd + nb * nd : 0);
  for (int e = 0; e < dd.neff; ++e) {
    if (!((mask >> dd.eff_src[e]) & 1ull)) continue;
    const bool anti = dd.eff_anti[e];
    const double *D = dd.dc + (size_t)e * ncart2;
    const size_t xo = jk_copy_offset(dd);
    double *J = dd.jc + xo + (size_t)e * ncart2;
    double *K = dd.kc ? dd.kc + xo + (size_t)e * ncart2 : nullptr;
    for (int o = tl; o < NO; o += G) {
      double v = 0.0;
      int row, col;
      double *dst;
      if (o < na * nb) {  // Jp[a,b]
        if (anti) continue;
        const int a = o / nb, bb = o % nb;
        for (int c = 0; c < nc; ++c)
          for (int d = 0; d < nd; ++d)
            v += blk[((a * nb + bb) * nc + c) * nd + d] * D[(c0 + c) * ncrt + d0 + d];
        v *= 2.0 * f;
        row = a0 + a;
        col = b0 + bb;
        dst = J;
      } else if (o < NJ) {  // Jp[c,d]
        if (anti) continue;
        const int q = o - na * nb, c = q / nd, d = q % nd;
        for (int a = 0; a < na; ++a)
          for (int bb = 0; bb < nb; ++bb)
            v += blk[((a * nb + bb) * nc + c) * nd + d] * D[(a0 + a) * ncrt + b0 + bb];
        v *= 2.0 * f;
        row = c0 + c;
        col = d0 + d;
        dst = J;
      } else {
        int q = o - NJ;
        if (q < na * nc) {  // Kp[a,c] += g D[b,d]
          const int a = q / nc, c = q % nc;
          for (int bb = 0; bb < nb; ++bb)
            for (int d = 0; d < nd; ++d)
              v += blk[((a * nb + bb) * nc + c) * nd + d] * D[(b0 + bb) * ncrt + d0 + d];
          row = a0 + a;
          col = c0 + c;
        } else if ((q -= na * nc) < nb * nc) {  // Kp[b,c] += g D[a,d]
          const int bb = q / nc, c = q % nc;
          for (int a = 0; a < na; ++a)
            for (int d = 0; d < nd; ++d)
              v += blk[((a * nb + bb) * nc + c) * nd + d] * D[(a0 + a) * ncrt + d0 + d];
          row = b0 + bb;
          col = c0 + c;
        } else if ((q -= nb * nc) < na * nd) {  // Kp[a,d] += g D[b,c]
          const int a = q / nd, d = q % nd;
          for (int bb = 0; bb < nb; ++bb)
            for (int c = 0; c < nc; ++c)
              v += blk[((a * nb + bb) * nc + c) * nd + d] * D[(b0 + bb) * ncrt + c0 + c];
          row = a0 + a;
          col = d0 + d;
        } else {  // Kp[b,d] += g D[a,c]
          q -= na * nd;
          const int bb = q / nd, d = q % nd;
          for (int a = 0; a < na; ++a)
            for (int c = 0; c < nc; ++c)
              v += blk[((a * nb + bb) * nc + c) * nd + d] * D[(a0 + a) * ncrt + c0 + c];
          row = b0 + bb;
          col = d0 + d;
        }
        v *= f;
        dst = K;
      }
      jk_add(dd, dst + (size_t)row * ncrt + col, v);
    }
  }
}

// Compile-time element -> per-axis 2D-table offsets for the register path.
template <class C>
struct ElemTab {
  int x[C::N4], y[C::N4], z[C::N4];
  constexpr ElemTab() : x(), y(), z() {
    for (int e = 0; e < C::N4; ++e) {
      int id = e % C::nd, r = e / C::nd, ic = r % C::nc;
      r /= C::nc;
      int ib = r % C::nb, ia = r / C::nb;
      int p[4][3] = {};
      const int ls[4] = {C::LA, C::LB, C::LC, C::LD};
      const int is[4] = {ia, ib, ic, id};
      for (int q = 0; q < 4; ++q) {  // lx descending, ly descending (am.rs:23-31)
        int d = 0;
        while ((d + 1) * (d + 2) / 2 <= is[q]) ++d;
        p[q][0] = ls[q] - d;
        p[q][2] = is[q] - d * (d + 1) / 2;
        p[q][1] = d - p[q][2];
      }
      int o[3] = {};
      for (int ax = 0; ax < 3; ++ax)
        o[ax] = ((p[0][ax] * (C::LB + 1) + p[1][ax]) * (C::LC + 1) + p[2][ax]) * (C::LD + 1) + p[3][ax];
      x[e] = o[0];
      y[e] = o[1];
      z[e] = o[2];
    }
  }
};

template <class C>
__device__ __forceinline__ void init_offsets(int tl, int *offs) {
#pragma unroll
  for (int k = 0; k < C::E; ++k) {
    int e = tl + C::G * k;
    offs[k] = e < C::N4 ? elem_offsets<C>(e) : 0;
  }
}

// Class pairs with pair_class(la,lb) >= pair_class(lc,ld), shells up to d:
// the kernel orientation the plan guarantees.
#define IG_FOR_CLASSES(X)                                                                    \
  X(0, 0, 0, 0)                                                                              \
  X(1, 0, 0, 0) X(1, 0, 1, 0)                                                                \
  X(1, 1, 0, 0) X(1, 1, 1, 0) X(1, 1, 1, 1)                                                  \
  X(2, 0, 0, 0) X(2, 0, 1, 0) X(2, 0, 1, 1) X(2, 0, 2, 0)                                    \
  X(2, 1, 0, 0) X(2, 1, 1, 0) X(2, 1, 1, 1) X(2, 1, 2, 0) X(2, 1, 2, 1)                      \
  X(2, 2, 0, 0) X(2, 2, 1, 0) X(2, 2, 1, 1) X(2, 2, 2, 0) X(2, 2, 2, 1) X(2, 2, 2, 2)

// Class pairs with an f pair on the bra and blocks of at most 3600 Cartesian
// elements: compiled J/K team kernels (the other f kernels, ERI, Schwarz and
// gradient, stay on the runtime-l kernels).
#define IG_FOR_F_TEAM_CLASSES(X)                                                             \
  X(3, 0, 0, 0) X(3, 0, 1, 0) X(3, 0, 1, 1) X(3, 0, 2, 0) X(3, 0, 2, 1) X(3, 0, 2, 2)        \
  X(3, 0, 3, 0)                                                                              \
  X(3, 1, 0, 0) X(3, 1, 1, 0) X(3, 1, 1, 1) X(3, 1, 2, 0) X(3, 1, 2, 1) X(3, 1, 2, 2)        \
  X(3, 1, 3, 0) X(3, 1, 3, 1)                                                                \
  X(3, 2, 0, 0) X(3, 2, 1, 0) X(3, 2, 1, 1) X(3, 2, 2, 0) X(3, 2, 2, 1) X(3, 2, 2, 2)        \
  X(3, 2, 3, 0) X(3, 2, 3, 1) X(3, 2, 3, 2)                                                  \
  X(3, 3, 0, 0) X(3, 3, 1, 0) X(3, 3, 1, 1) X(3, 3, 2, 0) X(3, 3, 2, 1) X(3, 3, 2, 2)        \
  X(3, 3, 3, 0) X(3, 3, 3, 1)

// Classes with only the team kernel with contracted-space transfers: none.
// (ff|fd) and (ff|ff) built that way (blocks of 6000 and 10000 elements in
// LDS) missed the 1e-12 bar on a small f basis (3.6e-12: the degree-3
// transfers on both sides amplify the contracted intermediates), so they
// stay on the runtime-l kernels, which transfer per root.
#define IG_FOR_F_HRR_ONLY_CLASSES(X)

// The f class pairs of at most 270 elements (four register slices) also have
// the lane-per-quartet kernels (register, row sums; FP64 and FP32; batched).
#define IG_FOR_F_LANE_CLASSES(X)                                                             \
  X(3, 0, 0, 0) X(3, 0, 1, 0) X(3, 0, 1, 1) X(3, 0, 2, 0) X(3, 0, 2, 1) X(3, 0, 3, 0)        \
  X(3, 1, 0, 0) X(3, 1, 1, 0) X(3, 1, 1, 1) X(3, 1, 2, 0)                                    \
  X(3, 2, 0, 0) X(3, 2, 1, 0)                                                                \
  X(3, 3, 0, 0)

}  // namespace ig
