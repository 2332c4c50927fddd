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
  static constexpr int WORK = 2 * NR + 3 * NR * T;  // roots/weights + tables
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

// Contracted Cartesian block of quartet (pb | pk) accumulated into acc[]
// (element tl + G*k).  `work` is this team's LDS area (C::WORK doubles).
template <class C>
__device__ __forceinline__ void quartet_block(const PairsDev &p, const RysDev &R, int pb, int pk,
                                              const QuartetGeom &g, double omega, int tl,
                                              const int *offs, double *acc,
                                              double *__restrict__ work) {
  constexpr int G = C::G, NR = C::NR, T = C::T;
  double *rw = work;
  double *tab = work + 2 * NR;
  double AB[3], CD[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    AB[x] = g.A[x] - g.B[x];
    CD[x] = g.C[x] - g.D[x];
  }
  const int ob = p.pp_off[pb], nbp = p.pp_n[pb];
  const int ok = p.pp_off[pk], nkp = p.pp_n[pk];
  for (int ib = 0; ib < nbp; ++ib) {
    const double pz = p.zeta[ob + ib];
    const double P[3] = {p.px[ob + ib], p.py[ob + ib], p.pz[ob + ib]};
    const double kzp = kTwoPi52 * p.kz[ob + ib], i2p = p.i2z[ob + ib];
    for (int ik = 0; ik < nkp; ++ik) {
      const double qz = p.zeta[ok + ik];
      const double Q[3] = {p.px[ok + ik], p.py[ok + ik], p.pz[ok + ik]};
      const double kzq = p.kz[ok + ik], i2q = p.i2z[ok + ik];
      const double pq = pz + qz;
      const double rs = rsqrt(pq), ipq = rs * rs;
      const double rho = pz * qz * ipq;
      double PQ[3], r2 = 0.0;
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        PQ[x] = P[x] - Q[x];
        r2 += PQ[x] * PQ[x];
      }
      double Tq = rho * r2;
      double pref = kzp * kzq * rs;
      double s_att = 1.0;
      if (omega > 0.0) {  // erf(w r)/r: roots/weights at sT, x -> s x, w -> sqrt(s) w
        s_att = omega * omega / (rho + omega * omega);
        Tq *= s_att;
        pref *= sqrt(s_att);
      }
      // phase 1: roots and weights
      for (int idx = tl; idx < 2 * NR; idx += G) {
        double v = rys_series(R, NR, Tq, idx);
        rw[idx] = (idx < NR) ? v * s_att : v;
      }
      team_sync();
      // phase 2: per (root, axis) 2D tables
      for (int task = tl; task < 3 * NR; task += G) {
        const int r = task / 3, ax = task - 3 * r;
        // per-axis scalars by select, keeping the small arrays in registers
        auto pick = [ax](const double *v) { return ax == 0 ? v[0] : (ax == 1 ? v[1] : v[2]); };
        const double x = rw[r];
        const double fq = x * qz * ipq, fp = x * pz * ipq;
        const double b00 = 0.5 * x * ipq;
        const double b10 = (1.0 - fq) * i2p;
        const double b01 = (1.0 - fp) * i2q;
        const double pqa = pick(PQ);
        const double c00 = (pick(P) - pick(g.A)) - fq * pqa;
        const double cp0 = (pick(Q) - pick(g.C)) + fp * pqa;
        // the z table carries the weight and prefactor
        rys_axis_table<C::LA, C::LB, C::LC, C::LD>(c00, cp0, b10, b01, b00, pick(AB), pick(CD),
                                                   tab + task * T,
                                                   ax == 2 ? pref * rw[NR + r] : 1.0);
      }
      team_sync();
      // phase 3: assembly, one root at a time so only 3E table reads are live
#pragma unroll 1
      for (int r = 0; r < NR; ++r) {
        const double *tx = tab + (3 * r) * T, *ty = tx + T, *tz = ty + T;
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
// elements: compiled J/K team kernels ((ff|fd) and (ff|ff) stay on the
// runtime-l kernels, as do every other f kernel: ERI, Schwarz, gradient).
#define IG_FOR_F_TEAM_CLASSES(X)                                                             \
  X(3, 0, 0, 0) X(3, 0, 1, 0) X(3, 0, 1, 1) X(3, 0, 2, 0) X(3, 0, 2, 1) X(3, 0, 2, 2)        \
  X(3, 0, 3, 0)                                                                              \
  X(3, 1, 0, 0) X(3, 1, 1, 0) X(3, 1, 1, 1) X(3, 1, 2, 0) X(3, 1, 2, 1) X(3, 1, 2, 2)        \
  X(3, 1, 3, 0) X(3, 1, 3, 1)                                                                \
  X(3, 2, 0, 0) X(3, 2, 1, 0) X(3, 2, 1, 1) X(3, 2, 2, 0) X(3, 2, 2, 1) X(3, 2, 2, 2)        \
  X(3, 2, 3, 0) X(3, 2, 3, 1) X(3, 2, 3, 2)                                                  \
  X(3, 3, 0, 0) X(3, 3, 1, 0) X(3, 3, 1, 1) X(3, 3, 2, 0) X(3, 3, 2, 1) X(3, 3, 2, 2)        \
  X(3, 3, 3, 0) X(3, 3, 3, 1)

// The f class pairs of at most 270 elements (four register slices) also have
// the lane-per-quartet kernels (register, row sums; FP64 and FP32; batched).
#define IG_FOR_F_LANE_CLASSES(X)                                                             \
  X(3, 0, 0, 0) X(3, 0, 1, 0) X(3, 0, 1, 1) X(3, 0, 2, 0) X(3, 0, 2, 1) X(3, 0, 3, 0)        \
  X(3, 1, 0, 0) X(3, 1, 1, 0) X(3, 1, 1, 1) X(3, 1, 2, 0)                                    \
  X(3, 2, 0, 0) X(3, 2, 1, 0)                                                                \
  X(3, 3, 0, 0)

}  // namespace ig
