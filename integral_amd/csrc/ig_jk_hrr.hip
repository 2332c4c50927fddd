// ig_jk_hrr.hip -- team J/K kernel with the horizontal transfers taken out of
// the primitive loop (J/K variant 4).
//
// The team kernel (ig_kernels.hip) builds, per primitive quartet and root,
// the fully transferred 2D tables I(i,j|k,l) and assembles the whole
// Cartesian block (ab|cd).  Here each primitive quartet only runs the
// vertical recurrence, to I(e,0|f,0) with e <= la+lb and f <= lc+ld, and the
// team assembles the contracted intermediate
//   [E|F] = sum_prim sum_root w I_x(ex,fx) I_y(ey,fy) I_z(ez,fz)
// over the Cartesian triples E of degree la..la+lb and F of degree
// lc..lc+ld.  The horizontal transfers are linear and independent of the
// exponents, so they are applied once per quartet to the contracted [E|F]:
//   (a b| = sum_{t <= b} C(b,t) AB^(b-t) [a+t|      (bra, per F)
//   (ab|c d) = sum_{s <= d} C(d,s) CD^(d-s) (ab|c+s] (ket, per ab)
// with componentwise binomials and powers (the closed form of the recurrence
// (a, b+1_i| = (a+1_i, b| + AB_i (a, b|, engine/rys.rs:234-308 and
// os_eri.rs:3040-3143 apply it level by level).  Per primitive quartet this
// replaces N4 block elements by |E|x|F| (0.6-0.8 of N4 for d and f classes)
// and the per-root transfers by the short vertical tables.
#include "ig_cls.h"

namespace ig {

constexpr int deg_ncart(int lo, int hi) {
  int s = 0;
  for (int l = lo; l <= hi; ++l) s += ncart(l);
  return s;
}
constexpr int binom_small(int n, int k) {
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}
// Cartesian powers of component i of degree l (lx descending, ly descending)
constexpr void cart_triple(int l, int i, int *t) {
  int d = 0;
  while ((d + 1) * (d + 2) / 2 <= i) ++d;
  t[0] = l - d;
  t[2] = i - d * (d + 1) / 2;
  t[1] = d - t[2];
}
// index of triple (x, y, z) in the space of degrees lo, lo+1, ...
constexpr int space_index(int lo, int x, int y, int z) {
  const int d = y + z;
  return deg_ncart(lo, x + y + z - 1) + d * (d + 1) / 2 + z;
}
constexpr void space_triple(int lo, int u, int *t) {
  int l = lo;
  while (u >= ncart(l)) {
    u -= ncart(l);
    ++l;
  }
  cart_triple(l, u, t);
}

template <class C>
struct HrrCls {
  static constexpr int LA = C::LA, LB = C::LB, LC = C::LC, LD = C::LD;
  static constexpr int NB = LA + LB, NK = LC + LD;
  static constexpr int TV = (NB + 1) * (NK + 1);  // one vertical table
  static constexpr int NE = deg_ncart(LA, NB), NF = deg_ncart(LC, NK);
  static constexpr int NEF = NE * NF;
  static constexpr int NR = C::NR;
  // team size of the plain team kernel: the per-quartet stages (transfers,
  // digestion) keep as many lanes busy as there
  static constexpr int G = pick_team(cmax(NEF, C::N4));
  static constexpr int E = (NEF + G - 1) / G;
  static constexpr int TEAMS = 64 / G;
  static constexpr int NH1 = C::na * C::nb * NF;  // bra-transferred (ab|F]
  static constexpr int NT = 8;                     // terms of one transfer, at most
  static constexpr bool useful = LB > 0 || LD > 0;
  // LDS of one team: the vertical tables during the primitive loop; after it
  // R1 (the contracted [E|F], later the block) and R2 ((ab|F], or the block
  // when the bra needs no transfer), then the transfer coefficients
  static constexpr int WORK = 3 * NR * TV + rec_chunk(G) * rec_size(NR);
  static constexpr int R1 = cmax(NEF, C::N4);
  static constexpr int R2 = LB > 0 ? NH1 : (LD > 0 ? C::N4 : 0);
  static constexpr int CO = NT * (C::nb + C::nd);
  static constexpr int PER = cmax(WORK, R1 + R2) + CO;
};

template <class C>
struct HrrTab {
  using H = HrrCls<C>;
  static constexpr int na = C::na, nb = C::nb, nc = C::nc, nd = C::nd, NT = H::NT;
  int vofs[H::NEF];  // vertical-table offsets of [E|F], x | y << 8 | z << 16
  int bn[nb], bexp[nb][NT], bcoef[nb][NT], bsrc[na * nb][NT];
  int kn[nd], kexp[nd][NT], kcoef[nd][NT], ksrc[nc * nd][NT];
  constexpr HrrTab()
      : vofs(), bn(), bexp(), bcoef(), bsrc(), kn(), kexp(), kcoef(), ksrc() {
    for (int u = 0; u < H::NEF; ++u) {
      int e[3] = {}, f[3] = {};
      space_triple(C::LA, u / H::NF, e);
      space_triple(C::LC, u % H::NF, f);
      int o[3] = {};
      for (int ax = 0; ax < 3; ++ax) o[ax] = e[ax] * (H::NK + 1) + f[ax];
      vofs[u] = o[0] | (o[1] << 8) | (o[2] << 16);
    }
    for (int j = 0; j < nb; ++j) {
      int B[3] = {};
      cart_triple(C::LB, j, B);
      int n = 0;
      for (int tx = 0; tx <= B[0]; ++tx)
        for (int ty = 0; ty <= B[1]; ++ty)
          for (int tz = 0; tz <= B[2]; ++tz) {
            bexp[j][n] = (B[0] - tx) | ((B[1] - ty) << 4) | ((B[2] - tz) << 8);
            bcoef[j][n] = binom_small(B[0], tx) * binom_small(B[1], ty) * binom_small(B[2], tz);
            for (int i = 0; i < na; ++i) {
              int A[3] = {};
              cart_triple(C::LA, i, A);
              bsrc[i * nb + j][n] = space_index(C::LA, A[0] + tx, A[1] + ty, A[2] + tz);
            }
            ++n;
          }
      bn[j] = n;
    }
    for (int l = 0; l < nd; ++l) {
      int D[3] = {};
      cart_triple(C::LD, l, D);
      int n = 0;
      for (int sx = 0; sx <= D[0]; ++sx)
        for (int sy = 0; sy <= D[1]; ++sy)
          for (int sz = 0; sz <= D[2]; ++sz) {
            kexp[l][n] = (D[0] - sx) | ((D[1] - sy) << 4) | ((D[2] - sz) << 8);
            kcoef[l][n] = binom_small(D[0], sx) * binom_small(D[1], sy) * binom_small(D[2], sz);
            for (int k = 0; k < nc; ++k) {
              int Cc[3] = {};
              cart_triple(C::LC, k, Cc);
              ksrc[k * nd + l][n] = space_index(C::LC, Cc[0] + sx, Cc[1] + sy, Cc[2] + sz);
            }
            ++n;
          }
      kn[l] = n;
    }
  }
};

// one copy per class in constant memory (indexed at run time)
template <class C>
__constant__ const HrrTab<C> kHrrTab = HrrTab<C>();

// Vertical recurrence of one axis and root: I(n,0|m,0), n <= NB, m <= NK,
// written to out[n*(NK+1)+m] (the first half of rys_axis_table).
template <int NB, int NK>
__device__ __forceinline__ void rys_axis_vrr(double c00, double cp0, double b10, double b01,
                                             double b00, double *__restrict__ out, double g00) {
  double g[NB + 1][NK + 1];
  g[0][0] = g00;
  if constexpr (NB >= 1) g[1][0] = c00 * g00;
#pragma unroll
  for (int n = 1; n < NB; ++n) g[n + 1][0] = c00 * g[n][0] + double(n) * b10 * g[n - 1][0];
#pragma unroll
  for (int m = 0; m < NK; ++m) {
#pragma unroll
    for (int n = 0; n <= NB; ++n) {
      double t = cp0 * g[n][m];
      if (m >= 1) t += double(m) * b01 * g[n][m - 1];
      if (n >= 1) t += double(n) * b00 * g[n - 1][m];
      g[n][m + 1] = t;
    }
  }
#pragma unroll
  for (int n = 0; n <= NB; ++n)
#pragma unroll
    for (int m = 0; m <= NK; ++m) out[n * (NK + 1) + m] = g[n][m];
}

__device__ __forceinline__ double pow_small(double x, int e) {
  double r = 1.0;
  for (int k = 0; k < e; ++k) r *= x;
  return r;
}

template <class C>
__global__ void __launch_bounds__(64) jk_hrr_kernel(BasisDev b, PairsDev p, RysDev R,
                                                    QuartetSet qs, DensDev dd, double omega,
                                                    unsigned long long *counters) {
  using H = HrrCls<C>;
  constexpr int G = H::G, NR = H::NR, TV = H::TV;
  constexpr int nb = C::nb, nc = C::nc, nd = C::nd, NF = H::NF, NT = H::NT;
  const HrrTab<C> &TAB = kHrrTab<C>;
  __shared__ double lds[H::TEAMS * H::PER];
  const int lane = threadIdx.x, team = lane / G, tl = lane % G;
  double *work = lds + team * H::PER;
  double *tab = work, *rec = work + 3 * NR * TV;
  constexpr int PQ = rec_chunk(G), RS = rec_size(NR);
  double *r1 = work, *r2 = work + H::R1;
  double *bco = work + cmax(H::WORK, H::R1 + H::R2), *kco = bco + NT * nb;
  int offs[H::E];
#pragma unroll
  for (int k = 0; k < H::E; ++k) {
    const int u = tl + G * k;
    offs[k] = u < H::NEF ? TAB.vofs[u] : 0;
  }
  unsigned long long my_quartets = 0, my_prims = 0;
  for (long long t = (long long)blockIdx.x * H::TEAMS + team; t < qs.total;
       t += (long long)gridDim.x * H::TEAMS) {
    int P, Q;
    decode_quartet(qs, t, P, Q);
    const int pb = qs.x_off + P, pk = qs.y_off + Q;
    QuartetGeom g;
    load_geom(b, p.sh1[pb], p.sh2[pb], p.sh1[pk], p.sh2[pk], g);
    // density-weighted Schwarz mask (direct.rs:176-190) on symmetrised bounds
    const double pair_bound = p.q[pb] * p.q[pk];
    unsigned long long mask = 0;
    const int nsh = b.nsh;
    for (int d = 0; d < dd.ndens; ++d) {
      const double *bd = dd.bound + (size_t)d * nsh * nsh;
      double m = bd[g.sA * nsh + g.sB];
      m = fmax(m, bd[g.sC * nsh + g.sD]);
      m = fmax(m, bd[g.sA * nsh + g.sC]);
      m = fmax(m, bd[g.sA * nsh + g.sD]);
      m = fmax(m, bd[g.sB * nsh + g.sC]);
      m = fmax(m, bd[g.sB * nsh + g.sD]);
      if (pair_bound * m >= dd.tau) mask |= 1ull << d;
    }
    if (!mask) continue;
    double acc[H::E];
#pragma unroll
    for (int k = 0; k < H::E; ++k) acc[k] = 0.0;
    const int ob = p.pp_off[pb], nbp = p.pp_n[pb];
    const int ok = p.pp_off[pk], nkp = p.pp_n[pk];
    const int npq = nbp * nkp;
    for (int c0 = 0; c0 < npq; c0 += PQ) {
      const int nc = npq - c0 < PQ ? npq - c0 : PQ;
      prim_records<NR, G>(p, R, ob, ok, nkp, c0, nc, omega, tl, rec);
      for (int q = 0; q < nc; ++q) {
        const double *r = rec + q * RS;
        for (int task = tl; task < 3 * NR; task += G) {
          const int root = task / 3, ax = task - 3 * root;
          const AxisTerms a = axis_terms<NR>(r, root, ax, g);
          rys_axis_vrr<H::NB, H::NK>(a.c00, a.cp0, a.b10, a.b01, a.b00, tab + task * TV, a.g00);
        }
        team_sync();
#pragma unroll 1
        for (int root = 0; root < NR; ++root) {  // z tables carry the weight and prefactor
          const double *tx = tab + (3 * root) * TV, *ty = tx + TV, *tz = ty + TV;
#pragma unroll
          for (int k = 0; k < H::E; ++k) {
            if (tl + G * k < H::NEF) {
              const int o = offs[k];
              acc[k] += tx[o & 255] * ty[(o >> 8) & 255] * tz[o >> 16];
            }
          }
        }
        team_sync();
      }
    }
    if (tl == 0) {
      ++my_quartets;
      my_prims += (unsigned long long)nbp * nkp;
    }
    // contracted [E|F] -> R1; transfer coefficients C(b,t) AB^(b-t), C(d,s) CD^(d-s)
#pragma unroll
    for (int k = 0; k < H::E; ++k)
      if (tl + G * k < H::NEF) r1[tl + G * k] = acc[k];
    for (int o = tl; o < NT * (nb + nd); o += G) {
      const bool bra = o < NT * nb;
      const int q = bra ? o : o - NT * nb, j = q / NT, n = q % NT;
      double v = 0.0;
      if (bra ? n < TAB.bn[j] : n < TAB.kn[j]) {
        const int e = bra ? TAB.bexp[j][n] : TAB.kexp[j][n];
        const double dx = bra ? g.A[0] - g.B[0] : g.C[0] - g.D[0];
        const double dy = bra ? g.A[1] - g.B[1] : g.C[1] - g.D[1];
        const double dz = bra ? g.A[2] - g.B[2] : g.C[2] - g.D[2];
        v = double(bra ? TAB.bcoef[j][n] : TAB.kcoef[j][n]) * pow_small(dx, e & 15) *
            pow_small(dy, (e >> 4) & 15) * pow_small(dz, e >> 8);
      }
      bco[o] = v;
    }
    team_sync();
    const double *h1 = r1;  // (ab|F] with ab over na x nb
    if constexpr (C::LB > 0) {
      for (int o = tl; o < H::NH1; o += G) {
        const int ab = o / NF, f = o % NF, j = ab % nb;
        double v = 0.0;
        for (int n = 0; n < TAB.bn[j]; ++n) v += bco[j * NT + n] * r1[TAB.bsrc[ab][n] * NF + f];
        r2[o] = v;
      }
      h1 = r2;
      team_sync();
    }
    const double *blk = h1;
    if constexpr (C::LD > 0) {
      double *out = C::LB > 0 ? r1 : r2;
      for (int o = tl; o < C::N4; o += G) {
        const int ab = o / (nc * nd), cd = o % (nc * nd), l = cd % nd;
        double v = 0.0;
        for (int n = 0; n < TAB.kn[l]; ++n) v += kco[l * NT + n] * h1[ab * NF + TAB.ksrc[cd][n]];
        out[o] = v;
      }
      blk = out;
      team_sync();
    }
    team_digest<C, G>(b, dd, g, pb == pk, mask, blk, tl);
    team_sync();
  }
  if (counters && tl == 0) {
    atomicAdd(counters, my_quartets);
    atomicAdd(counters + 1, my_prims);
  }
}

namespace {

template <int LA, int LB, int LC, int LD>
hipError_t jk_hrr_one(const BasisDev &b, const PairsDev &p, const RysDev &r, const QuartetSet &qs,
                      const DensDev &d, double omega, unsigned long long *ctr, hipStream_t s) {
  using C = Cls<LA, LB, LC, LD>;
  using H = HrrCls<C>;
  if constexpr (!H::useful) {
    return hipErrorInvalidValue;
  } else {
    static_assert(H::TEAMS * H::PER * sizeof(double) <= 160 * 1024, "team LDS over 160 KB");
    if (qs.total <= 0) return hipSuccess;
    long long blocks = (qs.total + H::TEAMS - 1) / H::TEAMS;
    if (blocks > 256 * 64) blocks = 256 * 64;
    hipLaunchKernelGGL(jk_hrr_kernel<C>, dim3(int(blocks)), dim3(64), 0, s, b, p, r, qs, d, omega,
                       ctr);
    return hipGetLastError();
  }
}

}  // namespace

// offered for the classes the team kernel is meant for (blocks over two
// register slices); on smaller ones one lane per quartet is faster
bool jk_hrr_compiled(int la, int lb, int lc, int ld) {
#define IG_CASE(A, B, Cc, D) \
  if (la == A && lb == B && lc == Cc && ld == D) return true;
  IG_FOR_F_HRR_ONLY_CLASSES(IG_CASE)
#undef IG_CASE
  return jk_team_compiled(la, lb, lc, ld) && (lb > 0 || ld > 0) &&
         ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) > 2 * kRegBlockMax;
}

bool jk_compiled(int la, int lb, int lc, int ld) {
  return jk_team_compiled(la, lb, lc, ld) || jk_hrr_compiled(la, lb, lc, ld);
}

hipError_t launch_jk_hrr(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                         const RysDev &r, const QuartetSet &qs, const DensDev &d, double omega,
                         unsigned long long *ctr, hipStream_t s) {
#define IG_CASE(A, B, Cc, D)                       \
  if (la == A && lb == B && lc == Cc && ld == D) \
    return jk_hrr_one<A, B, Cc, D>(b, p, r, qs, d, omega, ctr, s);
  IG_FOR_CLASSES(IG_CASE)
  IG_FOR_F_TEAM_CLASSES(IG_CASE)
  IG_FOR_F_HRR_ONLY_CLASSES(IG_CASE)
#undef IG_CASE
  return hipErrorInvalidValue;
}

}  // namespace ig
