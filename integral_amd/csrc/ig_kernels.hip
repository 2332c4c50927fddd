// ig_kernels.hip -- gfx950 shell-quartet kernels (Rys quadrature) with fused
// epilogues: Schwarz diagonal, integral-direct J/K digestion, dense tensor
// scatter.  One workgroup is one wavefront; a wavefront holds 64/G teams and
// a team of G lanes owns one shell quartet at a time:
//   phase 1  the team's lanes evaluate the 2n Rys root/weight series   -> LDS
//   phase 2  the lanes build the 3n per-axis 2D integral tables         -> LDS
//   phase 3  every lane assembles its slice of the Cartesian block
//            (sum over roots of w Ix Iy Iz) into registers
// repeated over the primitive quartets (early contraction), then the block
// goes to LDS for the epilogue.  G is the smallest power of two that keeps
// the block at <= 32 doubles per lane.
#include "ig_cls.h"

namespace ig {
// ---------------------------------------------------------------------------
// J/K: Jp[ab] += 2f g D[cd], Jp[cd] += 2f g D[ab], Kp[ac] += f g D[bd],
// Kp[bc] += f g D[ad], Kp[ad] += f g D[bc], Kp[bd] += f g D[ac]  with
// f = 1/|stabiliser| of the shell quartet; J = Jp + Jp^T, K = Kp +- Kp^T.
// This equals the element scatter of direct.rs:284-404 summed over the
// canonical quartets (exact for any density; the antisymmetric part of a
// non-symmetric D only enters K).

template <class C>
__global__ void __launch_bounds__(64) jk_kernel(BasisDev b, PairsDev p, RysDev R, QuartetSet qs,
                                                DensDev dd, double omega,
                                                unsigned long long *counters) {
  constexpr int G = C::G, PER = cmax(C::WORK, C::N4);
  __shared__ double lds[C::TEAMS * PER];
  const int lane = threadIdx.x, team = lane / G, tl = lane % G;
  double *work = lds + team * PER;
  int offs[C::E];
  init_offsets<C>(tl, offs);
  unsigned long long my_quartets = 0, my_prims = 0;
  for (long long t = (long long)blockIdx.x * C::TEAMS + team; t < qs.total;
       t += (long long)gridDim.x * C::TEAMS) {
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
    double acc[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) acc[k] = 0.0;
    quartet_block<C>(p, R, pb, pk, g, omega, tl, offs, acc, work);
    if (tl == 0) {
      ++my_quartets;
      my_prims += (unsigned long long)p.pp_n[pb] * p.pp_n[pk];
    }
    // block -> LDS
    double *blk = work;
#pragma unroll
    for (int k = 0; k < C::E; ++k)
      if (tl + G * k < C::N4) blk[tl + G * k] = acc[k];
    team_sync();
    team_digest<C, G>(b, dd, g, pb == pk, mask, blk, tl);
    team_sync();
  }
  if (counters && tl == 0) {
    atomicAdd(counters, my_quartets);
    atomicAdd(counters + 1, my_prims);
  }
}

// ---------------------------------------------------------------------------
// Schwarz: Q[P] = sqrt(max_{mu,nu} |(mu nu|mu nu)|) in function space
// (integrals.rs:842-862) from the Cartesian (P|P) block.
template <class C>
__global__ void __launch_bounds__(64) schwarz_kernel(BasisDev b, PairsDev p, RysDev R, int first,
                                                     int count, double *qout) {
  static_assert(C::LA == C::LC && C::LB == C::LD, "diagonal class");
  constexpr int G = C::G, PER = cmax(C::WORK, C::N4 + G);
  __shared__ double lds[C::TEAMS * PER];
  const int lane = threadIdx.x, team = lane / G, tl = lane % G;
  double *work = lds + team * PER;
  int offs[C::E];
  init_offsets<C>(tl, offs);
  for (int t = blockIdx.x * C::TEAMS + team; t < count; t += gridDim.x * C::TEAMS) {
    const int pr = first + t;
    QuartetGeom g;
    load_geom(b, p.sh1[pr], p.sh2[pr], p.sh1[pr], p.sh2[pr], g);
    double acc[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) acc[k] = 0.0;
    quartet_block<C>(p, R, pr, pr, g, 0.0, tl, offs, acc, work);
    double *blk = work;
#pragma unroll
    for (int k = 0; k < C::E; ++k)
      if (tl + G * k < C::N4) blk[tl + G * k] = acc[k];
    team_sync();
    constexpr int na = C::na, nb = C::nb;
    const bool sa = b.kind[g.sA], sb = b.kind[g.sB];
    const int nfa = sa ? 2 * C::LA + 1 : na, nfb = sb ? 2 * C::LB + 1 : nb;
    const double *Ma = b.c2s + b.c2s_off[C::LA], *Mb = b.c2s + b.c2s_off[C::LB];
    double mx = 0.0;
    for (int o = tl; o < nfa * nfb; o += G) {
      const int mu = o / nfb, nu = o % nfb;
      double v = 0.0;
      for (int a = 0; a < na; ++a) {
        const double ma = sa ? Ma[mu * na + a] : (a == mu ? 1.0 : 0.0);
        if (ma == 0.0) continue;
        for (int bb = 0; bb < nb; ++bb) {
          const double mb = sb ? Mb[nu * nb + bb] : (bb == nu ? 1.0 : 0.0);
          if (mb == 0.0) continue;
          for (int c = 0; c < na; ++c) {
            const double mc = sa ? Ma[mu * na + c] : (c == mu ? 1.0 : 0.0);
            if (mc == 0.0) continue;
            for (int d = 0; d < nb; ++d) {
              const double md = sb ? Mb[nu * nb + d] : (d == nu ? 1.0 : 0.0);
              if (md == 0.0) continue;
              v += ma * mb * mc * md * blk[((a * nb + bb) * na + c) * nb + d];
            }
          }
        }
      }
      mx = fmax(mx, fabs(v));
    }
    double *red = work + C::N4;
    team_sync();
    red[tl] = mx;
    team_sync();
    if (tl == 0) {
      for (int i = 1; i < G; ++i) mx = fmax(mx, red[i]);
      qout[pr] = sqrt(mx);
    }
    team_sync();
  }
}

// ---------------------------------------------------------------------------
// Dense tensor: c2s in LDS, then each distinct shell permutation of the
// quartet written once (integrals.rs:1271-1312).
__device__ __constant__ int kPerms8[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2},
                                             {1, 0, 3, 2}, {2, 3, 0, 1}, {2, 3, 1, 0},
                                             {3, 2, 0, 1}, {3, 2, 1, 0}};

template <class C>
__global__ void __launch_bounds__(64) eri_kernel(BasisDev b, PairsDev p, RysDev R, QuartetSet qs,
                                                 EriOut eo) {
  double *out = eo.out;
  constexpr int G = C::G, PER = cmax(C::WORK, 2 * C::N4);
  __shared__ double lds[C::TEAMS * PER];
  const int lane = threadIdx.x, team = lane / G, tl = lane % G;
  double *work = lds + team * PER;
  int offs[C::E];
  init_offsets<C>(tl, offs);
  const size_t nao = (size_t)b.nao;
  for (long long t = (long long)blockIdx.x * C::TEAMS + team; t < qs.total;
       t += (long long)gridDim.x * C::TEAMS) {
    int P, Q;
    decode_quartet(qs, t, P, Q);
    const int pb = qs.x_off + P, pk = qs.y_off + Q;
    QuartetGeom g;
    load_geom(b, p.sh1[pb], p.sh2[pb], p.sh1[pk], p.sh2[pk], g);
    double acc[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) acc[k] = 0.0;
    quartet_block<C>(p, R, pb, pk, g, eo.omega, tl, offs, acc, work);
    double *cur = work, *nxt = work + C::N4;
#pragma unroll
    for (int k = 0; k < C::E; ++k)
      if (tl + G * k < C::N4) cur[tl + G * k] = acc[k];
    team_sync();
    const int sh[4] = {g.sA, g.sB, g.sC, g.sD};
    const int ls[4] = {C::LA, C::LB, C::LC, C::LD};
    int dims[4] = {C::na, C::nb, C::nc, C::nd};
    for (int ax = 0; ax < 4; ++ax) {
      if (!b.kind[sh[ax]]) continue;
      const int nin = dims[ax], nout = 2 * ls[ax] + 1;
      const double *M = b.c2s + b.c2s_off[ls[ax]];
      int outer = 1, inner = 1;
      for (int x = 0; x < ax; ++x) outer *= dims[x];
      for (int x = ax + 1; x < 4; ++x) inner *= dims[x];
      for (int o = tl; o < outer * nout * inner; o += G) {
        const int i = o % inner, rq = o / inner, q = rq % nout, oo = rq / nout;
        double v = 0.0;
        for (int j = 0; j < nin; ++j) v += M[q * nin + j] * cur[(oo * nin + j) * inner + i];
        nxt[o] = v;
      }
      dims[ax] = nout;
      double *tmp = cur;
      cur = nxt;
      nxt = tmp;
      team_sync();
    }
    const int nf = dims[0] * dims[1] * dims[2] * dims[3];
    if (eo.mode == kDf3c || eo.mode == kDf2c) {
      for (int e = tl; e < nf; e += G) {
        const int id = e % dims[3], r1 = e / dims[3], ic = r1 % dims[2], r2 = r1 / dims[2],
                  ibb = r2 % dims[1], ia = r2 / dims[1];
        const int idx[4] = {ia, ibb, ic, id};
        df_store(b, eo, sh, idx, cur[e]);
      }
      team_sync();
      continue;
    }
    if (eo.start) {  // packed: the canonical element only
      for (int e = tl; e < nf; e += G) {
        const int id = e % dims[3], r1 = e / dims[3], ic = r1 % dims[2], r2 = r1 / dims[2],
                  ibb = r2 % dims[1], ia = r2 / dims[1];
        packed_store(b, eo.start, eo.row0, out, sh[0], ia, sh[1], ibb, sh[2], ic, sh[3], id,
                     cur[e]);
      }
      team_sync();
      continue;
    }
    int offf[4];
    for (int x = 0; x < 4; ++x) offf[x] = b.off_func[sh[x]];
    int seen[8][4];
    int nseen = 0;
    for (int pi = 0; pi < 8; ++pi) {
      int tup[4];
      for (int q = 0; q < 4; ++q) tup[q] = sh[kPerms8[pi][q]];
      bool dup = false;
      for (int s = 0; s < nseen; ++s)
        dup |= (seen[s][0] == tup[0] && seen[s][1] == tup[1] && seen[s][2] == tup[2] &&
                seen[s][3] == tup[3]);
      if (dup) continue;
      for (int q = 0; q < 4; ++q) seen[nseen][q] = tup[q];
      ++nseen;
      if (eo.rows_end >= 0) {  // rows owned by the bra-pair range only
        const long long hi = tup[0] > tup[1] ? tup[0] : tup[1], lo = tup[0] + tup[1] - hi;
        const long long cp = hi * (hi + 1) / 2 + lo;
        if (cp < eo.rows_begin || cp >= eo.rows_end) continue;
      }
      size_t stride[4];
      size_t base = 0;
      for (int q = 0; q < 4; ++q) {
        const int src = kPerms8[pi][q];
        size_t st = 1;
        for (int e = 0; e < 3 - q; ++e) st *= nao;
        stride[src] = st;
        base += (size_t)offf[src] * st;
      }
      for (int e = tl; e < nf; e += G) {
        const int id = e % dims[3], r1 = e / dims[3], ic = r1 % dims[2], r2 = r1 / dims[2],
                  ibb = r2 % dims[1], ia = r2 / dims[1];
        out[base + ia * stride[0] + ibb * stride[1] + ic * stride[2] + id * stride[3]] = cur[e];
      }
    }
    team_sync();
  }
}



// ---------------------------------------------------------------------------
// dispatch tables over (la, lb, lc, ld) with la >= lb, lc >= ld, l <= kMaxKernelL

namespace {

template <int LA, int LB, int LC, int LD>
inline int grid_for(long long total) {
  using C = Cls<LA, LB, LC, LD>;
  long long blocks = (total + C::TEAMS - 1) / C::TEAMS;
  const long long cap = 256 * 64;
  if (blocks > cap) blocks = cap;
  if (blocks < 1) blocks = 1;
  return int(blocks);
}

template <int LA, int LB, int LC, int LD>
hipError_t jk_team_one(const BasisDev &b, const PairsDev &p, const RysDev &r, const QuartetSet &qs,
                       const DensDev &d, double omega, unsigned long long *ctr, hipStream_t s) {
  using C = Cls<LA, LB, LC, LD>;
  if (qs.total <= 0) return hipSuccess;
  hipLaunchKernelGGL(jk_kernel<C>, dim3(grid_for<LA, LB, LC, LD>(qs.total)), dim3(64), 0, s, b, p,
                     r, qs, d, omega, ctr);
  return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
hipError_t eri_one(const BasisDev &b, const PairsDev &p, const RysDev &r, const QuartetSet &qs,
                   const EriOut &o, hipStream_t s) {
  using C = Cls<LA, LB, LC, LD>;
  if (qs.total <= 0) return hipSuccess;
  hipLaunchKernelGGL(eri_kernel<C>, dim3(grid_for<LA, LB, LC, LD>(qs.total)), dim3(64), 0, s, b,
                     p, r, qs, o);
  return hipGetLastError();
}

template <int LA, int LB>
hipError_t schwarz_one(const BasisDev &b, const PairsDev &p, const RysDev &r, int first,
                       int count, double *q, hipStream_t s) {
  using C = Cls<LA, LB, LA, LB>;
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(schwarz_kernel<C>, dim3(grid_for<LA, LB, LA, LB>(count)), dim3(64), 0, s, b,
                     p, r, first, count, q);
  return hipGetLastError();
}

}  // namespace

bool class_supported(int la, int lb, int lc, int ld) {
  return la <= kMaxKernelL && lb <= la && lc <= kMaxKernelL && ld <= lc &&
         pair_class(la, lb) >= pair_class(lc, ld);
}

bool jk_team_compiled(int la, int lb, int lc, int ld) {
  if (class_supported(la, lb, lc, ld)) return true;
#define IG_CASE(A, B, Cc, D) \
  if (la == A && lb == B && lc == Cc && ld == D) return true;
  IG_FOR_F_TEAM_CLASSES(IG_CASE)
#undef IG_CASE
  return false;
}

bool jk_lane_compiled(int la, int lb, int lc, int ld) {
  if (class_supported(la, lb, lc, ld)) return true;
#define IG_CASE(A, B, Cc, D) \
  if (la == A && lb == B && lc == Cc && ld == D) return true;
  IG_FOR_F_LANE_CLASSES(IG_CASE)
#undef IG_CASE
  return false;
}

int jk_default_variant(int la, int lb, int lc, int ld) {
  // the autotuner's usual choice: register kernels with row sums on s,p,d
  // classes up to eight element slices (each slice rebuilds the quartet's 2D
  // tables; (dd|dd) goes to the team kernel), and on f classes up to two
  // (measured on s,p,d,f molecules: the contracted-space transfers win on
  // the largest blocks with a d or f second ket shell, (fd|fd), (ff|dd),
  // (fd|dd); the autotuner finds the others)
  // total angular momentum 2 ((ps|ps), (pp|ss), (ds|ss)) and (pp|ps): the
  // OS/HGP block with row sums, 0.5-0.9 ms faster per class than the Rys one
  // on the nao 520 s,p,d workload (no roots; (pp|ps) contracts its 9x9
  // primitive quartets in [e|f] space and transfers once); (ds|ps) and
  // (dp|ss) time the same or slower there and stay on Rys
  if (la + lb + lc + ld == 2 || (la == 1 && lb == 1 && lc == 1 && ld == 0)) return kJkHgp;
  const int n4 = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
  if (class_supported(la, lb, lc, ld) && n4 <= 8 * kRegBlockMax) return kJkRowAcc;
  if (n4 <= 2 * kRegBlockMax && jk_lane_compiled(la, lb, lc, ld)) return kJkRowAcc;
  if (!jk_team_compiled(la, lb, lc, ld)) return kJkTeamHrr;
  return n4 >= 2000 && ld >= 2 && jk_hrr_compiled(la, lb, lc, ld) ? kJkTeamHrr : kJkTeam;
}

hipError_t launch_jk(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                     const RysDev &r, const QuartetSet &qs, const DensDev &d, double omega,
                     int variant, unsigned long long *ctr, hipStream_t s) {
  if (variant == kJkTeamHrr)  // contracted-space transfers (ig_jk_hrr.hip)
    return launch_jk_hrr(la, lb, lc, ld, b, p, r, qs, d, omega, ctr, s);
  if (variant != kJkTeam)  // lane-per-quartet kernels (ig_jk_lane*.hip)
    return launch_jk_lane(la, lb, lc, ld, b, p, r, qs, d, omega, variant, ctr, s);
#define IG_CASE(A, B, Cc, D)                       \
  if (la == A && lb == B && lc == Cc && ld == D) \
    return jk_team_one<A, B, Cc, D>(b, p, r, qs, d, omega, ctr, s);
  IG_FOR_CLASSES(IG_CASE)
  IG_FOR_F_TEAM_CLASSES(IG_CASE)
#undef IG_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_eri(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                      const RysDev &r, const QuartetSet &qs, const EriOut &o, hipStream_t s) {
#define IG_CASE(A, B, Cc, D) \
  if (la == A && lb == B && lc == Cc && ld == D) return eri_one<A, B, Cc, D>(b, p, r, qs, o, s);
  IG_FOR_CLASSES(IG_CASE)
#undef IG_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_schwarz(int l1, int l2, const BasisDev &b, const PairsDev &p, const RysDev &r,
                          int first, int count, double *q, hipStream_t s) {
  if (l1 == 0 && l2 == 0) return schwarz_one<0, 0>(b, p, r, first, count, q, s);
  if (l1 == 1 && l2 == 0) return schwarz_one<1, 0>(b, p, r, first, count, q, s);
  if (l1 == 1 && l2 == 1) return schwarz_one<1, 1>(b, p, r, first, count, q, s);
  if (l1 == 2 && l2 == 0) return schwarz_one<2, 0>(b, p, r, first, count, q, s);
  if (l1 == 2 && l2 == 1) return schwarz_one<2, 1>(b, p, r, first, count, q, s);
  if (l1 == 2 && l2 == 2) return schwarz_one<2, 2>(b, p, r, first, count, q, s);
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// dense helpers

// shell of function-space / Cartesian index via binary search on offsets
__device__ __forceinline__ int shell_of(const int *off, int nsh, int idx) {
  int lo = 0, hi = nsh - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= idx)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// M^T X M per shell block: Cartesian (nao_cart^2) from function space (nao^2)
__global__ void to_cart_kernel(BasisDev b, const double *__restrict__ mf, int nmats,
                               double *__restrict__ mc) {
  const size_t nc2 = (size_t)b.nao_cart * b.nao_cart;
  const size_t total = nc2 * nmats;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const int m = int(t / nc2);
    const int rem = int(t % nc2);
    const int I = rem / b.nao_cart, J = rem % b.nao_cart;
    const int S = shell_of(b.off_cart, b.nsh, I), U = shell_of(b.off_cart, b.nsh, J);
    const int a = I - b.off_cart[S], c = J - b.off_cart[U];
    const int lS = b.l[S], lU = b.l[U];
    const int nS = b.kind[S] ? 2 * lS + 1 : ncart(lS), nU = b.kind[U] ? 2 * lU + 1 : ncart(lU);
    const double *MS = b.c2s + b.c2s_off[lS], *MU = b.c2s + b.c2s_off[lU];
    const double *X = mf + (size_t)m * b.nao * b.nao;
    double v = 0.0;
    for (int mu = 0; mu < nS; ++mu) {
      const double ms = b.kind[S] ? MS[mu * ncart(lS) + a] : (mu == a ? 1.0 : 0.0);
      if (ms == 0.0) continue;
      const double *row = X + (size_t)(b.off_func[S] + mu) * b.nao + b.off_func[U];
      for (int nu = 0; nu < nU; ++nu) {
        const double mu_ = b.kind[U] ? MU[nu * ncart(lU) + c] : (nu == c ? 1.0 : 0.0);
        v += ms * row[nu] * mu_;
      }
    }
    mc[t] = v;
  }
}

// symmetric / antisymmetric split of densities (function space)
__global__ void split_kernel(int nao, const double *__restrict__ d, int ndens, int has_anti,
                             double *__restrict__ ds, double *__restrict__ da) {
  const size_t n2 = (size_t)nao * nao;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n2 * ndens;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t m = t / n2, r = t % n2, i = r / nao, j = r % nao;
    const double x = d[t], y = d[m * n2 + j * nao + i];
    ds[t] = 0.5 * (x + y);
    if (has_anti) da[t] = 0.5 * (x - y);
  }
}

// max |X| per shell block, symmetrised over (i,j)/(j,i); X in function space
__global__ void block_max_kernel(BasisDev b, const double *__restrict__ x, int nmats,
                                 double *__restrict__ out) {
  const int nsh = b.nsh;
  const size_t total = (size_t)nsh * nsh * nmats;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const int m = int(t / ((size_t)nsh * nsh));
    const int r = int(t % ((size_t)nsh * nsh));
    const int i = r / nsh, j = r % nsh;
    const double *X = x + (size_t)m * b.nao * b.nao;
    const int ni = b.kind[i] ? 2 * b.l[i] + 1 : ncart(b.l[i]);
    const int nj = b.kind[j] ? 2 * b.l[j] + 1 : ncart(b.l[j]);
    double mx = 0.0;
    for (int a = 0; a < ni; ++a)
      for (int c = 0; c < nj; ++c) {
        mx = fmax(mx, fabs(X[(size_t)(b.off_func[i] + a) * b.nao + b.off_func[j] + c]));
        mx = fmax(mx, fabs(X[(size_t)(b.off_func[j] + c) * b.nao + b.off_func[i] + a]));
      }
    out[t] = mx;
  }
}

// J_d += M (Jp + Jp^T) M^T, K_d += M (Kp_s + Kp_s^T + Kp_a - Kp_a^T) M^T; one
// thread per (density, mu, nu) sums every effective density of that source.
__global__ void jk_finish_kernel(BasisDev b, DensDev d, double *__restrict__ J,
                                 double *__restrict__ K) {
  const size_t n2 = (size_t)b.nao * b.nao;
  const size_t nc2 = (size_t)b.nao_cart * b.nao_cart;
  const int ncrt = b.nao_cart;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n2 * d.ndens;
       t += (size_t)gridDim.x * blockDim.x) {
    const int src = int(t / n2);
    const int r = int(t % n2);
    const int mu = r / b.nao, nu = r % b.nao;
    const int S = shell_of(b.off_func, b.nsh, mu), U = shell_of(b.off_func, b.nsh, nu);
    const int p = mu - b.off_func[S], q = nu - b.off_func[U];
    const int lS = b.l[S], lU = b.l[U], nS = ncart(lS), nU = ncart(lU);
    const double *MS = b.c2s + b.c2s_off[lS], *MU = b.c2s + b.c2s_off[lU];
    double vj = 0.0, vk = 0.0;
    for (int e = 0; e < d.neff; ++e) {
      if (d.eff_src[e] != src) continue;
      const bool anti = d.eff_anti[e];
      const double *Jp = d.jc + (size_t)e * nc2;
      const double *Kp = d.kc ? d.kc + (size_t)e * nc2 : nullptr;
      for (int a = 0; a < nS; ++a) {
        const double ms = b.kind[S] ? MS[p * nS + a] : (a == p ? 1.0 : 0.0);
        if (ms == 0.0) continue;
        const size_t I = (size_t)b.off_cart[S] + a;
        for (int c = 0; c < nU; ++c) {
          const double mu_ = b.kind[U] ? MU[q * nU + c] : (c == q ? 1.0 : 0.0);
          if (mu_ == 0.0) continue;
          const size_t Jx = (size_t)b.off_cart[U] + c;
          const double w = ms * mu_;
          if (!anti) vj += w * (Jp[I * ncrt + Jx] + Jp[Jx * ncrt + I]);
          if (Kp)
            vk += w * (anti ? Kp[I * ncrt + Jx] - Kp[Jx * ncrt + I]
                            : Kp[I * ncrt + Jx] + Kp[Jx * ncrt + I]);
        }
      }
    }
    J[t] += vj;
    if (d.kc && K) K[t] += vk;
  }
}

__global__ void fill_kernel(double *__restrict__ x, size_t n, double v) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    x[t] = v;
}

static int grid1d(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  return int(g < 1 ? 1 : g);
}

hipError_t launch_to_cart(const BasisDev &b, const double *mf, int nmats, double *mc,
                          hipStream_t s) {
  size_t n = (size_t)b.nao_cart * b.nao_cart * nmats;
  hipLaunchKernelGGL(to_cart_kernel, dim3(grid1d(n)), dim3(256), 0, s, b, mf, nmats, mc);
  return hipGetLastError();
}

// One index of a 4-index tensor from function to Cartesian space:
// out[p, I, q] = sum_mu M_S(mu, I) in[p, off_func(S) + mu, q], S the shell of
// Cartesian index I and M_S its c2s rows (the identity on a Cartesian
// shell); in is [pre, nao, post], out [pre, nao_cart, post].
__global__ void axis_to_cart_kernel(BasisDev b, const double *__restrict__ in,
                                    double *__restrict__ out, size_t pre, size_t post) {
  const size_t nc = (size_t)b.nao_cart, total = pre * nc * post;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t q = t % post, r = t / post, p = r / nc;
    const int I = int(r % nc);
    const int S = shell_of(b.off_cart, b.nsh, I);
    const int a = I - b.off_cart[S], lS = b.l[S];
    const double *src = in + (p * (size_t)b.nao + b.off_func[S]) * post + q;
    double v;
    if (!b.kind[S]) {
      v = src[(size_t)a * post];
    } else {
      const double *MS = b.c2s + b.c2s_off[lS];
      v = 0.0;
      for (int mu = 0; mu < 2 * lS + 1; ++mu) v += MS[mu * ncart(lS) + a] * src[(size_t)mu * post];
    }
    out[t] = v;
  }
}

// Dense two-particle density (nao^4, function space) to Cartesian space, one
// index per pass, last index first.  t0 and t1 hold nao_cart^4 doubles each;
// the result is left in t1.
hipError_t launch_gamma_to_cart(const BasisDev &b, const double *g, double *t0, double *t1,
                                hipStream_t s) {
  const size_t n = (size_t)b.nao, c = (size_t)b.nao_cart;
  const double *src[4] = {g, t0, t1, t0};
  double *dst[4] = {t0, t1, t0, t1};
  const size_t pre[4] = {n * n * n, n * n, n, 1}, post[4] = {1, c, c * c, c * c * c};
  for (int k = 0; k < 4; ++k) {
    hipLaunchKernelGGL(axis_to_cart_kernel, dim3(grid1d(pre[k] * c * post[k])), dim3(256), 0, s,
                       b, src[k], dst[k], pre[k], post[k]);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

__global__ void asym_flag_kernel(const double *__restrict__ d, int n, int ndens, int *flag) {
  const size_t n2 = (size_t)n * n;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n2 * ndens;
       t += (size_t)gridDim.x * blockDim.x) {
    const size_t m = t / n2, r = t % n2, i = r / n, j = r % n;
    if (d[t] != d[m * n2 + j * n + i]) atomicOr(flag, 1);
  }
}

hipError_t launch_asym_flag(const double *d, int n, int ndens, int *flag, hipStream_t s) {
  hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
  if (e != hipSuccess) return e;
  size_t total = (size_t)n * n * ndens;
  hipLaunchKernelGGL(asym_flag_kernel, dim3(grid1d(total)), dim3(256), 0, s, d, n, ndens, flag);
  return hipGetLastError();
}

hipError_t launch_fill(double *x, size_t n, double v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fill_kernel, dim3(grid1d(n)), dim3(256), 0, s, x, n, v);
  return hipGetLastError();
}

hipError_t launch_density_prep(const BasisDev &b, const double *d_func, int ndens, int has_anti,
                               double *ds, double *da, double *dc, double *bound,
                               hipStream_t s) {
  const size_t n2 = (size_t)b.nao * b.nao, nc2 = (size_t)b.nao_cart * b.nao_cart;
  hipLaunchKernelGGL(split_kernel, dim3(grid1d(n2 * ndens)), dim3(256), 0, s, b.nao, d_func,
                     ndens, has_anti, ds, da);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = launch_to_cart(b, ds, ndens, dc, s)) != hipSuccess) return e;
  if (has_anti && (e = launch_to_cart(b, da, ndens, dc + (size_t)ndens * nc2, s)) != hipSuccess)
    return e;
  size_t nb = (size_t)b.nsh * b.nsh * ndens;
  hipLaunchKernelGGL(block_max_kernel, dim3(grid1d(nb)), dim3(256), 0, s, b, d_func, ndens,
                     bound);
  return hipGetLastError();
}

hipError_t launch_block_max(const BasisDev &b, const double *m, int nmats, double *bmax,
                            hipStream_t s) {
  size_t nb = (size_t)b.nsh * b.nsh * nmats;
  hipLaunchKernelGGL(block_max_kernel, dim3(grid1d(nb)), dim3(256), 0, s, b, m, nmats, bmax);
  return hipGetLastError();
}

__global__ void fold_copies_kernel(double *__restrict__ x, size_t n, size_t stride, int ncopy) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    double v = x[t];
    for (int k = 1; k < ncopy; ++k) v += x[k * stride + t];
    x[t] = v;
  }
}

__global__ void xcc_probe_kernel(unsigned *hist) {
  if (threadIdx.x == 0) atomicAdd(hist + xcc_id(), 1u);
}

hipError_t launch_fold_copies(const DensDev &d, size_t nc2, hipStream_t s) {
  if (d.ncopy <= 1) return hipSuccess;
  const size_t n = nc2 * d.neff;
  hipLaunchKernelGGL(fold_copies_kernel, dim3(grid1d(n)), dim3(256), 0, s, d.jc, n,
                     d.copy_stride, d.ncopy);
  if (d.kc)
    hipLaunchKernelGGL(fold_copies_kernel, dim3(grid1d(n)), dim3(256), 0, s, d.kc, n,
                       d.copy_stride, d.ncopy);
  return hipGetLastError();
}

__global__ void boys_eval_kernel(RysDev R, int mmax, const double *__restrict__ T, size_t n,
                                 double *__restrict__ out) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    double F[kBoysMaxM + 1];
    boys_eval(R.boys, mmax, T[t], F);
    for (int m = 0; m <= mmax; ++m) out[t * (mmax + 1) + m] = F[m];
  }
}

__global__ void rys_eval_kernel(RysDev R, int nr, const double *__restrict__ T, size_t n,
                                double *__restrict__ x, double *__restrict__ w) {
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x)
    for (int i = 0; i < nr; ++i) {
      x[t * nr + i] = rys_series(R, nr, T[t], i);
      w[t * nr + i] = rys_series(R, nr, T[t], nr + i);
    }
}

hipError_t launch_boys_eval(const RysDev &r, int mmax, const double *T, size_t n, double *out,
                            hipStream_t s) {
  hipLaunchKernelGGL(boys_eval_kernel, dim3(grid1d(n)), dim3(256), 0, s, r, mmax, T, n, out);
  return hipGetLastError();
}

hipError_t launch_rys_eval(const RysDev &r, int nr, const double *T, size_t n, double *x,
                           double *w, hipStream_t s) {
  hipLaunchKernelGGL(rys_eval_kernel, dim3(grid1d(n)), dim3(256), 0, s, r, nr, T, n, x, w);
  return hipGetLastError();
}

hipError_t launch_xcc_probe(unsigned *hist, int groups, hipStream_t s) {
  hipLaunchKernelGGL(xcc_probe_kernel, dim3(groups), dim3(64), 0, s, hist);
  return hipGetLastError();
}

hipError_t launch_jk_finish(const BasisDev &b, const DensDev &d, double *J, double *K,
                            hipStream_t s) {
  size_t n = (size_t)b.nao * b.nao * d.ndens;
  hipLaunchKernelGGL(jk_finish_kernel, dim3(grid1d(n)), dim3(256), 0, s, b, d, J, K);
  return hipGetLastError();
}

}  // namespace ig
