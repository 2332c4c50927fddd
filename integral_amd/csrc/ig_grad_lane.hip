// ig_grad_lane.hip -- lane-per-quartet adjoint gradient kernel: one shell
// quartet per lane, the contraction of the ERI first derivatives with the
// density weights of direct_grad.rs:394-547 without forming any block.
//
// Per lane and element slice [LO, HI): the weight W_e of every Cartesian
// element (Coulomb and exchange terms, direct_grad.rs:490-547) in registers;
// then per primitive quartet and Rys root the slice's weights are contracted
// with the unraised 2D tables into per-axis power-tuple sums, and those sums
// with the derivative tables of centres A, B (one table raised on A) and C
// (one raised on C).  D follows by translational invariance.  The 12 per-atom
// sums of a quartet go to an LDS copy of the gradient (natom x 3) with LDS
// atomics; each workgroup adds its copy to global memory once.
#include "ig_cls.h"

namespace ig {

#include "ig_reg.h"

namespace {

constexpr int kGradWaves = 4;
constexpr int kGradThreads = 64 * kGradWaves;
constexpr int kGradSlice = 27;     // elements per slice
constexpr int kGradSliceWide = 54;  // elements per slice of the large one-wavefront classes

// Cartesian powers of the four shells of every element of class C.
template <class C>
struct GradPow {
  int a[C::N4][3], b[C::N4][3], c[C::N4][3], d[C::N4][3];
  constexpr GradPow() : a(), b(), c(), d() {
    for (int e = 0; e < C::N4; ++e) {
      const int id = e % C::nd, r1 = e / C::nd, ic = r1 % C::nc, r2 = r1 / C::nc;
      const int ib = r2 % C::nb, ia = r2 / C::nb;
      pw(C::LA, ia, a[e]);
      pw(C::LB, ib, b[e]);
      pw(C::LC, ic, c[e]);
      pw(C::LD, id, d[e]);
    }
  }
  static constexpr void pw(int l, int i, int *o) {  // lx descending, ly descending
    int k = 0;
    while ((k + 1) * (k + 2) / 2 <= i) ++k;
    o[0] = l - k;
    o[2] = i - k * (k + 1) / 2;
    o[1] = k - o[2];
  }
};

// Offset of (a, b, c, d) in a table with extents (EA, EB, EC, ED).
__host__ __device__ constexpr int toff(int a, int b, int c, int d, int EB, int EC, int ED) {
  return ((a * EB + b) * EC + c) * ED + d;
}

// Where each element of the slice [LO, HI) sits in the unraised per-axis
// table (extents (LA+1, LB+1, LC+1, LD+1)), and which table entries (power
// tuples (a_x, b_x, c_x, d_x) of one axis) some element of the slice uses.
template <class C, int LO, int HI>
struct GradMap {
  static constexpr int EA = C::LA + 1, EB = C::LB + 1, EC = C::LC + 1, ED = C::LD + 1;
  static constexpr int T0 = EA * EB * EC * ED;
  int o[HI - LO][3];
  bool used[3][T0];
  constexpr GradMap() : o(), used() {
    GradPow<C> Pw{};
    for (int e = LO; e < HI; ++e)
      for (int ax = 0; ax < 3; ++ax) {
        const int t = toff(Pw.a[e][ax], Pw.b[e][ax], Pw.c[e][ax], Pw.d[e][ax], EB, EC, ED);
        o[e - LO][ax] = t;
        used[ax][t] = true;
      }
  }
};

// Derivative sums of one Rys root.  With Z_x[t] = sum over the slice's
// elements e whose x powers are the tuple t of W_e I_y[e] I_z[e] (and
// cyclically), the centre-X derivative along x is
//   sum_t Z_x[t] (2 alpha_X I_x[t + 1_X] - t_X I_x[t - 1_X]),
// so the elements are visited once per root, not once per centre.  The A and
// B derivatives both come from the table raised on A, using the transfer
// I(a, b+1) = I(a+1, b) + (A-B) I(a, b); the C derivative from the table
// raised on C.  D follows by translational invariance.
template <class C, int LO, int HI, int WS>
__device__ __forceinline__ void grad_root(const double (&c00)[3], const double (&cp0)[3],
                                          double b10, double b01, double b00,
                                          const double (&AB)[3], const double (&CD)[3],
                                          double ta, double tb, double tc, double pw,
                                          const double *W, double *g) {
  using M = GradMap<C, LO, HI>;
  constexpr M Mp{};
  constexpr int T0 = M::T0, EB = M::EB, EC = M::EC, ED = M::ED;
  constexpr int SA = EB * EC * ED, SB = EC * ED;  // strides of a and b (both layouts)
  double Z[3][T0];
  {
    double I[3][T0];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
      rys_axis_table<C::LA, C::LB, C::LC, C::LD>(c00[ax], cp0[ax], b10, b01, b00, AB[ax],
                                                 CD[ax], I[ax]);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
      for (int t = 0; t < T0; ++t) Z[ax][t] = 0.0;
#pragma unroll
    for (int e = LO; e < HI; ++e) {
      const int ox = Mp.o[e - LO][0], oy = Mp.o[e - LO][1], oz = Mp.o[e - LO][2];
      const double w = W[(e - LO) * WS];
      const double wx = w * I[0][ox];
      Z[0][ox] += (w * I[1][oy]) * I[2][oz];
      Z[1][oy] += wx * I[2][oz];
      Z[2][oz] += wx * I[1][oy];
    }
  }
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    double sa = 0.0, sb = 0.0, sc = 0.0;
    {
      double R[(C::LA + 2) * SA];  // raised on A: same offsets as the unraised layout
      rys_axis_table<C::LA + 1, C::LB, C::LC, C::LD>(c00[ax], cp0[ax], b10, b01, b00, AB[ax],
                                                     CD[ax], R);
#pragma unroll
      for (int t = 0; t < T0; ++t) {
        if (!Mp.used[ax][t]) continue;
        const int a = t / SA, b = (t / SB) % EB;
        const double up = R[t + SA];
        double da = ta * up, db = tb * (up + AB[ax] * R[t]);
        if (a > 0) da -= double(a) * R[t - SA];
        if (b > 0) db -= double(b) * R[t - SB];
        sa += Z[ax][t] * da;
        sb += Z[ax][t] * db;
      }
    }
    {
      constexpr int RC = C::LC + 2;
      double R[(C::LA + 1) * EB * RC * ED];  // raised on C
      rys_axis_table<C::LA, C::LB, C::LC + 1, C::LD>(c00[ax], cp0[ax], b10, b01, b00, AB[ax],
                                                     CD[ax], R);
#pragma unroll
      for (int t = 0; t < T0; ++t) {
        if (!Mp.used[ax][t]) continue;
        const int a = t / SA, b = (t / SB) % EB, c = (t / ED) % EC, d = t % ED;
        const int o = toff(a, b, c, d, EB, RC, ED);
        double dc = tc * R[o + ED];
        if (c > 0) dc -= double(c) * R[o - ED];
        sc += Z[ax][t] * dc;
      }
    }
    g[ax] += pw * sa;
    g[3 + ax] += pw * sb;
    g[6 + ax] += pw * sc;
  }
}

// Wavefronts per SIMD the kernel is built for: two for total angular
// momentum 3 and 4, and 5 with at most one d shell (their weights in LDS
// keep them there without spills); one elsewhere (few elements, or tables
// too large for two wavefronts).  Chosen from per-class kernel times.
template <class C>
constexpr int grad_waves() {
  constexpr int L = C::LA + C::LB + C::LC + C::LD;
  constexpr int nd = (C::LA == 2) + (C::LB == 2) + (C::LC == 2) + (C::LD == 2);
  return (L >= 3 && L <= 4) || (L == 5 && nd <= 1) ? 2 : 1;
}

// Elements per slice: 27, or 54 for the one-wavefront classes of 108
// elements and more, which halves the per-slice work (weights, screening,
// roots and tables are redone in every slice).  The two-wavefront (d p|p p)
// keeps 27: 54 would cost it its second wavefront (measured 10% slower).
template <class C>
constexpr int grad_slice_cap() {
  return C::N4 >= 108 && grad_waves<C>() == 1 ? kGradSliceWide : kGradSlice;
}

// Where the weights of a lane's quartet live: in LDS (slot stride
// kGradThreads, so a wavefront's reads are conflict-free) for the
// two-wavefront classes and the wide slices, in VGPRs otherwise.
template <class C>
constexpr bool grad_w_in_lds() {
  return grad_waves<C>() == 2 || grad_slice_cap<C>() == kGradSliceWide;
}

// LDS a workgroup may take: two workgroups per CU at two wavefronts per
// SIMD, the whole CU at one.
template <class C>
constexpr size_t grad_lds_max() {
  return grad_waves<C>() == 2 ? 80 * 1024 : 160 * 1024;
}

// With the weights in LDS, the slot offset passes through an empty asm
// statement at every root so that the reads stay inside the root loop
// instead of being hoisted back into registers.
template <class C, int LO, int HI, bool WL = grad_w_in_lds<C>(), int WPE = grad_waves<C>()>
__global__ void __launch_bounds__(kGradThreads) __attribute__((amdgpu_waves_per_eu(WPE)))
    grad_lane_kernel(BasisDev b, PairsDev p, RysDev R, QuartetSet qs, GradDev gd,
                     unsigned long long *counters, int lds_grad) {
  constexpr int NR = (C::LA + C::LB + C::LC + C::LD + 1) / 2 + 1;
  extern __shared__ double lds[];
  stage_rys<NR>(R, lds);  // ends with a barrier
  const int nrys = 2 * R.nsub[NR] * NR * kRysTerms;
  double *Wl = lds + nrys;  // WL: (HI - LO) slots of stride kGradThreads
  double *G = lds + nrys + (WL ? (HI - LO) * kGradThreads : 0);  // natom x 3 (when lds_grad)
  if (lds_grad) {
    for (int i = threadIdx.x; i < lds_grad; i += blockDim.x) G[i] = 0.0;
    __syncthreads();
  }
  const int nsh = b.nsh, ncrt = b.nao_cart;
  const size_t nc2 = (size_t)ncrt * ncrt;
  const int nterm = gd.ncoul + gd.nexch;
  unsigned long long my_q = 0, my_p = 0;
  const long long chunk = (qs.total + gridDim.x - 1) / gridDim.x;
  const long long begin = (long long)blockIdx.x * chunk;
  const long long end = begin + chunk < qs.total ? begin + chunk : qs.total;
  RowCursor cur;
  for (long long base = begin; base < end; base += blockDim.x) {
    const long long t = base + threadIdx.x;
    if (t >= end) break;  // no barrier below: lanes may leave independently
    int P, Q;
    cur.next(qs, t, P, Q);
    const int pb = qs.x_off + P, pk = qs.y_off + Q;
    QuartetGeom g;
    load_geom(b, p.sh1[pb], p.sh2[pb], p.sh1[pk], p.sh2[pk], g);
    const int atA = b.atom[g.sA], atB = b.atom[g.sB], atC = b.atom[g.sC], atD = b.atom[g.sD];
    if (atA == atB && atB == atC && atC == atD) continue;  // direct_grad.rs:356-358
    const double deg =
        (g.sA == g.sB ? 1.0 : 2.0) * (g.sC == g.sD ? 1.0 : 2.0) * (pb == pk ? 1.0 : 2.0);
    double bound = 0.0;  // density block bound (direct_grad.rs:626-644)
    {
      const int i = g.sA, j = g.sB, k = g.sC, l = g.sD;
      for (int tm = 0; tm < nterm; ++tm) {
        const double *Lm = gd.bmax + (size_t)(2 * tm) * nsh * nsh;
        const double *Rm = Lm + (size_t)nsh * nsh;
        if (tm < gd.ncoul)
          bound += 0.5 * fabs(gd.fac[tm]) *
                   (Lm[i * nsh + j] * Rm[k * nsh + l] + Lm[k * nsh + l] * Rm[i * nsh + j]);
        else
          bound += 0.25 * fabs(gd.fac[tm]) *
                   (Lm[i * nsh + k] * Rm[j * nsh + l] + Lm[j * nsh + k] * Rm[i * nsh + l] +
                    Lm[i * nsh + l] * Rm[j * nsh + k] + Lm[j * nsh + l] * Rm[i * nsh + k]);
      }
    }
    if (p.q[pb] * p.q[pk] * deg * bound < gd.tau) continue;
    // weights of the slice's elements
    double Wr[WL ? 1 : HI - LO];
    const size_t a0 = b.off_cart[g.sA], b0 = b.off_cart[g.sB], c0 = b.off_cart[g.sC],
                 d0 = b.off_cart[g.sD];
#pragma unroll
    for (int e = LO; e < HI; ++e) {
      const int id = e % C::nd, r1 = e / C::nd, ic = r1 % C::nc, r2 = r1 / C::nc;
      const int ib = r2 % C::nb, ia = r2 / C::nb;
      const size_t A = a0 + ia, B = b0 + ib, Cc = c0 + ic, D = d0 + id;
      double w = 0.0;
      if (gd.gamma)
        w = gamma_weight(gd.gamma, size_t(ncrt), A, B, Cc, D, g.sA != g.sB, g.sC != g.sD, pb != pk);
      for (int tm = 0; tm < nterm; ++tm) {
        const double *Lc = gd.mats + (size_t)(2 * tm) * nc2, *Rc = Lc + nc2;
        if (tm < gd.ncoul)
          w += 0.5 * deg * gd.fac[tm] *
               (Lc[A * ncrt + B] * Rc[Cc * ncrt + D] + Rc[A * ncrt + B] * Lc[Cc * ncrt + D]);
        else
          w += 0.25 * deg * gd.fac[tm] *
               (Lc[A * ncrt + Cc] * Rc[B * ncrt + D] + Lc[B * ncrt + Cc] * Rc[A * ncrt + D] +
                Lc[A * ncrt + D] * Rc[B * ncrt + Cc] + Lc[B * ncrt + D] * Rc[A * ncrt + Cc]);
      }
      if constexpr (WL)
        Wl[(e - LO) * kGradThreads + threadIdx.x] = w;
      else
        Wr[e - LO] = w;
    }
    double gs[9];
#pragma unroll
    for (int x = 0; x < 9; ++x) gs[x] = 0.0;
    double AB[3], CD[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      AB[x] = g.A[x] - g.B[x];
      CD[x] = g.C[x] - g.D[x];
    }
    const int ob = p.pp_off[pb], nbp = p.pp_n[pb], ok = p.pp_off[pk], nkp = p.pp_n[pk];
    for (int ik = 0; ik < nkp; ++ik) {
      const double qz = p.zeta[ok + ik], gam = p.alpha[ok + ik];
      const double Qv[3] = {p.px[ok + ik], p.py[ok + ik], p.pz[ok + ik]};
      const double kzq = kTwoPi52 * p.kz[ok + ik], i2q = p.i2z[ok + ik];
      for (int ib = 0; ib < nbp; ++ib) {
        const double pz = p.zeta[ob + ib], al = p.alpha[ob + ib], be = p.beta[ob + ib];
        const double Pv[3] = {p.px[ob + ib], p.py[ob + ib], p.pz[ob + ib]};
        const double kzp = p.kz[ob + ib], i2p = p.i2z[ob + ib];
        const double pq = pz + qz, rs = rsqrt(pq), ipq = rs * rs, rho = pz * qz * ipq;
        double PQ[3], r2 = 0.0;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          PQ[x] = Pv[x] - Qv[x];
          r2 += PQ[x] * PQ[x];
        }
        double Tq = rho * r2;
        double pref = kzp * kzq * rs;
        double s_att = 1.0;
        if (gd.omega > 0.0) {  // erf(w r)/r: roots/weights at sT, x -> s x, w -> sqrt(s) w
          s_att = gd.omega * gd.omega / (rho + gd.omega * gd.omega);
          Tq *= s_att;
          pref *= sqrt(s_att);
        }
        double rx[NR], rwt[NR];
        rys_all_lds<NR>(R, lds, Tq, rx, rwt);
#pragma unroll 1
        for (int r = 0; r < NR; ++r) {
          const double x = rx[r] * s_att;
          const double fq = x * qz * ipq, fp = x * pz * ipq;
          const double b00 = 0.5 * x * ipq, b10 = (1.0 - fq) * i2p, b01 = (1.0 - fp) * i2q;
          double c00[3], cp0[3];
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            c00[ax] = (Pv[ax] - g.A[ax]) - fq * PQ[ax];
            cp0[ax] = (Qv[ax] - g.C[ax]) + fp * PQ[ax];
          }
          const double pw = pref * rwt[r];
          if constexpr (WL) {
            int wo = threadIdx.x;
            asm volatile("" : "+v"(wo));
            grad_root<C, LO, HI, kGradThreads>(c00, cp0, b10, b01, b00, AB, CD, 2.0 * al,
                                               2.0 * be, 2.0 * gam, pw, Wl + wo, gs);
          } else {
            grad_root<C, LO, HI, 1>(c00, cp0, b10, b01, b00, AB, CD, 2.0 * al, 2.0 * be,
                                    2.0 * gam, pw, Wr, gs);
          }
        }
      }
    }
    if (LO == 0) {
      ++my_q;
      my_p += (unsigned long long)nbp * nkp;
    }
    const int at[4] = {atA, atB, atC, atD};
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      const double gdv = -(gs[x] + gs[3 + x] + gs[6 + x]);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const double v = k < 3 ? gs[3 * k + x] : gdv;
        if (lds_grad)
          atomicAdd(G + 3 * at[k] + x, v);
        else
          atomicAdd(gd.out + 3 * at[k] + x, v);
      }
    }
  }
  if (lds_grad) {
    __syncthreads();
    for (int i = threadIdx.x; i < lds_grad; i += blockDim.x)
      if (G[i] != 0.0) atomicAdd(gd.out + i, G[i]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    my_q += __shfl_down(my_q, off);
    my_p += __shfl_down(my_p, off);
  }
  if (counters && (threadIdx.x & 63) == 0 && LO == 0) {
    atomicAdd(counters, my_q);
    atomicAdd(counters + 1, my_p);
  }
}

template <class C>
struct GradSlices {
  static constexpr int NS = (C::N4 + grad_slice_cap<C>() - 1) / grad_slice_cap<C>();
  static constexpr int E = (C::N4 + NS - 1) / NS;
};

template <class C, int S>
hipError_t launch_grad_slices(const BasisDev &b, const PairsDev &p, const RysDev &r,
                              const QuartetSet &qs, const GradDev &g, unsigned long long *ctr,
                              hipStream_t s, int blocks, size_t lds, int lds_grad) {
  constexpr int NS = GradSlices<C>::NS, E = GradSlices<C>::E;
  if constexpr (S < NS) {
    constexpr int LO = S * E, HI = (S + 1) * E < C::N4 ? (S + 1) * E : C::N4;
    const size_t wl = grad_w_in_lds<C>() ? size_t(HI - LO) * kGradThreads * sizeof(double) : 0;
    hipLaunchKernelGGL((grad_lane_kernel<C, LO, HI>), dim3(blocks), dim3(kGradThreads), lds + wl, s,
                       b, p, r, qs, g, ctr, lds_grad);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_grad_slices<C, S + 1>(b, p, r, qs, g, ctr, s, blocks, lds, lds_grad);
  } else {
    return hipSuccess;
  }
}

template <int LA, int LB, int LC, int LD>
hipError_t grad_lane_one(const BasisDev &b, const PairsDev &p, const RysDev &r,
                         const QuartetSet &qs, const GradDev &g, int natom,
                         unsigned long long *ctr, hipStream_t s) {
  using C = Cls<LA, LB, LC, LD>;
  if (qs.total <= 0) return hipSuccess;
  constexpr int NR = (LA + LB + LC + LD + 1) / 2 + 1;
  const size_t rys = size_t(2) * r.nsub[NR] * NR * kRysTerms * sizeof(double);
  const size_t gl = size_t(3) * natom * sizeof(double);
  const size_t wl =  // widest slice
      grad_w_in_lds<C>() ? size_t(GradSlices<C>::E) * kGradThreads * sizeof(double) : 0;
  const int lds_grad = rys + wl + gl <= grad_lds_max<C>() ? 3 * natom : 0;
  long long blocks = (qs.total + kGradThreads - 1) / kGradThreads;
  if (blocks > 2048) blocks = 2048;
  return launch_grad_slices<C, 0>(b, p, r, qs, g, ctr, s, int(blocks),
                                  rys + (lds_grad ? gl : 0), lds_grad);
}

}  // namespace

hipError_t launch_grad_lane(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                            const RysDev &r, const QuartetSet &qs, const GradDev &g, int natom,
                            unsigned long long *ctr, hipStream_t s) {
#define IG_CASE(A, B, Cc, D)                       \
  if (la == A && lb == B && lc == Cc && ld == D) \
    return grad_lane_one<A, B, Cc, D>(b, p, r, qs, g, natom, ctr, s);
  IG_FOR_CLASSES(IG_CASE)
#undef IG_CASE
  return hipErrorInvalidValue;
}

}  // namespace ig
