// ig_reg.h -- lane-per-quartet J/K path (included inside namespace ig by
// ig_lane.h).  Rys roots and weights come from the polynomial tables
// staged in LDS once per workgroup; per-root 2D tables and the contracted block live
// in registers with compile-time indices; J/K are digested from registers.
//
// Blocks larger than one lane's register budget are cut into NS element
// slices [LO, HI) of at most kRegBlockMax elements.  Each slice is a separate
// launch: every wavefront of a slice launch walks the same quartet range,
// rebuilds the (cheap) 2D tables in registers and accumulates and digests only
// its slice, so slices never communicate and the slice index is uniform per
// wavefront (no divergence).
//
// Each wavefront walks a contiguous range of screened quartets; lanes of a
// wavefront usually share the bra pair, so the bra-side Coulomb updates are
// reduced across the wavefront before one atomic per element.

// Rys coefficient tables of NR roots copied to LDS: roots then weights,
// nsub * NR * kRysTerms doubles each.
template <int NR>
__device__ __forceinline__ void stage_rys(const RysDev &R, double *lds) {
  const int n = R.nsub[NR] * NR * kRysTerms;
  const double *r = R.buf + R.root_off[NR], *w = R.buf + R.wt_off[NR];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    lds[i] = r[i];
    lds[n + i] = w[i];
  }
  __syncthreads();
}

template <int NR>
__device__ __forceinline__ void rys_all_lds(const RysDev &R, const double *lds, double T,
                                            double *x, double *w) {
  const double hi = 33.0 + 2.5 * double(2 * NR - 1);
  if (T >= hi) {
    const double *L = R.buf + R.lag_off[NR];
    const double rs = rsqrt(T), it = rs * rs, sc = 0.5 * rs;
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      x[i] = L[i] * it;
      w[i] = L[NR + i] * sc;
    }
    return;
  }
  double xi;
  const int s = rys_cell(R, NR, T, xi);
  const double *cr = lds + s * NR * kRysTerms;
  const double *cw = cr + R.nsub[NR] * NR * kRysTerms;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    x[i] = rys_poly(cr + i * kRysTerms, xi);
    w[i] = rys_poly(cw + i * kRysTerms, xi);
  }
}

// Elements [LO, HI) of the contracted Cartesian block of (pb | pk) into
// acc[e - LO].  Ket primitives in the outer loop (per-lane data), bra
// primitives inside (usually uniform across the wavefront).
// F is the arithmetic type of the tables and the block (float for the
// mixed-precision quartets); primitive-quartet scalars and the Rys roots and
// weights are always evaluated in double and rounded once.
template <class C, int LO, int HI, typename F>
__device__ __forceinline__ void quartet_block_reg(const PairsDev &p, const RysDev &R,
                                                  const double *rys_lds, int pb, int pk,
                                                  const QuartetGeom &g, double omega, F *acc) {
  constexpr int NR = C::NR, T = C::T;
  constexpr ElemTab<C> O{};
  double AB[3], CD[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    AB[x] = g.A[x] - g.B[x];
    CD[x] = g.C[x] - g.D[x];
  }
  const int ob = p.pp_off[pb], nbp = p.pp_n[pb];
  const int ok = p.pp_off[pk], nkp = p.pp_n[pk];
  for (int ik = 0; ik < nkp; ++ik) {
    const double qz = p.zeta[ok + ik];
    const double Q[3] = {p.px[ok + ik], p.py[ok + ik], p.pz[ok + ik]};
    const double kzq = kTwoPi52 * p.kz[ok + ik], i2q = p.i2z[ok + ik];
    double QC[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) QC[x] = Q[x] - g.C[x];
    for (int ib = 0; ib < nbp; ++ib) {
      const double pz = p.zeta[ob + ib];
      const double P[3] = {p.px[ob + ib], p.py[ob + ib], p.pz[ob + ib]};
      const double kzp = p.kz[ob + ib], i2p = p.i2z[ob + ib];
      const double pq = pz + qz;
      // 2 pi^{5/2} K_ab K_cd / (zeta_p zeta_q sqrt(zeta_p + zeta_q)) from one rsqrt
      const double rs = rsqrt(pq);
      const double ipq = rs * rs;
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
      double rx[NR], rwt[NR];
      rys_all_lds<NR>(R, rys_lds, Tq, rx, rwt);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const double x = rx[r] * s_att;
        const double fq = x * qz * ipq, fp = x * pz * ipq;
        const F b00 = F(0.5 * x * ipq);
        const F b10 = F((1.0 - fq) * i2p);
        const F b01 = F((1.0 - fp) * i2q);
        // the weight and prefactor ride on the z table (its I(0,0|0,0))
        const F pw = F(pref * rwt[r]);
        F I[3][T];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
          rys_axis_table<C::LA, C::LB, C::LC, C::LD, F>(
              F((P[ax] - g.A[ax]) - fq * PQ[ax]), F(QC[ax] + fp * PQ[ax]), b10, b01, b00,
              F(AB[ax]), F(CD[ax]), I[ax], ax == 2 ? pw : F(1));
#pragma unroll
        for (int e = LO; e < HI; ++e) acc[e - LO] += I[0][O.x[e]] * I[1][O.y[e]] * I[2][O.z[e]];
      }
    }
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// J/K digestion of the block elements [LO, HI) held in registers.  Called by
// every lane of the wavefront (`on` false for lanes without a quartet).  When
// all lanes with work share the bra shell pair, the Jp[a,b] sums are reduced
// across the wavefront first and written by `leader`, the first lane with
// work (its geometry names that shared bra pair; a lane without work may hold
// another one).
template <class C, int LO, int HI, typename F>
__device__ __forceinline__ void digest_reg(const BasisDev &b, const DensDev &dd,
                                           const QuartetGeom &g, bool same_pair,
                                           unsigned long long mask, bool on, bool bra_uniform,
                                           int leader, const F *acc) {
  constexpr int nb = C::nb, nc = C::nc, nd = C::nd;
  constexpr int A0 = LO / (nb * nc * nd), A1 = (HI - 1) / (nb * nc * nd) + 1;  // a rows touched
#define IG_E(a, bb, c, d) (((a) * nb + (bb)) * nc + (c)) * nd + (d)
#define IG_IN(e) ((e) >= LO && (e) < HI)
  const double f = (g.sA == g.sB ? 0.5 : 1.0) * (g.sC == g.sD ? 0.5 : 1.0) * (same_pair ? 0.5 : 1.0);
  const size_t n = (size_t)b.nao_cart, n2 = n * n;
  const size_t a0 = b.off_cart[g.sA], b0 = b.off_cart[g.sB], c0 = b.off_cart[g.sC],
               d0 = b.off_cart[g.sD];
  const int lane = threadIdx.x & 63;
  const size_t xo = jk_copy_offset(dd);
  for (int e = 0; e < dd.neff; ++e) {
    const bool use = on && ((mask >> dd.eff_src[e]) & 1ull);
    const bool anti = dd.eff_anti[e];  // uniform
    const double *D = dd.dc + (size_t)e * n2;
    if (!anti) {
      double *J = dd.jc + xo + (size_t)e * n2;
      double jcd[nc * nd];
#pragma unroll
      for (int q = 0; q < nc * nd; ++q) jcd[q] = 0.0;
#pragma unroll
      for (int a = A0; a < A1; ++a)
#pragma unroll
        for (int bb = 0; bb < nb; ++bb) {
          // skip (a,b) rows that do not overlap the slice
          if (IG_E(a, bb, 0, 0) >= HI || IG_E(a, bb, nc - 1, nd - 1) < LO) continue;
          double v = 0.0;
          if (use) {
            const double dab = D[(a0 + a) * n + b0 + bb];
#pragma unroll
            for (int c = 0; c < nc; ++c)
#pragma unroll
              for (int d = 0; d < nd; ++d) {
                const int ee = IG_E(a, bb, c, d);
                if (IG_IN(ee)) {
                  const double gv = double(acc[ee - LO]);
                  v += gv * D[(c0 + c) * n + d0 + d];
                  jcd[c * nd + d] += gv * dab;
                }
              }
          }
          if (bra_uniform) {  // f depends on the lane's ket: scale before summing
            v = wave_sum(2.0 * f * v);
            if (lane == leader && v != 0.0) jk_add(dd, J + (a0 + a) * n + b0 + bb, v);
          } else if (use) {
            jk_add(dd, J + (a0 + a) * n + b0 + bb, 2.0 * f * v);
          }
        }
      if (use) {
#pragma unroll
        for (int c = 0; c < nc; ++c)
#pragma unroll
          for (int d = 0; d < nd; ++d)
            jk_add(dd, J + (c0 + c) * n + d0 + d, 2.0 * f * jcd[c * nd + d]);
      }
    }
    if (use && dd.kc && dd.want_k) {
      double *K = dd.kc + xo + (size_t)e * n2;
#pragma unroll
      for (int a = A0; a < A1; ++a)
#pragma unroll
        for (int c = 0; c < nc; ++c) {  // Kp[a,c] += g D[b,d]
          double v = 0.0;
          bool any = false;
#pragma unroll
          for (int bb = 0; bb < nb; ++bb)
#pragma unroll
            for (int d = 0; d < nd; ++d)
              if (IG_IN(IG_E(a, bb, c, d))) {
                v += double(acc[IG_E(a, bb, c, d) - LO]) * D[(b0 + bb) * n + d0 + d];
                any = true;
              }
          if (any) jk_add(dd, K + (a0 + a) * n + c0 + c, f * v);
        }
#pragma unroll
      for (int bb = 0; bb < nb; ++bb)
#pragma unroll
        for (int c = 0; c < nc; ++c) {  // Kp[b,c] += g D[a,d]
          double v = 0.0;
          bool any = false;
#pragma unroll
          for (int a = A0; a < A1; ++a)
#pragma unroll
            for (int d = 0; d < nd; ++d)
              if (IG_IN(IG_E(a, bb, c, d))) {
                v += double(acc[IG_E(a, bb, c, d) - LO]) * D[(a0 + a) * n + d0 + d];
                any = true;
              }
          if (any) jk_add(dd, K + (b0 + bb) * n + c0 + c, f * v);
        }
#pragma unroll
      for (int a = A0; a < A1; ++a)
#pragma unroll
        for (int d = 0; d < nd; ++d) {  // Kp[a,d] += g D[b,c]
          double v = 0.0;
          bool any = false;
#pragma unroll
          for (int bb = 0; bb < nb; ++bb)
#pragma unroll
            for (int c = 0; c < nc; ++c)
              if (IG_IN(IG_E(a, bb, c, d))) {
                v += double(acc[IG_E(a, bb, c, d) - LO]) * D[(b0 + bb) * n + c0 + c];
                any = true;
              }
          if (any) jk_add(dd, K + (a0 + a) * n + d0 + d, f * v);
        }
#pragma unroll
      for (int bb = 0; bb < nb; ++bb)
#pragma unroll
        for (int d = 0; d < nd; ++d) {  // Kp[b,d] += g D[a,c]
          double v = 0.0;
          bool any = false;
#pragma unroll
          for (int a = A0; a < A1; ++a)
#pragma unroll
            for (int c = 0; c < nc; ++c)
              if (IG_IN(IG_E(a, bb, c, d))) {
                v += double(acc[IG_E(a, bb, c, d) - LO]) * D[(a0 + a) * n + c0 + c];
                any = true;
              }
          if (any) jk_add(dd, K + (b0 + bb) * n + d0 + d, f * v);
        }
    }
  }
#undef IG_E
#undef IG_IN
}

// Lane-per-quartet J/K kernel for block slice [LO, HI).  Every lane of a
// wavefront runs the same number of iterations (lanes past the range or with
// an empty density mask compute nothing) so wavefront reductions see all 64
// lanes.  Only slice 0 counts quartets.
template <class C, int LO, int HI, typename F>
__global__ void __launch_bounds__(64) jk_reg_kernel(BasisDev b, PairsDev p, RysDev R,
                                                    QuartetSet qs, DensDev dd, double omega,
                                                    unsigned long long *counters) {
  extern __shared__ double rys_lds[];
  stage_rys<C::NR>(R, rys_lds);
  unsigned long long my_quartets = 0, my_prims = 0;
  const int nsh = b.nsh;
  const int lane = threadIdx.x & 63;
  long long begin, end;
  wave_range(qs.total, begin, end);
  RowCursor cur;
  for (long long base = begin; base < end; base += 64) {
    const long long t = base + lane;
    bool on = t < end;
    int P = 0, Q = 0;
    if (on) cur.next(qs, t, P, Q);
    const int pb = qs.x_off + P, pk = qs.y_off + Q;
    QuartetGeom g;
    load_geom(b, p.sh1[pb], p.sh2[pb], p.sh1[pk], p.sh2[pk], g);
    unsigned long long mask = 0;
    if (on) {
      const double pair_bound = p.q[pb] * p.q[pk];
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
      on = mask != 0;
    }
    F acc[HI - LO];
#pragma unroll
    for (int e = 0; e < HI - LO; ++e) acc[e] = F(0);
    if (on) {
      quartet_block_reg<C, LO, HI, F>(p, R, rys_lds, pb, pk, g, omega, acc);
      if (LO == 0) {
        ++my_quartets;
        my_prims += (unsigned long long)p.pp_n[pb] * p.pp_n[pk];
      }
    }
    // bra pair shared by every lane with work?  (no lane with work: leader -1,
    // nothing is written)
    const int leader = __ffsll((unsigned long long)__ballot(on)) - 1;
    const int pb_first = __shfl(pb, leader < 0 ? 0 : leader);
    const bool bra_uniform = __all(!on || pb == pb_first);
    digest_reg<C, LO, HI, F>(b, dd, g, pb == pk, mask, on, bra_uniform, leader, acc);
  }
  if (LO == 0) {
    for (int off = 32; off > 0; off >>= 1) {
      my_quartets += __shfl_down(my_quartets, off);
      my_prims += __shfl_down(my_prims, off);
    }
    if (counters && lane == 0) {
      atomicAdd(counters, my_quartets);
      atomicAdd(counters + 1, my_prims);
    }
  }
}

// number of register slices for a class and the launch of all of them
template <class C, typename F>
struct RegSlices {
  static constexpr int CAP = kRegBlockMax;  // float blocks too: larger ones spill
  static constexpr int NS = (C::N4 + CAP - 1) / CAP;
  static constexpr int E = (C::N4 + NS - 1) / NS;
};

template <class C, int S, typename F>
hipError_t launch_jk_reg_slices(const BasisDev &b, const PairsDev &p, const RysDev &r,
                                const QuartetSet &qs, const DensDev &d, double omega,
                                unsigned long long *ctr, hipStream_t s, int blocks, size_t lds) {
  constexpr int NS = RegSlices<C, F>::NS, E = RegSlices<C, F>::E;
  if constexpr (S < NS) {
    constexpr int LO = S * E, HI = (S + 1) * E < C::N4 ? (S + 1) * E : C::N4;
    hipLaunchKernelGGL((jk_reg_kernel<C, LO, HI, F>), dim3(blocks), dim3(64), lds, s, b, p, r,
                       qs, d, omega, ctr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_jk_reg_slices<C, S + 1, F>(b, p, r, qs, d, omega, ctr, s, blocks, lds);
  } else {
    return hipSuccess;
  }
}
