// ig_hgp.h -- Obara-Saika / Head-Gordon-Pople quartet block for the classes
// of total angular momentum <= 2 (included inside namespace ig by ig_lane.h).
//
// Per primitive quartet: the Boys values F_0..F_L (boys_eval), the bra
// vertical recurrence
//   [e+1_i|00]^(m) = PA_i [e|00]^(m) + WP_i [e|00]^(m+1)
//                    + e_i/(2 zeta) ([e-1_i|00]^(m) - rho/zeta [e-1_i|00]^(m+1))
// and, for a p ket, the ket transfer
//   [e|f+1_j]^(0) = QC_j [e|f]^(0) + WQ_j [e|f]^(1) + e_j/(2(zeta+eta)) [e-1_j|f]^(1)
// (engine/os_eri.rs:455-697, 925-1215 recurrences), contracted over the
// primitives; then one horizontal transfer (a, b+1_j| = (a+1_j, b| + AB_j (a, b|
// in contracted space for a (p p| bra (os_eri.rs:3040-3143).  These classes are
// 40% of the quartets of an s,p,d basis; with one or two Boys orders they need
// no Rys roots.
constexpr int hgp_total(int la, int lb, int lc, int ld) { return la + lb + lc + ld; }

template <class C>
struct HgpClass {
  static constexpr int L = hgp_total(C::LA, C::LB, C::LC, C::LD);
  static constexpr bool ok = L <= 2;
};

// d Cartesian component index (lx desc, ly desc) of p_i + p_j
__device__ __forceinline__ constexpr int d_of(int i, int j) {
  // xx xy xz yy yz zz
  return i == 0 ? j : (i == 1 ? (j == 0 ? 1 : (j == 1 ? 3 : 4)) : (j == 0 ? 2 : (j == 1 ? 4 : 5)));
}

template <class C, int LO, int HI, typename F>
__device__ __forceinline__ void quartet_block_hgp(const PairsDev &p, const RysDev &R, int pb,
                                                  int pk, const QuartetGeom &g, F *acc) {
  static_assert(HgpClass<C>::ok, "OS/HGP block: total angular momentum <= 2");
  static_assert(LO == 0 && HI == C::N4, "OS/HGP block is never sliced");
  constexpr int L = HgpClass<C>::L;
  constexpr int LA = C::LA, LB = C::LB, LC = C::LC;
  // contracted [e|f]: e of degree LA + LB (bra), f of degree LC (ket); for a
  // (p p| bra also [p|00] for the horizontal transfer
  constexpr int NE = ncart(LA + LB), NF = ncart(LC);
  double ce[NE * NF];
  double cp[3];
#pragma unroll
  for (int k = 0; k < NE * NF; ++k) ce[k] = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) cp[k] = 0.0;
  const int ob = p.pp_off[pb], nbp = p.pp_n[pb];
  const int ok = p.pp_off[pk], nkp = p.pp_n[pk];
  for (int ik = 0; ik < nkp; ++ik) {
    const double qz = p.zeta[ok + ik];
    const double Q[3] = {p.px[ok + ik], p.py[ok + ik], p.pz[ok + ik]};
    const double kzq = kTwoPi52 * p.kz[ok + ik];
    for (int ib = 0; ib < nbp; ++ib) {
      const double pz = p.zeta[ob + ib];
      const double P[3] = {p.px[ob + ib], p.py[ob + ib], p.pz[ob + ib]};
      const double kzp = p.kz[ob + ib], i2p = p.i2z[ob + ib];
      const double pq = pz + qz, rs = rsqrt(pq), ipq = rs * rs;
      const double rho = pz * qz * ipq;
      double PQ[3], r2 = 0.0;
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        PQ[x] = P[x] - Q[x];
        r2 += PQ[x] * PQ[x];
      }
      double Fm[L + 1];
      boys_eval(R.boys, L, rho * r2, Fm);
      const double pref = kzp * kzq * rs;
#pragma unroll
      for (int m = 0; m <= L; ++m) Fm[m] *= pref;
      if constexpr (L == 0) {
        ce[0] += Fm[0];
      } else {
        double PA[3], WP[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          PA[x] = P[x] - g.A[x];
          WP[x] = -qz * ipq * PQ[x];
        }
        // [p_i|00]^(m), m = 0 .. L-1
        double p0[3], p1[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          p0[i] = PA[i] * Fm[0] + WP[i] * Fm[1];
          if constexpr (L >= 2) p1[i] = PA[i] * Fm[1] + WP[i] * Fm[2];
        }
        if constexpr (L == 1) {  // (p s|s s)
#pragma unroll
          for (int i = 0; i < 3; ++i) ce[i] += p0[i];
        } else if constexpr (LC == 1) {  // (p s|p s): ket transfer
          const double QC[3] = {Q[0] - g.C[0], Q[1] - g.C[1], Q[2] - g.C[2]};
          const double half = 0.5 * ipq;
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              double v = QC[j] * p0[i] + (pz * ipq * PQ[j]) * p1[i];
              if (i == j) v += half * Fm[1];
              ce[i * 3 + j] += v;
            }
        } else {  // (d s|s s) and (p p|s s): [d|00]^(0)
          const double fd = i2p * (Fm[0] - rho * 2.0 * i2p * Fm[1]);
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = i; j < 3; ++j) {
              double v = PA[j] * p0[i] + WP[j] * p1[i];
              if (i == j) v += fd;
              ce[d_of(i, j)] += v;
            }
          if constexpr (LB == 1) {
#pragma unroll
            for (int i = 0; i < 3; ++i) cp[i] += p0[i];
          }
        }
      }
    }
  }
  // block elements in (a, b, c, d) order, d fastest
  if constexpr (LB == 1 && LA == 1) {  // (p_i p_j|ss) = [p_i + p_j|00] + AB_j [p_i|00]
    const double AB[3] = {g.A[0] - g.B[0], g.A[1] - g.B[1], g.A[2] - g.B[2]};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) acc[i * 3 + j] += F(ce[d_of(i < j ? i : j, i < j ? j : i)] + AB[j] * cp[i]);
  } else {
#pragma unroll
    for (int k = 0; k < NE * NF; ++k) acc[k] += F(ce[k]);
  }
}
