// ig_hgp.h -- Obara-Saika / Head-Gordon-Pople quartet block for the classes
// of total angular momentum <= 2, and of total 3 with an (s s| or (p s| ket
// and a second bra shell of at most p (included inside namespace ig by
// ig_lane.h).
//
// Per primitive quartet: the Boys values F_0..F_L (boys_eval), the bra
// vertical recurrence
//   [e+1_i|00]^(m) = PA_i [e|00]^(m) + WP_i [e|00]^(m+1)
//                    + e_i/(2 zeta) ([e-1_i|00]^(m) - rho/zeta [e-1_i|00]^(m+1))
// and, for a p ket, the ket transfer
//   [e|f+1_j]^(0) = QC_j [e|f]^(0) + WQ_j [e|f]^(1) + e_j/(2(zeta+eta)) [e-1_j|f]^(1)
// (engine/os_eri.rs:455-697, 925-1215 recurrences), contracted over the
// primitives; then one horizontal transfer (a, b+1_j| = (a+1_j, b| + AB_j (a, b|
// in contracted space for a (p p| bra (os_eri.rs:3040-3143).  The total-2
// classes are 40% of the quartets of an s,p,d basis; with one or two Boys
// orders they need no Rys roots.
constexpr int hgp_total(int la, int lb, int lc, int ld) { return la + lb + lc + ld; }

// Total angular momentum 3 is covered for kets (p s| or (s s| and a bra
// whose second shell is s or p: (pp|ps), (ds|ps), (dp|ss) of an s,p,d basis
// (quartet_block_hgp3).
constexpr bool hgp_covered(int la, int lb, int lc, int ld) {
  return hgp_total(la, lb, lc, ld) <= 2 ||
         (hgp_total(la, lb, lc, ld) == 3 && ld == 0 && lc <= 1 && lb <= 1);
}

template <class C>
struct HgpClass {
  static constexpr int L = hgp_total(C::LA, C::LB, C::LC, C::LD);
  static constexpr bool ok = hgp_covered(C::LA, C::LB, C::LC, C::LD);
};

// Cartesian component i of degree l (lx descending, then ly descending):
// its power along axis ax, the index of the component with powers (x, y, z),
// and the axis the recurrence steps down along (the first nonzero power).
__host__ __device__ constexpr int hgp_pow(int l, int i, int ax) {
  int d = 0;
  while ((d + 1) * (d + 2) / 2 <= i) ++d;
  const int z = i - d * (d + 1) / 2;
  return ax == 0 ? l - d : (ax == 1 ? d - z : z);
}
__host__ __device__ constexpr int hgp_index(int y, int z) {
  const int d = y + z;
  return d * (d + 1) / 2 + z;
}
__host__ __device__ constexpr int hgp_dir(int l, int i) {
  return hgp_pow(l, i, 0) > 0 ? 0 : (hgp_pow(l, i, 1) > 0 ? 1 : 2);
}
// component of degree l - 1 below component i of degree l along axis ax
__host__ __device__ constexpr int hgp_down(int l, int i, int ax) {
  return hgp_index(hgp_pow(l, i, 1) - (ax == 1), hgp_pow(l, i, 2) - (ax == 2));
}
// component of degree l + 1 above component i of degree l along axis ax
__host__ __device__ constexpr int hgp_up(int l, int i, int ax) {
  return hgp_index(hgp_pow(l, i, 1) + (ax == 1), hgp_pow(l, i, 2) + (ax == 2));
}

// Slots of the bra vertical-recurrence values [e|00]^(m): degree n <= NB,
// Boys order m <= L - n.
template <int L, int NB>
struct HgpVrr {
  static constexpr int base(int n) {
    int s = 0;
    for (int k = 0; k < n; ++k) s += ncart(k) * (L + 1 - k);
    return s;
  }
  static constexpr int at(int n, int c, int m) { return base(n) + c * (L + 1 - n) + m; }
  static constexpr int size = base(NB + 1);
};

// The total-L = 3 classes of hgp_covered: per primitive quartet the Boys
// values F_0..F_3, the bra vertical recurrence
//   [e+1_i|00]^(m) = PA_i [e|00]^(m) + WP_i [e|00]^(m+1)
//                    + e_i/(2 zeta) ([e-1_i|00]^(m) - rho/zeta [e-1_i|00]^(m+1))
// to degree la + lb, and for a (p s| ket the transfer
//   [e|1_j]^(0) = QC_j [e|00]^(0) + WQ_j [e|00]^(1) + e_j/(2(zeta+eta)) [e-1_j|00]^(1)
// (os_eri.rs:455-697, 925-1215), contracted over the primitives in [e|f]
// space (degrees la and la + lb); then one horizontal transfer
// (a, 1_j| = (a+1_j| + AB_j (a| per quartet (os_eri.rs:3040-3143).
template <class C, typename F>
__device__ __forceinline__ void quartet_block_hgp3(const PairsDev &p, const RysDev &R, int pb,
                                                   int pk, const QuartetGeom &g, F *acc) {
  constexpr int LA = C::LA, LB = C::LB, LC = C::LC, NB = LA + LB, L = NB + LC;
  static_assert(L == 3 && C::LD == 0 && LC <= 1 && LB <= 1, "OS/HGP block of total L = 3");
  using V = HgpVrr<L, NB>;
  constexpr int NLO = ncart(LA), NHI = LB ? ncart(NB) : 0, NF = ncart(LC);
  double ce[(NLO + NHI) * NF];  // [e|f], e of degree LA then of degree NB
#pragma unroll
  for (int k = 0; k < (NLO + NHI) * NF; ++k) ce[k] = 0.0;
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
      double PA[3], WP[3];
#pragma unroll
      for (int x = 0; x < 3; ++x) {
        PA[x] = P[x] - g.A[x];
        WP[x] = -qz * ipq * PQ[x];
      }
      const double r2p = rho * 2.0 * i2p;  // rho / zeta
      double v[V::size];
#pragma unroll
      for (int m = 0; m <= L; ++m) v[V::at(0, 0, m)] = Fm[m] * pref;
#pragma unroll
      for (int n = 1; n <= NB; ++n)
#pragma unroll
        for (int c = 0; c < ncart(n); ++c) {
          const int i = hgp_dir(n, c), ei = hgp_pow(n, c, i);
          const int c1 = hgp_down(n, c, i);
#pragma unroll
          for (int m = 0; m <= L - n; ++m) {
            double t = PA[i] * v[V::at(n - 1, c1, m)] + WP[i] * v[V::at(n - 1, c1, m + 1)];
            if (ei >= 2) {
              const int c2 = hgp_down(n - 1, c1, i);
              t += double(ei - 1) * i2p * (v[V::at(n - 2, c2, m)] - r2p * v[V::at(n - 2, c2, m + 1)]);
            }
            v[V::at(n, c, m)] = t;
          }
        }
      if constexpr (LC == 0) {
#pragma unroll
        for (int c = 0; c < NLO; ++c) ce[c] += v[V::at(LA, c, 0)];
        if constexpr (LB > 0) {
#pragma unroll
          for (int c = 0; c < NHI; ++c) ce[NLO + c] += v[V::at(NB, c, 0)];
        }
      } else {
        double QC[3], WQ[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          QC[x] = Q[x] - g.C[x];
          WQ[x] = pz * ipq * PQ[x];
        }
        const double half = 0.5 * ipq;
#pragma unroll
        for (int h = 0; h < (LB ? 2 : 1); ++h) {
          const int deg = h ? NB : LA, nco = ncart(deg), off = h ? NLO : 0;
#pragma unroll
          for (int c = 0; c < nco; ++c)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
              double t = QC[j] * v[V::at(deg, c, 0)] + WQ[j] * v[V::at(deg, c, 1)];
              const int ej = hgp_pow(deg, c, j);
              if (ej >= 1) t += double(ej) * half * v[V::at(deg - 1, hgp_down(deg, c, j), 1)];
              ce[(off + c) * NF + j] += t;
            }
        }
      }
    }
  }
  // block elements (a, b, c, d), d fastest (nd = 1)
  if constexpr (LB == 0) {
#pragma unroll
    for (int k = 0; k < NLO * NF; ++k) acc[k] += F(ce[k]);
  } else {  // (a 1_j| = (a+1_j| + AB_j (a|
    const double AB[3] = {g.A[0] - g.B[0], g.A[1] - g.B[1], g.A[2] - g.B[2]};
#pragma unroll
    for (int a = 0; a < NLO; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int f = 0; f < NF; ++f)
          acc[(a * 3 + j) * NF + f] +=
              F(ce[(NLO + hgp_up(LA, a, j)) * NF + f] + AB[j] * ce[a * NF + f]);
  }
}

// d Cartesian component index (lx desc, ly desc) of p_i + p_j
__device__ __forceinline__ constexpr int d_of(int i, int j) {
  // xx xy xz yy yz zz
  return i == 0 ? j : (i == 1 ? (j == 0 ? 1 : (j == 1 ? 3 : 4)) : (j == 0 ? 2 : (j == 1 ? 4 : 5)));
}

// The classes of total L <= 2.
template <class C, typename F>
__device__ __forceinline__ void quartet_block_hgp2(const PairsDev &p, const RysDev &R, int pb,
                                                   int pk, const QuartetGeom &g, F *acc) {
  constexpr int L = HgpClass<C>::L;
  static_assert(L <= 2, "OS/HGP block of total L <= 2");
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

// OS/HGP block of quartet (pb | pk) into acc (all N4 elements: never sliced).
template <class C, int LO, int HI, typename F>
__device__ __forceinline__ void quartet_block_hgp(const PairsDev &p, const RysDev &R, int pb,
                                                  int pk, const QuartetGeom &g, F *acc) {
  static_assert(HgpClass<C>::ok, "OS/HGP block: class not covered (hgp_covered)");
  static_assert(LO == 0 && HI == C::N4, "OS/HGP block is never sliced");
  if constexpr (HgpClass<C>::L == 3)
    quartet_block_hgp3<C, F>(p, R, pb, pk, g, acc);
  else
    quartet_block_hgp2<C, F>(p, R, pb, pk, g, acc);
}
