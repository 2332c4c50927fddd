// ig_rowacc.h -- lane-per-quartet J/K path with per-row LDS accumulation
// (included inside namespace ig by ig_lane.h, after ig_reg.h).
//
// Device-scope FP64 atomics are the ceiling of the plain register path: every
// quartet sends na*nc + nb*nc + na*nd + nb*nd exchange updates (and the
// ket-side Coulomb ones) to memory, and a whole build issues them at a fixed
// rate whatever the class.  Here a workgroup of kRowWaves wavefronts walks a
// contiguous range of the screened quartet list, 64 quartets per wavefront per
// step, and keeps the running sums of the bra row under accumulation (the row
// of the first quartet of the step) in LDS:
//   JB[e][a][b]                   bra-side Coulomb
//   KA[e][a][col], KB[e][b][col]  exchange rows of the bra's two shells
// `col` runs over a compact column space: the Cartesian functions of every
// shell with the ket class's first momentum, then (when it differs) of every
// shell with its second.  Updates go to LDS with LDS atomics; when the row
// changes the workgroup adds the nonzero entries to global memory with one
// atomic each and clears them.  A row's kets share few shells, so each entry
// collects many quartets.  Quartets of another row inside a step (row
// boundaries) and the ket-side Coulomb updates take the global path of
// jk_reg_kernel.

constexpr int kRowWaves = 4;

// Compact column space of the ket class (momenta LC >= LD) of class C.
template <class C>
struct RowCols {
  int W, dbase;
  __device__ explicit RowCols(const BasisDev &b) {
    const int nC = b.lshell_off[C::LC + 1] - b.lshell_off[C::LC];
    const int nD = b.lshell_off[C::LD + 1] - b.lshell_off[C::LD];
    dbase = (C::LC == C::LD) ? 0 : nC * C::nc;
    W = nC * C::nc + (C::LC == C::LD ? 0 : nD * C::nd);
  }
  __device__ int col_c(const BasisDev &b, int sC, int c) const { return b.rank[sC] * C::nc + c; }
  __device__ int col_d(const BasisDev &b, int sD, int d) const {
    return dbase + b.rank[sD] * C::nd + d;
  }
  // column -> Cartesian function index
  __device__ int cart(const BasisDev &b, int col) const {
    if (C::LC == C::LD || col < dbase)
      return b.off_cart[b.lshells[b.lshell_off[C::LC] + col / C::nc]] + col % C::nc;
    col -= dbase;
    return b.off_cart[b.lshells[b.lshell_off[C::LD] + col / C::nd]] + col % C::nd;
  }
};

// Doubles of LDS the accumulators of class C need for `neff` densities and W
// columns (host and device agree on this layout: JB, KA, KB).
__host__ __device__ constexpr size_t rowacc_doubles(int na, int nb, int neff, int W) {
  return size_t(neff) * (size_t(na) * nb + size_t(na + nb) * W);
}

// Add the accumulated row `row` to global J'/K' and clear the accumulators.
template <class C>
__device__ void rowacc_flush(const BasisDev &b, const PairsDev &p, const QuartetSet &qs,
                             const DensDev &dd, const RowCols<C> &rc, int row, double *JB,
                             size_t xo) {
  constexpr int na = C::na, nb = C::nb;
  const int pb = qs.x_off + row;
  const size_t a0 = b.off_cart[p.sh1[pb]], b0 = b.off_cart[p.sh2[pb]];
  const size_t n = b.nao_cart, n2 = n * n;
  const int neff = dd.neff, W = rc.W;
  const int nJ = neff * na * nb, nKA = neff * na * W;
  const int ntot = int(rowacc_doubles(na, nb, neff, W));
  for (int i = threadIdx.x; i < ntot; i += blockDim.x) {
    const double v = JB[i];
    if (v == 0.0) continue;
    JB[i] = 0.0;
    double *dst;
    if (i < nJ) {
      const int e = i / (na * nb), r = i % (na * nb);
      dst = dd.jc + xo + e * n2 + (a0 + r / nb) * n + b0 + r % nb;
    } else if (i < nJ + nKA) {
      const int k = i - nJ, e = k / (na * W), r = k % (na * W);
      dst = dd.kc + xo + e * n2 + (a0 + r / W) * n + rc.cart(b, r % W);
    } else {
      const int k = i - nJ - nKA, e = k / (nb * W), r = k % (nb * W);
      dst = dd.kc + xo + e * n2 + (b0 + r / W) * n + rc.cart(b, r % W);
    }
    atomicAdd(dst, v);
  }
  __syncthreads();
}

// Digest of the block elements [LO, HI) of lanes in the accumulated row:
// bra-side Coulomb and exchange into LDS, ket-side Coulomb to global memory.
// Called by every lane (`in` false for lanes that do not take part).
template <class C, int LO, int HI, typename F>
__device__ __forceinline__ void digest_row(const BasisDev &b, const DensDev &dd,
                                           const QuartetGeom &g, bool same_pair,
                                           unsigned long long mask, bool in, bool all_in,
                                           const RowCols<C> &rc, double *JB, size_t xo,
                                           const F *acc) {
  constexpr int nb = C::nb, nc = C::nc, nd = C::nd;
  constexpr int na = C::na;
  constexpr int A0 = LO / (nb * nc * nd), A1 = (HI - 1) / (nb * nc * nd) + 1;
#define IG_E(a, bb, c, d) (((a) * nb + (bb)) * nc + (c)) * nd + (d)
#define IG_IN(e) ((e) >= LO && (e) < HI)
  const double f = (g.sA == g.sB ? 0.5 : 1.0) * (g.sC == g.sD ? 0.5 : 1.0) * (same_pair ? 0.5 : 1.0);
  const size_t n = (size_t)b.nao_cart, n2 = n * n;
  const size_t a0 = b.off_cart[g.sA], b0 = b.off_cart[g.sB], c0 = b.off_cart[g.sC],
               d0 = b.off_cart[g.sD];
  const int lane = threadIdx.x & 63;
  const int W = rc.W;
  const int colc = in ? rc.col_c(b, g.sC, 0) : 0, cold = in ? rc.col_d(b, g.sD, 0) : 0;
  for (int e = 0; e < dd.neff; ++e) {
    const bool use = in && ((mask >> dd.eff_src[e]) & 1ull);
    const bool anti = dd.eff_anti[e];  // uniform
    const double *D = dd.dc + (size_t)e * n2;
    if (!anti) {
      double *J = dd.jc + xo + (size_t)e * n2;
      double *jb = JB + e * na * nb;
      double jcd[nc * nd];
#pragma unroll
      for (int q = 0; q < nc * nd; ++q) jcd[q] = 0.0;
#pragma unroll
      for (int a = A0; a < A1; ++a)
#pragma unroll
        for (int bb = 0; bb < nb; ++bb) {
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
          if (all_in) {  // one row for the whole wavefront: reduce, one LDS add
            v = wave_sum(2.0 * f * v);
            if (lane == 0 && v != 0.0) atomicAdd(jb + a * nb + bb, v);
          } else if (use) {
            atomicAdd(jb + a * nb + bb, 2.0 * f * v);
          }
        }
      if (use) {
#pragma unroll
        for (int c = 0; c < nc; ++c)
#pragma unroll
          for (int d = 0; d < nd; ++d)
            atomicAdd(J + (c0 + c) * n + d0 + d, 2.0 * f * jcd[c * nd + d]);
      }
    }
    if (use && dd.kc && dd.want_k) {
      double *ka = JB + dd.neff * na * nb + (size_t)e * na * W;
      double *kb = JB + dd.neff * (na * nb + na * W) + (size_t)e * nb * W;
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
          if (any) atomicAdd(ka + a * W + colc + c, f * v);
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
          if (any) atomicAdd(kb + bb * W + colc + c, f * v);
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
          if (any) atomicAdd(ka + a * W + cold + d, f * v);
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
          if (any) atomicAdd(kb + bb * W + cold + d, f * v);
        }
    }
  }
#undef IG_E
#undef IG_IN
}

// Quartets [begin, end) of one basis's set: the steps, the row sums and the
// final flush.  Called uniformly by the whole workgroup; JB holds zeros on
// entry and on return.  ENG 0: Rys quartet block (quartet_block_reg); ENG 1:
// the OS/HGP block of ig_hgp.h (total angular momentum <= 2, Coulomb only).
template <class C, int LO, int HI, typename F, int ENG>
__device__ __forceinline__ void rowacc_range(const BasisDev &b, const PairsDev &p,
                                             const RysDev &R, const QuartetSet &qs,
                                             const DensDev &dd, double omega, long long begin,
                                             long long end, const double *rys_lds, double *JB,
                                             unsigned long long &my_quartets,
                                             unsigned long long &my_prims) {
  const RowCols<C> rc(b);
  const int nsh = b.nsh;
  const size_t xo = jk_copy_offset(dd);
  RowCursor cur, bcur;
  int racc = -1;  // accumulated row, uniform over the workgroup
  for (long long base = begin; base < end; base += blockDim.x) {
    const long long t = base + threadIdx.x;
    bool on = t < end;
    int P = 0, Q = 0;
    if (on) cur.next(qs, t, P, Q);
    // the step's row is that of its first quartet, which every thread
    // tracks itself (a second cursor on `base`), so the workgroup agrees on
    // it without exchanging it; wavefronts only meet at a row change, where
    // the finished row's LDS sums are flushed (the sums themselves are LDS
    // atomics, safe between wavefronts at different steps)
    int row, Qb;
    bcur.next(qs, base, row, Qb);
    if (row != racc) {
      if (racc >= 0) {
        __syncthreads();
        rowacc_flush<C>(b, p, qs, dd, rc, racc, JB, xo);
      }
      racc = row;
    }
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
      if constexpr (ENG == 0)
        quartet_block_reg<C, LO, HI, F>(p, R, rys_lds, pb, pk, g, omega, acc);
      else
        quartet_block_hgp<C, LO, HI, F>(p, R, pb, pk, g, acc);
      if (LO == 0) {
        ++my_quartets;
        my_prims += (unsigned long long)p.pp_n[pb] * p.pp_n[pk];
      }
    }
    const bool in = on && P == racc;
    const bool out = on && P != racc;
    if (__any(in))
      digest_row<C, LO, HI, F>(b, dd, g, pb == pk, mask, in, __all(!on || in), rc, JB, xo, acc);
    if (__any(out)) {
      // other rows: the global path, wavefront-reduced when they share a bra
      const int leader = __ffsll((unsigned long long)__ballot(out)) - 1;
      const int pb_first = __shfl(pb, leader);
      const bool bra_uniform = __all(!out || pb == pb_first);
      digest_reg<C, LO, HI, F>(b, dd, g, pb == pk, mask, out, bra_uniform, leader, acc);
    }
  }
  __syncthreads();
  if (racc >= 0) rowacc_flush<C>(b, p, qs, dd, rc, racc, JB, xo);
}

template <int LO>
__device__ __forceinline__ void rowacc_count(unsigned long long q, unsigned long long pr,
                                             unsigned long long *counters) {
  if (LO != 0) return;
  for (int off = 32; off > 0; off >>= 1) {
    q += __shfl_down(q, off);
    pr += __shfl_down(pr, off);
  }
  if (counters && (threadIdx.x & 63) == 0) {
    atomicAdd(counters, q);
    atomicAdd(counters + 1, pr);
  }
}

template <class C, int LO, int HI, typename F, int ENG = 0>
__global__ void __launch_bounds__(64 * kRowWaves)
    jk_rowacc_kernel(BasisDev b, PairsDev p, RysDev R, QuartetSet qs, DensDev dd, double omega,
                     unsigned long long *counters) {
  extern __shared__ double lds[];
  if constexpr (ENG == 0) stage_rys<C::NR>(R, lds);
  const int nrys = ENG == 0 ? 2 * R.nsub[C::NR] * C::NR * kRysTerms : 0;
  double *JB = lds + nrys;
  const int nacc = int(rowacc_doubles(C::na, C::nb, dd.neff, RowCols<C>(b).W));
  for (int i = threadIdx.x; i < nacc; i += blockDim.x) JB[i] = 0.0;
  __syncthreads();  // zeros (and the staged Rys tables) before any use
  unsigned long long my_quartets = 0, my_prims = 0;
  const long long chunk = (qs.total + gridDim.x - 1) / gridDim.x;
  const long long begin = (long long)blockIdx.x * chunk;
  const long long end = begin + chunk < qs.total ? begin + chunk : qs.total;
  rowacc_range<C, LO, HI, F, ENG>(b, p, R, qs, dd, omega, begin, end, lds, JB, my_quartets,
                                  my_prims);
  rowacc_count<LO>(my_quartets, my_prims, counters);
}

// Many bases in one launch: the quartet sets of class C of every basis are
// one concatenated list (seg: nb + 1 prefix sums of their totals).
// Workgroups take chunks of it from a device work queue (one atomic counter,
// zeroed before the launch), so molecules of uneven size balance.  A chunk
// may span bases; each part runs with that basis's views (read from the
// device arrays bs, ps, sets, dds).  The LDS accumulators hold nacc doubles:
// the host's size for the widest basis that has quartets of this class (the
// launch's dynamic LDS is sized from the same number).
template <class C, int LO, int HI, typename F, int ENG = 0>
__global__ void __launch_bounds__(64 * kRowWaves)
    jk_rowacc_batch_kernel(const BasisDev *bs, const PairsDev *ps, RysDev R,
                           const QuartetSet *sets, const DensDev *dds, const long long *seg,
                           int nb, unsigned *queue, long long chunk, int nacc,
                           unsigned long long *counters) {
  extern __shared__ double lds[];
  __shared__ long long s_item[2];
  if constexpr (ENG == 0) stage_rys<C::NR>(R, lds);
  const int nrys = ENG == 0 ? 2 * R.nsub[C::NR] * C::NR * kRysTerms : 0;
  double *JB = lds + nrys;
  for (int i = threadIdx.x; i < nacc; i += blockDim.x) JB[i] = 0.0;
  unsigned long long my_quartets = 0, my_prims = 0;
  const long long total = seg[nb];
  for (int it = 0;; ++it) {
    if (threadIdx.x == 0) s_item[it & 1] = (long long)atomicAdd(queue, 1u);
    __syncthreads();  // also orders the accumulator zeroing before first use
    long long g0 = s_item[it & 1] * chunk;
    if (g0 >= total) break;  // uniform: every workgroup leaves once the list is done
    const long long g1 = g0 + chunk < total ? g0 + chunk : total;
    while (g0 < g1) {
      int lo = 0, hi = nb - 1;  // basis holding g0: seg[m] <= g0 < seg[m + 1]
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= g0)
          lo = mid;
        else
          hi = mid - 1;
      }
      const long long e = g1 < seg[lo + 1] ? g1 : seg[lo + 1];
      const BasisDev b = bs[lo];
      const PairsDev p = ps[lo];
      const QuartetSet qs = sets[lo];
      const DensDev dd = dds[lo];
      rowacc_range<C, LO, HI, F, ENG>(b, p, R, qs, dd, 0.0, g0 - seg[lo], e - seg[lo], lds, JB,
                                      my_quartets, my_prims);
      g0 = e;
    }
  }
  rowacc_count<LO>(my_quartets, my_prims, counters);
}

template <class C, int S, typename F, int ENG = 0>
hipError_t launch_jk_rowacc_slices(const BasisDev &b, const PairsDev &p, const RysDev &r,
                                   const QuartetSet &qs, const DensDev &d, double omega,
                                   unsigned long long *ctr, hipStream_t s, int blocks,
                                   size_t lds) {
  constexpr int NS = RegSlices<C, F>::NS, E = RegSlices<C, F>::E;
  if constexpr (S < NS) {
    constexpr int LO = S * E, HI = (S + 1) * E < C::N4 ? (S + 1) * E : C::N4;
    // about four rounds of the workgroups that fit at once: one workgroup per
    // CU for the register-heavy classes, several for the small ones
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void *>(&jk_rowacc_kernel<C, LO, HI, F, ENG>),
            64 * kRowWaves, lds) != hipSuccess ||
        per_cu < 1)
      per_cu = 1;
    const long long cap = 4LL * device_cu_count() * per_cu;
    const int grid = int(blocks < cap ? blocks : cap);
    hipLaunchKernelGGL((jk_rowacc_kernel<C, LO, HI, F, ENG>), dim3(grid), dim3(64 * kRowWaves),
                       lds, s, b, p, r, qs, d, omega, ctr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_jk_rowacc_slices<C, S + 1, F, ENG>(b, p, r, qs, d, omega, ctr, s, blocks, lds);
  } else {
    return hipSuccess;
  }
}

template <class C, int S, typename F, int ENG = 0>
hipError_t launch_jk_rowacc_batch_slices(const BasisDev *bs, const PairsDev *ps, const RysDev &r,
                                         const QuartetSet *sets, const DensDev *dds,
                                         const long long *seg, int nb, unsigned *queues,
                                         long long chunk, int nacc, unsigned long long *ctr,
                                         hipStream_t s, int blocks, size_t lds) {
  constexpr int NS = RegSlices<C, F>::NS, E = RegSlices<C, F>::E;
  if constexpr (S < NS) {
    constexpr int LO = S * E, HI = (S + 1) * E < C::N4 ? (S + 1) * E : C::N4;
    hipLaunchKernelGGL((jk_rowacc_batch_kernel<C, LO, HI, F, ENG>), dim3(blocks),
                       dim3(64 * kRowWaves), lds, s, bs, ps, r, sets, dds, seg, nb, queues + S,
                       chunk, nacc, ctr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_jk_rowacc_batch_slices<C, S + 1, F, ENG>(bs, ps, r, sets, dds, seg, nb, queues,
                                                          chunk, nacc, ctr, s, blocks, lds);
  } else {
    return hipSuccess;
  }
}
