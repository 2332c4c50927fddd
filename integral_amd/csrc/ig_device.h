// ig_device.h -- device building blocks shared by the quartet kernels:
// Rys roots/weights, per-axis Rys 2D integral tables, wave-local team
// synchronisation and screened-quartet decoding.
#pragma once

#include "ig_internal.h"

namespace ig {

constexpr double kTwoPi52 = 34.98683665524972;  // 2 pi^{5/2}

// XCD executing this wavefront: hardware register XCC_ID, field [3:0]
// (s_getreg_b32 hwreg(HW_REG_XCC_ID = 20, offset 0, size 4)).
__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 15; }

// Offset of this wavefront's accumulator copy (see DensDev).
__device__ __forceinline__ size_t jk_copy_offset(const DensDev &d) {
  return d.ncopy > 1 ? size_t(xcc_id() % d.ncopy) * d.copy_stride : 0;
}

// J'/K' accumulation (device-scope FP64 atomic add).
__device__ __forceinline__ void jk_add(const DensDev &, double *p, double v) { atomicAdd(p, v); }

// Weight of the Cartesian element (A B | C D) of a canonical shell quartet
// against a dense two-particle density G (nc^4, row-major): G summed over the
// ordered quartets the canonical one stands for, i.e. over the distinct
// shell-level permutations (swap within the bra when its shells differ,
// within the ket likewise, bra with ket when the pairs differ), so a
// canonical-quartet sweep gives the all-ordered-quartets contraction of
// grad.rs:426-466 for any G, symmetric or not.
__device__ __forceinline__ double gamma_weight(const double *__restrict__ G, size_t nc, size_t A,
                                               size_t B, size_t C, size_t D, bool ab, bool cd,
                                               bool pq) {
  auto at = [&](size_t i, size_t j, size_t k, size_t l) {
    return G[((i * nc + j) * nc + k) * nc + l];
  };
  double w = at(A, B, C, D);
  if (ab) w += at(B, A, C, D);
  if (cd) w += at(A, B, D, C);
  if (ab && cd) w += at(B, A, D, C);
  if (pq) {
    w += at(C, D, A, B);
    if (cd) w += at(D, C, A, B);
    if (ab) w += at(C, D, B, A);
    if (ab && cd) w += at(D, C, B, A);
  }
  return w;
}

// Packed storage of one function-space element (ab|cd) of shells (sA, sB |
// sC, sD) with function indices fa..fd (packed_eri.rs:10-21 layout): each
// AO pair is put in canonical order (shell i >= j; inside one shell a >= b),
// numbered start[i(i+1)/2 + j] + (a(a+1)/2 + b on a diagonal shell pair,
// a*n_j + b otherwise), and (t|u) stored at t(t+1)/2 + u with u <= t.
// Where a block holds one packed integral twice -- (ab| and (ba| of one
// shell, or (t|u) and (u|t) of one shell pair -- only the copy the
// reference stores is written (a >= b, c >= d, u <= t; store_block,
// packed_eri.rs:300-330), so the result is exactly that element of the block.
__device__ __forceinline__ long long packed_ao_pair(const BasisDev &b, const long long *start,
                                                    int si, int a, int sj, int bb) {
  if (si < sj || (si == sj && a < bb)) {
    const int s = si, x = a;
    si = sj;
    a = bb;
    sj = s;
    bb = x;
  }
  const int nj = b.kind[sj] ? 2 * b.l[sj] + 1 : ncart(b.l[sj]);
  return start[(long long)si * (si + 1) / 2 + sj] +
         (si == sj ? (long long)a * (a + 1) / 2 + bb : (long long)a * nj + bb);
}
__device__ __forceinline__ void packed_store(const BasisDev &b, const long long *start,
                                             long long row0, double *out, int sA, int fa, int sB,
                                             int fb, int sC, int fc, int sD, int fd, double v) {
  if ((sA == sB && fa < fb) || (sC == sD && fc < fd)) return;
  long long t = packed_ao_pair(b, start, sA, fa, sB, fb);
  long long u = packed_ao_pair(b, start, sC, fc, sD, fd);
  const bool same_pair = (sA == sC && sB == sD) || (sA == sD && sB == sC);
  if (same_pair && t < u) return;
  if (t < u) {
    const long long x = t;
    t = u;
    u = x;
  }
  out[t * (t + 1) / 2 + u - row0] = v;
}

// Density-fitting stores of one function-space element of a (.. | ..) block
// of the combined basis (modes kDf3c, kDf2c; see EriOut): sh are the four
// shells, idx the element's function index in each.
__device__ __forceinline__ void df_store(const BasisDev &b, const EriOut &eo, const int *sh,
                                         const int *idx, double v) {
  if (eo.mode == kDf3c) {
    const int m = (sh[0] < eo.nmain_sh && sh[1] < eo.nmain_sh) ? 0 : 2, x = 2 - m;
    const int ap = sh[x] >= eo.dummy_sh0 ? x + 1 : x;
    const long long mu = b.off_func[sh[m]] + idx[m], nu = b.off_func[sh[m + 1]] + idx[m + 1];
    const long long p = b.off_func[sh[ap]] - eo.nao_m + idx[ap];
    eo.out[(mu * eo.nao_m + nu) * eo.naux + p] = v;
    if (sh[m] != sh[m + 1]) eo.out[(nu * eo.nao_m + mu) * eo.naux + p] = v;
  } else {  // kDf2c: (P, 1s | Q, 1s), mirrored; a diagonal shell pair keeps P >= Q
    const int p1 = sh[0] >= eo.dummy_sh0 ? 1 : 0, p2 = sh[2] >= eo.dummy_sh0 ? 3 : 2;
    if (sh[p1] == sh[p2] && idx[p1] < idx[p2]) return;
    const long long P = b.off_func[sh[p1]] - eo.nao_m + idx[p1];
    const long long Q = b.off_func[sh[p2]] - eo.nao_m + idx[p2];
    eo.out[P * eo.naux + Q] = v;
    eo.out[Q * eo.naux + P] = v;
  }
}

// Lanes of one team always sit in one wavefront, so ordering LDS traffic
// between phases only needs a wavefront-scope fence plus a scheduling barrier.
__device__ __forceinline__ void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Degree-12 polynomial of one Rys series on one cell, power basis in the
// cell variable xi in [-1, 1], by Horner.
__device__ __forceinline__ double rys_poly(const double *__restrict__ c, double xi) {
  double r = c[kRysTerms - 1];
#pragma unroll
  for (int k = kRysTerms - 2; k >= 0; --k) r = fma(r, xi, c[k]);
  return r;
}

// Cell index and cell variable of T on the n-point table: v = T * nsub / hi,
// cell s = floor(v) (clamped to the last cell), xi = 2 (v - s) - 1.
__device__ __forceinline__ int rys_cell(const RysDev &R, int n, double T, double &xi) {
  const int nsub = R.nsub[n];
  const double v = T * (double(nsub) / (33.0 + 2.5 * double(2 * n - 1)));
  int s = int(v);
  s = s < nsub ? s : nsub - 1;
  xi = 2.0 * (v - double(s)) - 1.0;
  return s;
}

// Boys function F_0..F_mmax (mmax <= kBoysMaxM - 8) at T, the reference's
// method (math/boys.rs:146-193): below T = 93.05 the nearest grid point T0 of
// the table, a 9-term Taylor series for the top order
// F_m(T0 + d) = sum_k F_{m+k}(T0) (-d)^k / k!, e^-T = e^-T0 e^-d (Taylor), and
// the downward recurrence F_{m-1} = (2T F_m + e^-T) / (2m-1); above it the
// asymptotic F_0 = sqrt(pi/T)/2 and F_m = F_{m-1} (2m-1) / (2T) (e^-T < 1e-40
// is dropped).
__device__ __forceinline__ void boys_eval(const double *__restrict__ tab, int mmax, double T,
                                          double *F) {
  if (T < 93.05) {
    const int i = int(T * 10.0 + 0.5);
    const double x = 0.1 * double(i) - T;  // -d
    const double *row = tab + i * kBoysCols;
    constexpr double inv_fact[9] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720,
                                    1.0 / 5040, 1.0 / 40320};
    double f = row[mmax + 8] * inv_fact[8];
#pragma unroll
    for (int k = 7; k >= 0; --k) f = fma(f, x, row[mmax + k] * inv_fact[k]);
    double e = 1.0 / 362880.0;  // e^x, |x| <= 0.05, 10 terms
#pragma unroll
    for (int k = 8; k >= 0; --k) e = fma(e, x, inv_fact[k]);
    const double eT = row[kBoysMaxM + 1] * e;
    F[mmax] = f;
    for (int m = mmax; m > 0; --m) F[m - 1] = (2.0 * T * F[m] + eT) * (1.0 / double(2 * m - 1));
  } else {
    const double rs = rsqrt(T);
    F[0] = 0.88622692545275801365 * rs;  // sqrt(pi)/2
    const double h = 0.5 * rs * rs;
    for (int m = 1; m <= mmax; ++m) F[m] = F[m - 1] * double(2 * m - 1) * h;
  }
}

// Series `idx` of the n-point Rys rule at T: idx < n is squared root idx,
// otherwise weight idx - n.  Finite branch: polynomial cell lookup; T at or
// beyond 33 + 2.5(2n-1): scaled Gauss-Laguerre(-1/2) (math/rys.rs:174-185).
__device__ __forceinline__ double rys_series(const RysDev &R, int n, double T, int idx) {
  const double hi = 33.0 + 2.5 * double(2 * n - 1);
  const bool is_w = idx >= n;
  const int i = is_w ? idx - n : idx;
  if (T >= hi) {
    const double *L = R.buf + R.lag_off[n];
    const double rs = rsqrt(T);
    return is_w ? L[n + i] * (0.5 * rs) : L[i] * (rs * rs);
  }
  double xi;
  const int s = rys_cell(R, n, T, xi);
  const double *c = R.buf + (is_w ? R.wt_off[n] : R.root_off[n]) + (s * n + i) * kRysTerms;
  return rys_poly(c, xi);
}

// Rys 2D integrals of one axis for one root: I[i][j][k][l], i<=LA, j<=LB,
// k<=LC, l<=LD, from the vertical recurrence on (bra, ket) powers and the two
// horizontal transfers (engine/rys.rs:234-308 recurrences), written to `out`
// as ((i*(LB+1)+j)*(LC+1)+k)*(LD+1)+l.  Rolling levels keep the register
// footprint at two (NB+1)x(NK+1) slabs.
// F: arithmetic type of the tables (double, or float for the mixed-precision
// quartets).  g00 is I(0,0|0,0): 1, or a factor every entry should carry
// (the recurrences are linear and homogeneous in it), e.g. the quadrature
// weight and prefactor folded into one axis at no cost.
template <int LA, int LB, int LC, int LD, typename F = double>
__device__ __forceinline__ void rys_axis_table(F c00, F cp0, F b10, F b01, F b00, F ab, F cd,
                                               F *__restrict__ out, F g00 = F(1)) {
  constexpr int NB = LA + LB, NK = LC + LD;
  F g[NB + 1][NK + 1];
  g[0][0] = g00;
  if constexpr (NB >= 1) g[1][0] = c00 * g00;
#pragma unroll
  for (int n = 1; n < NB; ++n) g[n + 1][0] = c00 * g[n][0] + F(n) * b10 * g[n - 1][0];
#pragma unroll
  for (int m = 0; m < NK; ++m) {
#pragma unroll
    for (int n = 0; n <= NB; ++n) {
      F t = cp0 * g[n][m];
      if (m >= 1) t += F(m) * b01 * g[n][m - 1];
      if (n >= 1) t += F(n) * b00 * g[n - 1][m];
      g[n][m + 1] = t;
    }
  }
  // bra transfer level by level (j = 0..LB); each finished (i, j) row feeds
  // the ket transfer immediately.
#pragma unroll
  for (int j = 0; j <= LB; ++j) {
    if (j > 0) {
#pragma unroll
      for (int i = 0; i <= NB - j; ++i)
#pragma unroll
        for (int m = 0; m <= NK; ++m) g[i][m] = g[i + 1][m] + ab * g[i][m];
    }
#pragma unroll
    for (int i = 0; i <= LA; ++i) {
      F t[NK + 1];
#pragma unroll
      for (int k = 0; k <= NK; ++k) t[k] = g[i][k];
#pragma unroll
      for (int l = 0; l <= LD; ++l) {
        if (l > 0) {
#pragma unroll
          for (int k = 0; k <= NK - l; ++k) t[k] = t[k + 1] + cd * t[k];
        }
#pragma unroll
        for (int k = 0; k <= LC; ++k)
          out[((i * (LB + 1) + j) * (LC + 1) + k) * (LD + 1) + l] = t[k];
      }
    }
  }
}

// Incremental decoder for a lane that walks t upward in steps: one binary
// search at the first index, then rows advance by comparing against the cached
// end of the current row (implicit sets), or a direct list read (explicit).
struct RowCursor {
  int seg = 0, P = 0, k = 0;  // segment, its bra row and ket bucket
  long long seg_end = -1;     // rowptr[seg + 1]
  long long seg_begin = 0;
  int qb = 0;                 // first ket of the segment
  __device__ __forceinline__ void next(const QuartetSet &qs, long long t, int &Pout, int &Qout) {
    if (qs.list) {
      int2 v = qs.list[t];
      Pout = v.x;
      Qout = v.y;
      return;
    }
    if (seg_end < 0) {
      int lo = 0, hi = qs.x_n * qs.nseg - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (qs.rowptr[mid] <= t)
          lo = mid;
        else
          hi = mid - 1;
      }
      seg = lo;
      P = seg / qs.nseg;
      k = seg - P * qs.nseg;
      seg_begin = qs.rowptr[seg];
      seg_end = qs.rowptr[seg + 1];
      qb = qs.qbase[seg];
    }
    if (t >= seg_end) {
      while (t >= seg_end) {
        ++seg;
        if (++k == qs.nseg) {
          k = 0;
          ++P;
        }
        seg_begin = seg_end;
        seg_end = qs.rowptr[seg + 1];
      }
      qb = qs.qbase[seg];
    }
    Pout = P;
    Qout = qb + int(t - seg_begin);
  }
};

// Contiguous range of flat quartet indices owned by one wavefront.
__device__ __forceinline__ void wave_range(long long total, long long &begin, long long &end) {
  const long long waves = (long long)gridDim.x * (blockDim.x >> 6);
  const long long w = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  long long chunk = (total + waves - 1) / waves;
  begin = w * chunk;
  end = begin + chunk < total ? begin + chunk : total;
}

// Decode flat quartet index t of a QuartetSet into (P, Q) (relative indices).
__device__ __forceinline__ void decode_quartet(const QuartetSet &qs, long long t, int &P, int &Q) {
  if (qs.list) {
    int2 v = qs.list[t];
    P = v.x;
    Q = v.y;
    return;
  }
  int lo = 0, hi = qs.x_n * qs.nseg - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (qs.rowptr[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  P = lo / qs.nseg;
  Q = int(t - qs.rowptr[lo]) + qs.qbase[lo];
}

}  // namespace ig
