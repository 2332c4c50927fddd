// ig_big.hip -- runtime-angular-momentum quartet kernels for the classes the
// compiled (template) kernels do not cover: shells up to l = 6, and gradient
// tables raised by one on three centres for shells up to l = 5.
//
// One 256-thread workgroup owns one shell quartet at a time:
//   * lanes evaluate the n Rys roots/weights (n <= 13) into LDS;
//   * per batch of RB roots (RB chosen on the host so the batch's 2D tables
//     and their recurrence work areas fit the dynamic LDS budget), 3*RB lanes
//     build the x, y, z tables (runtime extents) in LDS;
//   * all lanes accumulate their strided slice of the Cartesian block into the
//     workgroup's global scratch slot (blocks reach 28^4 elements, beyond LDS);
//   * the epilogue (J/K digestion, dense scatter, Schwarz diagonal or the
//     adjoint gradient contraction) reads the block back from the slot.
// The grid is persistent (at most kBigMaxWG workgroups, one scratch slot each)
// and walks the quartet set with a grid stride; every workgroup leaves the
// loop once its stride passes the set's total.
#include "ig_device.h"

namespace ig {

namespace {

constexpr int kBigThreads = 256;
constexpr int kBigWaves = kBigThreads / 64;
constexpr int kBigMaxWG = 1024;
constexpr size_t kBigLdsBudget = 64 * 1024;         // dynamic LDS per workgroup
// jk_big_kernel's root batches: up to kBigLdsTwo bytes keep two workgroups on
// a CU (measured on a 20-atom s..g + l = 6 basis: as fast as or faster than
// fewer, larger batches at one workgroup per CU for every class).
constexpr size_t kBigLdsTwo = 78 * 1024;
constexpr size_t kBigScratchBytes = size_t(1) << 30;  // all slots of one launch

struct BigStatic {
  double rw[2 * kMaxRys];
  double red[9 * kBigWaves];
};

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// Work-area size (doubles) of one (root, axis) table task.
__host__ __device__ inline int table_work(int la, int lb, int lc, int ld) {
  const int NB = la + lb, NK = lc + ld;
  return (NB + 1) * (NK + 1) + (NK + 1);
}
__host__ __device__ inline int table_size(int la, int lb, int lc, int ld) {
  return (la + 1) * (lb + 1) * (lc + 1) * (ld + 1);
}

// Same recurrences as rys_axis_table (engine/rys.rs:234-308), runtime extents:
// vertical recurrence on (bra, ket) combined powers into g, then the bra and
// ket horizontal transfers; out[((i*(lb+1)+j)*(lc+1)+k)*(ld+1)+l].
// With swap_b the bra recurrence builds on the second shell's centre (c00
// from P - B, ab = B - A) and the transfer moves momentum to the first; with
// swap_k likewise on the ket (cp0 from Q - D, cd = D - C).  The output keeps
// the standard layout either way.  The caller builds on the centre of the
// tighter primitive: the transfer onto a tight shell across a long AB
// cancels catastrophically (an l = 4 shell of exponent 5.3 paired with a
// diffuse l = 6 one 5.6 bohr away: building on the l = 6 shell loses 5e-7
// of the block, building on the tight one 5e-14, against a 40-digit
// McMurchie-Davidson value).
__device__ void table_rt(int la, int lb, int lc, int ld, bool swap_b, bool swap_k, double c00,
                         double cp0, double b10, double b01, double b00, double ab, double cd,
                         double *g, double *t, double *out) {
  const int NB = la + lb, NK = lc + ld, W = NK + 1;
  const int ea = swap_b ? lb : la, eb = swap_b ? la : lb;  // built / received bra momenta
  const int ec = swap_k ? ld : lc, ed = swap_k ? lc : ld;
  g[0] = 1.0;
  if (NB >= 1) g[W] = c00;
  for (int n = 1; n < NB; ++n) g[(n + 1) * W] = c00 * g[n * W] + double(n) * b10 * g[(n - 1) * W];
  for (int m = 0; m < NK; ++m)
    for (int n = 0; n <= NB; ++n) {
      double v = cp0 * g[n * W + m];
      if (m >= 1) v += double(m) * b01 * g[n * W + m - 1];
      if (n >= 1) v += double(n) * b00 * g[(n - 1) * W + m];
      g[n * W + m + 1] = v;
    }
  for (int j = 0; j <= eb; ++j) {
    if (j > 0)  // bra transfer: I(i, j) = I(i+1, j-1) + AB I(i, j-1), in place over i
      for (int i = 0; i <= NB - j; ++i)
        for (int m = 0; m <= NK; ++m) g[i * W + m] = g[(i + 1) * W + m] + ab * g[i * W + m];
    for (int i = 0; i <= ea; ++i) {
      for (int k = 0; k <= NK; ++k) t[k] = g[i * W + k];
      const int a = swap_b ? j : i, bb = swap_b ? i : j;
      for (int l = 0; l <= ed; ++l) {
        if (l > 0)
          for (int k = 0; k <= NK - l; ++k) t[k] = t[k + 1] + cd * t[k];
        for (int k = 0; k <= ec; ++k) {
          const int c = swap_k ? l : k, d = swap_k ? k : l;
          out[((a * (lb + 1) + bb) * (lc + 1) + c) * (ld + 1) + d] = t[k];
        }
      }
    }
  }
}

struct BigClass {
  int la, lb, lc, ld, na, nb, nc, nd, N4, NR;
  __device__ BigClass(int a, int b, int c, int d, int raise)
      : la(a), lb(b), lc(c), ld(d), na(ncart(a)), nb(ncart(b)), nc(ncart(c)), nd(ncart(d)) {
    N4 = na * nb * nc * nd;
    NR = (a + b + c + d + raise) / 2 + 1;
  }
  // element e -> per-axis powers of the four shells
  __device__ void powers(int e, int pa[3], int pb[3], int pc[3], int pd[3]) const {
    const int id = e % nd;
    int r = e / nd;
    const int ic = r % nc;
    r /= nc;
    const int ib = r % nb, ia = r / nb;
    cart_powers(la, ia, pa[0], pa[1], pa[2]);
    cart_powers(lb, ib, pb[0], pb[1], pb[2]);
    cart_powers(lc, ic, pc[0], pc[1], pc[2]);
    cart_powers(ld, id, pd[0], pd[1], pd[2]);
  }
};

// Primitive-quartet scalars (same expressions as quartet_block in ig_kernels.hip).
struct PrimQ {
  double pz, qz, ipq, i2p, i2q, P[3], Q[3], PQ[3], pref, T, s_att, al, be, ga, de;
};

__device__ __forceinline__ void prim_quartet(const PairsDev &p, int ob, int ib, int ok, int ik,
                                             double omega, PrimQ &q) {
  q.pz = p.zeta[ob + ib];
  q.qz = p.zeta[ok + ik];
  q.P[0] = p.px[ob + ib];
  q.P[1] = p.py[ob + ib];
  q.P[2] = p.pz[ob + ib];
  q.Q[0] = p.px[ok + ik];
  q.Q[1] = p.py[ok + ik];
  q.Q[2] = p.pz[ok + ik];
  q.al = p.alpha[ob + ib];
  q.be = p.beta[ob + ib];
  q.ga = p.alpha[ok + ik];
  q.de = p.beta[ok + ik];
  const double rs = rsqrt(q.pz + q.qz);
  q.ipq = rs * rs;
  q.i2p = p.i2z[ob + ib];
  q.i2q = p.i2z[ok + ik];
  const double rho = q.pz * q.qz * q.ipq;
  double r2 = 0.0;
  for (int x = 0; x < 3; ++x) {
    q.PQ[x] = q.P[x] - q.Q[x];
    r2 += q.PQ[x] * q.PQ[x];
  }
  q.T = rho * r2;
  q.pref = kTwoPi52 * p.kz[ob + ib] * p.kz[ok + ik] * rs;
  q.s_att = 1.0;
  if (omega > 0.0) {  // erf(w r)/r: roots/weights at sT, x -> s x, w -> sqrt(s) w
    q.s_att = omega * omega / (rho + omega * omega);
    q.T *= q.s_att;
    q.pref *= sqrt(q.s_att);
  }
}

struct BigGeom {
  int sA, sB, sC, sD;
  double A[3], C[3], AB[3], CD[3];
  __device__ void load(const BasisDev &b, int a, int bb, int c, int d) {
    sA = a;
    sB = bb;
    sC = c;
    sD = d;
    for (int x = 0; x < 3; ++x) {
      A[x] = b.xyz[3 * a + x];
      C[x] = b.xyz[3 * c + x];
      AB[x] = A[x] - b.xyz[3 * bb + x];
      CD[x] = C[x] - b.xyz[3 * d + x];
    }
  }
};

// Roots/weights of one primitive quartet into S.rw (ends with a barrier).
__device__ __forceinline__ void roots_rt(const RysDev &R, int NR, const PrimQ &q, BigStatic &S) {
  for (int idx = threadIdx.x; idx < 2 * NR; idx += blockDim.x) {
    const double v = rys_series(R, NR, q.T, idx);
    S.rw[idx] = (idx < NR) ? v * q.s_att : v;
  }
  __syncthreads();
}

// Tables of roots [r0, r0+nr) for extents (la..ld): task t = 3*(r-r0)+axis
// writes tab + t*T (ends with a barrier).
__device__ __forceinline__ void tables_rt(int la, int lb, int lc, int ld, int r0, int nr,
                                          const PrimQ &q, const BigGeom &g, const BigStatic &S,
                                          double *tab, double *work) {
  const int T = table_size(la, lb, lc, ld), GW = table_work(la, lb, lc, ld);
  const int NW = (la + lb + 1) * (lc + ld + 1);
  for (int task = threadIdx.x; task < 3 * nr; task += blockDim.x) {
    const int r = r0 + task / 3, ax = task % 3;
    const double x = S.rw[r];
    const double fq = x * q.qz * q.ipq, fp = x * q.pz * q.ipq;
    const double b00 = 0.5 * x * q.ipq;
    const double b10 = (1.0 - fq) * q.i2p;
    const double b01 = (1.0 - fp) * q.i2q;
    // build each pair on the centre of its tighter primitive (uniform over
    // the workgroup: one primitive quartet at a time)
    const bool swap_b = q.be > q.al, swap_k = q.de > q.ga;
    const double c00 =
        (q.P[ax] - (swap_b ? g.A[ax] - g.AB[ax] : g.A[ax])) - fq * q.PQ[ax];
    const double cp0 =
        (q.Q[ax] - (swap_k ? g.C[ax] - g.CD[ax] : g.C[ax])) + fp * q.PQ[ax];
    double *w = work + task * GW;
    table_rt(la, lb, lc, ld, swap_b, swap_k, c00, cp0, b10, b01, b00,
             swap_b ? -g.AB[ax] : g.AB[ax], swap_k ? -g.CD[ax] : g.CD[ax], w, w + NW,
             tab + task * T);
  }
  __syncthreads();
}

// 512 lanes at about 250 VGPRs each fill a CU's register file: one workgroup
// per CU, which may then take most of the LDS for its root batches (fewer
// batches, so fewer table builds and barriers per work item).
constexpr int kRtThreads = 512;
constexpr int kRtN = 2 * kMaxL + 1;  // vertical-recurrence extent (n, m <= 12)
constexpr int kRtX = kMaxL + 1;      // transfer extent (received momentum <= 6)
constexpr size_t kRtLdsBudget = 150 * 1024;  // dynamic LDS of a root batch
constexpr int kRtMaxWG = 4096;
// Blocks of at least this many elements take jk_rt_kernel, smaller ones
// jk_big_kernel (measured on a 20-atom s..g + l = 6 basis: the sliced kernel
// is 1.3-2.4x faster for (ii|ig) and (ii|ii), 329280 and 614656 elements;
// the per-quartet one 1.4-1.8x faster from 70000 to 219520 elements, where
// building the tables once per quartet outweighs the global slot).
constexpr int kRtMinElems = 262144;

// Doubles per (root, axis) of the table steps: vertical-recurrence table,
// bra-transferred table, final table.
struct RtLayout {
  int G, H, T;
  __host__ __device__ RtLayout(int la, int lb, int lc, int ld)
      : G((la + lb + 1) * (lc + ld + 1)), H((la + 1) * (lb + 1) * (lc + ld + 1)),
        T((la + 1) * (lb + 1) * (lc + 1) * (ld + 1)) {}
  __host__ __device__ int per_root() const { return 3 * (G + H + T); }
};

// Vertical recurrence of one (root, axis) into gout[n * (NK + 1) + m],
// n <= NB, m <= NK (the recurrences of table_rt, column by column).
__device__ __forceinline__ void vrr_rt(int NB, int NK, double c00, double cp0, double b10,
                                       double b01, double b00, double *gout) {
  const int W = NK + 1;
  double prev[kRtN], cur[kRtN];
#pragma unroll
  for (int n = 0; n < kRtN; ++n) {
    prev[n] = 0.0;
    cur[n] = 0.0;
  }
  cur[0] = 1.0;
  if (NB >= 1) cur[1] = c00;
#pragma unroll
  for (int n = 1; n + 1 < kRtN; ++n)
    if (n < NB) cur[n + 1] = c00 * cur[n] + double(n) * b10 * cur[n - 1];
#pragma unroll
  for (int n = 0; n < kRtN; ++n)
    if (n <= NB) gout[n * W] = cur[n];
  for (int m = 0; m < NK; ++m) {
    double nxt[kRtN];
#pragma unroll
    for (int n = 0; n < kRtN; ++n) {
      double v = cp0 * cur[n] + double(m) * b01 * prev[n];
      if (n >= 1) v += double(n) * b00 * cur[n - 1];
      nxt[n] = v;
    }
#pragma unroll
    for (int n = 0; n < kRtN; ++n) {
      if (n <= NB) gout[n * W + m + 1] = nxt[n];
      prev[n] = cur[n];
      cur[n] = nxt[n];
    }
  }
}

// Tables of roots [r0, r0+nr) of primitive quartet q for class (la..ld) in
// two steps over the workgroup: the vertical recurrence of each (root, axis)
// into gt (vrr_rt, columns in VGPRs), then one lane per (root, axis, u, v)
// for built / received bra momenta (u, v): the bra transfer of the rows
// g(u..u+v, m) and the ket transfer of the result, both in VGPRs, in the
// operation order of table_rt.  Each pair is built on the centre of its
// tighter primitive and transferred to the other.  Table of (root r, axis)
// at out + (3*(r-r0)+axis)*T in the standard layout, without weight or
// prefactor; gt holds 3*nr*G doubles.  Ends with a barrier.
__device__ __forceinline__ void tables_fused(int la, int lb, int lc, int ld, const RtLayout &L,
                                             int r0, int nr, const PrimQ &q, const BigGeom &g,
                                             const BigStatic &S, double *out, double *gt) {
  const int NB = la + lb, NK = lc + ld, W = NK + 1;
  // uniform over the workgroup: one primitive quartet at a time
  const bool swap_b = q.be > q.al, swap_k = q.de > q.ga;
  const int ea = swap_b ? lb : la, eb = swap_b ? la : lb;  // built / received
  const int ec = swap_k ? ld : lc, ed = swap_k ? lc : ld;
  for (int task = threadIdx.x; task < 3 * nr; task += blockDim.x) {
    const int r = r0 + task / 3, ax = task % 3;
    const double x = S.rw[r];
    const double fq = x * q.qz * q.ipq, fp = x * q.pz * q.ipq;
    const double b00 = 0.5 * x * q.ipq;
    const double b10 = (1.0 - fq) * q.i2p;
    const double b01 = (1.0 - fp) * q.i2q;
    const double c00 = (q.P[ax] - (swap_b ? g.A[ax] - g.AB[ax] : g.A[ax])) - fq * q.PQ[ax];
    const double cp0 = (q.Q[ax] - (swap_k ? g.C[ax] - g.CD[ax] : g.C[ax])) + fp * q.PQ[ax];
    vrr_rt(NB, NK, c00, cp0, b10, b01, b00, gt + task * L.G);
  }
  __syncthreads();
  const int nuv = (ea + 1) * (eb + 1);
  for (int idx = threadIdx.x; idx < 3 * nr * nuv; idx += blockDim.x) {
    const int uv = idx % nuv, task = idx / nuv, u = uv / (eb + 1), v = uv % (eb + 1);
    const int ax = task % 3;
    const double sab = swap_b ? -g.AB[ax] : g.AB[ax];
    const double scd = swap_k ? -g.CD[ax] : g.CD[ax];
    const double *gg = gt + task * L.G + u * W;
    // bra transfer, column by column: h(u, v; m) from g(u..u+v, m)
    double t[kRtN];
#pragma unroll
    for (int m = 0; m < kRtN; ++m) {
      double vec[kRtX];
#pragma unroll
      for (int i = 0; i < kRtX; ++i) vec[i] = (m <= NK && i <= v) ? gg[i * W + m] : 0.0;
#pragma unroll
      for (int st = 1; st < kRtX; ++st)
        if (st <= v)
#pragma unroll
          for (int i = 0; i + 1 < kRtX; ++i) vec[i] = vec[i + 1] + sab * vec[i];
      t[m] = vec[0];
    }
    // ket transfer in place over t, the outputs after each step
    const int ia = swap_b ? v : u, ib = swap_b ? u : v;
    double *oo = out + task * L.T + (ia * (lb + 1) + ib) * (lc + 1) * (ld + 1);
#pragma unroll
    for (int l = 0; l < kRtX; ++l) {
      if (l > ed) break;
      if (l > 0)
#pragma unroll
        for (int k = 0; k + 1 < kRtN; ++k) t[k] = t[k + 1] + scd * t[k];
#pragma unroll
      for (int k = 0; k < kRtX; ++k)
        if (k <= ec) oo[(swap_k ? l : k) * (ld + 1) + (swap_k ? k : l)] = t[k];
    }
  }
  __syncthreads();
}

// Contracted Cartesian block of quartet (pb | pk) into blk (N4 doubles).
// Table offsets of every block element of class (la..ld), packed 21 bits
// per axis (x | y << 21 | z << 42): one table for the whole launch, read with
// coalesced loads instead of decoding the element's Cartesian powers.
__global__ void big_offsets_kernel(int la, int lb, int lc, int ld,
                                   unsigned long long *__restrict__ out) {
  const BigClass K(la, lb, lc, ld, 0);
  const int SA = (lb + 1) * (lc + 1) * (ld + 1), SB = (lc + 1) * (ld + 1), SC = ld + 1;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < K.N4; e += gridDim.x * blockDim.x) {
    int pa[3], pb[3], pc[3], pd[3];
    K.powers(e, pa, pb, pc, pd);
    unsigned long long v = 0;
    for (int ax = 0; ax < 3; ++ax)
      v |= (unsigned long long)(pa[ax] * SA + pb[ax] * SB + pc[ax] * SC + pd[ax]) << (21 * ax);
    out[e] = v;
  }
}

__device__ void big_block(const BigClass &K, int rb, const PairsDev &p, const RysDev &R, int pb,
                          int pk, const BigGeom &g, double omega, BigStatic &S, double *dyn,
                          const unsigned long long *__restrict__ eoff, double *__restrict__ blk) {
  const RtLayout L(K.la, K.lb, K.lc, K.ld);
  const int T = L.T;
  double *tab = dyn, *gt = dyn + 3 * rb * T;
  const int ob = p.pp_off[pb], nbp = p.pp_n[pb], ok = p.pp_off[pk], nkp = p.pp_n[pk];
  if (nbp * nkp == 0)  // no primitive quartet: an empty block
    for (int e = threadIdx.x; e < K.N4; e += blockDim.x) blk[e] = 0.0;
  constexpr unsigned long long m21 = (1ull << 21) - 1;
  for (int ib = 0; ib < nbp; ++ib)
    for (int ik = 0; ik < nkp; ++ik) {
      PrimQ q;
      prim_quartet(p, ob, ib, ok, ik, omega, q);
      roots_rt(R, K.NR, q, S);
      for (int r0 = 0; r0 < K.NR; r0 += rb) {
        const int nr = K.NR - r0 < rb ? K.NR - r0 : rb;
        tables_fused(K.la, K.lb, K.lc, K.ld, L, r0, nr, q, g, S, tab, gt);
        // the first batch of the first primitive quartet writes every element
        // of the slot (no read of it, no separate clearing pass)
        const bool first = ib == 0 && ik == 0 && r0 == 0;
        // eight elements per lane and pass: their offset loads, table reads
        // and slot updates are independent and issue together
        for (int e0 = threadIdx.x; e0 < K.N4; e0 += 8 * kBigThreads) {
          int ox[8], oy[8], oz[8];
          double v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * kBigThreads;
            const unsigned long long o = e < K.N4 ? eoff[e] : 0ull;  // 0: in-range dummy
            ox[u] = int(o & m21);
            oy[u] = int((o >> 21) & m21);
            oz[u] = int(o >> 42);
            v[u] = 0.0;
          }
          for (int k = 0; k < nr; ++k) {
            const double *tx = tab + 3 * k * T;
            const double w = S.rw[K.NR + r0 + k];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] += w * tx[ox[u]] * tx[T + oy[u]] * tx[2 * T + oz[u]];
          }
          if (first) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int e = e0 + u * kBigThreads;
              if (e < K.N4) blk[e] = q.pref * v[u];
            }
          } else {
            double old[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int e = e0 + u * kBigThreads;
              old[u] = e < K.N4 ? blk[e] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int e = e0 + u * kBigThreads;
              if (e < K.N4) blk[e] = old[u] + q.pref * v[u];
            }
          }
        }
        __syncthreads();  // tables are rebuilt for the next batch
      }
    }
}

__device__ __forceinline__ void next_quartet(const QuartetSet &qs, long long t, int &pb, int &pk) {
  int P, Q;
  decode_quartet(qs, t, P, Q);
  pb = qs.x_off + P;
  pk = qs.y_off + Q;
}

// ---------------------------------------------------------------------------
// Direct J/K for the runtime-l classes with blocks under kRtMinElems elements:
// one workgroup per quartet, the block in its global slot (big_block), then
// the digest read back from the slot; the screening and digestion of
// jk_reg_kernel (direct.rs:131-218 contract, Cartesian digestion of the
// canonical quartet with weight f).  Few elements per quartet leave the
// sliced kernel below with little work between its barriers.
__global__ void __launch_bounds__(kBigThreads)
    jk_big_kernel(int la, int lb, int lc, int ld, int rb, BasisDev b, PairsDev p, RysDev R,
                  QuartetSet qs, DensDev dd, double omega, const unsigned long long *eoff,
                  double *scratch, size_t slot, unsigned long long *counters) {
  __shared__ BigStatic S;
  extern __shared__ double dyn[];
  const BigClass K(la, lb, lc, ld, 0);
  double *blk = scratch + (size_t)blockIdx.x * slot;
  const int nsh = b.nsh, ncrt = b.nao_cart;
  const size_t n2 = (size_t)ncrt * ncrt;
  unsigned long long my_q = 0, my_p = 0;
  for (long long t = blockIdx.x; t < qs.total; t += gridDim.x) {
    int pb, pk;
    next_quartet(qs, t, pb, pk);
    BigGeom g;
    g.load(b, p.sh1[pb], p.sh2[pb], p.sh1[pk], p.sh2[pk]);
    const double pair_bound = p.q[pb] * p.q[pk];
    unsigned long long mask = 0;
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
    if (!mask) continue;  // uniform over the workgroup
    big_block(K, rb, p, R, pb, pk, g, omega, S, dyn, eoff, blk);
    __syncthreads();  // the epilogue reads elements written by other lanes
    if (threadIdx.x == 0) {
      ++my_q;
      my_p += (unsigned long long)p.pp_n[pb] * p.pp_n[pk];
    }
    const double f =
        (g.sA == g.sB ? 0.5 : 1.0) * (g.sC == g.sD ? 0.5 : 1.0) * (pb == pk ? 0.5 : 1.0);
    const int a0 = b.off_cart[g.sA], b0 = b.off_cart[g.sB], c0 = b.off_cart[g.sC],
              d0 = b.off_cart[g.sD];
    const int na = K.na, nb = K.nb, nc = K.nc, nd = K.nd, ncd = nc * nd, nab = na * nb;
    const int nds = nab + ncd + na * nc + na * nd + nb * nc + nb * nd;
    const int NJ = nab + ncd;
    const bool wk = dd.kc && dd.want_k;
    const int NO = NJ + (wk ? na * nc + nb * nc + na * nd + nb * nd : 0);
    // the quartet's six density sub-blocks in LDS (the tables' LDS is free:
    // every lane passed big_block's last barrier)
    double *Dab = dyn, *Dcd = Dab + nab, *Dac = Dcd + ncd, *Dad = Dac + na * nc,
           *Dbc = Dad + na * nd, *Dbd = Dbc + nb * nc;
    const size_t xo = jk_copy_offset(dd);
    for (int e = 0; e < dd.neff; ++e) {
      if (!((mask >> dd.eff_src[e]) & 1ull)) continue;  // uniform
      const bool anti = dd.eff_anti[e];                 // antisymmetric parts feed K only
      if (anti && !wk) continue;
      const double *D = dd.dc + (size_t)e * n2;
      double *J = dd.jc + xo + (size_t)e * n2;
      double *Kp = wk ? dd.kc + xo + (size_t)e * n2 : nullptr;
      __syncthreads();  // the previous density's sums are done with the sub-blocks
      for (int i = threadIdx.x; i < nds; i += blockDim.x) {
        int q = i, row, col;
        if (q < nab) {
          row = a0 + q / nb;
          col = b0 + q % nb;
        } else if ((q -= nab) < ncd) {
          row = c0 + q / nd;
          col = d0 + q % nd;
        } else if ((q -= ncd) < na * nc) {
          row = a0 + q / nc;
          col = c0 + q % nc;
        } else if ((q -= na * nc) < na * nd) {
          row = a0 + q / nd;
          col = d0 + q % nd;
        } else if ((q -= na * nd) < nb * nc) {
          row = b0 + q / nc;
          col = c0 + q % nc;
        } else {
          q -= nb * nc;
          row = b0 + q / nd;
          col = d0 + q % nd;
        }
        Dab[i] = D[(size_t)row * ncrt + col];
      }
      __syncthreads();
      // one lane per J'/K' entry: the entry sums n1 x m2 terms
      // blk[base + i*S1 + j*S2] * Ds[i*m2 + j] (Ds the matching sub-block),
      // sixteen (then eight) terms per step with independent loads and
      // eight partial sums
      for (int o = threadIdx.x; o < NO; o += blockDim.x) {
        if (anti && o < NJ) continue;
        int n1, m2, S1, S2, base, row, col;
        const double *Ds;
        double w;
        double *dst;
        if (o < nab) {  // J[a,b] += 2f sum_cd g D[c,d]
          n1 = nc, m2 = nd, S1 = nd, S2 = 1, base = o * ncd, Ds = Dcd;
          row = a0 + o / nb, col = b0 + o % nb, w = 2.0 * f, dst = J;
        } else if (o < NJ) {  // J[c,d] += 2f sum_ab g D[a,b]
          const int q = o - nab;
          n1 = na, m2 = nb, S1 = nb * ncd, S2 = ncd, base = q, Ds = Dab;
          row = c0 + q / nd, col = d0 + q % nd, w = 2.0 * f, dst = J;
        } else {
          int q = o - NJ;
          if (q < na * nc) {  // Kp[a,c] += f sum_bd g D[b,d]
            const int a = q / nc, c = q % nc;
            n1 = nb, m2 = nd, S1 = ncd, S2 = 1, base = a * nb * ncd + c * nd, Ds = Dbd;
            row = a0 + a, col = c0 + c;
          } else if ((q -= na * nc) < nb * nc) {  // Kp[b,c] += f sum_ad g D[a,d]
            const int bb = q / nc, c = q % nc;
            n1 = na, m2 = nd, S1 = nb * ncd, S2 = 1, base = bb * ncd + c * nd, Ds = Dad;
            row = b0 + bb, col = c0 + c;
          } else if ((q -= nb * nc) < na * nd) {  // Kp[a,d] += f sum_bc g D[b,c]
            const int a = q / nd, d = q % nd;
            n1 = nb, m2 = nc, S1 = ncd, S2 = nd, base = a * nb * ncd + d, Ds = Dbc;
            row = a0 + a, col = d0 + d;
          } else {  // Kp[b,d] += f sum_ac g D[a,c]
            q -= na * nd;
            const int bb = q / nd, d = q % nd;
            n1 = na, m2 = nc, S1 = nb * ncd, S2 = nd, base = bb * ncd + d, Ds = Dac;
            row = b0 + bb, col = d0 + d;
          }
          w = f, dst = Kp;
        }
        const int nt = n1 * m2;
        double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        int i = 0, j = 0, t = 0;  // term t = i*m2 + j
        for (; t + 16 <= nt; t += 16) {
          double gv[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            gv[u] = blk[base + i * S1 + j * S2];
            if (++j == m2) j = 0, ++i;
          }
#pragma unroll
          for (int u = 0; u < 16; ++u) acc[u & 7] += gv[u] * Ds[t + u];
        }
        for (; t + 8 <= nt; t += 8) {
          double gv[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            gv[u] = blk[base + i * S1 + j * S2];
            if (++j == m2) j = 0, ++i;
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) acc[u] += gv[u] * Ds[t + u];
        }
        for (; t < nt; ++t) {
          acc[0] += blk[base + i * S1 + j * S2] * Ds[t];
          if (++j == m2) j = 0, ++i;
        }
        const double v = w * (((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7])));
        if (v != 0.0) jk_add(dd, dst + (size_t)row * ncrt + col, v);
      }
    }
    __syncthreads();  // the slot is overwritten by the next quartet
  }
  if (counters && threadIdx.x == 0) {
    atomicAdd(counters, my_q);
    atomicAdd(counters + 1, my_p);
  }
}

// Sum of blk[el(t) - s0] * Ds[t] over terms t = t0, t0 + step, ... < t1, with
// el(t) = base + (t / m2) * S1 + (t % m2) * S2; terms whose element lies
// outside the slice [s0, s0 + ns) count zero.  Eight terms per step, the
// last step predicated: their LDS loads are independent and their products
// go to separate sums.
__device__ __forceinline__ double slice_terms(const double *blk, int s0, int ns,
                                              const double *Ds, int base, int m2, int S1,
                                              int S2, int t0, int t1, int step) {
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = t0; t < t1; t += 8 * step) {
    double gv[8], dv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int tt = t + u * step, i = tt / m2;
      const int el = base + i * S1 + (tt - i * m2) * S2 - s0;
      const bool ok = tt < t1 && el >= 0 && el < ns;
      gv[u] = blk[ok ? el : 0];
      dv[u] = Ds[ok ? tt : 0];
      gv[u] = ok ? gv[u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += gv[u] * dv[u];
  }
  return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

// ---------------------------------------------------------------------------
// Direct J/K for the runtime-l classes: the screening and digestion of
// jk_reg_kernel (direct.rs:131-218 contract, Cartesian digestion of the
// canonical quartet with weight f).
//
// Work item = (quartet, slice of kRtThreads * EPL block elements in (a, b,
// c, d) order).  Lane l owns elements s0 + l + kRtThreads * k (k < EPL) and
// accumulates them in registers over primitives and roots, so the block never
// goes to memory.  Per primitive quartet and batch of roots the workgroup
// builds the 2D tables in LDS in three parallel steps:
//   vertical recurrence g(n, m): one lane per (root, axis), columns in VGPRs;
//   bra transfer h(u, v; m): one lane per (root, axis, u, m), every v;
//   ket transfer I(u, v; uk, vk): one lane per (root, axis, u, v, uk), every vk;
// each pair built on the centre of its tighter primitive and transferred to
// the other (as table_rt).  The Rys weight and the prefactor ride on the z
// table.  The digest copies the slice to LDS and sums every J'/K' entry the
// slice touches against the quartet's six density sub-blocks, staged in LDS,
// with one device atomic per entry and slice.
template <int EPL>
__global__ void __launch_bounds__(kRtThreads)
    jk_rt_kernel(int la, int lb, int lc, int ld, int rb, int nsl, BasisDev b,
                 PairsDev p, RysDev R, QuartetSet qs, DensDev dd, double omega,
                 const unsigned long long *__restrict__ eoff, unsigned long long *counters) {
  __shared__ BigStatic S;
  extern __shared__ double dyn[];
  const BigClass K(la, lb, lc, ld, 0);
  const RtLayout L(la, lb, lc, ld);
  const int NB = la + lb, NK = lc + ld, W = NK + 1;
  const int nsh = b.nsh, ncrt = b.nao_cart;
  const size_t n2 = (size_t)ncrt * ncrt;
  // a contiguous range of work items per workgroup: the quartet index only
  // grows, so the row cursor decodes it without a search per item
  const long long work = qs.total * nsl;
  const long long per_wg = (work + gridDim.x - 1) / gridDim.x;
  const long long w_begin = (long long)blockIdx.x * per_wg;
  const long long w_end = w_begin + per_wg < work ? w_begin + per_wg : work;
  constexpr unsigned long long m21 = (1ull << 21) - 1;
  unsigned long long my_q = 0, my_p = 0;
  RowCursor cur;
  for (long long w = w_begin; w < w_end; ++w) {
    const long long t = w / nsl;
    const int sl = int(w - t * nsl);
    int pb, pk;
    {
      int P, Q;
      cur.next(qs, t, P, Q);
      pb = qs.x_off + P;
      pk = qs.y_off + Q;
    }
    BigGeom g;
    g.load(b, p.sh1[pb], p.sh2[pb], p.sh1[pk], p.sh2[pk]);
    const double pair_bound = p.q[pb] * p.q[pk];
    unsigned long long mask = 0;
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
    if (!mask) continue;  // uniform over the workgroup
    // this lane's elements of the slice and their table offsets
    const int s0 = sl * kRtThreads * EPL;
    int ox[EPL], oy[EPL], oz[EPL];
    double acc[EPL];
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const int e = s0 + int(threadIdx.x) + kRtThreads * k;
      ox[k] = oy[k] = oz[k] = 0;
      acc[k] = 0.0;
      if (e < K.N4) {
        const unsigned long long o = eoff[e];
        ox[k] = int(o & m21);
        oy[k] = int((o >> 21) & m21);
        oz[k] = int(o >> 42);
      }
    }
    const int ob = p.pp_off[pb], nbp = p.pp_n[pb], ok = p.pp_off[pk], nkp = p.pp_n[pk];
    for (int ib = 0; ib < nbp; ++ib)
      for (int ik = 0; ik < nkp; ++ik) {
        PrimQ q;
        prim_quartet(p, ob, ib, ok, ik, omega, q);
        roots_rt(R, K.NR, q, S);
        // build each pair on the centre of its tighter primitive (uniform
        // over the workgroup: one primitive quartet at a time)
        const bool swap_b = q.be > q.al, swap_k = q.de > q.ga;
        const int ea = swap_b ? lb : la, eb = swap_b ? la : lb;  // built / received
        const int ec = swap_k ? ld : lc, ed = swap_k ? lc : ld;
        for (int r0 = 0; r0 < K.NR; r0 += rb) {
          const int nr = K.NR - r0 < rb ? K.NR - r0 : rb;
          double *out = dyn, *gt = dyn + 3 * rb * L.T, *ht = gt + 3 * rb * L.G;
          for (int task = threadIdx.x; task < 3 * nr; task += kRtThreads) {
            const int r = r0 + task / 3, ax = task % 3;
            const double x = S.rw[r];
            const double fq = x * q.qz * q.ipq, fp = x * q.pz * q.ipq;
            const double b00 = 0.5 * x * q.ipq;
            const double b10 = (1.0 - fq) * q.i2p;
            const double b01 = (1.0 - fp) * q.i2q;
            const double c00 =
                (q.P[ax] - (swap_b ? g.A[ax] - g.AB[ax] : g.A[ax])) - fq * q.PQ[ax];
            const double cp0 =
                (q.Q[ax] - (swap_k ? g.C[ax] - g.CD[ax] : g.C[ax])) + fp * q.PQ[ax];
            vrr_rt(NB, NK, c00, cp0, b10, b01, b00, gt + task * L.G);
          }
          __syncthreads();
          // bra transfer: h(u, v; m) = h(u+1, v-1; m) + sab h(u, v-1; m)
          const int nbra = 3 * nr * (ea + 1) * W;
          for (int idx = threadIdx.x; idx < nbra; idx += kRtThreads) {
            const int m = idx % W, r1 = idx / W, u = r1 % (ea + 1), task = r1 / (ea + 1);
            const double sab = swap_b ? -g.AB[task % 3] : g.AB[task % 3];
            const double *gg = gt + task * L.G + u * W + m;
            double vec[kRtX];
#pragma unroll
            for (int i = 0; i < kRtX; ++i) vec[i] = i <= eb ? gg[i * W] : 0.0;
            double *hh = ht + task * L.H + u * (eb + 1) * W + m;
#pragma unroll
            for (int v = 0; v < kRtX; ++v) {
              if (v > eb) break;
              hh[v * W] = vec[0];
#pragma unroll
              for (int i = 0; i + 1 < kRtX; ++i) vec[i] = vec[i + 1] + sab * vec[i];
            }
          }
          __syncthreads();
          // ket transfer: I(.., uk, vk) = I(.., uk+1, vk-1) + scd I(.., uk, vk-1);
          // the z table takes the weight and prefactor
          const int nab = (la + 1) * (lb + 1);
          const int nket = 3 * nr * nab * (ec + 1);
          for (int idx = threadIdx.x; idx < nket; idx += kRtThreads) {
            const int uk = idx % (ec + 1), r1 = idx / (ec + 1), uv = r1 % nab, task = r1 / nab;
            const int u = uv / (eb + 1), v = uv % (eb + 1);
            const int i = swap_b ? v : u, j = swap_b ? u : v;
            const int ax = task % 3;
            const double scd = swap_k ? -g.CD[ax] : g.CD[ax];
            const double scale = ax == 2 ? q.pref * S.rw[K.NR + r0 + task / 3] : 1.0;
            const double *hh = ht + task * L.H + uv * W + uk;
            double vec[kRtX];
#pragma unroll
            for (int c = 0; c < kRtX; ++c) vec[c] = c <= ed ? hh[c] : 0.0;
            double *oo = out + task * L.T + (i * (lb + 1) + j) * (lc + 1) * (ld + 1);
#pragma unroll
            for (int vk = 0; vk < kRtX; ++vk) {
              if (vk > ed) break;
              const int kk = swap_k ? vk : uk, ll = swap_k ? uk : vk;
              oo[kk * (ld + 1) + ll] = scale * vec[0];
#pragma unroll
              for (int c = 0; c + 1 < kRtX; ++c) vec[c] = vec[c + 1] + scd * vec[c];
            }
          }
          __syncthreads();
          for (int rr = 0; rr < nr; ++rr) {
            const double *X = out + 3 * rr * L.T, *Y = X + L.T, *Z = Y + L.T;
#pragma unroll
            for (int k = 0; k < EPL; ++k) acc[k] += X[ox[k]] * Y[oy[k]] * Z[oz[k]];
          }
          __syncthreads();  // the next batch rebuilds the tables
        }
      }
    if (sl == 0 && threadIdx.x == 0) {
      ++my_q;
      my_p += (unsigned long long)nbp * nkp;
    }
    // digest (the tables' LDS is free: every lane passed the last barrier):
    // the slice to LDS once, then per density the six density sub-blocks of
    // the quartet and the J'/K' entries the slice touches.  Entries with many
    // terms (J[a,b], K[a,c], K[a,d]) take a group of lanes each (the lanes
    // stride over the terms, then an in-group sum); entries with few (J[c,d],
    // K[b,c], K[b,d]) a lane each.  Every sum runs over the rows (a, b) the
    // slice holds.
    const double f =
        (g.sA == g.sB ? 0.5 : 1.0) * (g.sC == g.sD ? 0.5 : 1.0) * (pb == pk ? 0.5 : 1.0);
    const int a0 = b.off_cart[g.sA], b0 = b.off_cart[g.sB], c0 = b.off_cart[g.sC],
              d0 = b.off_cart[g.sD];
    const int na = K.na, nb = K.nb, nc = K.nc, nd = K.nd, ncd = nc * nd;
    const int nds = na * nb + ncd + na * nc + na * nd + nb * nc + nb * nd;
    double *Dab = dyn, *Dcd = Dab + na * nb, *Dac = Dcd + ncd, *Dad = Dac + na * nc,
           *Dbc = Dad + na * nd, *Dbd = Dbc + nb * nc, *blk = Dab + nds;
    const int s1 = s0 + kRtThreads * EPL < K.N4 ? s0 + kRtThreads * EPL : K.N4;
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
      const int el = s0 + int(threadIdx.x) + kRtThreads * k;
      if (el < s1) blk[el - s0] = acc[k];
    }
    const int r0 = s0 / ncd, r1 = (s1 - 1) / ncd + 1;  // rows (a, b) of the slice
    const int ns = s1 - s0;
    const int A0 = r0 / nb, A1 = (r1 - 1) / nb + 1, nA = A1 - A0;
    const bool wk = dd.kc && dd.want_k;
    const size_t xo = jk_copy_offset(dd);
    for (int e = 0; e < dd.neff; ++e) {
      if (!((mask >> dd.eff_src[e]) & 1ull)) continue;  // uniform
      const bool anti = dd.eff_anti[e];                 // antisymmetric parts feed K only
      if (anti && !wk) continue;
      const double *D = dd.dc + (size_t)e * n2;
      double *J = dd.jc + xo + (size_t)e * n2;
      double *Kp = wk ? dd.kc + xo + (size_t)e * n2 : nullptr;
      __syncthreads();  // the slice is in LDS; the previous density's sums are done
      for (int i = threadIdx.x; i < nds; i += kRtThreads) {
        int q = i, row, col;
        if (q < na * nb) {
          row = a0 + q / nb;
          col = b0 + q % nb;
        } else if ((q -= na * nb) < ncd) {
          row = c0 + q / nd;
          col = d0 + q % nd;
        } else if ((q -= ncd) < na * nc) {
          row = a0 + q / nc;
          col = c0 + q % nc;
        } else if ((q -= na * nc) < na * nd) {
          row = a0 + q / nd;
          col = d0 + q % nd;
        } else if ((q -= na * nd) < nb * nc) {
          row = b0 + q / nc;
          col = c0 + q % nc;
        } else {
          q -= nb * nc;
          row = b0 + q / nd;
          col = d0 + q % nd;
        }
        Dab[i] = D[(size_t)row * ncrt + col];
      }
      __syncthreads();
      // group entries: J[a,b] rows, then K[a,c], then K[a,d]; each entry
      // takes a group of G lanes (a power of two, about eight terms per
      // lane), the group's lanes stride over its terms and sum in-group
      for (int kind = 0; kind < 3; ++kind) {
        const int nent = kind == 0 ? (anti ? 0 : r1 - r0) : (wk ? nA * (kind == 1 ? nc : nd) : 0);
        const int maxt = kind == 0 ? ncd : nb * (kind == 1 ? nd : nc);
        int G = 1;
        while (G < 64 && 8 * G < maxt) G <<= 1;
        const int gl = threadIdx.x & (G - 1);
        for (int o = threadIdx.x / G; o < nent; o += kRtThreads / G) {
          double v;
          double *dst;
          int row, col;
          if (kind == 0) {  // J[a,b] += 2f sum_cd g D[c,d]
            const int rr = r0 + o;
            const int lo = s0 > rr * ncd ? s0 - rr * ncd : 0;
            const int hi = s1 < (rr + 1) * ncd ? s1 - rr * ncd : ncd;
            v = 2.0 * f * slice_terms(blk, s0, ns, Dcd, rr * ncd, 1, 1, 0, lo + gl, hi, G);
            dst = J;
            row = a0 + rr / nb;
            col = b0 + rr % nb;
          } else {
            const bool ac = kind == 1;
            const int a = A0 + o / (ac ? nc : nd), x = o % (ac ? nc : nd);
            const int bl = r0 - a * nb > 0 ? r0 - a * nb : 0;
            const int bh = r1 - a * nb < nb ? r1 - a * nb : nb;
            const int ny = ac ? nd : nc;  // summed index of the other ket shell
            if (ac)  // Kp[a,c] += f sum_bd g D[b,d]
              v = slice_terms(blk, s0, ns, Dbd, a * nb * ncd + x * nd, ny, ncd, 1, bl * ny + gl,
                              bh * ny, G);
            else  // Kp[a,d] += f sum_bc g D[b,c]
              v = slice_terms(blk, s0, ns, Dbc, a * nb * ncd + x, ny, ncd, nd, bl * ny + gl,
                              bh * ny, G);
            v *= f;
            dst = Kp;
            row = a0 + a;
            col = (ac ? c0 : d0) + x;
          }
          // in-group sum: the xor partners of a lane stay inside its group,
          // whose lanes share the entry and so the loop trip
          for (int m = G >> 1; m > 0; m >>= 1) v += __shfl_xor(v, m);
          if (gl == 0 && v != 0.0) jk_add(dd, dst + (size_t)row * ncrt + col, v);
        }
      }
      // lane entries: J[c,d], then K[b,c], then K[b,d]
      const int nl0 = anti ? 0 : ncd, nl1 = wk ? nb * nc : 0, nl2 = wk ? nb * nd : 0;
      for (int o = threadIdx.x; o < nl0 + nl1 + nl2; o += kRtThreads) {
        double v;
        double *dst;
        int row, col;
        if (o < nl0) {  // J[c,d] += 2f sum_ab g D[a,b] over the slice's rows
          v = 2.0 * f * slice_terms(blk, s0, ns, Dab, o, 1, ncd, 0, r0, r1, 1);
          dst = J;
          row = c0 + o / nd;
          col = d0 + o % nd;
        } else {
          int q = o - nl0;
          const bool bc = q < nl1;
          if (!bc) q -= nl1;
          const int bb = q / (bc ? nc : nd), x = q % (bc ? nc : nd);
          const int ny = bc ? nd : nc;
          // a with row (a, bb) in the slice
          int al = (r0 - bb + nb - 1) / nb;
          if (r0 - bb <= 0) al = 0;
          const int ah = (r1 - 1 - bb) >= 0 ? (r1 - 1 - bb) / nb + 1 : 0;
          if (bc)  // Kp[b,c] += f sum_ad g D[a,d]
            v = slice_terms(blk, s0, ns, Dad, bb * ncd + x * nd, ny, nb * ncd, 1, al * ny,
                            ah * ny, 1);
          else  // Kp[b,d] += f sum_ac g D[a,c]
            v = slice_terms(blk, s0, ns, Dac, bb * ncd + x, ny, nb * ncd, nd, al * ny, ah * ny,
                            1);
          v *= f;
          dst = Kp;
          row = b0 + bb;
          col = (bc ? c0 : d0) + x;
        }
        if (v != 0.0) jk_add(dd, dst + (size_t)row * ncrt + col, v);
      }
    }
    __syncthreads();  // the next work item rebuilds the tables in this LDS
  }
  if (counters && threadIdx.x == 0) {
    atomicAdd(counters, my_q);
    atomicAdd(counters + 1, my_p);
  }
}

// ---------------------------------------------------------------------------
// Dense ERI: c2s per spherical shell in the slot (two halves, ping-pong), then
// scatter to every distinct permutation of the canonical quartet.
__device__ __constant__ int kBigPerms8[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2},
                                                {1, 0, 3, 2}, {2, 3, 0, 1}, {2, 3, 1, 0},
                                                {3, 2, 0, 1}, {3, 2, 1, 0}};

__global__ void __launch_bounds__(kBigThreads)
    eri_big_kernel(int la, int lb, int lc, int ld, int rb, BasisDev b, PairsDev p, RysDev R,
                   QuartetSet qs, EriOut eo, const unsigned long long *eoff, double *scratch,
                   size_t slot) {
  double *out = eo.out;
  __shared__ BigStatic S;
  extern __shared__ double dyn[];
  const BigClass K(la, lb, lc, ld, 0);
  double *half0 = scratch + (size_t)blockIdx.x * slot, *half1 = half0 + K.N4;
  const size_t nao = (size_t)b.nao;
  for (long long t = blockIdx.x; t < qs.total; t += gridDim.x) {
    int pb, pk;
    next_quartet(qs, t, pb, pk);
    BigGeom g;
    g.load(b, p.sh1[pb], p.sh2[pb], p.sh1[pk], p.sh2[pk]);
    big_block(K, rb, p, R, pb, pk, g, eo.omega, S, dyn, eoff, half0);
    __syncthreads();
    double *c0 = half0, *c1 = half1;
    const int sh[4] = {g.sA, g.sB, g.sC, g.sD};
    const int ls[4] = {la, lb, lc, ld};
    int dims[4] = {K.na, K.nb, K.nc, K.nd};
    for (int ax = 0; ax < 4; ++ax) {
      if (!b.kind[sh[ax]]) continue;
      const int nin = dims[ax], nout = 2 * ls[ax] + 1;
      const double *M = b.c2s + b.c2s_off[ls[ax]];
      int outer = 1, inner = 1;
      for (int x = 0; x < ax; ++x) outer *= dims[x];
      for (int x = ax + 1; x < 4; ++x) inner *= dims[x];
      for (int o = threadIdx.x; o < outer * nout * inner; o += blockDim.x) {
        const int i = o % inner, rq = o / inner, m = rq % nout, oo = rq / nout;
        double v = 0.0;
        for (int j = 0; j < nin; ++j) v += M[m * nin + j] * c0[(oo * nin + j) * inner + i];
        c1[o] = v;
      }
      dims[ax] = nout;
      double *tmp = c0;
      c0 = c1;
      c1 = tmp;
      __syncthreads();
    }
    const int nf = dims[0] * dims[1] * dims[2] * dims[3];
    if (eo.mode == kDf3c || eo.mode == kDf2c) {
      for (int e = threadIdx.x; e < nf; e += blockDim.x) {
        const int id = e % dims[3], r1 = e / dims[3], ic = r1 % dims[2], r2 = r1 / dims[2],
                  ibb = r2 % dims[1], ia = r2 / dims[1];
        const int idx[4] = {ia, ibb, ic, id};
        df_store(b, eo, sh, idx, c0[e]);
      }
      __syncthreads();
      continue;
    }
    if (eo.start) {  // packed: the canonical element only
      for (int e = threadIdx.x; e < nf; e += blockDim.x) {
        const int id = e % dims[3], r1 = e / dims[3], ic = r1 % dims[2], r2 = r1 / dims[2],
                  ibb = r2 % dims[1], ia = r2 / dims[1];
        packed_store(b, eo.start, eo.row0, out, sh[0], ia, sh[1], ibb, sh[2], ic, sh[3], id,
                     c0[e]);
      }
      __syncthreads();
      continue;
    }
    int seen[8][4];
    int nseen = 0;
    for (int pi = 0; pi < 8; ++pi) {
      int tup[4];
      for (int m = 0; m < 4; ++m) tup[m] = sh[kBigPerms8[pi][m]];
      bool dup = false;
      for (int s = 0; s < nseen; ++s)
        dup |= (seen[s][0] == tup[0] && seen[s][1] == tup[1] && seen[s][2] == tup[2] &&
                seen[s][3] == tup[3]);
      if (dup) continue;
      for (int m = 0; m < 4; ++m) seen[nseen][m] = tup[m];
      ++nseen;
      if (eo.rows_end >= 0) {  // rows owned by the bra-pair range only
        const long long hi = tup[0] > tup[1] ? tup[0] : tup[1], lo = tup[0] + tup[1] - hi;
        const long long cp = hi * (hi + 1) / 2 + lo;
        if (cp < eo.rows_begin || cp >= eo.rows_end) continue;
      }
      // position m of the permuted index takes source axis kBigPerms8[pi][m]
      size_t stride[4], base = 0;
      for (int m = 0; m < 4; ++m) {
        const int src = kBigPerms8[pi][m];
        size_t st = 1;
        for (int x = 0; x < 3 - m; ++x) st *= nao;
        stride[src] = st;
        base += (size_t)b.off_func[sh[src]] * st;
      }
      for (int e = threadIdx.x; e < nf; e += blockDim.x) {
        const int id = e % dims[3], r1 = e / dims[3], ic = r1 % dims[2], r2 = r1 / dims[2],
                  ibb = r2 % dims[1], ia = r2 / dims[1];
        out[base + ia * stride[0] + ibb * stride[1] + ic * stride[2] + id * stride[3]] = c0[e];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Schwarz diagonal: Q[ab] = sqrt(max |(ab|ab)|) over the pair's functions
// (screening.rs:17-52), spherical shells transformed on the fly.
__global__ void __launch_bounds__(kBigThreads)
    schwarz_big_kernel(int la, int lb, int rb, BasisDev b, PairsDev p, RysDev R, int first,
                       int count, double *qout, const unsigned long long *eoff, double *scratch,
                       size_t slot) {
  __shared__ BigStatic S;
  extern __shared__ double dyn[];
  const BigClass K(la, lb, la, lb, 0);
  double *blk = scratch + (size_t)blockIdx.x * slot;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = blockIdx.x; t < count; t += gridDim.x) {
    const int pr = first + t;
    BigGeom g;
    g.load(b, p.sh1[pr], p.sh2[pr], p.sh1[pr], p.sh2[pr]);
    big_block(K, rb, p, R, pr, pr, g, 0.0, S, dyn, eoff, blk);
    __syncthreads();
    const bool sa = b.kind[g.sA], sb = b.kind[g.sB];
    const int na = K.na, nb = K.nb;
    const int nfa = sa ? 2 * la + 1 : na, nfb = sb ? 2 * lb + 1 : nb;
    const double *Ma = b.c2s + b.c2s_off[la], *Mb = b.c2s + b.c2s_off[lb];
    double mx = 0.0;
    for (int o = threadIdx.x; o < nfa * nfb; o += blockDim.x) {
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
    mx = wave_max_d(mx);
    if (lane == 0) S.red[wave] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < kBigWaves; ++w) mx = fmax(mx, S.red[w]);
      qout[pr] = sqrt(mx);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Adjoint gradient: the same contraction as grad_kernel in ig_kernels.hip
// (grad.rs direct gradient): per element the density weight W (Coulomb and
// exchange terms), then per root the derivative tables from the raised 2D
// integrals, D's derivative by translational invariance.
__global__ void __launch_bounds__(kBigThreads)
    grad_big_kernel(int la, int lb, int lc, int ld, int rb, BasisDev b, PairsDev p, RysDev R,
                    QuartetSet qs, GradDev gd, double *scratch, size_t slot,
                    unsigned long long *counters) {
  __shared__ BigStatic S;
  extern __shared__ double dyn[];
  const BigClass K(la, lb, lc, ld, 1);
  double *W = scratch + (size_t)blockIdx.x * slot;
  const int nsh = b.nsh, ncrt = b.nao_cart;
  const size_t nc2 = (size_t)ncrt * ncrt;
  const int nterm = gd.ncoul + gd.nexch;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // raised extents: A, B, C up by one, D unchanged
  const int RA = la + 1, RB = lb + 1, RC = lc + 1, RD = ld;
  const int TR = table_size(RA, RB, RC, RD);
  const int SA = (RB + 1) * (RC + 1) * (RD + 1), SB = (RC + 1) * (RD + 1), SC = RD + 1;
  double *tab = dyn, *work = dyn + 3 * rb * TR;
  unsigned long long my_q = 0, my_p = 0;
  for (long long t = blockIdx.x; t < qs.total; t += gridDim.x) {
    int pb, pk;
    next_quartet(qs, t, pb, pk);
    BigGeom g;
    g.load(b, p.sh1[pb], p.sh2[pb], p.sh1[pk], p.sh2[pk]);
    const int atA = b.atom[g.sA], atB = b.atom[g.sB], atC = b.atom[g.sC], atD = b.atom[g.sD];
    if (atA == atB && atB == atC && atC == atD) continue;  // one-centre: zero by invariance
    const double deg =
        (g.sA == g.sB ? 1.0 : 2.0) * (g.sC == g.sD ? 1.0 : 2.0) * (pb == pk ? 1.0 : 2.0);
    double bound = 0.0;
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
    const int a0 = b.off_cart[g.sA], b0 = b.off_cart[g.sB], c0 = b.off_cart[g.sC],
              d0 = b.off_cart[g.sD];
    for (int e = threadIdx.x; e < K.N4; e += blockDim.x) {
      const int id = e % K.nd, r1 = e / K.nd, ic = r1 % K.nc, r2 = r1 / K.nc, ib = r2 % K.nb,
                ia = r2 / K.nb;
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
      W[e] = w;  // each lane reads back only its own elements
    }
    double gs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int ob = p.pp_off[pb], nbp = p.pp_n[pb], ok = p.pp_off[pk], nkp = p.pp_n[pk];
    for (int ib = 0; ib < nbp; ++ib)
      for (int ik = 0; ik < nkp; ++ik) {
        PrimQ q;
        prim_quartet(p, ob, ib, ok, ik, gd.omega, q);
        roots_rt(R, K.NR, q, S);
        const double ta = 2.0 * q.al, tb = 2.0 * q.be, tc = 2.0 * q.ga;
        for (int r0 = 0; r0 < K.NR; r0 += rb) {
          const int nr = K.NR - r0 < rb ? K.NR - r0 : rb;
          tables_rt(RA, RB, RC, RD, r0, nr, q, g, S, tab, work);
          for (int e = threadIdx.x; e < K.N4; e += blockDim.x) {
            int pa[3], pbp[3], pc[3], pd[3];
            K.powers(e, pa, pbp, pc, pd);
            const double we = q.pref * W[e];
            for (int k = 0; k < nr; ++k) {
              const double s = we * S.rw[K.NR + r0 + k];
              double v0[3], dA[3], dB[3], dC[3];
              for (int ax = 0; ax < 3; ++ax) {
                const double *tx = tab + (3 * k + ax) * TR;
                const int oo = pa[ax] * SA + pbp[ax] * SB + pc[ax] * SC + pd[ax];
                v0[ax] = tx[oo];
                // d/dA x^a = 2 alpha x^{a+1} - a x^{a-1}; the lowered read stays
                // in bounds (offset 0 when a = 0, weighted by a = 0)
                dA[ax] = ta * tx[oo + SA] - double(pa[ax]) * tx[oo - (pa[ax] ? SA : 0)];
                dB[ax] = tb * tx[oo + SB] - double(pbp[ax]) * tx[oo - (pbp[ax] ? SB : 0)];
                dC[ax] = tc * tx[oo + SC] - double(pc[ax]) * tx[oo - (pc[ax] ? SC : 0)];
              }
              const double yz = v0[1] * v0[2], xz = v0[0] * v0[2], xy = v0[0] * v0[1];
              gs[0] += s * dA[0] * yz;
              gs[1] += s * dA[1] * xz;
              gs[2] += s * dA[2] * xy;
              gs[3] += s * dB[0] * yz;
              gs[4] += s * dB[1] * xz;
              gs[5] += s * dB[2] * xy;
              gs[6] += s * dC[0] * yz;
              gs[7] += s * dC[1] * xz;
              gs[8] += s * dC[2] * xy;
            }
          }
          __syncthreads();
        }
      }
    if (threadIdx.x == 0) {
      ++my_q;
      my_p += (unsigned long long)nbp * nkp;
    }
    for (int x = 0; x < 9; ++x) {
      const double v = wave_sum_d(gs[x]);
      if (lane == 0) S.red[x * kBigWaves + wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
      double s = 0.0;
      for (int w = 0; w < kBigWaves; ++w) s += S.red[threadIdx.x * kBigWaves + w];
      const int which = threadIdx.x / 3, ax = threadIdx.x % 3;
      const int at = which == 0 ? atA : (which == 1 ? atB : atC);
      atomicAdd(gd.out + 3 * at + ax, s);
      atomicAdd(gd.out + 3 * atD + ax, -s);
    }
    __syncthreads();
  }
  if (counters && threadIdx.x == 0) {
    atomicAdd(counters, my_q);
    atomicAdd(counters + 1, my_p);
  }
}

// Workgroups of one launch: one per quartet up to kBigMaxWG, fewer when the
// slots (slot doubles each) would exceed kBigScratchBytes.
int big_grid(long long total, size_t slot) {
  long long cap = (long long)(kBigScratchBytes / (slot * sizeof(double)));
  cap = cap < 1 ? 1 : (cap > kBigMaxWG ? kBigMaxWG : cap);
  const long long g = total < cap ? total : cap;
  return int(g < 1 ? 1 : g);
}

// Roots per table batch and the dynamic LDS bytes for extents (la..ld) and
// nr roots: as many roots as fit the budget, at least one.
void root_batch(int la, int lb, int lc, int ld, int nr, int &rb, size_t &bytes) {
  const size_t per = size_t(3) * (table_size(la, lb, lc, ld) + table_work(la, lb, lc, ld)) *
                     sizeof(double);
  rb = int(kBigLdsBudget / per);
  rb = rb < 1 ? 1 : (rb > nr ? nr : rb);
  bytes = per * rb;
}

// LDS bytes per root of big_block's tables for class (la..ld): the x, y, z
// tables and their vertical-recurrence tables.
size_t block_root_bytes(int la, int lb, int lc, int ld) {
  const RtLayout L(la, lb, lc, ld);
  return size_t(3) * (L.T + L.G) * sizeof(double);
}

// Roots per table batch and the dynamic LDS bytes of big_block for class
// (la..ld) in a budget: as many roots as it holds, at least one.
void block_batch(int la, int lb, int lc, int ld, int &rb, size_t &bytes) {
  const int nr = (la + lb + lc + ld) / 2 + 1;
  const size_t per = block_root_bytes(la, lb, lc, ld);
  rb = int(kBigLdsBudget / per);
  rb = rb < 1 ? 1 : (rb > nr ? nr : rb);
  bytes = per * rb;
}

// Roots per table batch and dynamic LDS bytes of jk_big_kernel for class
// (la..ld): as many roots as kBigLdsTwo holds, at least one.
void jk_big_batch(int la, int lb, int lc, int ld, int &rb, size_t &bytes) {
  const int nr = (la + lb + lc + ld) / 2 + 1;
  const size_t per = block_root_bytes(la, lb, lc, ld);
  rb = int(kBigLdsTwo / per);
  rb = rb < 1 ? 1 : (rb > nr ? nr : rb);
  bytes = per * rb;
}

}  // namespace

// Scratch layout of one launch: the element offset table (N4 64-bit words,
// rounded up to 32 for alignment), then one block slot per workgroup.
size_t big_scratch_elems(int la, int lb, int lc, int ld, long long total, int two_halves) {
  const size_t n4 = size_t(ncart(la)) * ncart(lb) * ncart(lc) * ncart(ld);
  const size_t slot = n4 * (two_halves ? 2 : 1);
  return (n4 + 31) / 32 * 32 + size_t(big_grid(total, slot)) * slot;
}

namespace {
// offset table at the front of scratch; returns the first slot
double *big_prepare(int la, int lb, int lc, int ld, double *scratch, hipStream_t s) {
  const int n4 = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
  hipLaunchKernelGGL(big_offsets_kernel, dim3((n4 + 255) / 256), dim3(256), 0, s, la, lb, lc, ld,
                     reinterpret_cast<unsigned long long *>(scratch));
  return scratch + (n4 + 31) / 32 * 32;
}

// Dynamic LDS and launch shape of jk_rt_kernel for class (la..ld): roots per
// table batch (as many as the budget holds, at least one), the digest's
// staged sub-blocks and slice, elements per lane and slices per quartet.
struct RtLaunch {
  int rb, epl, nsl;
  size_t lds;
  RtLaunch(int la, int lb, int lc, int ld) {
    const RtLayout L(la, lb, lc, ld);
    const int nr = (la + lb + lc + ld) / 2 + 1;
    const size_t per = size_t(L.per_root()) * sizeof(double);
    const size_t budget = per > kRtLdsBudget ? per : kRtLdsBudget;
    rb = int(budget / per);
    rb = rb < 1 ? 1 : (rb > nr ? nr : rb);
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    const int nds = na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd;
    const int n4 = na * nb * nc * nd;
    epl = 16;
    const int slice = n4 < kRtThreads * epl ? n4 : kRtThreads * epl;
    const size_t dig = size_t(nds + slice) * sizeof(double);
    lds = per * rb > dig ? per * rb : dig;
    nsl = (n4 + kRtThreads * epl - 1) / (kRtThreads * epl);
  }
};
}  // namespace

// J/K scratch of the runtime-l classes: the element offset table, and for
// the blocks under kRtMinElems one block slot per workgroup.
size_t jk_big_scratch_elems(int la, int lb, int lc, int ld, long long total) {
  const size_t n4 = size_t(ncart(la)) * ncart(lb) * ncart(lc) * ncart(ld);
  if (n4 < size_t(kRtMinElems)) return big_scratch_elems(la, lb, lc, ld, total, 0);
  return (n4 + 31) / 32 * 32;
}

hipError_t launch_jk_big(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                         const RysDev &r, const QuartetSet &qs, const DensDev &d, double omega,
                         double *scratch, unsigned long long *ctr, hipStream_t s) {
  if (qs.total <= 0) return hipSuccess;
  const size_t n4 = size_t(ncart(la)) * ncart(lb) * ncart(lc) * ncart(ld);
  if (n4 < size_t(kRtMinElems)) {  // one workgroup per quartet, block slots in scratch
    int rb;
    size_t lds;
    jk_big_batch(la, lb, lc, ld, rb, lds);
    // the epilogue stages the quartet's density sub-blocks in the same LDS
    const size_t nds = size_t(ncart(la)) * ncart(lb) + size_t(ncart(lc)) * ncart(ld) +
                       size_t(ncart(la) + ncart(lb)) * (ncart(lc) + ncart(ld));
    if (nds * sizeof(double) > lds) lds = nds * sizeof(double);
    double *slots = big_prepare(la, lb, lc, ld, scratch, s);
    hipLaunchKernelGGL(jk_big_kernel, dim3(big_grid(qs.total, n4)), dim3(kBigThreads), lds, s,
                       la, lb, lc, ld, rb, b, p, r, qs, d, omega,
                       reinterpret_cast<const unsigned long long *>(scratch), slots, n4, ctr);
    return hipGetLastError();
  }
  const RtLaunch RL(la, lb, lc, ld);
  const long long work = qs.total * RL.nsl;
  const int grid = int(work < kRtMaxWG ? work : kRtMaxWG);
  big_prepare(la, lb, lc, ld, scratch, s);  // the offset table at the front of scratch
  const auto *eoff = reinterpret_cast<const unsigned long long *>(scratch);
  hipLaunchKernelGGL(jk_rt_kernel<16>, dim3(grid), dim3(kRtThreads), RL.lds, s, la, lb, lc, ld,
                     RL.rb, RL.nsl, b, p, r, qs, d, omega, eoff, ctr);
  return hipGetLastError();
}

hipError_t launch_eri_big(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                          const RysDev &r, const QuartetSet &qs, const EriOut &o, double *scratch,
                          hipStream_t s) {
  if (qs.total <= 0) return hipSuccess;
  const size_t slot = 2 * size_t(ncart(la)) * ncart(lb) * ncart(lc) * ncart(ld);
  int rb;
  size_t lds;
  block_batch(la, lb, lc, ld, rb, lds);
  double *slots = big_prepare(la, lb, lc, ld, scratch, s);
  hipLaunchKernelGGL(eri_big_kernel, dim3(big_grid(qs.total, slot)), dim3(kBigThreads), lds, s, la,
                     lb, lc, ld, rb, b, p, r, qs, o,
                     reinterpret_cast<const unsigned long long *>(scratch), slots, slot);
  return hipGetLastError();
}

hipError_t launch_schwarz_big(int la, int lb, const BasisDev &b, const PairsDev &p,
                              const RysDev &r, int first, int count, double *q, double *scratch,
                              hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const size_t slot = size_t(ncart(la)) * ncart(lb) * ncart(la) * ncart(lb);
  int rb;
  size_t lds;
  block_batch(la, lb, la, lb, rb, lds);
  double *slots = big_prepare(la, lb, la, lb, scratch, s);
  hipLaunchKernelGGL(schwarz_big_kernel, dim3(big_grid(count, slot)), dim3(kBigThreads), lds, s,
                     la, lb, rb, b, p, r, first, count, q,
                     reinterpret_cast<const unsigned long long *>(scratch), slots, slot);
  return hipGetLastError();
}

hipError_t launch_grad_big(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                           const RysDev &r, const QuartetSet &qs, const GradDev &g,
                           double *scratch, unsigned long long *ctr, hipStream_t s) {
  if (qs.total <= 0) return hipSuccess;
  const size_t slot = size_t(ncart(la)) * ncart(lb) * ncart(lc) * ncart(ld);
  int rb;
  size_t lds;
  root_batch(la + 1, lb + 1, lc + 1, ld, (la + lb + lc + ld + 1) / 2 + 1, rb, lds);
  hipLaunchKernelGGL(grad_big_kernel, dim3(big_grid(qs.total, slot)), dim3(kBigThreads), lds, s, la,
                     lb, lc, ld, rb, b, p, r, qs, g, scratch, slot, ctr);
  return hipGetLastError();
}

}  // namespace ig
