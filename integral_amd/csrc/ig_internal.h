// ig_internal.h -- shared declarations of the HIP implementation (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace ig {

constexpr int kMaxL = 6;        // engine/os.rs:34
constexpr int kMaxGradL = 5;    // grad.rs:50
constexpr int kMaxRys = 13;     // math/rys.rs:104
constexpr int kRysTerms = 13;  // Rys table polynomials: degree 12 per cell
constexpr int kMaxDens = 64;    // direct.rs:145-148
constexpr int kMaxKernelL = 2;  // shells up to d have compiled quartet kernels in this build

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int pair_class(int l1, int l2) { return l1 * (l1 + 1) / 2 + l2; }
constexpr int kNumPairClass = pair_class(kMaxL, kMaxL) + 1;  // 28

// Cartesian component i of shell l: lx descending, then ly descending (am.rs:23-31).
__host__ __device__ inline void cart_powers(int l, int i, int &lx, int &ly, int &lz) {
  int d = 0;
  while ((d + 1) * (d + 2) / 2 <= i) ++d;  // d = l - lx
  lx = l - d;
  lz = i - d * (d + 1) / 2;
  ly = d - lz;
}

// ---------------------------------------------------------------------------
// host tables
struct RysHostTables {
  std::vector<double> data;  // all coefficients + Laguerre nodes/weights
  int nsub[kMaxRys + 1] = {};
  int root_off[kMaxRys + 1] = {}, wt_off[kMaxRys + 1] = {}, lag_off[kMaxRys + 1] = {};
  double worst[kMaxRys + 1] = {};
};
const RysHostTables &rys_host_tables();
// Boys table: kBoysRows grid points T0 = 0.1 i, columns F_0..F_kBoysMaxM, exp(-T0)
constexpr int kBoysMaxM = 33, kBoysRows = 931, kBoysCols = kBoysMaxM + 2;
const std::vector<double> &boys_host_table();
void rys_host_eval(int n, double T, double *x, double *w);
double shell_norm(double alpha, int l);
std::vector<double> c2s_scaled(int l);  // (2l+1) x ncart(l), includes sqrt(4pi/(2l+1))

// ---------------------------------------------------------------------------
// device-visible views (passed by value as kernel arguments)
struct RysDev {
  const double *boys;  // Boys table on the device (kBoysRows x kBoysCols)
  const double *buf;
  int nsub[kMaxRys + 1];
  int root_off[kMaxRys + 1], wt_off[kMaxRys + 1], lag_off[kMaxRys + 1];
};

struct BasisDev {
  int nsh, nao, nao_cart;
  const double *xyz;     // nsh x 3
  const int *l;          // nsh
  const int *kind;       // nsh (0 Cartesian, 1 spherical)
  const int *off_func;   // nsh
  const int *off_cart;   // nsh
  const int *atom;       // nsh
  const double *c2s;     // packed scaled c2s matrices for l=0..6
  int c2s_off[kMaxL + 1];
  // shells by angular momentum: lshells[lshell_off[l] + r] is the r-th shell
  // of momentum l, rank[s] the position of shell s in its group
  const int *rank, *lshells;
  int lshell_off[kMaxL + 2];
};

// Shell pairs in kernel orientation (first shell has the larger l), grouped
// by pair class and, inside a class, sorted by Schwarz bound (descending).
struct PairsDev {
  int n;
  const int *sh1, *sh2;  // shells, kernel orientation
  const int *canon;      // canonical pair index i(i+1)/2+j (i >= j)
  const double *q;       // Schwarz bound Q[i,j]
  const int *pp_off, *pp_n;  // primitive-pair range
  // primitive pairs, structure of arrays; kz = kc / zeta and i2z = 1 / (2 zeta)
  // spare the kernels their divisions
  const double *zeta, *px, *py, *pz, *kc, *alpha, *beta, *kz, *i2z;
};

// One launch covers the quartets {P, Q} with P in bra class X and Q in ket
// class Y (class(X) >= class(Y)); either implicit or an explicit (P, Q) list.
// Implicit mode: inside a class the pairs are grouped into buckets of equal
// primitive-pair count (descending), each sorted by Schwarz bound
// (descending); a bra row's kets in one bucket are a contiguous run of it
// (a Schwarz prefix, or for mixed precision the part of it above or below the
// FP32 cut).  Row r, bucket k is segment r * nseg + k: rowptr holds the
// x_n * nseg + 1 prefix sums of the segments' lengths and qbase the first
// ket (relative to y_off) of each segment.  Consecutive quartets of a row
// therefore share their ket primitive count, so the lanes of a wavefront
// run equal primitive loops.
struct QuartetSet {
  int x_off, x_n;  // bra pairs [x_off, x_off + x_n) of the sorted pair list
  int y_off, y_n;
  int same;        // X == Y
  int nseg;        // implicit mode: segments (ket buckets) per bra row
  const long long *rowptr;  // implicit mode: x_n * nseg + 1 prefix sums, or null
  const int *qbase;         // implicit mode: first ket of each segment
  const int2 *list;         // explicit (P, Q) pairs relative to x_off / y_off, or null
  long long total;
};

struct DensDev {
  int ndens;            // original densities
  int neff;             // effective densities (symmetric + antisymmetric parts)
  const double *dc;     // neff x nao_cart^2 Cartesian densities
  // accumulators (kc may be null): ncopy copies of neff x nao_cart^2, copy_stride
  // doubles apart; wavefronts add into the copy of the XCD they run on (its
  // XCC_ID), which spreads concurrent atomics to one element over ncopy
  // addresses.  The copies are summed before the finishing transform.
  double *jc, *kc;
  size_t copy_stride;
  int ncopy;
  const int *eff_src;   // neff: original density index
  const int *eff_anti;  // neff: 1 if antisymmetric part (K only)
  const double *bound;  // ndens x nsh^2, symmetrised max |D| per shell block
  double tau;
  int want_k;
};

struct GradDev {
  int ncoul, nexch;
  const double *fac;     // ncoul + nexch
  const double *mats;    // (ncoul+nexch) x 2 x nao_cart^2 Cartesian left/right
  const double *bmax;    // (ncoul+nexch) x 2 x nsh^2 symmetrised block maxima (function space)
  double *out;           // natom x 3 accumulator
  double tau;
  double omega;          // <= 0: Coulomb
  // dense two-particle density (nao_cart^4 Cartesian, row-major) instead of
  // the density terms (ncoul = nexch = 0 then), or null
  const double *gamma;
};

// ---------------------------------------------------------------------------
// kernel launchers (ig_kernels.hip)
// every compiled kernel (J/K, ERI, Schwarz, gradient): shells up to d
bool class_supported(int la, int lb, int lc, int ld);
// compiled J/K kernels: the team kernel (class_supported plus the f classes
// of IG_FOR_F_TEAM_CLASSES) and the lane kernels (plus IG_FOR_F_LANE_CLASSES)
bool jk_team_compiled(int la, int lb, int lc, int ld);
bool jk_lane_compiled(int la, int lb, int lc, int ld);
// the team kernel with contracted-space horizontal transfers (ig_jk_hrr.hip):
// team classes of more than two register slices with lb > 0 or ld > 0, and
// (ff|fd), (ff|ff)
bool jk_hrr_compiled(int la, int lb, int lc, int ld);
// any compiled J/K kernel
bool jk_compiled(int la, int lb, int lc, int ld);
hipError_t launch_jk_hrr(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                         const RysDev &r, const QuartetSet &qs, const DensDev &d, double omega,
                         unsigned long long *counters, hipStream_t s);
hipError_t launch_schwarz(int lcls1, int lcls2, const BasisDev &b, const PairsDev &p,
                          const RysDev &r, int first, int count, double *qout, hipStream_t s);
// J/K kernel variants of a compiled class: one quartet per team of lanes
// with LDS tables, one quartet per lane in registers (sliced blocks), or the
// register path with per-row LDS accumulation of the bra-side updates
// -- or, for total angular momentum <= 2, the row-sum kernel with an OS/HGP
// block from the Boys function instead of Rys roots
enum JkVariant { kJkTeam = 0, kJkReg = 1, kJkRowAcc = 2, kJkHgp = 3, kJkTeamHrr = 4, kJkNumVariants = 5 };
// dynamic LDS of one row-accumulating workgroup: the device's per-workgroup
// limit (160 KB on gfx950), read once (the register-heavy classes hold one
// workgroup per CU anyway; small accumulators leave room for several)
size_t rowacc_lds_max();
// compute units of the current device (sizes the row-sum kernels' grids)
int device_cu_count();
int jk_default_variant(int la, int lb, int lc, int ld);
// width of the row accumulators' compact column space for ket momenta lc >= ld
int rowacc_width(const BasisDev &b, int lc, int ld);
// LDS accumulator doubles of the row-sum kernels for class (la lb|lc ld)
size_t rowacc_acc_doubles(const BasisDev &b, int la, int lb, int lc, int ld, int neff);
// can `variant` run this class on this basis (row accumulation: implicit
// quartet sets only, and its LDS must fit)?
bool jk_variant_fits(int variant, int la, int lb, int lc, int ld, const BasisDev &b, int neff,
                     bool listed);
hipError_t launch_jk(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                     const RysDev &r, const QuartetSet &qs, const DensDev &d, double omega,
                     int variant, unsigned long long *counters, hipStream_t s);
// the lane-per-quartet variants (kJkReg, kJkRowAcc) of launch_jk, FP64 and
// with FP32 tables and block (mixed-precision quartets)
hipError_t launch_jk_lane(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                          const RysDev &r, const QuartetSet &qs, const DensDev &d, double omega,
                          int variant, unsigned long long *counters, hipStream_t s);
hipError_t launch_jk_lane_f32(int la, int lb, int lc, int ld, const BasisDev &b,
                              const PairsDev &p, const RysDev &r, const QuartetSet &qs,
                              const DensDev &d, double omega, int variant,
                              unsigned long long *counters, hipStream_t s);
// the same for the f classes (separate translation units; the launchers
// above hand classes they do not hold to these)
hipError_t launch_jk_lane_fcls(int la, int lb, int lc, int ld, const BasisDev &b,
                               const PairsDev &p, const RysDev &r, const QuartetSet &qs,
                               const DensDev &d, double omega, int variant,
                               unsigned long long *counters, hipStream_t s);
hipError_t launch_jk_lane_f32_fcls(int la, int lb, int lc, int ld, const BasisDev &b,
                                   const PairsDev &p, const RysDev &r, const QuartetSet &qs,
                                   const DensDev &d, double omega, int variant,
                                   unsigned long long *counters, hipStream_t s);
// the row-sum kernel (variant kJkRowAcc or kJkHgp) over the sets of nb bases
// in one launch per element slice: device arrays bs, ps, sets, dds (nb each),
// seg (nb + 1 prefix sums of the set totals), queues (one zeroed counter per
// slice); acc_doubles = LDS accumulator doubles of the widest basis
constexpr int kMaxSlices = 16;
hipError_t launch_jk_batch(int la, int lb, int lc, int ld, const BasisDev *bs, const PairsDev *ps,
                           const RysDev &r, const QuartetSet *sets, const DensDev *dds,
                           const long long *seg, int nb, long long total, unsigned *queues,
                           int variant, size_t acc_doubles, unsigned long long *counters,
                           hipStream_t s);
hipError_t launch_jk_batch_f32(int la, int lb, int lc, int ld, const BasisDev *bs,
                               const PairsDev *ps, const RysDev &r, const QuartetSet *sets,
                               const DensDev *dds, const long long *seg, int nb, long long total,
                               unsigned *queues, int variant, size_t acc_doubles,
                               unsigned long long *counters, hipStream_t s);
hipError_t launch_jk_batch_fcls(int la, int lb, int lc, int ld, const BasisDev *bs,
                                const PairsDev *ps, const RysDev &r, const QuartetSet *sets,
                                const DensDev *dds, const long long *seg, int nb, long long total,
                                unsigned *queues, int variant, size_t acc_doubles,
                                unsigned long long *counters, hipStream_t s);
hipError_t launch_jk_batch_f32_fcls(int la, int lb, int lc, int ld, const BasisDev *bs,
                                    const PairsDev *ps, const RysDev &r, const QuartetSet *sets,
                                    const DensDev *dds, const long long *seg, int nb,
                                    long long total, unsigned *queues, int variant,
                                    size_t acc_doubles, unsigned long long *counters,
                                    hipStream_t s);
// Output of the ERI kernels: the dense nao^4 tensor (every distinct
// permutation of each canonical quartet) or, with `start` set, the packed
// symmetry-unique triangle over shell-pair-blocked AO pairs (packed_eri.rs:
// 10-21): start[i(i+1)/2+j] is the first AO-pair index of canonical shell
// pair (i, j); element (t|u), u <= t, goes to t(t+1)/2 + u - row0.
// With rows_end >= 0 the dense scatter only writes the permutations whose
// first index pair (the output row) belongs to a canonical shell pair in
// [rows_begin, rows_end): the rows an EriBuilder bra-pair range owns.
// Density fitting (df.rs) runs the same kernels on a combined basis: main
// shells [0, nmain_sh), aux shells, then one unit-s dummy per aux shell from
// dummy_sh0 on (zero exponent, coefficient 1: the constant function).  Mode
// kDf3c stores (main pair | aux, dummy) blocks as [nao_m, nao_m, naux] with
// both bra orders; kDf2c stores (aux, dummy | aux, dummy) as the symmetric
// [naux, naux] metric (aux functions start at function index nao_m).
enum EriMode { kEriDense = 0, kEriPacked = 1, kDf3c = 2, kDf2c = 3 };
struct EriOut {
  double *out;
  const long long *start;  // packed: AO-pair starts (null otherwise)
  long long row0;
  long long rows_begin = 0, rows_end = -1;
  int mode = kEriDense;
  int nmain_sh = 0, dummy_sh0 = 0;
  long long nao_m = 0, naux = 0;
  double omega = 0.0;  // the launch's operator: erf(omega r)/r when > 0, else 1/r
};
hipError_t launch_eri(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                      const RysDev &r, const QuartetSet &qs, const EriOut &o, hipStream_t s);
// lane-per-quartet adjoint gradient (ig_grad_lane.hip); natom sizes its LDS
// copy of the gradient
hipError_t launch_grad_lane(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                            const RysDev &r, const QuartetSet &qs, const GradDev &g, int natom,
                            unsigned long long *counters, hipStream_t s);

// runtime-l launchers (ig_big.hip) for classes class_supported() rejects;
// `scratch` holds big_scratch_elems(...) doubles (one block slot per
// workgroup; two per workgroup for the dense ERI, which transforms in place)
size_t big_scratch_elems(int la, int lb, int lc, int ld, long long total, int two_halves);
// the same for the runtime-l J/K launcher (launch_jk_big), whose blocks of
// 262144 elements or more stay in registers (the element offset table only)
size_t jk_big_scratch_elems(int la, int lb, int lc, int ld, long long total);
hipError_t launch_schwarz_big(int la, int lb, const BasisDev &b, const PairsDev &p,
                              const RysDev &r, int first, int count, double *qout,
                              double *scratch, hipStream_t s);
hipError_t launch_jk_big(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                         const RysDev &r, const QuartetSet &qs, const DensDev &d, double omega,
                         double *scratch, unsigned long long *counters, hipStream_t s);
hipError_t launch_eri_big(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                          const RysDev &r, const QuartetSet &qs, const EriOut &o, double *scratch,
                          hipStream_t s);
hipError_t launch_grad_big(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                           const RysDev &r, const QuartetSet &qs, const GradDev &g,
                           double *scratch, unsigned long long *counters, hipStream_t s);

// small dense helpers
// ds/da: ndens x nao^2 symmetric/antisymmetric parts (da used only when
// has_anti); dc: (1 + has_anti) * ndens Cartesian matrices (symmetric parts
// first); bound: ndens x nsh^2 symmetrised block maxima of d_func.
hipError_t launch_density_prep(const BasisDev &b, const double *d_func, int ndens, int has_anti,
                               double *ds, double *da, double *dc, double *bound,
                               hipStream_t s);
hipError_t launch_fill(double *x, size_t n, double v, hipStream_t s);
// *flag = 1 if any of the ndens n x n matrices differs from its transpose
hipError_t launch_asym_flag(const double *d, int n, int ndens, int *flag, hipStream_t s);
hipError_t launch_block_max(const BasisDev &b, const double *m, int nmats, double *bmax,
                            hipStream_t s);
hipError_t launch_to_cart(const BasisDev &b, const double *m_func, int nmats, double *m_cart,
                          hipStream_t s);
// dense nao^4 two-particle density -> nao_cart^4 (t0, t1: nao_cart^4 doubles
// each; result in t1)
hipError_t launch_gamma_to_cart(const BasisDev &b, const double *gamma, double *t0, double *t1,
                                hipStream_t s);
hipError_t launch_jk_finish(const BasisDev &b, const DensDev &d, double *d_J, double *d_K,
                            hipStream_t s);
// sum the ncopy per-XCD accumulator copies of d into the first one
hipError_t launch_fold_copies(const DensDev &d, size_t nc2, hipStream_t s);
// Boys values F_0..F_mmax (out: n x (mmax+1)) and n-point Rys roots/weights
// (x, w: count x nr) at the points T, on the device
hipError_t launch_boys_eval(const RysDev &r, int mmax, const double *T, size_t n, double *out,
                            hipStream_t s);
hipError_t launch_rys_eval(const RysDev &r, int nr, const double *T, size_t n, double *x,
                           double *w, hipStream_t s);
// the fixed device array of the no-allocation entry points on the current
// device (its address from hipGetSymbolAddress, taken once per device) and
// its size in bytes
hipError_t noalloc_scratch(void **ptr, size_t *bytes);
// histogram (16 bins) of the XCC_ID seen by `groups` workgroups
hipError_t launch_xcc_probe(unsigned *hist, int groups, hipStream_t s);

}  // namespace ig
