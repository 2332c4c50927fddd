// ig_lane.h -- lane-per-quartet J/K kernels (one shell quartet per lane,
// block in registers): the plain register path (ig_reg.h) and the register
// path with per-row LDS sums (ig_rowacc.h), with their class dispatch, for the
// arithmetic type IG_LANE_T (double, or float for the mixed-precision
// quartets), exported as IG_LANE_FN, over the class list IG_LANE_CLASSES
// (classes it does not hold go on to IG_LANE_NEXT).  Each type and list is
// its own translation unit (ig_jk_lane{64,32}{,f}.hip) so they compile side
// by side.
#include "ig_cls.h"

namespace ig {

#include "ig_reg.h"
#include "ig_hgp.h"
#include "ig_rowacc.h"

#ifdef IG_LANE_HOST_HELPERS
int rowacc_width(const BasisDev &b, int lc, int ld) {
  const int nC = b.lshell_off[lc + 1] - b.lshell_off[lc];
  const int nD = b.lshell_off[ld + 1] - b.lshell_off[ld];
  return nC * ncart(lc) + (lc == ld ? 0 : nD * ncart(ld));
}

size_t rowacc_lds_max() {
  static const size_t v = [] {
    int dev = 0, bytes = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess ||
        bytes < 64 * 1024)
      return size_t(64 * 1024);
    return size_t(bytes);
  }();
  return v;
}

int device_cu_count() {
  static const int v = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n < 1)
      return 256;
    return n;
  }();
  return v;
}

size_t rowacc_acc_doubles(const BasisDev &b, int la, int lb, int lc, int ld, int neff) {
  return rowacc_doubles(ncart(la), ncart(lb), neff, rowacc_width(b, lc, ld));
}

bool jk_variant_fits(int variant, int la, int lb, int lc, int ld, const BasisDev &b, int neff,
                     bool listed) {
  if (variant == kJkTeam) return jk_team_compiled(la, lb, lc, ld);
  if (variant == kJkTeamHrr) return jk_hrr_compiled(la, lb, lc, ld);
  if (!jk_lane_compiled(la, lb, lc, ld)) return false;
  if (variant == kJkReg) return true;
  if (listed) return false;
  if (variant == kJkHgp && !hgp_covered(la, lb, lc, ld)) return false;
  const int nr = (la + lb + lc + ld) / 2 + 1;
  const size_t rys =
      variant == kJkHgp ? 0 : size_t(2) * rys_host_tables().nsub[nr] * nr * kRysTerms * sizeof(double);
  const size_t lds =
      rys + rowacc_doubles(ncart(la), ncart(lb), neff, rowacc_width(b, lc, ld)) * sizeof(double);
  return lds <= rowacc_lds_max();
}
#endif

namespace {

template <int LA, int LB, int LC, int LD, typename F>
hipError_t jk_lane_one(const BasisDev &b, const PairsDev &p, const RysDev &r, const QuartetSet &qs,
                       const DensDev &d, double omega, int variant, unsigned long long *ctr,
                       hipStream_t s) {
  using C = Cls<LA, LB, LC, LD>;
  if (qs.total <= 0) return hipSuccess;
  const size_t lds = size_t(2) * r.nsub[C::NR] * C::NR * kRysTerms * sizeof(double);
  if (variant == kJkHgp) {
    if constexpr (HgpClass<C>::ok) {
      const size_t acc = rowacc_doubles(C::na, C::nb, d.neff, rowacc_width(b, LC, LD));
      if (qs.list || acc * sizeof(double) > rowacc_lds_max()) return hipErrorInvalidValue;
      long long blocks = (qs.total + 64 * kRowWaves - 1) / (64 * kRowWaves);
      return launch_jk_rowacc_slices<C, 0, F, 1>(b, p, r, qs, d, omega, ctr, s, int(blocks),
                                                 acc * sizeof(double));
    } else {
      return hipErrorInvalidValue;
    }
  }
  if (variant == kJkRowAcc) {
    const size_t acc = rowacc_doubles(C::na, C::nb, d.neff, rowacc_width(b, LC, LD));
    if (qs.list || lds + acc * sizeof(double) > rowacc_lds_max()) return hipErrorInvalidValue;
    long long blocks = (qs.total + 64 * kRowWaves - 1) / (64 * kRowWaves);
    return launch_jk_rowacc_slices<C, 0, F>(b, p, r, qs, d, omega, ctr, s, int(blocks),
                                            lds + acc * sizeof(double));
  }
  long long blocks = (qs.total + 63) / 64;
  if (blocks > 8192) blocks = 8192;
  return launch_jk_reg_slices<C, 0, F>(b, p, r, qs, d, omega, ctr, s, int(blocks), lds);
}

}  // namespace

template <int LA, int LB, int LC, int LD, typename F>
hipError_t jk_lane_batch_one(const BasisDev *bs, const PairsDev *ps, const RysDev &r,
                             const QuartetSet *sets, const DensDev *dds, const long long *seg,
                             int nb, long long total, unsigned *queues, int variant,
                             size_t acc_doubles, unsigned long long *ctr, hipStream_t s) {
  using C = Cls<LA, LB, LC, LD>;
  if (total <= 0) return hipSuccess;
  const long long chunk = 4 * 64 * kRowWaves;  // quartets per work item
  long long blocks = (total + chunk - 1) / chunk;
  if (blocks > 1024) blocks = 1024;
  const size_t rys = size_t(2) * r.nsub[C::NR] * C::NR * kRysTerms * sizeof(double);
  if (variant == kJkHgp) {
    if constexpr (HgpClass<C>::ok)
      return launch_jk_rowacc_batch_slices<C, 0, F, 1>(bs, ps, r, sets, dds, seg, nb, queues,
                                                       chunk, int(acc_doubles), ctr, s,
                                                       int(blocks), acc_doubles * sizeof(double));
    else
      return hipErrorInvalidValue;
  }
  return launch_jk_rowacc_batch_slices<C, 0, F, 0>(bs, ps, r, sets, dds, seg, nb, queues, chunk,
                                                   int(acc_doubles), ctr, s, int(blocks),
                                                   rys + acc_doubles * sizeof(double));
}

hipError_t IG_LANE_BATCH_FN(int la, int lb, int lc, int ld, const BasisDev *bs, const PairsDev *ps,
                           const RysDev &r, const QuartetSet *sets, const DensDev *dds,
                           const long long *seg, int nb, long long total, unsigned *queues,
                           int variant, size_t acc_doubles, unsigned long long *ctr,
                           hipStream_t s) {
#define IG_CASE(A, B, Cc, D)                       \
  if (la == A && lb == B && lc == Cc && ld == D) \
    return jk_lane_batch_one<A, B, Cc, D, IG_LANE_T>(bs, ps, r, sets, dds, seg, nb, total, \
                                                     queues, variant, acc_doubles, ctr, s);
  IG_LANE_CLASSES(IG_CASE)
#undef IG_CASE
#ifdef IG_LANE_BATCH_NEXT
  return IG_LANE_BATCH_NEXT(la, lb, lc, ld, bs, ps, r, sets, dds, seg, nb, total, queues, variant,
                            acc_doubles, ctr, s);
#else
  return hipErrorInvalidValue;
#endif
}

hipError_t IG_LANE_FN(int la, int lb, int lc, int ld, const BasisDev &b, const PairsDev &p,
                     const RysDev &r, const QuartetSet &qs, const DensDev &d, double omega,
                     int variant, unsigned long long *ctr, hipStream_t s) {
#define IG_CASE(A, B, Cc, D)                       \
  if (la == A && lb == B && lc == Cc && ld == D) \
    return jk_lane_one<A, B, Cc, D, IG_LANE_T>(b, p, r, qs, d, omega, variant, ctr, s);
  IG_LANE_CLASSES(IG_CASE)
#undef IG_CASE
#ifdef IG_LANE_NEXT
  return IG_LANE_NEXT(la, lb, lc, ld, b, p, r, qs, d, omega, variant, ctr, s);
#else
  return hipErrorInvalidValue;
#endif
}

}  // namespace ig
