// ig_plan.cpp -- host side of the C ABI (include/integral_hip.h): basis
// validation, device plans (shell data, primitive-pair tables sorted by
// angular class and Schwarz bound), screened quartet sets, kernel launches.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <map>
#include <mutex>
#include <tuple>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/integral_hip.h"
#include "ig_internal.h"

using namespace ig;

// ---------------------------------------------------------------------------
// errors
namespace {
thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

struct DeviceError {
  int code;
  std::string msg;
};

void check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    // clear the runtime's last-error state, so the failure is reported once
    // and not again by a later call's launch check
    (void)hipGetLastError();
    throw DeviceError{IG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)};
  }
}

template <class F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const DeviceError &e) {
    return fail(e.code, e.msg);
  } catch (const std::bad_alloc &) {
    return fail(IG_ERR_DEVICE, "host allocation failed");
  } catch (const std::exception &e) {
    return fail(IG_ERR_DEVICE, e.what());
  }
}

// device buffer owning hipMalloc'd memory, or (fixed) a view of a region
// carved from memory it does not own, which never grows
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool fixed = false;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() {
    if (p && !fixed) (void)hipFree(p);
  }
  void view(void *q, size_t n) {
    if (p && !fixed) check(hipFree(p), "hipFree");
    p = q;
    bytes = n;
    fixed = true;
  }
  void reserve(size_t n) {
    if (n <= bytes) return;
    if (fixed) throw DeviceError{IG_ERR_DEVICE, "a fixed scratch region is too small"};
    if (p) check(hipFree(p), "hipFree");
    p = nullptr;
    bytes = 0;
    check(hipMalloc(&p, n ? n : 8), "hipMalloc");
    bytes = n;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
  template <class T>
  void upload(const std::vector<T> &v) {
    reserve(v.size() * sizeof(T));
    if (!v.empty()) check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
  }
};

int64_t pair_count(int nsh) { return int64_t(nsh) * (nsh + 1) / 2; }
void pair_of(int64_t p, int &i, int &j) {
  int64_t ii = int64_t((std::sqrt(8.0 * double(p) + 1.0) - 1.0) / 2.0);
  while (ii * (ii + 1) / 2 > p) --ii;
  while ((ii + 1) * (ii + 2) / 2 <= p) ++ii;
  i = int(ii);
  j = int(p - ii * (ii + 1) / 2);
}

}  // namespace

// ---------------------------------------------------------------------------
// basis
struct HostShell {
  int l, kind, nprim;
  double center[3];
  std::vector<double> exps, coefs, eff;
  int off_func, off_cart, atom;
  int nfunc() const { return kind ? 2 * l + 1 : ncart(l); }
};

struct ig_basis {
  std::vector<HostShell> shells;
  int nao = 0, nao_cart = 0, natom = 0;
};

extern "C" uint32_t integral_abi_version(void) { return 1u; }
extern "C" const char *ig_last_error_message(void) { return g_last_error.c_str(); }
extern "C" int ig_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int ig_basis_create(const ig_shell_desc *shells, size_t n, ig_basis **out) {
  if (!out || (n && !shells)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  return guarded([&] {
    auto b = std::make_unique<ig_basis>();
    std::vector<std::array<double, 3>> atoms;
    for (size_t s = 0; s < n; ++s) {
      const ig_shell_desc &d = shells[s];
      // Shell::with_kind order of checks (shell.rs:100-111)
      if (d.n_exponents <= 0)
        return fail(IG_ERR_EMPTY_CONTRACTION, "shell " + std::to_string(s) + ": empty contraction");
      if (d.n_exponents != d.n_coefficients)
        return fail(IG_ERR_MISMATCHED_CONTRACTION,
                    "shell " + std::to_string(s) + ": " + std::to_string(d.n_exponents) +
                        " exponents vs " + std::to_string(d.n_coefficients) + " coefficients");
      if (d.l < 0 || d.l > kMaxL)
        return fail(IG_ERR_L_TOO_HIGH, "shell " + std::to_string(s) + ": l=" +
                                           std::to_string(d.l) + " exceeds max 6");
      if (d.kind != IG_CARTESIAN && d.kind != IG_SPHERICAL)
        return fail(IG_ERR_INVALID_ARGUMENT, "bad shell kind");
      if (!d.exponents || !d.coefficients) return fail(IG_ERR_INVALID_ARGUMENT, "null primitives");
      HostShell h;
      h.l = d.l;
      h.kind = d.kind;
      h.nprim = d.n_exponents;
      std::memcpy(h.center, d.center, sizeof(h.center));
      h.exps.assign(d.exponents, d.exponents + h.nprim);
      h.coefs.assign(d.coefficients, d.coefficients + h.nprim);
      h.eff.resize(h.nprim);
      for (int i = 0; i < h.nprim; ++i)  // Shell::primitive_coeff, shell.rs:182-187
        h.eff[i] = h.exps[i] == 0.0 ? h.coefs[i] : h.coefs[i] * shell_norm(h.exps[i], h.l);
      h.off_func = b->nao;
      h.off_cart = b->nao_cart;
      b->nao += h.nfunc();
      b->nao_cart += ncart(h.l);
      int a = -1;
      for (size_t k = 0; k < atoms.size(); ++k)
        if (std::memcmp(atoms[k].data(), h.center, sizeof(h.center)) == 0) {
          a = int(k);
          break;
        }
      if (a < 0) {
        atoms.push_back({h.center[0], h.center[1], h.center[2]});
        a = int(atoms.size()) - 1;
      }
      h.atom = a;
      b->shells.push_back(std::move(h));
    }
    b->natom = int(atoms.size());
    *out = b.release();
    return int(IG_OK);
  });
}

extern "C" void ig_basis_destroy(ig_basis *b) { delete b; }
extern "C" size_t ig_basis_nao(const ig_basis *b) { return b ? size_t(b->nao) : 0; }
extern "C" size_t ig_basis_nao_cart(const ig_basis *b) { return b ? size_t(b->nao_cart) : 0; }
extern "C" size_t ig_basis_natom(const ig_basis *b) { return b ? size_t(b->natom) : 0; }
extern "C" size_t ig_basis_nshell(const ig_basis *b) { return b ? b->shells.size() : 0; }

// ---------------------------------------------------------------------------
// device plan: everything a quartet launch needs, built once per contractor
namespace {

struct RysDevice {
  DBuf buf, boys;
  RysDev view{};
};

RysDevice &rys_device() {
  // one copy per process on the current device (single-GPU library)
  static RysDevice *r = nullptr;
  static std::mutex m;
  std::lock_guard<std::mutex> g(m);
  if (!r) {
    auto *x = new RysDevice();
    const RysHostTables &t = rys_host_tables();
    x->buf.upload(t.data);
    x->view.buf = x->buf.as<double>();
    x->boys.upload(boys_host_table());
    x->view.boys = x->boys.as<double>();
    for (int n = 0; n <= kMaxRys; ++n) {
      x->view.nsub[n] = t.nsub[n];
      x->view.root_off[n] = t.root_off[n];
      x->view.wt_off[n] = t.wt_off[n];
      x->view.lag_off[n] = t.lag_off[n];
    }
    r = x;
  }
  return *r;
}

extern "C" int ig_boys_eval(int m_max, const double *T, size_t n, double *out) {
  if ((n && (!T || !out)) || m_max < 0 || m_max > kBoysMaxM - 8)
    return fail(IG_ERR_INVALID_ARGUMENT, "m_max must be in 0..25 and buffers non-null");
  return guarded([&] {
    if (!n) return int(IG_OK);
    DBuf dT, dout;
    dT.reserve(n * sizeof(double));
    dout.reserve(n * (m_max + 1) * sizeof(double));
    check(hipMemcpy(dT.p, T, n * sizeof(double), hipMemcpyHostToDevice), "copy T");
    check(launch_boys_eval(rys_device().view, m_max, dT.as<double>(), n, dout.as<double>(),
                           nullptr),
          "boys launch");
    check(hipMemcpy(out, dout.p, n * (m_max + 1) * sizeof(double), hipMemcpyDeviceToHost),
          "copy F");
    return int(IG_OK);
  });
}

extern "C" int ig_rys_eval(int nroots, const double *T, size_t n, double *roots,
                           double *weights) {
  if ((n && (!T || !roots || !weights)) || nroots < 1 || nroots > kMaxRys)
    return fail(IG_ERR_INVALID_ARGUMENT, "nroots must be in 1..13 and buffers non-null");
  return guarded([&] {
    if (!n) return int(IG_OK);
    DBuf dT, dx, dw;
    dT.reserve(n * sizeof(double));
    dx.reserve(n * nroots * sizeof(double));
    dw.reserve(n * nroots * sizeof(double));
    check(hipMemcpy(dT.p, T, n * sizeof(double), hipMemcpyHostToDevice), "copy T");
    check(launch_rys_eval(rys_device().view, nroots, dT.as<double>(), n, dx.as<double>(),
                          dw.as<double>(), nullptr),
          "rys launch");
    check(hipMemcpy(roots, dx.p, n * nroots * sizeof(double), hipMemcpyDeviceToHost), "copy x");
    check(hipMemcpy(weights, dw.p, n * nroots * sizeof(double), hipMemcpyDeviceToHost),
          "copy w");
    return int(IG_OK);
  });
}

// Number of per-XCD J'/K' accumulator copies: one per XCC_ID that a probe
// launch of many workgroups observes (the dispatcher spreads workgroups over
// every XCD of the device); 1 when the probe sees a single id.
int xcd_copies() {
  static int copies = 0;
  static std::mutex m;
  std::lock_guard<std::mutex> g(m);
  if (copies) return copies;
  DBuf hist;
  hist.reserve(16 * sizeof(unsigned));
  check(hipMemset(hist.p, 0, 16 * sizeof(unsigned)), "zero probe");
  check(launch_xcc_probe(hist.as<unsigned>(), 4096, nullptr), "xcc probe");
  unsigned h[16];
  check(hipMemcpy(h, hist.p, sizeof(h), hipMemcpyDeviceToHost), "copy probe");
  int distinct = 0, top = 0;
  for (int i = 0; i < 16; ++i)
    if (h[i]) {
      ++distinct;
      top = i + 1;
    }
  copies = distinct >= 2 ? top : 1;
  return copies;
}

struct ClassRange {
  int off = 0, n = 0, l1 = 0, l2 = 0;
  // buckets of equal primitive-pair count: bucket k holds the class's pairs
  // [off + bk[k], off + bk[k + 1])
  std::vector<int> bk;
};

struct Plan {
  const ig_basis *basis;  // shells are copied below; pointer only for sizes
  int nsh, nao, nao_cart, natom;
  std::vector<int> sh_l, sh_kind, sh_atom, sh_off_func, sh_off_cart, sh_nprim;
  // device basis
  DBuf d_xyz, d_l, d_kind, d_off_func, d_off_cart, d_atom, d_c2s, d_rank, d_lshells;
  BasisDev bdev{};
  // pairs, final (class, Q-descending) order
  int npairs = 0;
  std::vector<int> p_sh1, p_sh2, p_canon, p_ppoff, p_ppn;
  std::vector<double> p_q;
  std::vector<int> canon_to_pos;  // canonical pair index -> sorted position
  ClassRange cls[kNumPairClass];
  DBuf d_sh1, d_sh2, d_canon, d_q, d_ppoff, d_ppn;
  DBuf d_zeta, d_px, d_py, d_pz, d_kc, d_alpha, d_beta, d_kz, d_i2z;
  PairsDev pdev{};
  std::vector<double> schwarz;  // nsh x nsh (function space)
  std::mutex mu;                // serialises calls that use the scratch below
  DBuf s_dfunc, s_ds, s_da, s_dc, s_bound, s_jc, s_kc, s_J, s_K, s_eff, s_rowptr, s_qbase,
      s_list, s_counters, s_flag;
  DBuf s_big;  // block slots of the runtime-l kernels (classes beyond the compiled ones)
  // J/K kernel variant per (bra class, ket class): -1 = the built-in default,
  // else set by the autotuner or a loaded schedule (with its measured time)
  std::vector<int8_t> variant = std::vector<int8_t>(kNumPairClass * kNumPairClass, -1);
  std::vector<float> variant_ms = std::vector<float>(kNumPairClass * kNumPairClass, 0.0f);
  double cached_tau = NAN, cached_cut = NAN;
  std::vector<long long> cached_rowptr_host;
  std::vector<int> cached_qbase_host;
  // mixed precision: quartets whose Schwarz product Q_P Q_Q is below this
  // (and at or above the threshold) run with FP32 tables; 0 = FP64 only
  double fp32_below = 0.0;
  struct CP {
    int x, y;
    size_t rp_off;
    long long total;
    int fp32 = 0;         // 1: the FP32 part (the segments' kets below the cut)
    long long qb_off = -1;  // offset of its per-segment first kets in s_qbase
    int nseg = 1;         // segments (ket buckets) per bra row
    double prims = 0.0;   // primitive quartets before density screening
  };
  std::vector<CP> cached_cps;
  int eff_key = -1;
  // per-launch attribution of the last call (counters on the device, events
  // when profiling); finalised lazily on the stream the call used
  static constexpr int kMaxLaunch = kNumPairClass * (kNumPairClass + 1);  // FP64 + FP32 sets
  struct Launch {
    int la, lb, lc, ld, fp32;
  };
  std::vector<Launch> launches;
  std::vector<unsigned long long> launch_ctr;  // 2 per launch, host copy
  std::vector<float> launch_ms;
  bool profiling = false;
  std::vector<hipEvent_t> events;  // 2 per launch slot
  bool stats_pending = false;
  hipStream_t stats_stream = nullptr;
  int64_t last_stats[2] = {0, 0};
  // device-side order of the calls that use this plan's scratch: each call
  // waits on the stream for the previous call's last work (recorded here),
  // so calls on different streams never overlap on the scratch buffers
  hipEvent_t idle = nullptr;
  bool idle_recorded = false;
  // streams the class launches of one J/K call are spread over (the call's
  // own stream first), with the events that fork them from it and join them
  // back, so one launch's tail overlaps the others' work
  std::vector<hipStream_t> aux;
  hipEvent_t fork_ev = nullptr;
  std::vector<hipEvent_t> join_ev;
  hipStream_t cur_stream = nullptr;  // stream of the call in progress
  // guards the schedule and precision policy (variant, variant_ms,
  // fp32_below) that workspaces copy at each call
  std::mutex settings_mu;
  // scratch carved from a fixed region (the no-allocation entry): one
  // accumulator copy, no events, every upload an asynchronous copy
  bool fixed_scratch = false;
  ~Plan() {
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
    if (idle) (void)hipEventDestroy(idle);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    for (hipEvent_t e : join_ev) (void)hipEventDestroy(e);
    for (hipStream_t a : aux) {
      (void)hipStreamSynchronize(a);
      (void)hipStreamDestroy(a);
    }
  }
};

// Order a call on stream s after every earlier call on the plan's scratch.
void plan_enter(Plan &P, hipStream_t s) {
  P.cur_stream = s;
  if (P.idle_recorded) check(hipStreamWaitEvent(s, P.idle, 0), "wait for the previous call");
}

// Host data into a plan scratch buffer: a synchronous upload, or with fixed
// scratch an asynchronous copy on the call's stream (the host vector is the
// plan's own and outlives the copy's staging).
template <class T>
void upload_to(Plan &P, DBuf &buf, const std::vector<T> &v) {
  if (!P.fixed_scratch) {
    // a synchronous upload: the previous call's kernels may still read buf
    if (P.idle_recorded) check(hipEventSynchronize(P.idle), "wait for the previous call");
    buf.upload(v);
    return;
  }
  buf.reserve(v.size() * sizeof(T));
  if (!v.empty())
    check(hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice,
                         P.cur_stream),
          "copy to scratch");
}

// Mark the end of a call's device work on stream s.
void plan_leave(Plan &P, hipStream_t s) {
  if (!P.idle) check(hipEventCreateWithFlags(&P.idle, hipEventDisableTiming), "hipEventCreate");
  check(hipEventRecord(P.idle, s), "record call end");
  P.idle_recorded = true;
}

void begin_launches(Plan &P, hipStream_t s) {
  P.launches.clear();
  P.s_counters.reserve(2 * Plan::kMaxLaunch * sizeof(unsigned long long));
  check(launch_fill(P.s_counters.as<double>(), 2 * Plan::kMaxLaunch, 0.0, s), "zero counters");
  if (P.profiling && P.events.empty()) {
    P.events.resize(2 * Plan::kMaxLaunch);
    for (hipEvent_t &e : P.events) check(hipEventCreate(&e), "hipEventCreate");
  }
}

unsigned long long *next_launch(Plan &P, const ClassRange &X, const ClassRange &Y, hipStream_t s,
                                int fp32 = 0) {
  const int k = int(P.launches.size());
  if (k >= Plan::kMaxLaunch) throw DeviceError{IG_ERR_DEVICE, "too many class launches"};
  P.launches.push_back({X.l1, X.l2, Y.l1, Y.l2, fp32});
  if (P.profiling) check(hipEventRecord(P.events[2 * k], s), "event");
  return P.s_counters.as<unsigned long long>() + 2 * k;
}

void end_launch(Plan &P, hipStream_t s) {
  if (P.profiling) check(hipEventRecord(P.events[2 * (P.launches.size() - 1) + 1], s), "event");
}

void finalize_stats(Plan &P) {
  if (!P.stats_pending) return;
  const size_t n = P.launches.size();
  P.launch_ctr.assign(2 * std::max<size_t>(n, 1), 0);
  if (n)
    check(hipMemcpyAsync(P.launch_ctr.data(), P.s_counters.p, 2 * n * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, P.stats_stream),
          "copy counters");
  check(hipStreamSynchronize(P.stats_stream), "stats sync");
  P.last_stats[0] = P.last_stats[1] = 0;
  P.launch_ms.assign(n, 0.0f);
  for (size_t k = 0; k < n; ++k) {
    P.last_stats[0] += int64_t(P.launch_ctr[2 * k]);
    P.last_stats[1] += int64_t(P.launch_ctr[2 * k + 1]);
    if (P.profiling)
      check(hipEventElapsedTime(&P.launch_ms[k], P.events[2 * k], P.events[2 * k + 1]),
            "event time");
  }
  P.stats_pending = false;
}

void upload_basis(Plan &P, const ig_basis &b) {
  P.nsh = int(b.shells.size());
  P.nao = b.nao;
  P.nao_cart = b.nao_cart;
  P.natom = b.natom;
  std::vector<double> xyz;
  for (const HostShell &s : b.shells) {
    P.sh_l.push_back(s.l);
    P.sh_kind.push_back(s.kind);
    P.sh_atom.push_back(s.atom);
    P.sh_off_func.push_back(s.off_func);
    P.sh_off_cart.push_back(s.off_cart);
    P.sh_nprim.push_back(s.nprim);
    xyz.insert(xyz.end(), s.center, s.center + 3);
  }
  std::vector<double> c2s;
  for (int l = 0; l <= kMaxL; ++l) {
    P.bdev.c2s_off[l] = int(c2s.size());
    std::vector<double> m = c2s_scaled(l);
    c2s.insert(c2s.end(), m.begin(), m.end());
  }
  P.d_xyz.upload(xyz);
  P.d_l.upload(P.sh_l);
  P.d_kind.upload(P.sh_kind);
  P.d_off_func.upload(P.sh_off_func);
  P.d_off_cart.upload(P.sh_off_cart);
  P.d_atom.upload(P.sh_atom);
  P.d_c2s.upload(c2s);
  // shells grouped by angular momentum: rank of each shell inside its group
  // and the groups' shell lists (compact per-l column spaces)
  std::vector<int> rank(P.nsh), lshells;
  for (int l = 0; l <= kMaxL; ++l) {
    P.bdev.lshell_off[l] = int(lshells.size());
    for (int s = 0; s < P.nsh; ++s)
      if (P.sh_l[s] == l) {
        rank[s] = int(lshells.size()) - P.bdev.lshell_off[l];
        lshells.push_back(s);
      }
  }
  P.bdev.lshell_off[kMaxL + 1] = int(lshells.size());
  P.d_rank.upload(rank);
  P.d_lshells.upload(lshells);
  P.bdev.rank = P.d_rank.as<int>();
  P.bdev.lshells = P.d_lshells.as<int>();
  P.bdev.nsh = P.nsh;
  P.bdev.nao = P.nao;
  P.bdev.nao_cart = P.nao_cart;
  P.bdev.xyz = P.d_xyz.as<double>();
  P.bdev.l = P.d_l.as<int>();
  P.bdev.kind = P.d_kind.as<int>();
  P.bdev.off_func = P.d_off_func.as<int>();
  P.bdev.off_cart = P.d_off_cart.as<int>();
  P.bdev.atom = P.d_atom.as<int>();
  P.bdev.c2s = P.d_c2s.as<double>();
}

void upload_pairs(Plan &P) {
  P.d_sh1.upload(P.p_sh1);
  P.d_sh2.upload(P.p_sh2);
  P.d_canon.upload(P.p_canon);
  P.d_q.upload(P.p_q);
  P.d_ppoff.upload(P.p_ppoff);
  P.d_ppn.upload(P.p_ppn);
  P.pdev.n = P.npairs;
  P.pdev.sh1 = P.d_sh1.as<int>();
  P.pdev.sh2 = P.d_sh2.as<int>();
  P.pdev.canon = P.d_canon.as<int>();
  P.pdev.q = P.d_q.as<double>();
  P.pdev.pp_off = P.d_ppoff.as<int>();
  P.pdev.pp_n = P.d_ppn.as<int>();
}

// Orientation of a pair of equal angular momentum: should shell x be the
// first (the centre the vertical recurrences build on, from which the
// horizontal transfers move momentum to the second)?  The transfers
// (a, b+1| = (a+1, b| + AB (a, b| cancel badly when the second shell is the
// tight one and the product centre P sits next to it, far from A: on l = 6
// pairs with exponents 9.7 and 0.21, 6.4 bohr apart, the diffuse-first order
// loses 12-28% of the block against a 40-digit McMurchie-Davidson value, the
// tight-first order agrees to 4e-14.  So the first shell is the one whose
// worst primitive pair keeps P closest to it: smaller max over pairs of
// beta / (alpha + beta), beta the other shell's exponent.  Shells with a zero
// exponent (the unit-s dummies of density fitting) keep the canonical order.
bool tighter_first(const HostShell &x, const HostShell &y) {
  double xmin = INFINITY, xmax = 0.0, ymin = INFINITY, ymax = 0.0;
  for (double e : x.exps) {
    xmin = std::min(xmin, e);
    xmax = std::max(xmax, e);
  }
  for (double e : y.exps) {
    ymin = std::min(ymin, e);
    ymax = std::max(ymax, e);
  }
  if (!(xmin > 0.0 && ymin > 0.0)) return false;
  return ymax / (xmin + ymax) < xmax / (ymin + xmax);
}

// Build the plan: pairs in kernel orientation, primitive-pair SoA tables,
// Schwarz bounds on the device, then the final (class, Q desc) order.
std::unique_ptr<Plan> build_plan(const ig_basis &b, hipStream_t s) {
  auto P = std::make_unique<Plan>();
  P->basis = &b;
  upload_basis(*P, b);
  const int nsh = P->nsh;
  const int64_t np = pair_count(nsh);
  P->npairs = int(np);
  // primitive pairs (os_eri.rs:221-249 quantities; K c1 c2 folded together)
  std::vector<double> zeta, px, py, pz, kc, al, be, kz, i2z;
  struct Tmp {
    int sh1, sh2, canon, off, n, cls;
  };
  std::vector<Tmp> tmp(np);
  for (int64_t c = 0; c < np; ++c) {
    int i, j;
    pair_of(c, i, j);
    const HostShell *s1 = &b.shells[i], *s2 = &b.shells[j];
    int a1 = i, a2 = j;
    if (s1->l < s2->l || (s1->l == s2->l && tighter_first(*s2, *s1))) {
      std::swap(s1, s2);
      std::swap(a1, a2);
    }
    Tmp t{a1, a2, int(c), int(zeta.size()), 0, pair_class(s1->l, s2->l)};
    double d2 = 0.0;
    for (int x = 0; x < 3; ++x) d2 += (s1->center[x] - s2->center[x]) * (s1->center[x] - s2->center[x]);
    for (int p1 = 0; p1 < s1->nprim; ++p1)
      for (int p2 = 0; p2 < s2->nprim; ++p2) {
        double e1 = s1->exps[p1], e2 = s2->exps[p2], z = e1 + e2;
        // two unit-s dummies (zero exponents): a constant product function
        // that no density-fitting quartet uses (df.rs:1-39 pairs a dummy
        // with its aux shell only)
        if (z == 0.0) continue;
        double K = std::exp(-(e1 * e2 / z) * d2);
        double w = K * s1->eff[p1] * s2->eff[p2];
        if (K * std::fabs(s1->eff[p1] * s2->eff[p2]) < 1e-32) continue;  // os_eri.rs:211
        zeta.push_back(z);
        px.push_back((e1 * s1->center[0] + e2 * s2->center[0]) / z);
        py.push_back((e1 * s1->center[1] + e2 * s2->center[1]) / z);
        pz.push_back((e1 * s1->center[2] + e2 * s2->center[2]) / z);
        kc.push_back(w);
        kz.push_back(w / z);
        i2z.push_back(0.5 / z);
        al.push_back(e1);
        be.push_back(e2);
        ++t.n;
      }
    tmp[c] = t;
  }
  P->d_zeta.upload(zeta);
  P->d_px.upload(px);
  P->d_py.upload(py);
  P->d_pz.upload(pz);
  P->d_kc.upload(kc);
  P->d_alpha.upload(al);
  P->d_beta.upload(be);
  P->d_kz.upload(kz);
  P->d_i2z.upload(i2z);
  P->pdev.kz = P->d_kz.as<double>();
  P->pdev.i2z = P->d_i2z.as<double>();
  P->pdev.zeta = P->d_zeta.as<double>();
  P->pdev.px = P->d_px.as<double>();
  P->pdev.py = P->d_py.as<double>();
  P->pdev.pz = P->d_pz.as<double>();
  P->pdev.kc = P->d_kc.as<double>();
  P->pdev.alpha = P->d_alpha.as<double>();
  P->pdev.beta = P->d_beta.as<double>();
  // class order (stable by canonical index)
  std::vector<int> order(np);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return tmp[x].cls < tmp[y].cls; });
  auto fill = [&](const std::vector<int> &ord) {
    P->p_sh1.resize(np);
    P->p_sh2.resize(np);
    P->p_canon.resize(np);
    P->p_ppoff.resize(np);
    P->p_ppn.resize(np);
    for (int64_t k = 0; k < np; ++k) {
      const Tmp &t = tmp[ord[k]];
      P->p_sh1[k] = t.sh1;
      P->p_sh2[k] = t.sh2;
      P->p_canon[k] = t.canon;
      P->p_ppoff[k] = t.off;
      P->p_ppn[k] = t.n;
    }
  };
  fill(order);
  for (int c = 0; c < kNumPairClass; ++c) P->cls[c] = ClassRange{};
  for (int64_t k = 0; k < np; ++k) {
    int c = tmp[order[k]].cls;
    if (P->cls[c].n == 0) P->cls[c].off = int(k);
    P->cls[c].n++;
  }
  for (int l1 = 0; l1 <= kMaxL; ++l1)
    for (int l2 = 0; l2 <= l1; ++l2) {
      ClassRange &r = P->cls[pair_class(l1, l2)];
      r.l1 = l1;
      r.l2 = l2;
    }
  P->p_q.assign(np, 0.0);
  upload_pairs(*P);
  // Schwarz on the device, per pair class
  DBuf dq;
  dq.reserve(np * sizeof(double));
  RysDevice &R = rys_device();
  size_t big = 0;
  for (int c = 0; c < kNumPairClass; ++c) {
    const ClassRange &r = P->cls[c];
    if (r.n && !class_supported(r.l1, r.l2, r.l1, r.l2))
      big = std::max(big, big_scratch_elems(r.l1, r.l2, r.l1, r.l2, r.n, 0));
  }
  P->s_big.reserve(big * sizeof(double));
  for (int c = 0; c < kNumPairClass; ++c) {
    const ClassRange &r = P->cls[c];
    if (!r.n) continue;
    if (class_supported(r.l1, r.l2, r.l1, r.l2))
      check(launch_schwarz(r.l1, r.l2, P->bdev, P->pdev, R.view, r.off, r.n, dq.as<double>(), s),
            "schwarz launch");
    else
      check(launch_schwarz_big(r.l1, r.l2, P->bdev, P->pdev, R.view, r.off, r.n, dq.as<double>(),
                               P->s_big.as<double>(), s),
            "schwarz launch");
  }
  std::vector<double> q(np);
  check(hipMemcpyAsync(q.data(), dq.p, np * sizeof(double), hipMemcpyDeviceToHost, s), "copy Q");
  check(hipStreamSynchronize(s), "schwarz sync");
  // final order: class, then primitive-pair count descending (the ket
  // buckets of the quartet sets), then Q descending (stable)
  std::vector<int> order2(np);
  std::iota(order2.begin(), order2.end(), 0);
  std::stable_sort(order2.begin(), order2.end(), [&](int x, int y) {
    const Tmp &tx = tmp[order[x]], &ty = tmp[order[y]];
    if (tx.cls != ty.cls) return tx.cls < ty.cls;
    if (tx.n != ty.n) return tx.n > ty.n;
    return q[x] > q[y];
  });
  std::vector<int> final_order(np);
  std::vector<double> qs(np);
  for (int64_t k = 0; k < np; ++k) {
    final_order[k] = order[order2[k]];
    qs[k] = q[order2[k]];
  }
  fill(final_order);
  P->p_q = qs;
  for (int c = 0; c < kNumPairClass; ++c) {
    ClassRange &r = P->cls[c];
    r.bk.assign(1, 0);
    for (int k = 1; k < r.n; ++k)
      if (P->p_ppn[r.off + k] != P->p_ppn[r.off + k - 1]) r.bk.push_back(k);
    r.bk.push_back(r.n);
  }
  P->canon_to_pos.assign(np, 0);
  for (int64_t k = 0; k < np; ++k) P->canon_to_pos[P->p_canon[k]] = int(k);
  upload_pairs(*P);
  // nsh x nsh Schwarz matrix (symmetric; canonical orientation computed)
  P->schwarz.assign(size_t(nsh) * nsh, 0.0);
  for (int64_t k = 0; k < np; ++k) {
    int i, j;
    pair_of(P->p_canon[k], i, j);
    P->schwarz[size_t(i) * nsh + j] = P->p_q[k];
    P->schwarz[size_t(j) * nsh + i] = P->p_q[k];
  }
  return P;
}

// Number of kets of class y (Q-sorted) with qb*q >= tau; binary search on the
// same floating-point predicate the reference evaluates (direct.rs:172-175).
int count_kets(const double *qy, int ny, double qb, double tau) {
  int lo = 0, hi = ny;  // first index where qb*q < tau
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (qb * qy[mid] < tau)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Implicit quartet sets for every (X >= Y) class pair at threshold tau, on
// the host (build_rowptrs uploads them): pairs with Q == 0 contribute nothing
// (their integrals vanish) and are skipped.  With cut > tau each class pair
// has two sets: kets with Q_P Q_Q >= cut (a prefix of the row) in FP64 and
// the rest of the row, down to tau, as the FP32 set.
void rowptrs_host(Plan &P, double tau, double cut) {
  P.cached_cps.clear();
  P.cached_rowptr_host.clear();
  P.cached_qbase_host.clear();
  std::vector<int> lo, hi;
  for (int x = 0; x < kNumPairClass; ++x) {
    const ClassRange &X = P.cls[x];
    if (!X.n) continue;
    for (int y = 0; y <= x; ++y) {
      const ClassRange &Y = P.cls[y];
      if (!Y.n) continue;
      const int nseg = int(Y.bk.size()) - 1;
      lo.assign(size_t(X.n) * nseg, 0);
      hi.assign(size_t(X.n) * nseg, 0);
      for (int r = 0; r < X.n; ++r) {
        const double qb = P.p_q[X.off + r];
        if (qb == 0.0) continue;
        for (int k = 0; k < nseg; ++k) {
          const int b0 = Y.bk[k], bn = Y.bk[k + 1] - b0;
          const double *qy = P.p_q.data() + Y.off + b0;
          int cnt = count_kets(qy, bn, qb, tau);
          // one class: each unordered pair once, the ket at or before the
          // bra in the sorted order
          if (x == y) cnt = std::min(cnt, std::max(0, std::min(bn, r - b0 + 1)));
          while (cnt > 0 && qy[cnt - 1] == 0.0) --cnt;  // Q == 0 kets
          hi[size_t(r) * nseg + k] = cnt;
          lo[size_t(r) * nseg + k] =
              cut > tau ? std::min(cnt, count_kets(qy, bn, qb, cut)) : cnt;
        }
      }
      for (int part = 0; part < (cut > tau ? 2 : 1); ++part) {
        Plan::CP cp{x, y, P.cached_rowptr_host.size(), 0};
        cp.nseg = nseg;
        cp.qb_off = (long long)P.cached_qbase_host.size();
        long long acc = 0;
        P.cached_rowptr_host.push_back(0);
        for (int r = 0; r < X.n; ++r)
          for (int k = 0; k < nseg; ++k) {
            const size_t sg = size_t(r) * nseg + k;
            const int len = part == 0 ? lo[sg] : hi[sg] - lo[sg];
            acc += len;
            cp.prims += double(len) * P.p_ppn[X.off + r] * P.p_ppn[Y.off + Y.bk[k]];
            P.cached_rowptr_host.push_back(acc);
            P.cached_qbase_host.push_back(Y.bk[k] + (part == 0 ? 0 : lo[sg]));
          }
        cp.total = acc;
        cp.fp32 = part;
        P.cached_cps.push_back(cp);
      }
    }
  }
}

void build_rowptrs(Plan &P, double tau, double cut) {
  if (!(cut > tau)) cut = tau;
  if (P.cached_tau == tau && P.cached_cut == cut && !P.cached_cps.empty()) return;
  P.cached_tau = P.cached_cut = NAN;
  rowptrs_host(P, tau, cut);
  upload_to(P, P.s_rowptr, P.cached_rowptr_host);
  upload_to(P, P.s_qbase, P.cached_qbase_host);
  P.cached_tau = tau;
  P.cached_cut = cut;
}

void build_rowptrs(Plan &P, double tau) { build_rowptrs(P, tau, tau); }

QuartetSet implicit_set(Plan &P, const Plan::CP &cp) {
  QuartetSet qs{};
  qs.x_off = P.cls[cp.x].off;
  qs.x_n = P.cls[cp.x].n;
  qs.y_off = P.cls[cp.y].off;
  qs.y_n = P.cls[cp.y].n;
  qs.same = cp.x == cp.y;
  qs.nseg = cp.nseg;
  qs.rowptr = P.s_rowptr.as<long long>() + cp.rp_off;
  qs.qbase = P.s_qbase.as<int>() + cp.qb_off;
  qs.list = nullptr;
  qs.total = cp.total;
  return qs;
}

// explicit lists for the canonical bra pairs [b0, b1): kets k <= i,
// l <= (k == i ? j : k) passing Q_ij Q_kl >= tau (direct.rs:169-175)
struct ExplicitSets {
  std::vector<std::pair<int, int>> cls;  // (x, y)
  std::vector<std::vector<int2>> lists;
};

ExplicitSets explicit_sets(const Plan &P, int64_t b0, int64_t b1, double tau) {
  ExplicitSets out;
  int index[kNumPairClass][kNumPairClass];
  for (auto &r : index)
    for (int &v : r) v = -1;
  const int nsh = P.nsh;
  for (int64_t bc = b0; bc < b1; ++bc) {
    int i, j;
    pair_of(bc, i, j);
    double qab = P.schwarz[size_t(i) * nsh + j];
    if (qab == 0.0) continue;  // direct.rs:165-168
    int pb = P.canon_to_pos[bc];
    for (int k = 0; k <= i; ++k) {
      int ltop = (k == i) ? j : k;
      for (int l = 0; l <= ltop; ++l) {
        double qkl = P.schwarz[size_t(k) * nsh + l];
        if (qab * qkl < tau || qkl == 0.0) continue;
        int pk = P.canon_to_pos[int64_t(k) * (k + 1) / 2 + l];
        int cb = pair_class(P.sh_l[P.p_sh1[pb]], P.sh_l[P.p_sh2[pb]]);
        int ck = pair_class(P.sh_l[P.p_sh1[pk]], P.sh_l[P.p_sh2[pk]]);
        int x = cb, y = ck, px = pb, py = pk;
        if (cb < ck) {
          std::swap(x, y);
          std::swap(px, py);
        }
        if (index[x][y] < 0) {
          index[x][y] = int(out.cls.size());
          out.cls.push_back({x, y});
          out.lists.emplace_back();
        }
        out.lists[index[x][y]].push_back(int2{px - P.cls[x].off, py - P.cls[y].off});
      }
    }
  }
  return out;
}

}  // namespace

// ---------------------------------------------------------------------------
// public plan objects
struct ig_direct {
  std::unique_ptr<Plan> plan;
  int engine;
  // the no-allocation entry's plan: scratch carved from the fixed device
  // array (built on first use; host memory only)
  std::unique_ptr<Plan> fixed;
  std::mutex fixed_mu;
};

// DirectWorkspace (direct.rs:19-33): a caller's own scratch over a
// contractor's read-only plan data, so calls through different workspaces
// run concurrently (each on its own stream).
struct ig_direct_workspace {
  const ig_direct *owner;
  std::unique_ptr<Plan> plan;
};

struct ig_grad {
  std::unique_ptr<Plan> plan;
  int engine;
};

struct ig_grad_density {
  const ig_grad *owner;
  int ncoul = 0, nexch = 0;
  DBuf d_fac, d_mats, d_bmax;
  std::vector<double> fac;
};

namespace {

int check_engine(int engine) {
  if (engine != IG_ENGINE_AUTO && engine != IG_ENGINE_OS_HGP && engine != IG_ENGINE_RYS)
    return fail(IG_ERR_INVALID_ARGUMENT, "unknown engine");
  return IG_OK;
}

hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

// Classes beyond the compiled kernels run on the runtime-l kernels, whose
// block slots live in P.s_big; every launch of a call is sized before the
// first one is queued, so the buffer never moves under a running kernel.
size_t big_need(const ClassRange &X, const ClassRange &Y, long long total, int two_halves) {
  if (total <= 0 || class_supported(X.l1, X.l2, Y.l1, Y.l2)) return 0;
  return big_scratch_elems(X.l1, X.l2, Y.l1, Y.l2, total, two_halves);
}

// the same for J/K, whose compiled kernels also cover most f classes
size_t big_need_jk(const ClassRange &X, const ClassRange &Y, long long total) {
  if (total <= 0 || jk_compiled(X.l1, X.l2, Y.l1, Y.l2)) return 0;
  return jk_big_scratch_elems(X.l1, X.l2, Y.l1, Y.l2, total);
}

int jk_variant(const Plan &P, const ClassRange &X, const ClassRange &Y) {
  const int v = P.variant[pair_class(X.l1, X.l2) * kNumPairClass + pair_class(Y.l1, Y.l2)];
  return v >= 0 ? v : jk_default_variant(X.l1, X.l2, Y.l1, Y.l2);
}

// fp32: the set's quartets run with FP32 tables and block (lane kernels only;
// a class scheduled on the team kernel takes the register kernel, one on the
// OS/HGP block the register row-sum kernel, and the classes without lane
// kernels stay FP64).  The FP32 sets stay off the OS/HGP block: a mixed build
// of the nao 520 s,p,d workload at cut 1e-6 took an illegal memory access
// after (ss|ss), (ps|ss) and (pp|ps) moved to that block, and its cause is
// not found; every earlier mixed run at that size had them on Rys kernels.
hipError_t jk_launch(Plan &P, const ClassRange &X, const ClassRange &Y, const QuartetSet &qs,
                     const DensDev &dd, unsigned long long *ctr, hipStream_t s, int variant = -1,
                     bool fp32 = false) {
  RysDevice &R = rys_device();
  if (jk_compiled(X.l1, X.l2, Y.l1, Y.l2)) {
    const bool lane = jk_lane_compiled(X.l1, X.l2, Y.l1, Y.l2);
    int v = variant >= 0 ? variant : jk_variant(P, X, Y);
    const bool team = v == kJkTeam || v == kJkTeamHrr;
    if (fp32 && team && lane) v = kJkReg;
    if (fp32 && v == kJkHgp) v = kJkRowAcc;
    if (!jk_variant_fits(v, X.l1, X.l2, Y.l1, Y.l2, P.bdev, dd.neff, qs.list != nullptr))
      v = lane ? kJkReg : (jk_team_compiled(X.l1, X.l2, Y.l1, Y.l2) ? kJkTeam : kJkTeamHrr);
    if (fp32 && v != kJkTeam && v != kJkTeamHrr)
      return launch_jk_lane_f32(X.l1, X.l2, Y.l1, Y.l2, P.bdev, P.pdev, R.view, qs, dd, 0.0, v,
                                ctr, s);
    return launch_jk(X.l1, X.l2, Y.l1, Y.l2, P.bdev, P.pdev, R.view, qs, dd, 0.0, v, ctr, s);
  }
  return launch_jk_big(X.l1, X.l2, Y.l1, Y.l2, P.bdev, P.pdev, R.view, qs, dd, 0.0,
                       P.s_big.as<double>(), ctr, s);
}

hipError_t grad_launch(Plan &P, const ClassRange &X, const ClassRange &Y, const QuartetSet &qs,
                       const GradDev &gd, unsigned long long *ctr, hipStream_t s) {
  RysDevice &R = rys_device();
  if (class_supported(X.l1, X.l2, Y.l1, Y.l2))
    return launch_grad_lane(X.l1, X.l2, Y.l1, Y.l2, P.bdev, P.pdev, R.view, qs, gd, P.natom,
                            ctr, s);
  return launch_grad_big(X.l1, X.l2, Y.l1, Y.l2, P.bdev, P.pdev, R.view, qs, gd,
                         P.s_big.as<double>(), ctr, s);
}

// Density preparation of a J/K call: symmetric/antisymmetric split, Cartesian
// densities, shell-block bounds and zeroed Cartesian J'/K' accumulators.
// Buffers of a J/K call (host-side work only: allocations and the effective
// density table); the returned view is what the kernels use.
DensDev reserve_densities(Plan &P, int ndens, bool has_anti, double tau, bool want_k) {
  const size_t n2 = size_t(P.nao) * P.nao, nc2 = size_t(P.nao_cart) * P.nao_cart;
  const int neff = ndens * (has_anti ? 2 : 1);
  P.s_ds.reserve(n2 * ndens * sizeof(double));
  P.s_da.reserve(n2 * ndens * sizeof(double) * (has_anti ? 1 : 0) + 8);
  P.s_dc.reserve(nc2 * neff * sizeof(double));
  P.s_bound.reserve(size_t(P.nsh) * P.nsh * ndens * sizeof(double));
  // per-XCD copies while each stays under 1 GiB (large bases: one copy)
  const int ncopy = P.fixed_scratch ? 1
                    : nc2 * neff * xcd_copies() * sizeof(double) <= (size_t(1) << 30) ? xcd_copies()
                                                                                        : 1;
  P.s_jc.reserve(nc2 * neff * ncopy * sizeof(double));
  if (want_k) P.s_kc.reserve(nc2 * neff * ncopy * sizeof(double));
  if (P.eff_key != neff * 128 + ndens) {  // (source, antisymmetric) per effective density
    std::vector<int> eff(2 * neff);
    for (int e = 0; e < neff; ++e) {
      eff[e] = e % ndens;
      eff[neff + e] = e >= ndens;
    }
    upload_to(P, P.s_eff, eff);
    P.eff_key = neff * 128 + ndens;
  }
  DensDev dd{};
  dd.copy_stride = nc2 * neff;
  dd.ncopy = ncopy;
  dd.ndens = ndens;
  dd.neff = neff;
  dd.dc = P.s_dc.as<double>();
  dd.jc = P.s_jc.as<double>();
  dd.kc = want_k ? P.s_kc.as<double>() : nullptr;
  dd.eff_src = P.s_eff.as<int>();
  dd.eff_anti = P.s_eff.as<int>() + neff;
  dd.bound = P.s_bound.as<double>();
  dd.tau = tau;
  dd.want_k = want_k ? 1 : 0;
  return dd;
}

// Device work of the density preparation (kernels only, so it can be
// captured in a graph): split, Cartesian densities, block bounds, zeroed J'/K'.
void launch_densities(Plan &P, const DensDev &dd, const double *d_dens, bool has_anti,
                      hipStream_t s) {
  const size_t nc2 = size_t(P.nao_cart) * P.nao_cart;
  check(launch_density_prep(P.bdev, d_dens, dd.ndens, has_anti, P.s_ds.as<double>(),
                            P.s_da.as<double>(), P.s_dc.as<double>(), P.s_bound.as<double>(), s),
        "density prep");
  check(launch_fill(dd.jc, nc2 * dd.neff * dd.ncopy, 0.0, s), "zero J");
  if (dd.kc) check(launch_fill(dd.kc, nc2 * dd.neff * dd.ncopy, 0.0, s), "zero K");
}

DensDev prep_densities(Plan &P, const double *d_dens, int ndens, bool has_anti, double tau,
                       bool want_k, hipStream_t s) {
  const DensDev dd = reserve_densities(P, ndens, has_anti, tau, want_k);
  launch_densities(P, dd, d_dens, has_anti, s);
  return dd;
}

// Run J/K for ndens densities already on the device (d_dens, function space).
void run_jk(Plan &P, const double *d_dens, int ndens, bool has_anti, double tau, bool want_k,
            double *d_J, double *d_K, const ExplicitSets *ex, hipStream_t s) {
  begin_launches(P, s);
  const DensDev dd = prep_densities(P, d_dens, ndens, has_anti, tau, want_k, s);
  if (!ex) {
    build_rowptrs(P, tau, P.fp32_below);
    size_t big = 0;
    for (auto &cp : P.cached_cps) big = std::max(big, big_need_jk(P.cls[cp.x], P.cls[cp.y], cp.total));
    P.s_big.reserve(big * sizeof(double));
    // Class launches spread over kJkStreams streams, each launch to the
    // stream with the least estimated work so far, heavy classes first (the
    // long launches start early; the last ones fill the others' tails).
    // The runtime-l launches share one set of block slots (s_big) and all go
    // to the call's stream; per-class timing (profiling) and the fixed
    // scratch of the no-allocation entry keep everything on it.
    std::vector<hipStream_t> lanes{s};
    if (!P.fixed_scratch && !P.profiling) {
      constexpr int kJkStreams = 4;
      while (int(P.aux.size()) < kJkStreams - 1) {
        hipStream_t a = nullptr;
        check(hipStreamCreateWithFlags(&a, hipStreamNonBlocking), "stream");
        P.aux.push_back(a);
        hipEvent_t e = nullptr;
        check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        P.join_ev.push_back(e);
      }
      if (!P.fork_ev) check(hipEventCreateWithFlags(&P.fork_ev, hipEventDisableTiming), "hipEventCreate");
      check(hipEventRecord(P.fork_ev, s), "fork");
      for (hipStream_t a : P.aux) {
        check(hipStreamWaitEvent(a, P.fork_ev, 0), "fork wait");
        lanes.push_back(a);
      }
    }
    std::vector<double> load(lanes.size(), 0.0);
    for (auto it = P.cached_cps.rbegin(); it != P.cached_cps.rend(); ++it) {
      if (!it->total) continue;
      QuartetSet qs = implicit_set(P, *it);
      const ClassRange &X = P.cls[it->x], &Y = P.cls[it->y];
      const double n4 = double(ncart(X.l1) * ncart(X.l2) * ncart(Y.l1) * ncart(Y.l2));
      const double cost = n4 * (it->prims * ((X.l1 + X.l2 + Y.l1 + Y.l2) / 2 + 1) +
                                8.0 * double(it->total) * dd.neff);
      size_t k = 0;
      if (jk_compiled(X.l1, X.l2, Y.l1, Y.l2))
        for (size_t q = 1; q < lanes.size(); ++q)
          if (load[q] < load[k]) k = q;
      load[k] += cost;
      unsigned long long *ctr = next_launch(P, X, Y, lanes[k], it->fp32);
      check(jk_launch(P, X, Y, qs, dd, ctr, lanes[k], -1, it->fp32 != 0), "jk launch");
      end_launch(P, lanes[k]);
    }
    for (size_t q = 1; q < lanes.size(); ++q) {
      check(hipEventRecord(P.join_ev[q - 1], lanes[q]), "join");
      check(hipStreamWaitEvent(s, P.join_ev[q - 1], 0), "join wait");
    }
  } else {
    size_t total = 0;
    for (auto &l : ex->lists) total += l.size();
    std::vector<int2> flat;
    flat.reserve(total);
    for (auto &l : ex->lists) flat.insert(flat.end(), l.begin(), l.end());
    P.s_list.upload(flat);
    size_t big = 0;
    for (size_t c = 0; c < ex->cls.size(); ++c)
      big = std::max(big, big_need_jk(P.cls[ex->cls[c].first], P.cls[ex->cls[c].second],
                                      (long long)ex->lists[c].size()));
    P.s_big.reserve(big * sizeof(double));
    size_t off = 0;
    for (size_t c = 0; c < ex->cls.size(); ++c) {
      const ClassRange &X = P.cls[ex->cls[c].first], &Y = P.cls[ex->cls[c].second];
      QuartetSet qs{};
      qs.x_off = X.off;
      qs.x_n = X.n;
      qs.y_off = Y.off;
      qs.y_n = Y.n;
      qs.same = ex->cls[c].first == ex->cls[c].second;
      qs.list = P.s_list.as<int2>() + off;
      qs.total = (long long)ex->lists[c].size();
      off += ex->lists[c].size();
      unsigned long long *ctr = next_launch(P, X, Y, s);
      check(jk_launch(P, X, Y, qs, dd, ctr, s), "jk launch");
      end_launch(P, s);
    }
  }
  check(launch_fold_copies(dd, size_t(P.nao_cart) * P.nao_cart, s), "fold copies");
  check(launch_jk_finish(P.bdev, dd, d_J, d_K, s), "jk finish");
  P.stats_stream = s;
  P.stats_pending = true;
}

bool host_symmetric(const double *d, int n) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      if (d[size_t(i) * n + j] != d[size_t(j) * n + i]) return false;
  return true;
}

}  // namespace

extern "C" int ig_direct_create(const ig_basis *basis, int engine, ig_direct **out) {
  if (!basis || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (int rc = check_engine(engine)) return rc;
  return guarded([&] {
    auto d = std::make_unique<ig_direct>();
    d->engine = engine;
    d->plan = build_plan(*basis, nullptr);
    *out = d.release();
    return int(IG_OK);
  });
}

extern "C" void ig_direct_destroy(ig_direct *d) { delete d; }
extern "C" size_t ig_direct_nao(const ig_direct *d) { return d ? size_t(d->plan->nao) : 0; }
extern "C" size_t ig_direct_n_bra_pairs(const ig_direct *d) {
  return d ? size_t(d->plan->npairs) : 0;
}
extern "C" int ig_direct_bra_pairs(const ig_direct *d, int32_t *pairs, size_t n) {
  if (!d || (!pairs && n)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (n != size_t(d->plan->npairs)) return fail(IG_ERR_LENGTH_MISMATCH, "pairs length");
  for (size_t c = 0; c < n; ++c) {
    int i, j;
    pair_of(int64_t(c), i, j);
    pairs[2 * c] = i;
    pairs[2 * c + 1] = j;
  }
  return IG_OK;
}

extern "C" int ig_direct_density_shell_bounds(const ig_direct *d, const double *const *dens,
                                              size_t ndens, double *bounds, size_t len) {
  if (!d || !bounds || (ndens && !dens)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  const Plan &P = *d->plan;
  if (len != ndens * size_t(P.nsh) * P.nsh) return fail(IG_ERR_LENGTH_MISMATCH, "bounds length");
  // direct.rs:103-125 (ordered blocks, host computation)
  for (size_t k = 0; k < ndens; ++k) {
    if (!dens[k]) return fail(IG_ERR_INVALID_ARGUMENT, "null density");
    for (int i = 0; i < P.nsh; ++i)
      for (int j = 0; j < P.nsh; ++j) {
        int ni = P.sh_kind[i] ? 2 * P.sh_l[i] + 1 : ncart(P.sh_l[i]);
        int nj = P.sh_kind[j] ? 2 * P.sh_l[j] + 1 : ncart(P.sh_l[j]);
        double m = 0.0;
        for (int a = 0; a < ni; ++a)
          for (int c = 0; c < nj; ++c)
            m = std::max(m, std::fabs(dens[k][size_t(P.sh_off_func[i] + a) * P.nao +
                                               P.sh_off_func[j] + c]));
        bounds[(k * P.nsh + i) * P.nsh + j] = m;
      }
  }
  return IG_OK;
}

static int jk_host(Plan &P, const double *const *dens, size_t ndens, double tau, int want_k,
                   double *const *J, double *const *K, const ExplicitSets *ex, void *stream) {
  std::lock_guard<std::mutex> g(P.mu);
  const size_t n2 = size_t(P.nao) * P.nao;
  hipStream_t s = as_stream(stream);
  plan_enter(P, s);
  std::vector<double> host(n2 * ndens);
  bool sym = true;
  for (size_t k = 0; k < ndens; ++k) {
    std::memcpy(host.data() + k * n2, dens[k], n2 * sizeof(double));
    sym = sym && host_symmetric(dens[k], P.nao);
  }
  P.s_dfunc.upload(host);
  P.s_J.reserve(n2 * ndens * sizeof(double));
  P.s_K.reserve(n2 * ndens * sizeof(double));
  check(launch_fill(P.s_J.as<double>(), n2 * ndens, 0.0, s), "zero J");
  if (want_k) check(launch_fill(P.s_K.as<double>(), n2 * ndens, 0.0, s), "zero K");
  run_jk(P, P.s_dfunc.as<double>(), int(ndens), !sym, tau, want_k != 0, P.s_J.as<double>(),
         want_k ? P.s_K.as<double>() : nullptr, ex, s);
  std::vector<double> hj(n2 * ndens), hk(want_k ? n2 * ndens : 0);
  check(hipMemcpyAsync(hj.data(), P.s_J.p, hj.size() * sizeof(double), hipMemcpyDeviceToHost, s),
        "copy J");
  if (want_k)
    check(hipMemcpyAsync(hk.data(), P.s_K.p, hk.size() * sizeof(double), hipMemcpyDeviceToHost, s),
          "copy K");
  plan_leave(P, s);
  finalize_stats(P);  // synchronises the stream
  for (size_t k = 0; k < ndens; ++k) {
    for (size_t x = 0; x < n2; ++x) J[k][x] += hj[k * n2 + x];
    if (want_k)
      for (size_t x = 0; x < n2; ++x) K[k][x] += hk[k * n2 + x];
  }
  return int(IG_OK);
}

static int check_dens_args(const ig_direct *d, const double *const *dens, size_t ndens,
                           int want_k, double *const *J, double *const *K) {
  if (!d) return fail(IG_ERR_INVALID_ARGUMENT, "null contractor");
  if (ndens == 0) return fail(IG_ERR_NO_DENSITY, "direct contraction needs at least one density");
  if (ndens > size_t(kMaxDens))
    return fail(IG_ERR_TOO_MANY_DENSITIES, "at most 64 simultaneous densities are supported");
  if (!dens || !J || (want_k && !K)) return fail(IG_ERR_INVALID_ARGUMENT, "null buffer list");
  for (size_t k = 0; k < ndens; ++k)
    if (!dens[k] || !J[k] || (want_k && !K[k]))
      return fail(IG_ERR_INVALID_ARGUMENT, "null buffer");
  return IG_OK;
}

extern "C" int ig_direct_accumulate_bra_pair(const ig_direct *d, int32_t i, int32_t j,
                                             const double *const *dens, const double *bounds,
                                             size_t ndens, double threshold, int want_k,
                                             double *const *J, double *const *K, void *stream) {
  if (int rc = check_dens_args(d, dens, ndens, want_k, J, K)) return rc;
  const Plan &P = *d->plan;
  if (!(j >= 0 && j <= i && i < P.nsh))
    return fail(IG_ERR_NON_CANONICAL_BRA, "bra pair is not canonical (i >= j)");
  (void)bounds;  // the device recomputes the same bounds from the densities
  return guarded([&] {
    int64_t bc = int64_t(i) * (i + 1) / 2 + j;
    ExplicitSets ex = explicit_sets(P, bc, bc + 1, threshold);
    return jk_host(*d->plan, dens, ndens, threshold, want_k, J, K, &ex, stream);
  });
}

extern "C" int ig_direct_jk(const ig_direct *d, const double *const *dens, size_t ndens,
                            double threshold, int want_k, double *const *J, double *const *K,
                            int64_t b0, int64_t b1, void *stream) {
  if (int rc = check_dens_args(d, dens, ndens, want_k, J, K)) return rc;
  const Plan &P = *d->plan;
  const int64_t np = P.npairs;
  if (b1 < 0 || b1 > np) b1 = np;
  if (b0 < 0 || b0 > b1) return fail(IG_ERR_INVALID_ARGUMENT, "bad bra range");
  return guarded([&] {
    if (b0 == 0 && b1 == np)
      return jk_host(*d->plan, dens, ndens, threshold, want_k, J, K, nullptr, stream);
    ExplicitSets ex = explicit_sets(P, b0, b1, threshold);
    return jk_host(*d->plan, dens, ndens, threshold, want_k, J, K, &ex, stream);
  });
}

extern "C" int ig_direct_jk_device(const ig_direct *d, const double *d_dens, size_t ndens,
                                   double threshold, int want_k, double *d_J, double *d_K,
                                   void *stream) {
  if (!d || !d_dens || !d_J || (want_k && !d_K)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (ndens == 0) return fail(IG_ERR_NO_DENSITY, "direct contraction needs at least one density");
  if (ndens > size_t(kMaxDens))
    return fail(IG_ERR_TOO_MANY_DENSITIES, "at most 64 simultaneous densities are supported");
  return guarded([&] {
    Plan &P = *d->plan;
    std::lock_guard<std::mutex> g(P.mu);
    hipStream_t s = as_stream(stream);
    plan_enter(P, s);
    // exact symmetry test on the device; only a 4-byte flag comes back
    P.s_flag.reserve(sizeof(int));
    check(launch_asym_flag(d_dens, P.nao, int(ndens), P.s_flag.as<int>(), s), "symmetry check");
    int asym = 0;
    check(hipMemcpyAsync(&asym, P.s_flag.p, sizeof(int), hipMemcpyDeviceToHost, s), "copy flag");
    check(hipStreamSynchronize(s), "sync");
    run_jk(P, d_dens, int(ndens), asym != 0, threshold, want_k != 0, d_J, want_k ? d_K : nullptr,
           nullptr, s);
    plan_leave(P, s);
    return int(IG_OK);
  });
}

namespace {

// A caller's workspace plan: the source plan's read-only data (shells,
// pairs, Schwarz bounds and views of its device buffers, which it does not
// own) and scratch of its own.
std::unique_ptr<Plan> workspace_plan(const Plan &src) {
  auto W = std::make_unique<Plan>();
  W->basis = src.basis;
  W->nsh = src.nsh;
  W->nao = src.nao;
  W->nao_cart = src.nao_cart;
  W->natom = src.natom;
  W->sh_l = src.sh_l;
  W->sh_kind = src.sh_kind;
  W->sh_atom = src.sh_atom;
  W->sh_off_func = src.sh_off_func;
  W->sh_off_cart = src.sh_off_cart;
  W->sh_nprim = src.sh_nprim;
  W->bdev = src.bdev;
  W->npairs = src.npairs;
  W->p_sh1 = src.p_sh1;
  W->p_sh2 = src.p_sh2;
  W->p_canon = src.p_canon;
  W->p_ppoff = src.p_ppoff;
  W->p_ppn = src.p_ppn;
  W->p_q = src.p_q;
  W->canon_to_pos = src.canon_to_pos;
  std::copy(std::begin(src.cls), std::end(src.cls), std::begin(W->cls));
  W->pdev = src.pdev;
  W->schwarz = src.schwarz;
  return W;
}

// The schedule and precision policy of the source plan, copied into a
// workspace plan at the start of each of its calls.
void follow_settings(Plan &W, Plan &src) {
  std::lock_guard<std::mutex> g(src.settings_mu);
  W.variant = src.variant;
  W.variant_ms = src.variant_ms;
  W.fp32_below = src.fp32_below;
}

}  // namespace

extern "C" int ig_direct_workspace_create(const ig_direct *d, ig_direct_workspace **out) {
  if (!d || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  return guarded([&] {
    auto w = std::make_unique<ig_direct_workspace>();
    w->owner = d;
    w->plan = workspace_plan(*d->plan);
    *out = w.release();
    return int(IG_OK);
  });
}

extern "C" void ig_direct_workspace_destroy(ig_direct_workspace *w) {
  if (!w) return;
  if (w->plan->idle_recorded) (void)hipEventSynchronize(w->plan->idle);
  delete w;
}

static int check_workspace(const ig_direct *d, const ig_direct_workspace *w) {
  if (w && w->owner != d)
    return fail(IG_ERR_INVALID_ARGUMENT, "workspace was created for another contractor");
  return IG_OK;
}

// the plan a call runs on: the workspace's (settings copied from the
// contractor) or the contractor's own
static Plan &call_plan(const ig_direct *d, ig_direct_workspace *w) {
  if (!w) return *d->plan;
  follow_settings(*w->plan, *d->plan);
  return *w->plan;
}

extern "C" int ig_direct_accumulate_bra_pair_ws(const ig_direct *d, ig_direct_workspace *w,
                                                int32_t i, int32_t j, const double *const *dens,
                                                const double *bounds, size_t ndens,
                                                double threshold, int want_k, double *const *J,
                                                double *const *K, void *stream) {
  if (int rc = check_dens_args(d, dens, ndens, want_k, J, K)) return rc;
  if (int rc = check_workspace(d, w)) return rc;
  const Plan &P = *d->plan;
  if (!(j >= 0 && j <= i && i < P.nsh))
    return fail(IG_ERR_NON_CANONICAL_BRA, "bra pair is not canonical (i >= j)");
  (void)bounds;  // the device recomputes the same bounds from the densities
  return guarded([&] {
    int64_t bc = int64_t(i) * (i + 1) / 2 + j;
    ExplicitSets ex = explicit_sets(P, bc, bc + 1, threshold);
    return jk_host(call_plan(d, w), dens, ndens, threshold, want_k, J, K, &ex, stream);
  });
}

extern "C" int ig_direct_jk_ws(const ig_direct *d, ig_direct_workspace *w,
                               const double *const *dens, size_t ndens, double threshold,
                               int want_k, double *const *J, double *const *K, int64_t b0,
                               int64_t b1, void *stream) {
  if (int rc = check_dens_args(d, dens, ndens, want_k, J, K)) return rc;
  if (int rc = check_workspace(d, w)) return rc;
  const Plan &P = *d->plan;
  const int64_t np = P.npairs;
  if (b1 < 0 || b1 > np) b1 = np;
  if (b0 < 0 || b0 > b1) return fail(IG_ERR_INVALID_ARGUMENT, "bad bra range");
  return guarded([&] {
    Plan &W = call_plan(d, w);
    if (b0 == 0 && b1 == np) return jk_host(W, dens, ndens, threshold, want_k, J, K, nullptr, stream);
    ExplicitSets ex = explicit_sets(P, b0, b1, threshold);
    return jk_host(W, dens, ndens, threshold, want_k, J, K, &ex, stream);
  });
}

static int check_device_args(const ig_direct *d, const double *d_dens, size_t ndens, int want_k,
                             const double *d_J, const double *d_K) {
  if (!d || !d_dens || !d_J || (want_k && !d_K)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (ndens == 0) return fail(IG_ERR_NO_DENSITY, "direct contraction needs at least one density");
  if (ndens > size_t(kMaxDens))
    return fail(IG_ERR_TOO_MANY_DENSITIES, "at most 64 simultaneous densities are supported");
  return IG_OK;
}

extern "C" int ig_direct_jk_device_ws(const ig_direct *d, ig_direct_workspace *w,
                                      const double *d_dens, size_t ndens, double threshold,
                                      int want_k, double *d_J, double *d_K, int dens_symmetric,
                                      void *stream) {
  if (int rc = check_device_args(d, d_dens, ndens, want_k, d_J, d_K)) return rc;
  if (int rc = check_workspace(d, w)) return rc;
  return guarded([&] {
    Plan &P = call_plan(d, w);
    std::lock_guard<std::mutex> g(P.mu);
    hipStream_t s = as_stream(stream);
    plan_enter(P, s);
    run_jk(P, d_dens, int(ndens), dens_symmetric == 0, threshold, want_k != 0, d_J,
           want_k ? d_K : nullptr, nullptr, s);
    plan_leave(P, s);
    return int(IG_OK);
  });
}

extern "C" int ig_direct_workspace_last_stats(ig_direct_workspace *w, int64_t *st) {
  if (!w || !st) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  return guarded([&] {
    std::lock_guard<std::mutex> g(w->plan->mu);
    finalize_stats(*w->plan);
    st[0] = w->plan->last_stats[0];
    st[1] = w->plan->last_stats[1];
    return int(IG_OK);
  });
}

// ---------------------------------------------------------------------------
// J/K with no device allocation: the scratch of the call (densities, bounds,
// J'/K' accumulators, quartet-set rows, counters, runtime-l block slots) is
// carved from the fixed device array of ig_noalloc.hip.
extern "C" int ig_direct_jk_device_noalloc(const ig_direct *d, const double *d_dens,
                                           size_t ndens, double threshold, int want_k,
                                           double *d_J, double *d_K, int dens_symmetric,
                                           void *stream) {
  if (int rc = check_device_args(d, d_dens, ndens, want_k, d_J, d_K)) return rc;
  return guarded([&] {
    void *base = nullptr;
    size_t cap = 0;
    check(noalloc_scratch(&base, &cap), "scratch address");
    ig_direct *dm = const_cast<ig_direct *>(d);
    std::lock_guard<std::mutex> fg(dm->fixed_mu);
    if (!dm->fixed) {
      dm->fixed = workspace_plan(*d->plan);
      dm->fixed->fixed_scratch = true;
    }
    Plan &P = *dm->fixed;
    follow_settings(P, *d->plan);
    hipStream_t s = as_stream(stream);
    P.cur_stream = s;
    const bool has_anti = dens_symmetric == 0;
    const int nd = int(ndens), neff = nd * (has_anti ? 2 : 1);
    const size_t n2 = size_t(P.nao) * P.nao, nc2 = size_t(P.nao_cart) * P.nao_cart;
    // other plans' calls reuse the array: every upload is redone
    P.eff_key = -1;
    const double cut = P.fp32_below > threshold ? P.fp32_below : threshold;
    rowptrs_host(P, threshold, cut);
    size_t big = 0;
    for (auto &cp : P.cached_cps) big = std::max(big, big_need_jk(P.cls[cp.x], P.cls[cp.y], cp.total));
    struct Part {
      DBuf *buf;
      size_t bytes;
    };
    const Part parts[] = {
        {&P.s_ds, n2 * nd * sizeof(double)},
        {&P.s_da, (has_anti ? n2 * nd * sizeof(double) : 0) + 8},
        {&P.s_dc, nc2 * neff * sizeof(double)},
        {&P.s_bound, size_t(P.nsh) * P.nsh * nd * sizeof(double)},
        {&P.s_jc, nc2 * neff * sizeof(double)},
        {&P.s_kc, want_k ? nc2 * neff * sizeof(double) : 8},
        {&P.s_eff, 2 * size_t(neff) * sizeof(int)},
        {&P.s_rowptr, P.cached_rowptr_host.size() * sizeof(long long)},
        {&P.s_qbase, std::max<size_t>(1, P.cached_qbase_host.size()) * sizeof(int)},
        {&P.s_counters, 2 * Plan::kMaxLaunch * sizeof(unsigned long long)},
        {&P.s_big, big * sizeof(double)},
    };
    size_t need = 0;
    for (const Part &q : parts) need += (q.bytes + 255) / 256 * 256;
    if (need > cap)
      return fail(IG_ERR_UNSUPPORTED,
                  "this call needs " + std::to_string(need) +
                      " bytes of scratch; the no-allocation entry holds " + std::to_string(cap));
    size_t off = 0;
    for (const Part &q : parts) {
      q.buf->view(static_cast<char *>(base) + off, (q.bytes + 255) / 256 * 256);
      off += (q.bytes + 255) / 256 * 256;
    }
    upload_to(P, P.s_rowptr, P.cached_rowptr_host);
    upload_to(P, P.s_qbase, P.cached_qbase_host);
    P.cached_tau = threshold;
    P.cached_cut = cut;
    run_jk(P, d_dens, nd, has_anti, threshold, want_k != 0, d_J, want_k ? d_K : nullptr,
           nullptr, s);
    P.stats_pending = false;  // no counters are read back on this path
    return int(IG_OK);
  });
}

namespace {

// Time every J/K kernel variant of every compiled class pair of plan P on
// the device densities and keep the fastest (caller holds P.mu).
void autotune_plan(Plan &P, const double *d_dens, size_t ndens, double threshold, int want_k,
                   int reps, hipStream_t s) {
  finalize_stats(P);
  P.s_flag.reserve(sizeof(int));
  check(launch_asym_flag(d_dens, P.nao, int(ndens), P.s_flag.as<int>(), s), "symmetry check");
  int asym = 0;
  check(hipMemcpyAsync(&asym, P.s_flag.p, sizeof(int), hipMemcpyDeviceToHost, s), "copy flag");
  check(hipStreamSynchronize(s), "sync");
  // the accumulators of the trial launches are scratch: the caller's J/K
  // are not touched
  const DensDev dd = prep_densities(P, d_dens, int(ndens), asym != 0, threshold, want_k != 0, s);
  build_rowptrs(P, threshold);
  P.s_counters.reserve(2 * sizeof(unsigned long long));
  auto *ctr = P.s_counters.as<unsigned long long>();
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "hipEventCreate");
  check(hipEventCreate(&e1), "hipEventCreate");
  try {
    for (auto &cp : P.cached_cps) {
      const ClassRange &X = P.cls[cp.x], &Y = P.cls[cp.y];
      if (!cp.total || !jk_compiled(X.l1, X.l2, Y.l1, Y.l2)) continue;
      const QuartetSet qs = implicit_set(P, cp);
      int best = -1;
      float best_ms = 0.0f;
      for (int v = 0; v < kJkNumVariants; ++v) {
        if (!jk_variant_fits(v, X.l1, X.l2, Y.l1, Y.l2, P.bdev, dd.neff, false)) continue;
        float vmin = -1.0f;
        for (int r = 0; r < reps; ++r) {
          check(hipEventRecord(e0, s), "event");
          check(jk_launch(P, X, Y, qs, dd, ctr, s, v), "jk launch");
          check(hipEventRecord(e1, s), "event");
          check(hipEventSynchronize(e1), "event sync");
          float ms = 0.0f;
          check(hipEventElapsedTime(&ms, e0, e1), "event time");
          if (vmin < 0.0f || ms < vmin) vmin = ms;
        }
        if (best < 0 || vmin < best_ms) {
          best = v;
          best_ms = vmin;
        }
      }
      const int key = cp.x * kNumPairClass + cp.y;
      std::lock_guard<std::mutex> sg(P.settings_mu);
      P.variant[key] = int8_t(best);
      P.variant_ms[key] = best_ms;
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  P.launches.clear();
  P.stats_pending = false;
}

}  // namespace

extern "C" int ig_direct_autotune(ig_direct *d, const double *d_dens, size_t ndens,
                                  double threshold, int want_k, int reps, void *stream) {
  if (!d || !d_dens) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (ndens == 0) return fail(IG_ERR_NO_DENSITY, "direct contraction needs at least one density");
  if (ndens > size_t(kMaxDens))
    return fail(IG_ERR_TOO_MANY_DENSITIES, "at most 64 simultaneous densities are supported");
  if (reps < 1) return fail(IG_ERR_INVALID_ARGUMENT, "reps must be at least 1");
  return guarded([&] {
    Plan &P = *d->plan;
    std::lock_guard<std::mutex> g(P.mu);
    hipStream_t s = as_stream(stream);
    plan_enter(P, s);
    autotune_plan(P, d_dens, ndens, threshold, want_k, reps, s);
    plan_leave(P, s);
    return int(IG_OK);
  });
}

extern "C" int ig_direct_schedule_read(const ig_direct *d, ig_schedule_entry *out, size_t cap,
                                       size_t *n) {
  if (!d || !n || (cap && !out)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  return guarded([&] {
    Plan &P = *d->plan;
    std::lock_guard<std::mutex> g(P.mu);
    size_t k = 0;
    for (int x = 0; x < kNumPairClass; ++x)
      for (int y = 0; y <= x; ++y) {
        const ClassRange &X = P.cls[x], &Y = P.cls[y];
        if (!X.n || !Y.n || !jk_compiled(X.l1, X.l2, Y.l1, Y.l2)) continue;
        if (k < cap) {
          const int key = x * kNumPairClass + y;
          out[k].la = X.l1;
          out[k].lb = X.l2;
          out[k].lc = Y.l1;
          out[k].ld = Y.l2;
          out[k].variant = jk_variant(P, X, Y);
          out[k].tuned = P.variant[key] >= 0 ? 1 : 0;
          out[k].ms = P.variant_ms[key];
        }
        ++k;
      }
    *n = k;
    return int(IG_OK);
  });
}

namespace {

// validation of a schedule to load (IG_OK or the failure status)
int check_schedule(const ig_schedule_entry *in, size_t n) {
  for (size_t k = 0; k < n; ++k) {
    const ig_schedule_entry &e = in[k];
    const bool ok_l = e.la >= 0 && e.lb >= 0 && e.lc >= 0 && e.ld >= 0 && e.la <= kMaxL &&
                      e.lc <= kMaxL && e.lb <= e.la && e.ld <= e.lc;
    if (!ok_l || !jk_compiled(e.la, e.lb, e.lc, e.ld))
      return fail(IG_ERR_INVALID_ARGUMENT,
                  "schedule entry " + std::to_string(k) + " names a class without compiled kernels");
    if (e.variant < -1 || e.variant >= kJkNumVariants)
      return fail(IG_ERR_INVALID_ARGUMENT,
                  "schedule entry " + std::to_string(k) + " has an unknown variant");
  }
  return IG_OK;
}

void apply_schedule(Plan &P, const ig_schedule_entry *in, size_t n) {
  std::lock_guard<std::mutex> sg(P.settings_mu);
  for (size_t k = 0; k < n; ++k) {
    const int key = pair_class(in[k].la, in[k].lb) * kNumPairClass + pair_class(in[k].lc, in[k].ld);
    P.variant[key] = int8_t(in[k].variant);
    P.variant_ms[key] = float(in[k].variant >= 0 ? in[k].ms : 0.0);
  }
}

}  // namespace

extern "C" int ig_direct_schedule_write(ig_direct *d, const ig_schedule_entry *in, size_t n) {
  if (!d || (n && !in)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_schedule(in, n)) return rc;
  return guarded([&] {
    Plan &P = *d->plan;
    std::lock_guard<std::mutex> g(P.mu);
    apply_schedule(P, in, n);
    return int(IG_OK);
  });
}

extern "C" int ig_direct_set_mixed_precision(ig_direct *d, double fp32_below) {
  if (!d) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (!(std::isfinite(fp32_below) && fp32_below >= 0.0))
    return fail(IG_ERR_INVALID_ARGUMENT, "fp32_below must be finite and >= 0");
  std::lock_guard<std::mutex> g(d->plan->mu);
  std::lock_guard<std::mutex> sg(d->plan->settings_mu);
  d->plan->fp32_below = fp32_below;
  return IG_OK;
}

extern "C" int ig_direct_set_profiling(ig_direct *d, int enable) {
  if (!d) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> g(d->plan->mu);
  d->plan->profiling = enable != 0;
  return IG_OK;
}

extern "C" int ig_direct_profile_read(const ig_direct *d, ig_class_profile *out, size_t cap,
                                      size_t *n) {
  if (!d || !n || (cap && !out)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  return guarded([&] {
    Plan &P = *d->plan;
    std::lock_guard<std::mutex> g(P.mu);
    finalize_stats(P);
    *n = P.launches.size();
    for (size_t k = 0; k < P.launches.size() && k < cap; ++k) {
      out[k].fp32 = P.launches[k].fp32;
      out[k].la = P.launches[k].la;
      out[k].lb = P.launches[k].lb;
      out[k].lc = P.launches[k].lc;
      out[k].ld = P.launches[k].ld;
      out[k].quartets = int64_t(P.launch_ctr[2 * k]);
      out[k].prim_quartets = int64_t(P.launch_ctr[2 * k + 1]);
      out[k].ms = P.profiling && k < P.launch_ms.size() ? double(P.launch_ms[k]) : 0.0;
    }
    return int(IG_OK);
  });
}

extern "C" int ig_direct_last_stats(const ig_direct *d, int64_t *st) {
  if (!d || !st) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = guarded([&] {
        std::lock_guard<std::mutex> g(d->plan->mu);
        finalize_stats(*d->plan);
        return int(IG_OK);
      }))
    return rc;
  st[0] = d->plan->last_stats[0];
  st[1] = d->plan->last_stats[1];
  return IG_OK;
}

// ---------------------------------------------------------------------------
// dense tensor and Schwarz

extern "C" int ig_schwarz_bounds(const ig_basis *b, int engine, double *out, size_t len,
                                 void *stream) {
  if (!b || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  if (len != b->shells.size() * b->shells.size())
    return fail(IG_ERR_LENGTH_MISMATCH, "schwarz length");
  return guarded([&] {
    auto P = build_plan(*b, as_stream(stream));
    std::copy(P->schwarz.begin(), P->schwarz.end(), out);
    return int(IG_OK);
  });
}

static int eri_impl(const ig_basis *b, double tau, double *out, size_t len, int64_t *stats,
                    void *stream, double omega = 0.0) {
  const size_t nao = size_t(b->nao);
  if (len != nao * nao * nao * nao) return fail(IG_ERR_LENGTH_MISMATCH, "eri length");
  return guarded([&] {
    hipStream_t s = as_stream(stream);
    auto P = build_plan(*b, s);
    DBuf dout;
    dout.reserve(len * sizeof(double));
    check(launch_fill(dout.as<double>(), len, 0.0, s), "zero eri");
    build_rowptrs(*P, tau < 0 ? 0.0 : tau);
    RysDevice &R = rys_device();
    size_t big = 0;
    for (auto &cp : P->cached_cps)
      big = std::max(big, big_need(P->cls[cp.x], P->cls[cp.y], cp.total, 1));
    P->s_big.reserve(big * sizeof(double));
    EriOut eo{dout.as<double>(), nullptr, 0};
    eo.omega = omega;
    for (auto &cp : P->cached_cps) {
      if (!cp.total) continue;
      QuartetSet qs = implicit_set(*P, cp);
      const ClassRange &X = P->cls[cp.x], &Y = P->cls[cp.y];
      if (class_supported(X.l1, X.l2, Y.l1, Y.l2))
        check(launch_eri(X.l1, X.l2, Y.l1, Y.l2, P->bdev, P->pdev, R.view, qs, eo, s),
              "eri launch");
      else
        check(launch_eri_big(X.l1, X.l2, Y.l1, Y.l2, P->bdev, P->pdev, R.view, qs, eo,
                             P->s_big.as<double>(), s),
              "eri launch");
    }
    check(hipMemcpyAsync(out, dout.p, len * sizeof(double), hipMemcpyDeviceToHost, s), "copy eri");
    check(hipStreamSynchronize(s), "eri sync");
    if (stats) {  // integrals.rs:985-1005: canonical quartets and Schwarz skips
      int64_t np = P->npairs, total = np * (np + 1) / 2, kept = 0;
      std::vector<double> q = P->p_q;
      std::sort(q.begin(), q.end(), std::greater<double>());
      for (int64_t r = 0; r < np; ++r) {
        int c = count_kets(q.data(), int(np), q[r], tau);
        kept += std::min<int64_t>(c, r + 1);
      }
      stats[0] = total;
      stats[1] = total - kept;
    }
    return int(IG_OK);
  });
}

extern "C" int ig_eri_dense(const ig_basis *b, int engine, double *out, size_t len, void *stream) {
  if (!b || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  return eri_impl(b, 0.0, out, len, nullptr, stream);
}

extern "C" int ig_eri_screened(const ig_basis *b, int engine, double tau, double *out, size_t len,
                               int64_t *stats, void *stream) {
  if (!b || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  return eri_impl(b, tau, out, len, stats, stream);
}

// ---------------------------------------------------------------------------
// packed symmetry-unique ERIs: PackedEriBuilder (packed_eri.rs:67-240)
namespace {

int shell_nfunc(const HostShell &s) { return s.nfunc(); }
size_t tri_count(size_t n) { return n * (n + 1) / 2; }

// First AO-pair index of every canonical shell pair (i >= j, row-major), then
// the total (packed_eri.rs:99-110): inside one pair the function pairs (a, b)
// row-major, b <= a on a diagonal pair.
std::vector<long long> packed_pair_start(const ig_basis &b) {
  const int nsh = int(b.shells.size());
  std::vector<long long> start;
  start.reserve(size_t(nsh) * (nsh + 1) / 2 + 1);
  long long acc = 0;
  for (int i = 0; i < nsh; ++i)
    for (int j = 0; j <= i; ++j) {
      start.push_back(acc);
      const long long ni = shell_nfunc(b.shells[i]), nj = shell_nfunc(b.shells[j]);
      acc += i == j ? ni * (ni + 1) / 2 : ni * nj;
    }
  start.push_back(acc);
  return start;
}

// Rows of canonical bra pairs [b0, b1) of the packed triangle into `rows`
// (len = its length), quartets with Q_ij Q_kl < tau skipped (tau > 0).
int packed_impl(const ig_basis *b, double tau, int64_t b0, int64_t b1, double *rows, size_t len,
                void *stream) {
  const std::vector<long long> start = packed_pair_start(*b);
  const int64_t np = int64_t(start.size()) - 1;
  if (b1 < 0) b1 = np;
  if (b0 < 0 || b0 > b1 || b1 > np) return fail(IG_ERR_INVALID_ARGUMENT, "bra pair range");
  const size_t row0 = tri_count(size_t(start[b0])), row1 = tri_count(size_t(start[b1]));
  if (len != row1 - row0) return fail(IG_ERR_LENGTH_MISMATCH, "packed ERI rows length");
  if (!(tau >= 0.0) || std::isinf(tau)) return fail(IG_ERR_INVALID_ARGUMENT, "threshold");
  return guarded([&] {
    hipStream_t s = as_stream(stream);
    auto P = build_plan(*b, s);
    DBuf dstart, dout;
    dstart.upload(start);
    dout.reserve(std::max<size_t>(len, 1) * sizeof(double));
    if (len) check(launch_fill(dout.as<double>(), len, 0.0, s), "zero packed eri");
    const EriOut eo{dout.as<double>(), dstart.as<long long>(), (long long)row0};
    RysDevice &R = rys_device();
    auto launch = [&](const ClassRange &X, const ClassRange &Y, const QuartetSet &qs) {
      if (class_supported(X.l1, X.l2, Y.l1, Y.l2))
        check(launch_eri(X.l1, X.l2, Y.l1, Y.l2, P->bdev, P->pdev, R.view, qs, eo, s),
              "packed eri launch");
      else
        check(launch_eri_big(X.l1, X.l2, Y.l1, Y.l2, P->bdev, P->pdev, R.view, qs, eo,
                             P->s_big.as<double>(), s),
              "packed eri launch");
    };
    {  // explicit lists of the bra range, kets <= bra (packed_eri.rs:183): a
       // whole build and the fills of its parts evaluate every quartet in the
       // same orientation, so they agree bit for bit (packed_eri.rs:222-230)
      ExplicitSets ex = explicit_sets(*P, b0, b1, tau);
      std::vector<int2> flat;
      for (auto &l : ex.lists) flat.insert(flat.end(), l.begin(), l.end());
      P->s_list.upload(flat);
      size_t big = 0;
      for (size_t c = 0; c < ex.cls.size(); ++c)
        big = std::max(big, big_need(P->cls[ex.cls[c].first], P->cls[ex.cls[c].second],
                                     (long long)ex.lists[c].size(), 1));
      P->s_big.reserve(big * sizeof(double));
      size_t off = 0;
      for (size_t c = 0; c < ex.cls.size(); ++c) {
        const ClassRange &X = P->cls[ex.cls[c].first], &Y = P->cls[ex.cls[c].second];
        QuartetSet qs{};
        qs.x_off = X.off;
        qs.x_n = X.n;
        qs.y_off = Y.off;
        qs.y_n = Y.n;
        qs.same = ex.cls[c].first == ex.cls[c].second;
        qs.list = P->s_list.as<int2>() + off;
        qs.total = (long long)ex.lists[c].size();
        off += ex.lists[c].size();
        launch(X, Y, qs);
      }
    }
    if (len)
      check(hipMemcpyAsync(rows, dout.p, len * sizeof(double), hipMemcpyDeviceToHost, s),
            "copy packed eri");
    check(hipStreamSynchronize(s), "packed eri sync");
    return int(IG_OK);
  });
}

}  // namespace

// EriBuilder fill (eri_builder.rs:337-454): the dense rows owned by the
// canonical bra pairs [b0, b1).  Each unordered pair of shell pairs {P, Q}
// with P in the range is evaluated once -- Q <= P, or Q above the range --
// and its block written to every permutation slot whose row pair is owned.
static int eri_builder_impl(const ig_basis *b, int64_t b0, int64_t b1, double *out, size_t len,
                            void *stream, double omega = 0.0) {
  const size_t nao = size_t(b->nao);
  if (len != nao * nao * nao * nao) return fail(IG_ERR_LENGTH_MISMATCH, "eri length");
  const int nsh = int(b->shells.size());
  const int64_t np = int64_t(nsh) * (nsh + 1) / 2;
  if (b1 < 0) b1 = np;
  if (b0 < 0 || b0 > b1 || b1 > np) return fail(IG_ERR_INVALID_ARGUMENT, "bra pair range");
  return guarded([&] {
    hipStream_t s = as_stream(stream);
    auto P = build_plan(*b, s);
    DBuf dout;
    dout.reserve(len * sizeof(double));
    check(launch_fill(dout.as<double>(), len, 0.0, s), "zero eri");
    // quartet lists per (class, class), oriented as the kernels need
    int index[kNumPairClass][kNumPairClass];
    for (auto &r : index)
      for (int &v : r) v = -1;
    std::vector<std::pair<int, int>> cls;
    std::vector<std::vector<int2>> lists;
    for (int64_t bc = b0; bc < b1; ++bc) {
      int i, j;
      pair_of(bc, i, j);
      if (P->schwarz[size_t(i) * nsh + j] == 0.0) continue;  // no primitive pairs: zero
      const int pb = P->canon_to_pos[bc];
      for (int64_t kc = 0; kc < np; ++kc) {
        if (kc > bc && kc < b1) continue;  // evaluated with kc as the bra
        int k, l;
        pair_of(kc, k, l);
        if (P->schwarz[size_t(k) * nsh + l] == 0.0) continue;
        const int pk = P->canon_to_pos[kc];
        const int cb = pair_class(P->sh_l[P->p_sh1[pb]], P->sh_l[P->p_sh2[pb]]);
        const int ck = pair_class(P->sh_l[P->p_sh1[pk]], P->sh_l[P->p_sh2[pk]]);
        int x = cb, y = ck, px = pb, py = pk;
        if (cb < ck) {
          std::swap(x, y);
          std::swap(px, py);
        }
        if (index[x][y] < 0) {
          index[x][y] = int(cls.size());
          cls.push_back({x, y});
          lists.emplace_back();
        }
        lists[index[x][y]].push_back(int2{px - P->cls[x].off, py - P->cls[y].off});
      }
    }
    std::vector<int2> flat;
    for (auto &l : lists) flat.insert(flat.end(), l.begin(), l.end());
    P->s_list.upload(flat);
    size_t big = 0;
    for (size_t c = 0; c < cls.size(); ++c)
      big = std::max(big, big_need(P->cls[cls[c].first], P->cls[cls[c].second],
                                   (long long)lists[c].size(), 1));
    P->s_big.reserve(big * sizeof(double));
    EriOut eo{dout.as<double>(), nullptr, 0};
    eo.rows_begin = b0;
    eo.rows_end = b1;
    eo.omega = omega;
    RysDevice &R = rys_device();
    size_t off = 0;
    for (size_t c = 0; c < cls.size(); ++c) {
      const ClassRange &X = P->cls[cls[c].first], &Y = P->cls[cls[c].second];
      QuartetSet qs{};
      qs.x_off = X.off;
      qs.x_n = X.n;
      qs.y_off = Y.off;
      qs.y_n = Y.n;
      qs.same = cls[c].first == cls[c].second;
      qs.list = P->s_list.as<int2>() + off;
      qs.total = (long long)lists[c].size();
      off += lists[c].size();
      if (class_supported(X.l1, X.l2, Y.l1, Y.l2))
        check(launch_eri(X.l1, X.l2, Y.l1, Y.l2, P->bdev, P->pdev, R.view, qs, eo, s),
              "eri builder launch");
      else
        check(launch_eri_big(X.l1, X.l2, Y.l1, Y.l2, P->bdev, P->pdev, R.view, qs, eo,
                             P->s_big.as<double>(), s),
              "eri builder launch");
    }
    // owned rows back to the caller: for (i, j), rows (mu in i, nu in j) and
    // (nu, mu); for fixed mu the rows over nu are one contiguous run
    const size_t n2 = nao * nao;
    for (int64_t bc = b0; bc < b1; ++bc) {
      int i, j;
      pair_of(bc, i, j);
      const HostShell &si = b->shells[i], &sj = b->shells[j];
      for (int pass = 0; pass < (i == j ? 1 : 2); ++pass) {
        const HostShell &r = pass ? sj : si, &c = pass ? si : sj;
        for (int a = 0; a < r.nfunc(); ++a) {
          const size_t first = (size_t(r.off_func + a) * nao + c.off_func) * n2;
          check(hipMemcpyAsync(out + first, dout.as<double>() + first,
                               size_t(c.nfunc()) * n2 * sizeof(double), hipMemcpyDeviceToHost, s),
                "copy eri rows");
        }
      }
    }
    check(hipStreamSynchronize(s), "eri builder sync");
    return int(IG_OK);
  });
}

// ---------------------------------------------------------------------------
// density fitting (df.rs): 3-center (ij|P) and 2-center (P|Q) through the
// 4-center kernels with a unit-s dummy on the aux index (df.rs:1-39)
namespace {

// main shells, then the aux shells, then one unit-s dummy per aux shell
ig_basis df_combined(const ig_basis *main, const ig_basis &aux) {
  ig_basis c;
  std::vector<std::array<double, 3>> atoms;
  auto add = [&](HostShell h) {
    h.off_func = c.nao;
    h.off_cart = c.nao_cart;
    c.nao += h.nfunc();
    c.nao_cart += ncart(h.l);
    int a = -1;
    for (size_t k = 0; k < atoms.size(); ++k)
      if (std::memcmp(atoms[k].data(), h.center, sizeof(h.center)) == 0) a = int(k);
    if (a < 0) {
      atoms.push_back({h.center[0], h.center[1], h.center[2]});
      a = int(atoms.size()) - 1;
    }
    h.atom = a;
    c.shells.push_back(std::move(h));
  };
  if (main)
    for (const HostShell &s : main->shells) add(s);
  for (const HostShell &s : aux.shells) add(s);
  for (const HostShell &s : aux.shells) {
    HostShell d;  // unit_s (df.rs:33-37): exponent 0, coefficient 1, exempt from normalisation
    d.l = 0;
    d.kind = IG_CARTESIAN;
    d.nprim = 1;
    std::memcpy(d.center, s.center, sizeof(d.center));
    d.exps = {0.0};
    d.coefs = {1.0};
    d.eff = {1.0};
    add(std::move(d));
  }
  c.natom = int(atoms.size());
  return c;
}

// One DF evaluation on the combined basis: quartets (bra pair | ket pair)
// given as canonical pair indices of that basis, stored by mode.
void df_run(Plan &P, const std::vector<std::pair<int64_t, int64_t>> &quartets, EriOut eo,
            hipStream_t s) {
  int index[kNumPairClass][kNumPairClass];
  for (auto &r : index)
    for (int &v : r) v = -1;
  std::vector<std::pair<int, int>> cls;
  std::vector<std::vector<int2>> lists;
  const int nsh = P.nsh;
  for (const auto &q : quartets) {
    int i, j, k, l;
    pair_of(q.first, i, j);
    pair_of(q.second, k, l);
    if (P.schwarz[size_t(i) * nsh + j] == 0.0 || P.schwarz[size_t(k) * nsh + l] == 0.0)
      continue;  // no primitive pairs: the integrals are 0
    const int pb = P.canon_to_pos[q.first], pk = P.canon_to_pos[q.second];
    const int cb = pair_class(P.sh_l[P.p_sh1[pb]], P.sh_l[P.p_sh2[pb]]);
    const int ck = pair_class(P.sh_l[P.p_sh1[pk]], P.sh_l[P.p_sh2[pk]]);
    int x = cb, y = ck, px = pb, py = pk;
    if (cb < ck) {
      std::swap(x, y);
      std::swap(px, py);
    }
    if (index[x][y] < 0) {
      index[x][y] = int(cls.size());
      cls.push_back({x, y});
      lists.emplace_back();
    }
    lists[index[x][y]].push_back(int2{px - P.cls[x].off, py - P.cls[y].off});
  }
  std::vector<int2> flat;
  for (auto &l : lists) flat.insert(flat.end(), l.begin(), l.end());
  P.s_list.upload(flat);
  size_t big = 0;
  for (size_t c = 0; c < cls.size(); ++c)
    big = std::max(big, big_need(P.cls[cls[c].first], P.cls[cls[c].second],
                                 (long long)lists[c].size(), 1));
  P.s_big.reserve(big * sizeof(double));
  RysDevice &R = rys_device();
  size_t off = 0;
  for (size_t c = 0; c < cls.size(); ++c) {
    const ClassRange &X = P.cls[cls[c].first], &Y = P.cls[cls[c].second];
    QuartetSet qs{};
    qs.x_off = X.off;
    qs.x_n = X.n;
    qs.y_off = Y.off;
    qs.y_n = Y.n;
    qs.same = cls[c].first == cls[c].second;
    qs.list = P.s_list.as<int2>() + off;
    qs.total = (long long)lists[c].size();
    off += lists[c].size();
    if (class_supported(X.l1, X.l2, Y.l1, Y.l2))
      check(launch_eri(X.l1, X.l2, Y.l1, Y.l2, P.bdev, P.pdev, R.view, qs, eo, s), "df launch");
    else
      check(launch_eri_big(X.l1, X.l2, Y.l1, Y.l2, P.bdev, P.pdev, R.view, qs, eo,
                           P.s_big.as<double>(), s),
            "df launch");
  }
}

int64_t canon_index(int64_t i, int64_t j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

}  // namespace

static int eri_2c_impl(const ig_basis *aux, double *out, size_t len, void *stream, double omega) {
  const size_t naux = size_t(aux->nao);
  if (len != naux * naux) return fail(IG_ERR_LENGTH_MISMATCH, "2c metric length");
  return guarded([&] {
    hipStream_t s = as_stream(stream);
    const ig_basis comb = df_combined(nullptr, *aux);
    auto P = build_plan(comb, s);
    const int na = int(aux->shells.size());
    std::vector<std::pair<int64_t, int64_t>> qs;  // canonical aux pairs p >= q (df.rs:66-67)
    for (int p = 0; p < na; ++p)
      for (int q = 0; q <= p; ++q) qs.push_back({canon_index(na + p, p), canon_index(na + q, q)});
    DBuf dout;
    dout.reserve(std::max<size_t>(len, 1) * sizeof(double));
    if (len) check(launch_fill(dout.as<double>(), len, 0.0, s), "zero 2c");
    EriOut eo{dout.as<double>(), nullptr, 0};
    eo.mode = kDf2c;
    eo.nmain_sh = 0;
    eo.dummy_sh0 = na;
    eo.nao_m = 0;
    eo.naux = (long long)naux;
    eo.omega = omega;
    df_run(*P, qs, eo, s);
    if (len)
      check(hipMemcpyAsync(out, dout.p, len * sizeof(double), hipMemcpyDeviceToHost, s),
            "copy 2c");
    check(hipStreamSynchronize(s), "2c sync");
    return int(IG_OK);
  });
}

extern "C" int ig_eri_2c(const ig_basis *aux, int engine, double *out, size_t len, void *stream) {
  if (!aux || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  return eri_2c_impl(aux, out, len, stream, 0.0);
}

extern "C" int ig_schwarz_aux_bounds(const ig_basis *aux, int engine, double *out, size_t len,
                                     void *stream) {
  if (!aux || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  const size_t naux = size_t(aux->nao), na = aux->shells.size();
  if (len != na) return fail(IG_ERR_LENGTH_MISMATCH, "aux bound length");
  std::vector<double> m(naux * naux);
  if (int rc = ig_eri_2c(aux, engine, m.data(), m.size(), stream)) return rc;
  // sqrt(max_mu |(P mu|P mu)|) per aux shell (df.rs:257-280)
  for (size_t p = 0; p < na; ++p) {
    const HostShell &sp = aux->shells[p];
    double mx = 0.0;
    for (int a = 0; a < sp.nfunc(); ++a) {
      const size_t d = size_t(sp.off_func + a);
      mx = std::max(mx, std::fabs(m[d * naux + d]));
    }
    out[p] = std::sqrt(mx);
  }
  return IG_OK;
}

static int eri_3c_impl(const ig_basis *main, const ig_basis *aux, int64_t bra_begin,
                       int64_t bra_end, double *out, size_t len, void *stream, double omega) {
  const size_t nao = size_t(main->nao), naux = size_t(aux->nao);
  if (len != nao * nao * naux) return fail(IG_ERR_LENGTH_MISMATCH, "3c tensor length");
  const int nm = int(main->shells.size()), na = int(aux->shells.size());
  const int64_t np = int64_t(nm) * (nm + 1) / 2;
  if (bra_end < 0) bra_end = np;
  if (bra_begin < 0 || bra_begin > bra_end || bra_end > np)
    return fail(IG_ERR_INVALID_ARGUMENT, "bra pair range");
  return guarded([&] {
    hipStream_t s = as_stream(stream);
    const ig_basis comb = df_combined(main, *aux);
    auto P = build_plan(comb, s);
    std::vector<std::pair<int64_t, int64_t>> qs;  // every aux shell per bra pair (df.rs:475-500)
    for (int64_t bc = bra_begin; bc < bra_end; ++bc)
      for (int p = 0; p < na; ++p) qs.push_back({bc, canon_index(nm + na + p, nm + p)});
    DBuf dout;
    dout.reserve(std::max<size_t>(len, 1) * sizeof(double));
    if (len) check(launch_fill(dout.as<double>(), len, 0.0, s), "zero 3c");
    EriOut eo{dout.as<double>(), nullptr, 0};
    eo.mode = kDf3c;
    eo.nmain_sh = nm;
    eo.dummy_sh0 = nm + na;
    eo.nao_m = (long long)nao;
    eo.naux = (long long)naux;
    eo.omega = omega;
    df_run(*P, qs, eo, s);
    // the rows of the range's bra pairs: (mu in i, nu in j) and (nu, mu),
    // each mu's run over nu contiguous
    for (int64_t bc = bra_begin; bc < bra_end; ++bc) {
      int i, j;
      pair_of(bc, i, j);
      const HostShell &si = main->shells[i], &sj = main->shells[j];
      for (int pass = 0; pass < (i == j ? 1 : 2); ++pass) {
        const HostShell &r = pass ? sj : si, &c = pass ? si : sj;
        for (int a = 0; a < r.nfunc(); ++a) {
          const size_t first = (size_t(r.off_func + a) * nao + c.off_func) * naux;
          check(hipMemcpyAsync(out + first, dout.as<double>() + first,
                               size_t(c.nfunc()) * naux * sizeof(double), hipMemcpyDeviceToHost,
                               s),
                "copy 3c rows");
        }
      }
    }
    check(hipStreamSynchronize(s), "3c sync");
    return int(IG_OK);
  });
}

extern "C" int ig_eri_3c_fill(const ig_basis *main, const ig_basis *aux, int engine,
                              int64_t bra_begin, int64_t bra_end, double *out, size_t len,
                              void *stream) {
  if (!main || !aux || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  return eri_3c_impl(main, aux, bra_begin, bra_end, out, len, stream, 0.0);
}

// EriKernel argument of the *_kernel entry points: Coulomb, or erf with a
// finite omega > 0 (integrals.rs:316-321); returns the launch omega (0 for
// Coulomb) or a status code through rc.
static double kernel_omega(int kernel, double omega, int &rc) {
  rc = IG_OK;
  if (kernel == IG_KERNEL_COULOMB) return 0.0;
  if (kernel != IG_KERNEL_ERF) {
    rc = fail(IG_ERR_INVALID_ARGUMENT, "unknown kernel");
    return 0.0;
  }
  if (!(std::isfinite(omega) && omega > 0.0)) {
    rc = fail(IG_ERR_INVALID_OMEGA, "EriKernel::Erf requires a finite omega > 0");
    return 0.0;
  }
  return omega;
}

extern "C" int ig_eri_kernel(const ig_basis *b, int kernel, double omega, double *out, size_t len,
                             void *stream) {
  if (!b || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  int rc;
  const double w = kernel_omega(kernel, omega, rc);
  if (rc) return rc;
  return eri_impl(b, 0.0, out, len, nullptr, stream, w);
}

extern "C" int ig_eri_builder_fill_kernel(const ig_basis *b, int engine, int kernel, double omega,
                                          int64_t bra_begin, int64_t bra_end, double *out,
                                          size_t len, void *stream) {
  if (!b || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  int rc;
  const double w = kernel_omega(kernel, omega, rc);
  if (rc) return rc;
  return eri_builder_impl(b, bra_begin, bra_end, out, len, stream, w);
}

extern "C" int ig_eri_2c_kernel(const ig_basis *aux, int kernel, double omega, double *out,
                                size_t len, void *stream) {
  if (!aux || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  int rc;
  const double w = kernel_omega(kernel, omega, rc);
  if (rc) return rc;
  return eri_2c_impl(aux, out, len, stream, w);
}

extern "C" int ig_eri_3c_fill_kernel(const ig_basis *main, const ig_basis *aux, int engine,
                                     int kernel, double omega, int64_t bra_begin,
                                     int64_t bra_end, double *out, size_t len, void *stream) {
  if (!main || !aux || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  int rc;
  const double w = kernel_omega(kernel, omega, rc);
  if (rc) return rc;
  return eri_3c_impl(main, aux, bra_begin, bra_end, out, len, stream, w);
}

extern "C" size_t ig_packed_eri_n_ao_pairs(const ig_basis *b) {
  return b ? size_t(packed_pair_start(*b).back()) : 0;
}

extern "C" size_t ig_packed_eri_output_len(const ig_basis *b) {
  return b ? tri_count(ig_packed_eri_n_ao_pairs(b)) : 0;
}

extern "C" int ig_packed_eri_ao_pairs(const ig_basis *b, int32_t *pairs, size_t n) {
  if (!b || (n && !pairs)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  const int nsh = int(b->shells.size());
  size_t k = 0;
  for (int i = 0; i < nsh; ++i)
    for (int j = 0; j <= i; ++j) {
      const HostShell &si = b->shells[i], &sj = b->shells[j];
      for (int a = 0; a < si.nfunc(); ++a)
        for (int c = 0; c < (i == j ? a + 1 : sj.nfunc()); ++c) {
          if (k < n) {
            pairs[2 * k] = si.off_func + a;
            pairs[2 * k + 1] = sj.off_func + c;
          }
          ++k;
        }
    }
  if (k != n) return fail(IG_ERR_LENGTH_MISMATCH, "ao pair count is " + std::to_string(k));
  return IG_OK;
}

extern "C" int ig_packed_eri_rows(const ig_basis *b, int64_t bra_begin, int64_t bra_end,
                                  size_t *offset, size_t *len) {
  if (!b || !offset || !len) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  const std::vector<long long> start = packed_pair_start(*b);
  const int64_t np = int64_t(start.size()) - 1;
  if (bra_end < 0) bra_end = np;
  if (bra_begin < 0 || bra_begin > bra_end || bra_end > np)
    return fail(IG_ERR_INVALID_ARGUMENT, "bra pair range");
  *offset = tri_count(size_t(start[bra_begin]));
  *len = tri_count(size_t(start[bra_end])) - *offset;
  return IG_OK;
}

extern "C" int ig_packed_eri_fill(const ig_basis *b, int engine, int64_t bra_begin,
                                  int64_t bra_end, double threshold, double *rows, size_t len,
                                  void *stream) {
  if (!b || (len && !rows)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  return packed_impl(b, threshold, bra_begin, bra_end, rows, len, stream);
}

extern "C" int ig_eri_builder_fill(const ig_basis *b, int engine, int64_t bra_begin,
                                   int64_t bra_end, double *out, size_t len, void *stream) {
  if (!b || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  return eri_builder_impl(b, bra_begin, bra_end, out, len, stream);
}

extern "C" int ig_packed_eri_build(const ig_basis *b, int engine, double threshold, double *out,
                                   size_t len, void *stream) {
  if (!b || (len && !out)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  return packed_impl(b, threshold, 0, -1, out, len, stream);
}

// ---------------------------------------------------------------------------
// gradient (plan objects; contraction kernels in ig_kernels.hip)

extern "C" int ig_grad_create(const ig_basis *basis, int engine, ig_grad **out) {
  if (!basis || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (int rc = check_engine(engine)) return rc;
  for (size_t s = 0; s < basis->shells.size(); ++s)
    if (basis->shells[s].l > kMaxGradL)
      return fail(IG_ERR_L_TOO_HIGH_FOR_GRADIENT,
                  "shell " + std::to_string(s) + ": l=" + std::to_string(basis->shells[s].l) +
                      " exceeds the gradient maximum 5");
  return guarded([&] {
    auto g = std::make_unique<ig_grad>();
    g->engine = engine;
    g->plan = build_plan(*basis, nullptr);
    *out = g.release();
    return int(IG_OK);
  });
}
extern "C" void ig_grad_destroy(ig_grad *g) { delete g; }
extern "C" size_t ig_grad_natom(const ig_grad *g) { return g ? size_t(g->plan->natom) : 0; }
extern "C" size_t ig_grad_nao(const ig_grad *g) { return g ? size_t(g->plan->nao) : 0; }

extern "C" int ig_grad_density_create(const ig_grad *g, size_t nc, const double *cf,
                                      const double *const *cl, const double *const *cr,
                                      size_t nx, const double *xf, const double *const *xl,
                                      const double *const *xr, ig_grad_density **out) {
  if (!g || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if ((nc && (!cf || !cl || !cr)) || (nx && (!xf || !xl || !xr)))
    return fail(IG_ERR_INVALID_ARGUMENT, "null term list");
  for (size_t t = 0; t < nc; ++t)
    if (!cl[t] || !cr[t]) return fail(IG_ERR_INVALID_ARGUMENT, "null matrix");
  for (size_t t = 0; t < nx; ++t)
    if (!xl[t] || !xr[t]) return fail(IG_ERR_INVALID_ARGUMENT, "null matrix");
  return guarded([&] {
    Plan &P = *g->plan;
    auto d = std::make_unique<ig_grad_density>();
    d->owner = g;
    d->ncoul = int(nc);
    d->nexch = int(nx);
    const size_t n2 = size_t(P.nao) * P.nao, nc2 = size_t(P.nao_cart) * P.nao_cart;
    const size_t nterm = nc + nx, nm = 2 * nterm;
    std::vector<double> host(std::max<size_t>(nm, 1) * n2, 0.0);
    for (size_t t = 0; t < nterm; ++t) {
      const double *L = t < nc ? cl[t] : xl[t - nc];
      const double *R = t < nc ? cr[t] : xr[t - nc];
      std::memcpy(host.data() + (2 * t) * n2, L, n2 * sizeof(double));
      std::memcpy(host.data() + (2 * t + 1) * n2, R, n2 * sizeof(double));
      d->fac.push_back(t < nc ? cf[t] : xf[t - nc]);
    }
    if (d->fac.empty()) d->fac.push_back(0.0);
    DBuf func;
    func.upload(host);
    d->d_mats.reserve(std::max<size_t>(nm, 1) * nc2 * sizeof(double));
    d->d_bmax.reserve(std::max<size_t>(nm, 1) * size_t(P.nsh) * P.nsh * sizeof(double));
    if (nm) {
      check(launch_to_cart(P.bdev, func.as<double>(), int(nm), d->d_mats.as<double>(), nullptr),
            "density to Cartesian");
      check(launch_block_max(P.bdev, func.as<double>(), int(nm), d->d_bmax.as<double>(), nullptr),
            "density block maxima");
    }
    d->d_fac.upload(d->fac);
    check(hipDeviceSynchronize(), "density sync");
    *out = d.release();
    return int(IG_OK);
  });
}

extern "C" void ig_grad_density_destroy(ig_grad_density *d) { delete d; }

// DirectGradWorkspace (direct_grad.rs:51-67): a caller's own scratch over a
// gradient contractor's read-only plan data
struct ig_grad_workspace {
  const ig_grad *owner;
  std::unique_ptr<Plan> plan;
};

static int grad_run(const ig_grad *g, ig_grad_workspace *w, const ig_grad_density *dens,
                    int kernel, double omega, double tau, double *out, size_t len, int64_t b0,
                    int64_t b1, void *stream) {
  if (!g || !dens || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (w && w->owner != g)
    return fail(IG_ERR_INVALID_ARGUMENT, "workspace was created for another contractor");
  Plan &P = w ? *w->plan : *g->plan;
  if (dens->owner != g)
    return fail(IG_ERR_FOREIGN_DENSITY,
                "density was prepared by a plan with a different shell layout");
  if (len != size_t(P.natom) * 3) return fail(IG_ERR_LENGTH_MISMATCH, "out must hold natom*3");
  if (kernel != IG_KERNEL_COULOMB && kernel != IG_KERNEL_ERF)
    return fail(IG_ERR_INVALID_ARGUMENT, "unknown kernel");
  if (kernel == IG_KERNEL_ERF && !(std::isfinite(omega) && omega > 0.0))
    return fail(IG_ERR_INVALID_OMEGA, "EriKernel::Erf requires a finite omega > 0");
  return guarded([&] {
    std::lock_guard<std::mutex> lk(P.mu);
    hipStream_t s = as_stream(stream);
    plan_enter(P, s);
    P.s_J.reserve(len * sizeof(double));  // the device gradient accumulator
    check(launch_fill(P.s_J.as<double>(), len, 0.0, s), "zero grad");
    begin_launches(P, s);  // per-class counters (and events when profiling)
    GradDev gd{};
    gd.ncoul = dens->ncoul;
    gd.nexch = dens->nexch;
    gd.fac = dens->d_fac.as<double>();
    gd.mats = dens->d_mats.as<double>();
    gd.bmax = dens->d_bmax.as<double>();
    gd.out = P.s_J.as<double>();
    gd.tau = tau;
    gd.omega = kernel == IG_KERNEL_ERF ? omega : 0.0;
    const int64_t np = P.npairs;
    if (gd.ncoul + gd.nexch > 0) {
      if (b0 == 0 && b1 == np) {
        build_rowptrs(P, tau);
        size_t big = 0;
        for (auto &cp : P.cached_cps)
          big = std::max(big, big_need(P.cls[cp.x], P.cls[cp.y], cp.total, 0));
        P.s_big.reserve(big * sizeof(double));
        for (auto it = P.cached_cps.rbegin(); it != P.cached_cps.rend(); ++it) {
          if (!it->total) continue;
          QuartetSet qs = implicit_set(P, *it);
          const ClassRange &X = P.cls[it->x], &Y = P.cls[it->y];
          unsigned long long *ctr = next_launch(P, X, Y, s);
          check(grad_launch(P, X, Y, qs, gd, ctr, s), "grad launch");
          end_launch(P, s);
        }
      } else {
        ExplicitSets ex = explicit_sets(P, b0, b1, tau);
        std::vector<int2> flat;
        for (auto &l : ex.lists) flat.insert(flat.end(), l.begin(), l.end());
        P.s_list.upload(flat);
        size_t big = 0;
        for (size_t c = 0; c < ex.cls.size(); ++c)
          big = std::max(big, big_need(P.cls[ex.cls[c].first], P.cls[ex.cls[c].second],
                                       (long long)ex.lists[c].size(), 0));
        P.s_big.reserve(big * sizeof(double));
        size_t off = 0;
        for (size_t c = 0; c < ex.cls.size(); ++c) {
          const ClassRange &X = P.cls[ex.cls[c].first], &Y = P.cls[ex.cls[c].second];
          QuartetSet qs{};
          qs.x_off = X.off;
          qs.x_n = X.n;
          qs.y_off = Y.off;
          qs.y_n = Y.n;
          qs.same = ex.cls[c].first == ex.cls[c].second;
          qs.list = P.s_list.as<int2>() + off;
          qs.total = (long long)ex.lists[c].size();
          off += ex.lists[c].size();
          unsigned long long *ctr = next_launch(P, X, Y, s);
          check(grad_launch(P, X, Y, qs, gd, ctr, s), "grad launch");
          end_launch(P, s);
        }
      }
    }
    std::vector<double> h(len);
    check(hipMemcpyAsync(h.data(), P.s_J.p, len * sizeof(double), hipMemcpyDeviceToHost, s),
          "copy grad");
    plan_leave(P, s);
    P.stats_stream = s;
    P.stats_pending = true;
    finalize_stats(P);  // per-class counters and times; synchronises the stream
    for (size_t x = 0; x < len; ++x) out[x] += h[x];
    return int(IG_OK);
  });
}

extern "C" int ig_grad_accumulate_bra_pair(const ig_grad *g, const ig_grad_density *dens,
                                           int32_t i, int32_t j, int kernel, double omega,
                                           double threshold, double *out, size_t len,
                                           void *stream) {
  if (!g) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (!(j >= 0 && j <= i && i < g->plan->nsh))
    return fail(IG_ERR_NON_CANONICAL_BRA, "bra pair is not canonical (i >= j)");
  int64_t bc = int64_t(i) * (i + 1) / 2 + j;
  return grad_run(g, nullptr, dens, kernel, omega, threshold, out, len, bc, bc + 1, stream);
}

extern "C" int ig_grad_contract(const ig_grad *g, const ig_grad_density *dens, int kernel,
                                double omega, double threshold, double *out, size_t len,
                                void *stream) {
  if (!g) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  return grad_run(g, nullptr, dens, kernel, omega, threshold, out, len, 0, g->plan->npairs,
                  stream);
}

extern "C" int ig_grad_last_stats(const ig_grad *g, int64_t *st) {
  if (!g || !st) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(g->plan->mu);
  st[0] = g->plan->last_stats[0];
  st[1] = g->plan->last_stats[1];
  return IG_OK;
}

extern "C" int ig_grad_set_profiling(ig_grad *g, int enable) {
  if (!g) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  std::lock_guard<std::mutex> lk(g->plan->mu);
  g->plan->profiling = enable != 0;
  return IG_OK;
}

extern "C" int ig_grad_profile_read(const ig_grad *g, ig_class_profile *out, size_t cap,
                                    size_t *n) {
  if (!g || !n || (cap && !out)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  return guarded([&] {
    Plan &P = *g->plan;
    std::lock_guard<std::mutex> lk(P.mu);
    finalize_stats(P);
    *n = P.launches.size();
    for (size_t k = 0; k < P.launches.size() && k < cap; ++k) {
      out[k].fp32 = 0;
      out[k].la = P.launches[k].la;
      out[k].lb = P.launches[k].lb;
      out[k].lc = P.launches[k].lc;
      out[k].ld = P.launches[k].ld;
      out[k].quartets = int64_t(P.launch_ctr[2 * k]);
      out[k].prim_quartets = int64_t(P.launch_ctr[2 * k + 1]);
      out[k].ms = P.profiling && k < P.launch_ms.size() ? double(P.launch_ms[k]) : 0.0;
    }
    return int(IG_OK);
  });
}

extern "C" int ig_grad_workspace_create(const ig_grad *g, ig_grad_workspace **out) {
  if (!g || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  return guarded([&] {
    auto w = std::make_unique<ig_grad_workspace>();
    w->owner = g;
    w->plan = workspace_plan(*g->plan);
    *out = w.release();
    return int(IG_OK);
  });
}

extern "C" void ig_grad_workspace_destroy(ig_grad_workspace *w) {
  if (!w) return;
  if (w->plan->idle_recorded) (void)hipEventSynchronize(w->plan->idle);
  delete w;
}

extern "C" int ig_grad_accumulate_bra_pair_ws(const ig_grad *g, ig_grad_workspace *w,
                                              const ig_grad_density *dens, int32_t i, int32_t j,
                                              int kernel, double omega, double threshold,
                                              double *out, size_t len, void *stream) {
  if (!g) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (!(j >= 0 && j <= i && i < g->plan->nsh))
    return fail(IG_ERR_NON_CANONICAL_BRA, "bra pair is not canonical (i >= j)");
  int64_t bc = int64_t(i) * (i + 1) / 2 + j;
  return grad_run(g, w, dens, kernel, omega, threshold, out, len, bc, bc + 1, stream);
}

extern "C" int ig_grad_contract_ws(const ig_grad *g, ig_grad_workspace *w,
                                   const ig_grad_density *dens, int kernel, double omega,
                                   double threshold, double *out, size_t len, void *stream) {
  if (!g) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  return grad_run(g, w, dens, kernel, omega, threshold, out, len, 0, g->plan->npairs, stream);
}

// ---------------------------------------------------------------------------
// dense two-particle density: Basis::eri_grad_contract / _with / _kernel
// (grad.rs:417-466, 493-540).  The nao^4 gamma goes to Cartesian space once
// (four one-index passes), then every canonical quartet with Q > 0 runs
// through the gradient kernels with its weights read from gamma (summed over
// the ordered quartets it stands for) instead of the density terms.
extern "C" int ig_eri_grad_contract_device(const ig_basis *basis, int engine,
                                           const double *d_gamma, size_t gamma_len, int kernel,
                                           double omega, double *out, size_t out_len,
                                           void *stream) {
  if (!basis || !d_gamma || !out) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_engine(engine)) return rc;
  for (size_t sh = 0; sh < basis->shells.size(); ++sh)
    if (basis->shells[sh].l > kMaxGradL)
      return fail(IG_ERR_L_TOO_HIGH_FOR_GRADIENT,
                  "shell " + std::to_string(sh) + ": l=" + std::to_string(basis->shells[sh].l) +
                      " exceeds the gradient maximum 5");
  const size_t nao = size_t(basis->nao), n4 = nao * nao * nao * nao;
  if (gamma_len != n4)
    return fail(IG_ERR_GAMMA_LENGTH, "gamma must hold nao^4 = " + std::to_string(n4) +
                                         " elements, got " + std::to_string(gamma_len));
  if (kernel != IG_KERNEL_COULOMB && kernel != IG_KERNEL_ERF)
    return fail(IG_ERR_INVALID_ARGUMENT, "unknown kernel");
  if (kernel == IG_KERNEL_ERF && !(std::isfinite(omega) && omega > 0.0))
    return fail(IG_ERR_INVALID_OMEGA, "EriKernel::Erf requires a finite omega > 0");
  return guarded([&] {
    std::unique_ptr<Plan> Pp = build_plan(*basis, nullptr);
    Plan &P = *Pp;
    if (out_len != size_t(P.natom) * 3)
      return fail(IG_ERR_LENGTH_MISMATCH, "out must hold natom*3");
    hipStream_t s = as_stream(stream);
    const size_t nc = size_t(P.nao_cart), c4 = nc * nc * nc * nc;
    DBuf t0, t1, dout;
    t0.reserve(c4 * sizeof(double));
    t1.reserve(c4 * sizeof(double));
    dout.reserve(out_len * sizeof(double));
    check(launch_gamma_to_cart(P.bdev, d_gamma, t0.as<double>(), t1.as<double>(), s),
          "gamma to Cartesian");
    check(launch_fill(dout.as<double>(), out_len, 0.0, s), "zero grad");
    P.s_counters.reserve(2 * sizeof(unsigned long long));
    check(launch_fill(P.s_counters.as<double>(), 2, 0.0, s), "zero counters");
    GradDev gd{};
    gd.out = dout.as<double>();
    gd.tau = 0.0;  // every quartet, as the dense sweep
    gd.omega = kernel == IG_KERNEL_ERF ? omega : 0.0;
    gd.gamma = t1.as<double>();
    auto *ctr = P.s_counters.as<unsigned long long>();
    build_rowptrs(P, 0.0);
    size_t big = 0;
    for (auto &cp : P.cached_cps) big = std::max(big, big_need(P.cls[cp.x], P.cls[cp.y], cp.total, 0));
    P.s_big.reserve(big * sizeof(double));
    for (auto &cp : P.cached_cps) {
      if (!cp.total) continue;
      check(grad_launch(P, P.cls[cp.x], P.cls[cp.y], implicit_set(P, cp), gd, ctr, s),
            "grad launch");
    }
    std::vector<double> h(out_len);
    check(hipMemcpyAsync(h.data(), dout.p, out_len * sizeof(double), hipMemcpyDeviceToHost, s),
          "copy grad");
    check(hipStreamSynchronize(s), "grad sync");
    std::copy(h.begin(), h.end(), out);
    return int(IG_OK);
  });
}

// ---------------------------------------------------------------------------
// batched J/K over many bases (one symmetric density per basis)

// The batched graph's execution is switched off until the cause of an
// illegal memory access it took on the GPU (small mixed s,p,d,f molecules,
// after 64-molecule runs had passed) is found; ig_batch_jk_device fails with
// IG_ERR_UNSUPPORTED meanwhile.  Only the host-only build of the launch
// checker (tests/replay) defines IG_BATCH_GRAPH=1, to walk the capture and
// every graph launch on the CPU.
#ifndef IG_BATCH_GRAPH
#define IG_BATCH_GRAPH 0
#endif
constexpr bool kBatchGraphEnabled = IG_BATCH_GRAPH != 0;

struct ig_batch {
  std::vector<std::unique_ptr<Plan>> plans;
  std::mutex mu;
  hipStream_t stream = nullptr;  // the graph's own stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  // the configuration the graph was captured for
  double tau = NAN, cut = NAN;
  int want_k = -1;
  std::vector<const double *> dens;
  std::vector<double *> J, K;
  // device views and work lists of the captured graph
  DBuf d_bs, d_ps, d_dds, d_sets, d_seg, d_queue, d_ctr;
  size_t nlaunch = 0;
  double fp32_below = 0.0;
  int64_t last_stats[2] = {0, 0};
  bool stats_pending = false;
  ~ig_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev_in) (void)hipEventDestroy(ev_in);
    if (ev_out) (void)hipEventDestroy(ev_out);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Views and work lists of a batched J/K: what a capture bakes into the graph.
struct BatchKey {
  int x, y, fp32;
  bool operator<(const BatchKey &o) const {
    return std::tie(x, y, fp32) < std::tie(o.x, o.y, o.fp32);
  }
};
struct BatchLaunch {
  BatchKey k;
  bool batched;  // one row-sum launch over every basis (else per-basis launches)
  int variant;
  size_t acc;    // LDS accumulator doubles of the widest basis with the key
  long long total;
  size_t set_off, seg_off;
};
struct BatchPrep {
  std::vector<BasisDev> bs;
  std::vector<PairsDev> ps;
  std::vector<DensDev> dds;
  std::map<BatchKey, std::vector<std::pair<int, size_t>>> work;  // key -> (basis, cp index)
  std::vector<BatchLaunch> launches;
  std::vector<QuartetSet> sets;  // nb per batched launch (empty sets for bases without the key)
  std::vector<long long> seg;    // nb + 1 prefix sums per batched launch
};

// Host side of a batched J/K for (tau, cut, want_k): quartet sets, density
// buffers, device views and work lists uploaded, scratch reserved.  No
// kernel runs.
void batch_prepare(ig_batch &B, double tau, double cut, bool want_k, BatchPrep &bp) {
  const int nb = int(B.plans.size());
  std::vector<BasisDev> &bs = bp.bs;
  std::vector<PairsDev> &ps = bp.ps;
  std::vector<DensDev> &dds = bp.dds;
  bs.assign(nb, BasisDev{});
  ps.assign(nb, PairsDev{});
  dds.assign(nb, DensDev{});
  for (int m = 0; m < nb; ++m) {
    Plan &P = *B.plans[m];
    build_rowptrs(P, tau, cut);
    dds[m] = reserve_densities(P, 1, false, tau, want_k);
    bs[m] = P.bdev;
    ps[m] = P.pdev;
  }
  // work lists: every (x, y, fp32) key present in some basis
  auto &work = bp.work;
  work.clear();
  for (int m = 0; m < nb; ++m) {
    Plan &P = *B.plans[m];
    for (size_t c = 0; c < P.cached_cps.size(); ++c) {
      const Plan::CP &cp = P.cached_cps[c];
      if (cp.total) work[{cp.x, cp.y, cp.fp32}].push_back({m, c});
    }
  }
  auto &launches = bp.launches;
  auto &sets = bp.sets;
  auto &seg = bp.seg;
  launches.clear();
  sets.clear();
  seg.clear();
  for (auto &kv : work) {
    const BatchKey &k = kv.first;
    const Plan &P0 = *B.plans[kv.second[0].first];
    const ClassRange &X0 = P0.cls[k.x], &Y0 = P0.cls[k.y];
    BatchLaunch L{k, false, -1, 0, 0, sets.size(), seg.size()};
    if (jk_lane_compiled(X0.l1, X0.l2, Y0.l1, Y0.l2)) {
      int v = jk_variant(P0, X0, Y0);
      // FP32 sets take the register row-sum kernel, as in jk_launch
      if ((v == kJkTeam || v == kJkTeamHrr || v == kJkHgp) && k.fp32) v = kJkRowAcc;
      if (v == kJkRowAcc || v == kJkHgp) {
        bool fits = true;
        for (auto &mc : kv.second) {
          const Plan &P = *B.plans[mc.first];
          fits = fits && jk_variant_fits(v, X0.l1, X0.l2, Y0.l1, Y0.l2, P.bdev, 1, false);
          L.acc = std::max(L.acc, rowacc_acc_doubles(P.bdev, X0.l1, X0.l2, Y0.l1, Y0.l2, 1));
        }
        if (fits) {
          L.batched = true;
          L.variant = v;
        }
      }
    }
    if (L.batched) {  // every basis has a slot; those without the key hold an empty set
      std::vector<QuartetSet> row(nb, QuartetSet{});
      std::vector<long long> pre(nb + 1, 0);
      for (auto &mc : kv.second) row[mc.first] = implicit_set(*B.plans[mc.first],
                                                              B.plans[mc.first]->cached_cps[mc.second]);
      for (int m = 0; m < nb; ++m) pre[m + 1] = pre[m] + row[m].total;
      L.total = pre[nb];
      sets.insert(sets.end(), row.begin(), row.end());
      seg.insert(seg.end(), pre.begin(), pre.end());
    }
    launches.push_back(L);
  }
  B.d_bs.upload(bs);
  B.d_ps.upload(ps);
  B.d_dds.upload(dds);
  B.d_sets.upload(sets);
  B.d_seg.upload(seg);
  B.nlaunch = launches.size();
  B.d_queue.reserve(std::max<size_t>(1, launches.size()) * kMaxSlices * sizeof(unsigned));
  B.d_ctr.reserve(std::max<size_t>(1, launches.size()) * 2 * sizeof(unsigned long long));
  // runtime-l scratch of the per-basis launches, sized before capture
  for (auto &L : launches) {
    if (L.batched) continue;
    for (auto &mc : work[L.k]) {
      Plan &P = *B.plans[mc.first];
      const Plan::CP &cp = P.cached_cps[mc.second];
      const size_t need = big_need_jk(P.cls[cp.x], P.cls[cp.y], cp.total);
      if (need > P.s_big.bytes / sizeof(double)) P.s_big.reserve(need * sizeof(double));
    }
  }
}

// Capture the whole batched J/K into b.graph: per basis the density kernels,
// then per (class pair, precision) one launch of the row-sum kernel over every
// basis (work queue) -- or per-basis launches for classes that kernel cannot
// run -- then per basis the fold and finishing transform.
void batch_capture(ig_batch &B, double tau, double cut, bool want_k) {
  const int nb = int(B.plans.size());
  RysDevice &R = rys_device();
  BatchPrep bp;
  batch_prepare(B, tau, cut, want_k, bp);
  auto &work = bp.work;
  const auto &launches = bp.launches;
  const auto &dds = bp.dds;
  hipStream_t s = B.stream;
  check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    for (int m = 0; m < nb; ++m) launch_densities(*B.plans[m], dds[m], B.dens[m], false, s);
    // counters and queues cleared by kernels (no memset nodes in the graph)
    check(launch_fill(B.d_queue.as<double>(), B.d_queue.bytes / sizeof(double), 0.0, s), "zero queues");
    check(launch_fill(B.d_ctr.as<double>(), B.d_ctr.bytes / sizeof(double), 0.0, s), "zero counters");
    for (size_t li = 0; li < launches.size(); ++li) {
      const BatchLaunch &L = launches[li];
      unsigned long long *ctr = B.d_ctr.as<unsigned long long>() + 2 * li;
      const Plan &P0 = *B.plans[work[L.k][0].first];
      const ClassRange &X0 = P0.cls[L.k.x], &Y0 = P0.cls[L.k.y];
      if (L.batched) {
        auto fn = L.k.fp32 ? launch_jk_batch_f32 : launch_jk_batch;
        check(fn(X0.l1, X0.l2, Y0.l1, Y0.l2, B.d_bs.as<BasisDev>(), B.d_ps.as<PairsDev>(), R.view,
                 B.d_sets.as<QuartetSet>() + L.set_off, B.d_dds.as<DensDev>(),
                 B.d_seg.as<long long>() + L.seg_off, nb, L.total,
                 B.d_queue.as<unsigned>() + li * kMaxSlices, L.variant, L.acc, ctr, s),
              "batched jk launch");
      } else {
        for (auto &mc : work[L.k]) {
          Plan &P = *B.plans[mc.first];
          const Plan::CP &cp = P.cached_cps[mc.second];
          check(jk_launch(P, P.cls[cp.x], P.cls[cp.y], implicit_set(P, cp), dds[mc.first], ctr, s,
                          -1, cp.fp32 != 0),
                "jk launch");
        }
      }
    }
    for (int m = 0; m < nb; ++m) {
      Plan &P = *B.plans[m];
      check(launch_fold_copies(dds[m], size_t(P.nao_cart) * P.nao_cart, s), "fold copies");
      check(launch_jk_finish(P.bdev, dds[m], B.J[m], want_k ? B.K[m] : nullptr, s), "jk finish");
    }
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  if (B.exec) (void)hipGraphExecDestroy(B.exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  B.exec = nullptr;
  B.graph = nullptr;
  check(hipStreamEndCapture(s, &B.graph), "end capture");
  check(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0), "instantiate graph");
}


// Device array [p, p + n) back to the host (small tables; a synchronous copy).
template <class T>
std::vector<T> batch_copy_back(const T *p, size_t n) {
  std::vector<T> h(n);
  if (n) check(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost), "copy back");
  return h;
}

// Is [p, p + bytes) inside buffer b?
bool batch_inside(const DBuf &b, const void *p, size_t bytes) {
  const char *lo = static_cast<const char *>(b.p), *q = static_cast<const char *>(p);
  return b.p && q >= lo && q + bytes <= lo + b.bytes;
}

// Host walk of every quartet of every batched launch with the indexing of
// jk_rowacc_batch_kernel (chunk -> basis decode, row cursor, pair, shell and
// matrix indices, LDS row-sum columns): the first access outside its buffer,
// or "" when there is none.  Reads back from the device what the kernels
// would read (row pointers, segment kets, pair and shell tables) and checks
// every pointer the launch views hold against the plan buffers.
std::string batch_replay(ig_batch &B, const BatchPrep &bp) {
  const int nb = int(B.plans.size());
  std::ostringstream err;
  auto fail_at = [&](const BatchLaunch &L, int m, long long t, const std::string &what) {
    err << "class pair (" << L.k.x << "," << L.k.y << ") fp32 " << L.k.fp32 << ", basis " << m
        << ", quartet " << t << ": " << what;
    return err.str();
  };
  // per basis: density views against the plan buffers
  for (int m = 0; m < nb; ++m) {
    Plan &P = *B.plans[m];
    const DensDev &dd = bp.dds[m];
    const size_t n2 = size_t(P.nao_cart) * P.nao_cart;
    BatchLaunch none{};
    if (!batch_inside(P.s_dc, dd.dc, n2 * dd.neff * sizeof(double)))
      return fail_at(none, m, -1, "Cartesian densities outside s_dc");
    if (!batch_inside(P.s_jc, dd.jc, dd.copy_stride * dd.ncopy * sizeof(double)))
      return fail_at(none, m, -1, "J accumulators outside s_jc");
    if (dd.kc && !batch_inside(P.s_kc, dd.kc, dd.copy_stride * dd.ncopy * sizeof(double)))
      return fail_at(none, m, -1, "K accumulators outside s_kc");
    if (dd.copy_stride < n2 * size_t(dd.neff)) return fail_at(none, m, -1, "copy stride too small");
    if (!batch_inside(P.s_bound, dd.bound, size_t(dd.ndens) * P.nsh * P.nsh * sizeof(double)))
      return fail_at(none, m, -1, "density bounds outside s_bound");
    if (!batch_inside(P.s_eff, dd.eff_src, 2 * size_t(dd.neff) * sizeof(int)))
      return fail_at(none, m, -1, "effective-density table outside s_eff");
  }
  for (const BatchLaunch &L : bp.launches) {
    if (!L.batched) continue;
    const Plan &P0 = *B.plans[bp.work.at(L.k)[0].first];
    const ClassRange &X0 = P0.cls[L.k.x], &Y0 = P0.cls[L.k.y];
    const int la = X0.l1, lb = X0.l2, lc = Y0.l1, ld = Y0.l2;
    const int na = ncart(la), nb_ = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    // chunk -> basis decode of the kernel over the launch's prefix sums
    const long long *seg = bp.seg.data() + L.seg_off;
    if (seg[0] != 0 || seg[nb] != L.total) return fail_at(L, -1, -1, "prefix sums do not span the launch");
    for (int m = 0; m < nb; ++m)
      if (seg[m + 1] < seg[m]) return fail_at(L, m, -1, "prefix sums decrease");
    // the kernel's work items: chunks of 1024 quartets (4 per lane of its 256
    // lanes), each cut at the basis boundaries it crosses
    const long long chunk = 1024;
    for (long long c0 = 0; c0 < L.total; c0 += chunk) {
      long long g0 = c0;
      const long long g1 = c0 + chunk < L.total ? c0 + chunk : L.total;
      while (g0 < g1) {
        int lo = 0, hi = nb - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (seg[mid] <= g0) lo = mid; else hi = mid - 1;
        }
        if (!(seg[lo] <= g0 && g0 < seg[lo + 1])) return fail_at(L, lo, g0, "basis decode misses");
        const long long e = g1 < seg[lo + 1] ? g1 : seg[lo + 1];
        g0 = e;
      }
    }
    for (int m = 0; m < nb; ++m) {
      const QuartetSet &qs = bp.sets[L.set_off + m];
      if (qs.total != seg[m + 1] - seg[m]) return fail_at(L, m, -1, "set total differs from the prefix sums");
      if (!qs.total) continue;
      Plan &P = *B.plans[m];
      if (qs.list) return fail_at(L, m, -1, "explicit list in a batched launch");
      const ClassRange &X = P.cls[L.k.x], &Y = P.cls[L.k.y];
      if (X.l1 != la || X.l2 != lb || Y.l1 != lc || Y.l2 != ld) return fail_at(L, m, -1, "class momenta differ between bases");
      if (qs.x_off != X.off || qs.x_n != X.n || qs.y_off != Y.off || qs.y_n != Y.n)
        return fail_at(L, m, -1, "set ranges differ from the plan's class ranges");
      const size_t nrow = size_t(qs.x_n) * qs.nseg;
      if (qs.nseg < 1) return fail_at(L, m, -1, "no segments");
      if (!batch_inside(P.s_rowptr, qs.rowptr, (nrow + 1) * sizeof(long long)))
        return fail_at(L, m, -1, "row pointers outside s_rowptr");
      if (!batch_inside(P.s_qbase, qs.qbase, nrow * sizeof(int)))
        return fail_at(L, m, -1, "segment kets outside s_qbase");
      const std::vector<long long> rp = batch_copy_back(qs.rowptr, nrow + 1);
      const std::vector<int> qb = batch_copy_back(qs.qbase, nrow);
      if (rp[0] != 0 || rp[nrow] != qs.total) return fail_at(L, m, -1, "row pointers do not span the set");
      const std::vector<int> sh1 = batch_copy_back(P.pdev.sh1, size_t(P.npairs));
      const std::vector<int> sh2 = batch_copy_back(P.pdev.sh2, size_t(P.npairs));
      const std::vector<int> ppo = batch_copy_back(P.pdev.pp_off, size_t(P.npairs));
      const std::vector<int> ppn = batch_copy_back(P.pdev.pp_n, size_t(P.npairs));
      const std::vector<int> shl = batch_copy_back(P.bdev.l, size_t(P.nsh));
      const std::vector<int> offc = batch_copy_back(P.bdev.off_cart, size_t(P.nsh));
      const std::vector<int> rank = batch_copy_back(P.bdev.rank, size_t(P.nsh));
      const std::vector<int> lsh = batch_copy_back(P.bdev.lshells, size_t(P.nsh));
      const size_t nprim = P.d_zeta.bytes / sizeof(double);
      // row-sum columns of the ket class and the accumulators they index
      const int nC = P.bdev.lshell_off[lc + 1] - P.bdev.lshell_off[lc];
      const int nD = P.bdev.lshell_off[ld + 1] - P.bdev.lshell_off[ld];
      const int W = nC * nc + (lc == ld ? 0 : nD * nd);
      const int dbase = lc == ld ? 0 : nC * nc;
      const size_t ntot = rowacc_acc_doubles(P.bdev, la, lb, lc, ld, bp.dds[m].neff);
      if (ntot > L.acc) return fail_at(L, m, -1, "row sums need more LDS than the launch holds");
      for (int r = 0; r < nC; ++r) {  // the flush's column -> function map
        const int sh = lsh[P.bdev.lshell_off[lc] + r];
        if (sh < 0 || sh >= P.nsh || shl[sh] != lc) return fail_at(L, m, -1, "column shell of the wrong momentum");
      }
      for (int r = 0; r < nD; ++r) {
        const int sh = lsh[P.bdev.lshell_off[ld] + r];
        if (sh < 0 || sh >= P.nsh || shl[sh] != ld) return fail_at(L, m, -1, "column shell of the wrong momentum");
      }
      size_t sg = 0;
      for (long long t = 0; t < qs.total; ++t) {
        while (sg + 1 <= nrow && rp[sg + 1] <= t) ++sg;  // the row cursor's segment
        if (sg >= nrow) return fail_at(L, m, t, "segment past the set");
        const int Prow = int(sg / qs.nseg);
        const long long Q = qb[sg] + (t - rp[sg]);
        if (Prow >= qs.x_n || Q < 0 || Q >= qs.y_n) return fail_at(L, m, t, "pair index outside the class");
        const int pb = qs.x_off + Prow, pk = qs.y_off + int(Q);
        if (pb >= P.npairs || pk >= P.npairs) return fail_at(L, m, t, "pair index outside the pair table");
        if (ppo[pb] < 0 || size_t(ppo[pb]) + ppn[pb] > nprim || ppo[pk] < 0 || size_t(ppo[pk]) + ppn[pk] > nprim)
          return fail_at(L, m, t, "primitive pairs outside the primitive table");
        const int sA = sh1[pb], sB = sh2[pb], sC = sh1[pk], sD = sh2[pk];
        for (int sh : {sA, sB, sC, sD})
          if (sh < 0 || sh >= P.nsh) return fail_at(L, m, t, "shell index outside the basis");
        if (shl[sA] != la || shl[sB] != lb || shl[sC] != lc || shl[sD] != ld)
          return fail_at(L, m, t, "shell momentum differs from the class");
        const int ns[4] = {na, nb_, nc, nd};
        const int ss[4] = {sA, sB, sC, sD};
        for (int i = 0; i < 4; ++i)
          if (offc[ss[i]] < 0 || offc[ss[i]] + ns[i] > P.nao_cart)
            return fail_at(L, m, t, "Cartesian functions outside the matrices");
        if (rank[sC] * nc + nc > (lc == ld ? W : dbase) || dbase + rank[sD] * nd + nd > W)
          return fail_at(L, m, t, "row-sum column outside the LDS rows");
      }
    }
  }
  return std::string();
}

}  // namespace

extern "C" int ig_batch_create(const ig_basis *const *bases, size_t n, ig_batch **out) {
  if (!out || (n && !bases)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (n == 0) return fail(IG_ERR_INVALID_ARGUMENT, "a batch needs at least one basis");
  for (size_t m = 0; m < n; ++m)
    if (!bases[m]) return fail(IG_ERR_INVALID_ARGUMENT, "null basis in the batch");
  return guarded([&] {
    auto B = std::make_unique<ig_batch>();
    for (size_t m = 0; m < n; ++m) B->plans.push_back(build_plan(*bases[m], nullptr));
    check(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking), "stream");
    check(hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming), "event");
    check(hipEventCreateWithFlags(&B->ev_out, hipEventDisableTiming), "event");
    *out = B.release();
    return int(IG_OK);
  });
}

extern "C" void ig_batch_destroy(ig_batch *b) {
  if (b) (void)hipStreamSynchronize(b->stream);
  delete b;
}

extern "C" size_t ig_batch_size(const ig_batch *b) { return b ? b->plans.size() : 0; }

extern "C" int ig_batch_set_mixed_precision(ig_batch *b, double fp32_below) {
  if (!b) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (!(std::isfinite(fp32_below) && fp32_below >= 0.0))
    return fail(IG_ERR_INVALID_ARGUMENT, "fp32_below must be finite and >= 0");
  std::lock_guard<std::mutex> g(b->mu);
  b->fp32_below = fp32_below;
  return IG_OK;
}

extern "C" int ig_batch_jk_device(ig_batch *b, const double *const *d_dens, double threshold,
                                  int want_k, double *const *d_J, double *const *d_K,
                                  void *stream) {
  if (!b || !d_dens || !d_J || (want_k && !d_K)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (!kBatchGraphEnabled)
    return fail(IG_ERR_UNSUPPORTED,
                "batched J/K is switched off in this build: its graph took an illegal memory "
                "access on the GPU with small mixed s,p,d,f molecules and the cause is not found "
                "yet; use ig_direct_jk_device per molecule");
  const size_t nb = b->plans.size();
  for (size_t m = 0; m < nb; ++m)
    if (!d_dens[m] || !d_J[m] || (want_k && !d_K[m]))
      return fail(IG_ERR_INVALID_ARGUMENT, "null device pointer for basis " + std::to_string(m));
  return guarded([&] {
    std::lock_guard<std::mutex> g(b->mu);
    const double cut = b->fp32_below > threshold ? b->fp32_below : threshold;
    std::vector<const double *> dens(d_dens, d_dens + nb);
    std::vector<double *> J(d_J, d_J + nb), K(nb, nullptr);
    if (want_k) K.assign(d_K, d_K + nb);
    if (!b->exec || b->tau != threshold || b->cut != cut || b->want_k != (want_k != 0) ||
        b->dens != dens || b->J != J || b->K != K) {
      check(hipStreamSynchronize(b->stream), "sync");
      // the capture rewrites the device views and work lists an older graph
      // reads: drop that graph first, so a capture that fails leaves nothing
      // a later call could replay
      if (b->exec) (void)hipGraphExecDestroy(b->exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
      b->tau = NAN;  // set again only once the capture has succeeded
      b->dens = dens;
      b->J = J;
      b->K = K;
      b->cut = cut;
      b->want_k = want_k != 0;
      batch_capture(*b, threshold, cut, want_k != 0);
      b->tau = threshold;
    }
    hipStream_t s = as_stream(stream);
    check(hipEventRecord(b->ev_in, s), "event");
    check(hipStreamWaitEvent(b->stream, b->ev_in, 0), "wait");
    check(hipGraphLaunch(b->exec, b->stream), "graph launch");
    check(hipEventRecord(b->ev_out, b->stream), "event");
    check(hipStreamWaitEvent(s, b->ev_out, 0), "wait");
    b->stats_pending = true;
    return int(IG_OK);
  });
}

extern "C" int ig_batch_check(ig_batch *b, double threshold, int want_k, char *msg, size_t cap) {
  if (!b) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (!(std::isfinite(threshold) && threshold >= 0.0))
    return fail(IG_ERR_INVALID_ARGUMENT, "threshold must be finite and >= 0");
  return guarded([&] {
    std::lock_guard<std::mutex> g(b->mu);
    check(hipStreamSynchronize(b->stream), "sync");
    const double cut = b->fp32_below > threshold ? b->fp32_below : threshold;
    BatchPrep bp;
    batch_prepare(*b, threshold, cut, want_k != 0, bp);
    const std::string r = batch_replay(*b, bp);
    if (msg && cap) {
      const size_t n = std::min(cap - 1, r.size());
      std::memcpy(msg, r.data(), n);
      msg[n] = '\0';
    }
    if (!r.empty()) throw DeviceError{IG_ERR_DEVICE, "batched launch data out of range: " + r};
    return int(IG_OK);
  });
}

extern "C" int ig_batch_autotune(ig_batch *b, size_t molecule, const double *d_dens,
                                 double threshold, int want_k, int reps, void *stream) {
  if (!b || !d_dens) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (molecule >= b->plans.size()) return fail(IG_ERR_INVALID_ARGUMENT, "molecule out of range");
  if (reps < 1) return fail(IG_ERR_INVALID_ARGUMENT, "reps must be at least 1");
  return guarded([&] {
    std::lock_guard<std::mutex> g(b->mu);
    check(hipStreamSynchronize(b->stream), "sync");
    Plan &P = *b->plans[molecule];
    autotune_plan(P, d_dens, 1, threshold, want_k, reps, as_stream(stream));
    for (auto &Q : b->plans) {
      if (Q.get() == &P) continue;
      Q->variant = P.variant;
      Q->variant_ms = P.variant_ms;
    }
    // the trial launches moved plan buffers: capture again on the next call
    b->tau = NAN;
    return int(IG_OK);
  });
}

extern "C" int ig_batch_schedule_write(ig_batch *b, const ig_schedule_entry *in, size_t n) {
  if (!b || (n && !in)) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  if (int rc = check_schedule(in, n)) return rc;
  return guarded([&] {
    std::lock_guard<std::mutex> g(b->mu);
    check(hipStreamSynchronize(b->stream), "sync");
    for (auto &P : b->plans) apply_schedule(*P, in, n);
    b->tau = NAN;  // capture again with the new kernels
    return int(IG_OK);
  });
}

extern "C" int ig_batch_last_stats(ig_batch *b, int64_t *st) {
  if (!b || !st) return fail(IG_ERR_INVALID_ARGUMENT, "null argument");
  return guarded([&] {
    std::lock_guard<std::mutex> g(b->mu);
    if (b->stats_pending) {
      std::vector<unsigned long long> c(2 * std::max<size_t>(b->nlaunch, 1), 0);
      check(hipStreamSynchronize(b->stream), "sync");
      if (b->nlaunch)
        check(hipMemcpy(c.data(), b->d_ctr.p, 2 * b->nlaunch * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost),
              "copy counters");
      b->last_stats[0] = b->last_stats[1] = 0;
      for (size_t k = 0; k < b->nlaunch; ++k) {
        b->last_stats[0] += int64_t(c[2 * k]);
        b->last_stats[1] += int64_t(c[2 * k + 1]);
      }
      b->stats_pending = false;
    }
    st[0] = b->last_stats[0];
    st[1] = b->last_stats[1];
    return int(IG_OK);
  });
}
