// hip_replay.cpp -- host stand-in for the HIP runtime, for a CPU build of the
// library (its host code, compiled with --offload-host-only) that checks
// the device work the host code would launch.
//
// Device memory is host memory from an allocation table; streams and
// events do nothing; copies are memcpy.  A stream capture records the
// launches on the capturing stream (with a copy of their arguments) instead
// of running them, and each graph launch runs the recorded list against
// memory as it is then; allocations, frees, synchronous copies, memsets and
// launches on other streams while a stream captures are violations.  Kernel
// launches arrive by name (the launch stubs register their kernels' names
// at load time):
//   * the planning kernels whose results the host reads back run here:
//     Schwarz bounds (taken from a matrix the caller gives), the XCD probe,
//     the density symmetry flag, the fill kernel;
//   * every J/K launch (compiled register, row-sum, team and team-HRR
//     kernels, FP64 and FP32, Rys and OS/HGP; the runtime-l per-quartet and
//     sliced kernels) is checked: each pointer of its argument views must
//     lie in a live allocation of the size the views imply, the launch's
//     dynamic LDS must hold what the kernel stages (row sums; root-batch
//     tables and digest), the runtime-l kernels' offset table and block
//     slots must fit their scratch, and every quartet of its set, decoded as
//     the kernels' row cursor does, must name pairs, primitives, shells,
//     Cartesian functions and row-sum columns inside their tables;
//   * the gradient launches (compiled lane kernels and the runtime-l kernel)
//     get the same view and quartet checks, with their density views and
//     force accumulator in place of the J/K ones, and the ERI launches with
//     their output buffer;
//   * the batched row-sum launches get the checks of the single-basis
//     kernel per basis, through the device arrays of views they read, and
//     their zeroed work queue and LDS accumulator size;
//   * the density split, Cartesian transform, block-maximum, fold and
//     finishing kernels: every array against the sizes their loops cover;
//   * other kernels do nothing.
// A violation is recorded (the first few hundred, with the launch and the
// quartet) for the caller: replay_violation_count() / replay_violation(k).
// For the error-path tests an allocation or the launches of a named kernel
// can be made to fail (replay_fail_malloc_after, replay_fail_launch).
#include <cxxabi.h>

#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "ig_internal.h"

using namespace ig;

namespace {

struct Alloc {
  size_t bytes;
};
struct Config {
  dim3 grid, block;
  size_t shm = 0;
  hipStream_t stream = nullptr;
};
// A kernel launch made while a stream captures: not run then, but recorded
// with its configuration and a copy of its arguments (for the kernels whose
// argument lists the replay knows), and run at every launch of the graph.
struct Recorded {
  std::string name;
  Config c;
  bool known = false;
  // a launch of the host-executed kernels (tests/emu): its job, run at each
  // graph launch
  void (*run)(void *) = nullptr;
  void (*del)(void *) = nullptr;
  void *job = nullptr;
  std::vector<std::vector<char>> store;  // one copy per argument
  std::vector<void *> args;
};
// The tables are built on first use: the other objects' static initialisers
// register their kernels and device variables before this file's globals
// would be constructed.
struct State {
  std::map<uintptr_t, Alloc> alloc;
  std::recursive_mutex mu;
  std::vector<std::string> viol;
  std::unordered_map<const void *, std::string> kname;
  std::vector<double> schwarz;  // nsh x nsh
  int nsh = 0;
  long long jk_launches = 0, quartets = 0, other = 0;
  std::vector<Config> cfg;
  // failure injection for the error-path tests
  long long fail_malloc_after = -1;  // the k-th hipMalloc from now fails (-1: off)
  std::string fail_launch;           // launches of kernels whose name holds it fail
  hipError_t last = hipSuccess;      // the error hipGetLastError reports next
  long long mallocs = 0;             // hipMalloc calls since the last reset
  // graph capture: the batched row-sum launches recorded while a stream
  // captures, re-checked at every launch of the graph
  bool capturing = false;
  hipStream_t capture_stream = nullptr;
  std::vector<Recorded> *rec = nullptr;
  std::set<hipGraph_t> live_graphs;
  long long graph_launches = 0;
};
State &st() {
  static State s;
  return s;
}
#define g_alloc (st().alloc)
#define g_mu (st().mu)
#define g_viol (st().viol)
#define g_kname (st().kname)
#define g_schwarz (st().schwarz)
#define g_nsh (st().nsh)
#define g_jk_launches (st().jk_launches)
#define g_quartets (st().quartets)
#define g_other (st().other)
#define g_cfg (st().cfg)

void violation(const std::string &s) {
  if (g_viol.size() < 400) g_viol.push_back(s);
}

// [p, p + bytes) inside one live allocation
bool inside(const void *p, size_t bytes) {
  if (!p) return bytes == 0;
  const uintptr_t q = reinterpret_cast<uintptr_t>(p);
  auto it = g_alloc.upper_bound(q);
  if (it == g_alloc.begin()) return false;
  --it;
  return q >= it->first && q + bytes <= it->first + it->second.bytes;
}

void need(const std::string &launch, const char *what, const void *p, size_t bytes) {
  if (bytes && !inside(p, bytes)) {
    std::ostringstream s;
    s << launch << ": " << what << " (" << bytes << " bytes at " << p
      << ") outside every live allocation";
    violation(s.str());
  }
}

std::string demangle(const std::string &m) {
  int st = 0;
  char *d = abi::__cxa_demangle(m.c_str(), nullptr, nullptr, &st);
  std::string r = (st == 0 && d) ? d : m;
  std::free(d);
  return r;
}

// integers of the first "Cls<a, b, c, d>" and the template arguments after it
bool class_of(const std::string &n, int cls[4], std::vector<std::string> &rest) {
  const size_t i = n.find("Cls<");
  if (i == std::string::npos) return false;
  if (std::sscanf(n.c_str() + i, "Cls<%d, %d, %d, %d>", &cls[0], &cls[1], &cls[2], &cls[3]) != 4)
    return false;
  size_t j = n.find('>', i);  // end of Cls<...>
  size_t k = n.find('(', j);  // start of the parameter list
  std::string t = n.substr(j + 1, k - j - 1);
  rest.clear();
  std::string cur;
  for (char c : t) {
    if (c == ',' || c == '>') {
      if (!cur.empty()) rest.push_back(cur);
      cur.clear();
    } else if (c != ' ') {
      cur.push_back(c);
    }
  }
  if (!cur.empty()) rest.push_back(cur);
  return true;
}

int ncart_h(int l) { return (l + 1) * (l + 2) / 2; }

// The J/K launch of a kernel of class (la..ld) whose views (BasisDev,
// PairsDev, RysDev, QuartetSet, DensDev) are arguments v[0..4]: the views,
// the row-sum kernel's LDS (rest: its template arguments after the class)
// and every quartet.
// A gradient launch has a GradDev as v[4] instead (grad true).
void check_jk(const std::string &name, int la, int lb, int lc, int ld, void **v,
              const std::vector<std::string> &rest, const Config &c, bool grad = false,
              bool eri = false) {
  ++g_jk_launches;
  const BasisDev &b = *static_cast<const BasisDev *>(v[0]);
  const PairsDev &p = *static_cast<const PairsDev *>(v[1]);
  const RysDev &R = *static_cast<const RysDev *>(v[2]);
  const QuartetSet &qs = *static_cast<const QuartetSet *>(v[3]);
  static const DensDev none{};
  const DensDev &dd = grad ? none : *static_cast<const DensDev *>(v[4]);
  const int na = ncart_h(la), nb = ncart_h(lb), nc = ncart_h(lc), nd = ncart_h(ld);
  std::ostringstream tag;
  tag << name.substr(0, name.find('(')) << " grid " << c.grid.x << " lds " << c.shm;
  const std::string L = tag.str();
  const size_t nsh = size_t(b.nsh), ncrt = size_t(b.nao_cart), n2 = ncrt * ncrt;
  // argument views
  need(L, "shell centres", b.xyz, nsh * 3 * sizeof(double));
  need(L, "shell momenta", b.l, nsh * sizeof(int));
  need(L, "shell kinds", b.kind, nsh * sizeof(int));
  need(L, "function offsets", b.off_func, nsh * sizeof(int));
  need(L, "Cartesian offsets", b.off_cart, nsh * sizeof(int));
  need(L, "shell atoms", b.atom, nsh * sizeof(int));
  need(L, "shell ranks", b.rank, nsh * sizeof(int));
  need(L, "shells by momentum", b.lshells, nsh * sizeof(int));
  need(L, "pair shells 1", p.sh1, size_t(p.n) * sizeof(int));
  need(L, "pair shells 2", p.sh2, size_t(p.n) * sizeof(int));
  need(L, "pair bounds", p.q, size_t(p.n) * sizeof(double));
  need(L, "primitive offsets", p.pp_off, size_t(p.n) * sizeof(int));
  need(L, "primitive counts", p.pp_n, size_t(p.n) * sizeof(int));
  need(L, "Boys table", R.boys, size_t(kBoysRows) * kBoysCols * sizeof(double));
  if (!inside(b.off_cart, nsh * sizeof(int)) || !inside(p.pp_off, size_t(p.n) * sizeof(int)) ||
      !inside(p.pp_n, size_t(p.n) * sizeof(int)) || !inside(p.sh1, size_t(p.n) * sizeof(int)) ||
      !inside(p.sh2, size_t(p.n) * sizeof(int)) || !inside(b.l, nsh * sizeof(int)) ||
      !inside(b.rank, nsh * sizeof(int)) || !inside(b.lshells, nsh * sizeof(int)))
    return;  // reported; the tables below cannot be read
  size_t nprim = 0;
  for (int k = 0; k < p.n; ++k)
    nprim = std::max(nprim, size_t(std::max(0, p.pp_off[k])) + size_t(std::max(0, p.pp_n[k])));
  for (const double *a : {p.zeta, p.px, p.py, p.pz, p.kc, p.alpha, p.beta, p.kz, p.i2z})
    need(L, "primitive-pair table", a, nprim * sizeof(double));
  if (!grad) {
    need(L, "Cartesian densities", dd.dc, size_t(dd.neff) * n2 * sizeof(double));
    need(L, "J accumulators", dd.jc, dd.copy_stride * size_t(dd.ncopy) * sizeof(double));
    if (dd.kc)
      need(L, "K accumulators", dd.kc, dd.copy_stride * size_t(dd.ncopy) * sizeof(double));
    if (dd.copy_stride < size_t(dd.neff) * n2)
      violation(L + ": accumulator copy stride too small");
    need(L, "effective-density table", dd.eff_src, 2 * size_t(dd.neff) * sizeof(int));
    need(L, "density bounds", dd.bound, size_t(dd.ndens) * nsh * nsh * sizeof(double));
  } else if (!eri) {
    const GradDev &gd = *static_cast<const GradDev *>(v[4]);
    const size_t nm = size_t(gd.ncoul + gd.nexch);
    int natom = 0;
    if (inside(b.atom, nsh * sizeof(int)))
      for (size_t k = 0; k < nsh; ++k) natom = std::max(natom, b.atom[k] + 1);
    if (gd.gamma) {
      need(L, "two-particle density", gd.gamma, n2 * n2 * sizeof(double));
    } else {
      need(L, "density factors", gd.fac, nm * sizeof(double));
      need(L, "density matrices", gd.mats, nm * 2 * n2 * sizeof(double));
      need(L, "density block maxima", gd.bmax, nm * 2 * nsh * nsh * sizeof(double));
    }
    need(L, "force accumulator", gd.out, size_t(natom) * 3 * sizeof(double));
  }
  if (qs.total <= 0) return;
  if (qs.x_off < 0 || qs.x_n < 1 || qs.x_off + qs.x_n > p.n || qs.y_off < 0 || qs.y_n < 1 ||
      qs.y_off + qs.y_n > p.n) {
    violation(L + ": class ranges outside the pair table");
    return;
  }
  // row-sum kernels: LDS of the staged Rys tables and the row sums
  const bool rowacc = name.find("jk_rowacc_kernel") != std::string::npos ||
                      name.find("jk_rowacc_batch_kernel") != std::string::npos;
  int W = 0, dbase = 0;
  if (rowacc) {
    const int nC = b.lshell_off[lc + 1] - b.lshell_off[lc];
    const int nD = b.lshell_off[ld + 1] - b.lshell_off[ld];
    W = nC * nc + (lc == ld ? 0 : nD * nd);
    dbase = lc == ld ? 0 : nC * nc;
    const int eng = rest.size() >= 4 ? std::atoi(rest[3].c_str()) : 0;
    const int nr = (la + lb + lc + ld) / 2 + 1;
    const size_t nrys = eng == 0 ? size_t(2) * R.nsub[nr] * nr * kRysTerms : 0;
    const size_t nacc = size_t(dd.neff) * (size_t(na) * nb + size_t(na + nb) * W);
    if ((nrys + nacc) * sizeof(double) > c.shm) {
      std::ostringstream s;
      s << L << ": row-sum kernel stages " << (nrys + nacc) * sizeof(double)
        << " bytes of LDS, the launch gives " << c.shm;
      violation(s.str());
    }
    if (qs.list) violation(L + ": explicit list on the row-sum kernel");
  }
  // every quartet, decoded as the row cursor does
  auto qfail = [&](long long t, const std::string &what) {
    std::ostringstream s;
    s << L << ", quartet " << t << ": " << what;
    violation(s.str());
  };
  const size_t nrow = size_t(qs.x_n) * size_t(qs.nseg);
  if (!qs.list) {
    if (qs.nseg < 1) return violation(L + ": no segments");
    // REPLAY_PLANT=rowptr: the check asks for one row pointer more than the
    // set holds, so a correct build reports it (the checker's own test)
    const char *plant = std::getenv("REPLAY_PLANT");
    const size_t extra = plant && std::strcmp(plant, "rowptr") == 0 ? 1 : 0;
    need(L, "row pointers", qs.rowptr, (nrow + 1 + extra) * sizeof(long long));
    need(L, "segment kets", qs.qbase, nrow * sizeof(int));
    if (!inside(qs.rowptr, (nrow + 1) * sizeof(long long)) || !inside(qs.qbase, nrow * sizeof(int)))
      return;
    if (qs.rowptr[0] != 0 || qs.rowptr[nrow] != qs.total)
      return violation(L + ": row pointers do not span the set");
    for (size_t s = 0; s < nrow; ++s)
      if (qs.rowptr[s + 1] < qs.rowptr[s]) return violation(L + ": row pointers decrease");
  } else {
    need(L, "quartet list", qs.list, size_t(qs.total) * sizeof(int2));
    if (!inside(qs.list, size_t(qs.total) * sizeof(int2))) return;
  }
  size_t sg = 0;
  for (long long t = 0; t < qs.total; ++t) {
    ++g_quartets;
    int P, Q;
    if (qs.list) {
      P = qs.list[t].x;
      Q = qs.list[t].y;
    } else {
      while (sg < nrow && qs.rowptr[sg + 1] <= t) ++sg;
      if (sg >= nrow) return qfail(t, "segment past the set");
      P = int(sg / size_t(qs.nseg));
      Q = qs.qbase[sg] + int(t - qs.rowptr[sg]);
    }
    if (P < 0 || P >= qs.x_n || Q < 0 || Q >= qs.y_n) return qfail(t, "pair index outside the class");
    const int pb = qs.x_off + P, pk = qs.y_off + Q;
    for (int pr : {pb, pk})
      if (p.pp_off[pr] < 0 || size_t(p.pp_off[pr]) + size_t(p.pp_n[pr]) > nprim)
        return qfail(t, "primitive pairs outside the table");
    const int ss[4] = {p.sh1[pb], p.sh2[pb], p.sh1[pk], p.sh2[pk]};
    const int ls[4] = {la, lb, lc, ld}, ns[4] = {na, nb, nc, nd};
    for (int i = 0; i < 4; ++i) {
      if (ss[i] < 0 || size_t(ss[i]) >= nsh) return qfail(t, "shell index outside the basis");
      if (b.l[ss[i]] != ls[i]) return qfail(t, "shell momentum differs from the class");
      if (b.off_cart[ss[i]] < 0 || size_t(b.off_cart[ss[i]] + ns[i]) > ncrt)
        return qfail(t, "Cartesian functions outside the matrices");
    }
    if (rowacc) {
      const int colc = b.rank[ss[2]] * nc, cold = dbase + b.rank[ss[3]] * nd;
      if (colc < 0 || colc + nc > (lc == ld ? W : dbase) || cold < 0 || cold + nd > W)
        return qfail(t, "row-sum column outside the LDS rows");
    }
  }
  // the row-sum kernel's workgroups: contiguous ranges, each starting its
  // row cursor with a binary search over the segments
  if (rowacc && !qs.list) {
    const long long grid = c.grid.x ? c.grid.x : 1;
    const long long chunk = (qs.total + grid - 1) / grid;
    for (long long g = 0; g < grid; ++g) {
      const long long begin = g * chunk;
      if (begin >= qs.total) break;
      int lo = 0, hi = int(nrow) - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (qs.rowptr[mid] <= begin)
          lo = mid;
        else
          hi = mid - 1;
      }
      if (!(qs.rowptr[lo] <= begin && begin < qs.rowptr[lo + 1]))
        qfail(begin, "row cursor's first segment does not hold the range start");
    }
  }
}

// The runtime-l J/K launches: the element-offset table (and for the
// per-quartet kernel the block slots) in scratch, and the dynamic LDS of
// their root batches and digest, as the kernels lay them out; then the
// views and quartets as for the compiled kernels.
void check_jk_runtime_l(const std::string &name, void **args, const Config &c) {
  const int la = *static_cast<int *>(args[0]), lb = *static_cast<int *>(args[1]),
            lc = *static_cast<int *>(args[2]), ld = *static_cast<int *>(args[3]);
  const int rb = *static_cast<int *>(args[4]);
  const bool sliced = name.find("jk_rt_kernel") != std::string::npos;
  const int na = ncart_h(la), nb = ncart_h(lb), nc = ncart_h(lc), nd = ncart_h(ld);
  const size_t n4 = size_t(na) * nb * nc * nd;
  const size_t nds = size_t(na) * nb + size_t(nc) * nd + size_t(na + nb) * (nc + nd);
  const size_t G = size_t(la + lb + 1) * (lc + ld + 1), T = size_t(la + 1) * (lb + 1) * (lc + 1) * (ld + 1);
  const size_t H = size_t(la + 1) * (lb + 1) * (lc + ld + 1);
  const int nr = (la + lb + lc + ld) / 2 + 1;
  std::ostringstream tag;
  tag << name.substr(0, name.find('(')) << " (" << la << lb << "|" << lc << ld << ") grid "
      << c.grid.x << " lds " << c.shm;
  const std::string L = tag.str();
  if (rb < 1 || rb > nr) violation(L + ": root batch outside 1..n roots");
  size_t lds_need;
  void **views;
  const unsigned long long *eoff;
  if (sliced) {
    const int nsl = *static_cast<int *>(args[5]);
    const size_t slice = std::min(n4, size_t(512) * 16);
    if (size_t(nsl) * 512 * 16 < n4) violation(L + ": slices do not cover the block");
    lds_need = std::max(size_t(rb) * 3 * (G + H + T), nds + slice) * sizeof(double);
    views = args + 6;
    eoff = *static_cast<const unsigned long long **>(args[12]);
  } else {
    lds_need = std::max(size_t(rb) * 3 * (G + T), nds) * sizeof(double);
    views = args + 5;
    eoff = *static_cast<const unsigned long long **>(args[11]);
    const double *slots = *static_cast<double **>(args[12]);
    const size_t slot = *static_cast<size_t *>(args[13]);
    if (slot < n4) violation(L + ": block slot smaller than the block");
    need(L, "block slots", slots, size_t(c.grid.x) * slot * sizeof(double));
  }
  need(L, "element-offset table", eoff, n4 * sizeof(unsigned long long));
  if (lds_need > c.shm) {
    std::ostringstream s;
    s << L << ": tables and digest need " << lds_need << " bytes of LDS";
    violation(s.str());
  }
  check_jk(name, la, lb, lc, ld, views, {}, c);
}

// The basis views the density and finishing kernels read.
void need_basis(const std::string &L, const BasisDev &b) {
  const size_t nsh = size_t(b.nsh);
  need(L, "shell momenta", b.l, nsh * sizeof(int));
  need(L, "shell kinds", b.kind, nsh * sizeof(int));
  need(L, "function offsets", b.off_func, nsh * sizeof(int));
  need(L, "Cartesian offsets", b.off_cart, nsh * sizeof(int));
  int lmax = 0;
  if (inside(b.l, nsh * sizeof(int)))
    for (size_t k = 0; k < nsh; ++k) lmax = std::max(lmax, b.l[k]);
  if (lmax > kMaxL) violation(L + ": shell momentum beyond the c2s tables");
  else {
    const int nc = (lmax + 1) * (lmax + 2) / 2;
    need(L, "c2s matrices", b.c2s, size_t(b.c2s_off[lmax] + (2 * lmax + 1) * nc) * sizeof(double));
  }
}

// Density preparation, fold and finishing kernels: every array they index,
// by the sizes their loops run over.
void check_dens_kernel(const std::string &name, void **args, const Config &c) {
  auto has = [&](const char *k) { return name.find(k) != std::string::npos; };
  const std::string L = name.substr(0, name.find('('));
  const size_t D = sizeof(double);
  if (has("split_kernel")) {  // (nao, d, ndens, has_anti, ds, da)
    const size_t nao = size_t(*static_cast<int *>(args[0]));
    const size_t nd = size_t(*static_cast<int *>(args[2]));
    const int anti = *static_cast<int *>(args[3]);
    need(L, "densities", *static_cast<double **>(args[1]), nd * nao * nao * D);
    need(L, "symmetric part", *static_cast<double **>(args[4]), nd * nao * nao * D);
    if (anti) need(L, "antisymmetric part", *static_cast<double **>(args[5]), nd * nao * nao * D);
  } else if (has("to_cart_kernel")) {  // (b, mf, nmats, mc)
    const BasisDev &b = *static_cast<const BasisDev *>(args[0]);
    const size_t nm = size_t(*static_cast<int *>(args[2]));
    need_basis(L, b);
    need(L, "function-space matrices", *static_cast<double **>(args[1]),
         nm * size_t(b.nao) * b.nao * D);
    need(L, "Cartesian matrices", *static_cast<double **>(args[3]),
         nm * size_t(b.nao_cart) * b.nao_cart * D);
  } else if (has("block_max_kernel")) {  // (b, x, nmats, out)
    const BasisDev &b = *static_cast<const BasisDev *>(args[0]);
    const size_t nm = size_t(*static_cast<int *>(args[2]));
    need_basis(L, b);
    need(L, "matrices", *static_cast<double **>(args[1]), nm * size_t(b.nao) * b.nao * D);
    need(L, "block maxima", *static_cast<double **>(args[3]), nm * size_t(b.nsh) * b.nsh * D);
  } else if (has("fold_copies_kernel")) {  // (x, n, stride, ncopy)
    const size_t n = *static_cast<size_t *>(args[1]), stride = *static_cast<size_t *>(args[2]);
    const int ncopy = *static_cast<int *>(args[3]);
    if (stride < n) violation(L + ": copy stride below the copy size");
    need(L, "accumulator copies", *static_cast<double **>(args[0]),
         (size_t(std::max(ncopy, 1) - 1) * stride + n) * D);
  } else if (has("jk_finish_kernel")) {  // (b, d, J, K)
    const BasisDev &b = *static_cast<const BasisDev *>(args[0]);
    const DensDev &d = *static_cast<const DensDev *>(args[1]);
    const size_t n2 = size_t(b.nao) * b.nao, nc2 = size_t(b.nao_cart) * b.nao_cart;
    need_basis(L, b);
    need(L, "J output", *static_cast<double **>(args[2]), size_t(d.ndens) * n2 * D);
    double *K = *static_cast<double **>(args[3]);
    if (d.kc && K) need(L, "K output", K, size_t(d.ndens) * n2 * D);
    need(L, "J accumulators", d.jc, size_t(d.neff) * nc2 * D);
    if (d.kc) need(L, "K accumulators", d.kc, size_t(d.neff) * nc2 * D);
    need(L, "effective-density table", d.eff_src, 2 * size_t(d.neff) * sizeof(int));
    if (inside(d.eff_src, 2 * size_t(d.neff) * sizeof(int)))
      for (int e = 0; e < d.neff; ++e)
        if (d.eff_src[e] < 0 || d.eff_src[e] >= d.ndens) violation(L + ": effective density source out of range");
  }
  (void)c;
}

// One launch of jk_rowacc_batch_kernel(bs, ps, R, sets, dds, seg, nb, queue,
// chunk, nacc, counters): its device arrays of per-basis views, read at the
// time of the launch, and every basis's views and quartets.
void check_batch(const std::string &L, void **args, const Config &c) {
  const BasisDev *bs = *static_cast<const BasisDev **>(args[0]);
  const PairsDev *ps = *static_cast<const PairsDev **>(args[1]);
  const RysDev &R = *static_cast<const RysDev *>(args[2]);
  const QuartetSet *sets = *static_cast<const QuartetSet **>(args[3]);
  const DensDev *dds = *static_cast<const DensDev **>(args[4]);
  const long long *seg = *static_cast<const long long **>(args[5]);
  const size_t nb = size_t(*static_cast<int *>(args[6]));
  const unsigned *queue = *static_cast<unsigned **>(args[7]);
  const long long chunk = *static_cast<long long *>(args[8]);
  const int nacc = *static_cast<int *>(args[9]);
  need(L, "basis views", bs, nb * sizeof(BasisDev));
  need(L, "pair views", ps, nb * sizeof(PairsDev));
  need(L, "quartet sets", sets, nb * sizeof(QuartetSet));
  need(L, "density views", dds, nb * sizeof(DensDev));
  need(L, "set prefix sums", seg, (nb + 1) * sizeof(long long));
  need(L, "work queue", queue, sizeof(unsigned));
  if (chunk < 1) violation(L + ": chunk below one quartet");
  if (!inside(bs, nb * sizeof(BasisDev)) || !inside(ps, nb * sizeof(PairsDev)) ||
      !inside(sets, nb * sizeof(QuartetSet)) || !inside(dds, nb * sizeof(DensDev)) ||
      !inside(seg, (nb + 1) * sizeof(long long)))
    return;
  if (queue && inside(queue, sizeof(unsigned)) && *queue != 0)
    violation(L + ": work queue not zero at the launch");
  int cls[4];
  std::vector<std::string> rest;
  if (!class_of(L, cls, rest)) return violation("unparsed batched kernel name " + L);
  if (seg[0] != 0) violation(L + ": prefix sums do not start at 0");
  for (size_t m = 0; m < nb; ++m) {
    if (seg[m + 1] - seg[m] != sets[m].total) {
      std::ostringstream o;
      o << L << ": basis " << m << " set total differs from the prefix sums";
      violation(o.str());
    }
    if (sets[m].total <= 0) continue;
    // the LDS the kernel zeroes and the row sums of this basis
    const BasisDev &b = bs[m];
    const int nc = ncart_h(cls[2]), nd = ncart_h(cls[3]);
    {
      const int nC = b.lshell_off[cls[2] + 1] - b.lshell_off[cls[2]];
      const int nD = b.lshell_off[cls[3] + 1] - b.lshell_off[cls[3]];
      const long long W = (long long)nC * nc + (cls[2] == cls[3] ? 0 : (long long)nD * nd);
      const long long need_acc =
          (long long)dds[m].neff * ((long long)ncart_h(cls[0]) * ncart_h(cls[1]) +
                                     (long long)(ncart_h(cls[0]) + ncart_h(cls[1])) * W);
      if (need_acc > nacc) {
        std::ostringstream o;
        o << L << ": basis " << m << " needs " << need_acc << " accumulator doubles, the launch "
          << "zeroes " << nacc;
        violation(o.str());
      }
    }
    void *v[5] = {const_cast<BasisDev *>(&bs[m]), const_cast<PairsDev *>(&ps[m]),
                  const_cast<RysDev *>(&R), const_cast<QuartetSet *>(&sets[m]),
                  const_cast<DensDev *>(&dds[m])};
    // the row-sum checks of the single-basis kernel (the launch's LDS holds
    // the Rys tables and nacc doubles, the widest basis's row sums)
    check_jk(L + " basis " + std::to_string(m), cls[0], cls[1], cls[2], cls[3], v, rest, c);
  }
}

void run_schwarz(void **args, bool big) {
  // schwarz_kernel(BasisDev, PairsDev, RysDev, int first, int count, double *q);
  // schwarz_big_kernel(int la, int lb, int rb, BasisDev, PairsDev, RysDev,
  // int first, int count, double *q, ...)
  const int o = big ? 3 : 0;
  const PairsDev &p = *static_cast<const PairsDev *>(args[o + 1]);
  const int first = *static_cast<int *>(args[o + 3]), count = *static_cast<int *>(args[o + 4]);
  double *q = *static_cast<double **>(args[o + 5]);
  // REPLAY_UNIT_SCHWARZ: every pair's bound is 1 (nothing screened; for
  // plans over bases the caller's matrix does not describe)
  const bool unit = std::getenv("REPLAY_UNIT_SCHWARZ") != nullptr;
  for (int t = 0; t < count; ++t) {
    const int pr = first + t;
    const int i = p.sh1[pr], j = p.sh2[pr];
    q[pr] = unit ? 1.0 : (i < g_nsh && j < g_nsh) ? g_schwarz[size_t(i) * g_nsh + j] : 0.0;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// the caller's side
extern "C" void replay_set_schwarz(const double *q, int nsh) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  g_schwarz.assign(q, q + size_t(nsh) * nsh);
  g_nsh = nsh;
}
extern "C" int replay_violation_count() { return int(g_viol.size()); }
extern "C" const char *replay_violation(int k) {
  return k >= 0 && k < int(g_viol.size()) ? g_viol[size_t(k)].c_str() : "";
}
extern "C" long long replay_jk_launches() { return g_jk_launches; }
extern "C" long long replay_quartets() { return g_quartets; }
// Failure injection: the k-th hipMalloc from now returns
// hipErrorOutOfMemory (k < 0: off); every launch of a kernel whose name
// contains `name` fails with hipErrorLaunchFailure, reported by the next
// hipGetLastError (null or empty: off).
extern "C" void replay_fail_malloc_after(long long k) { st().fail_malloc_after = k; }
extern "C" void replay_fail_launch(const char *name) { st().fail_launch = name ? name : ""; }
extern "C" long long replay_live_allocations() { return (long long)g_alloc.size(); }
extern "C" long long replay_malloc_calls() { return st().mallocs; }
extern "C" void replay_reset() {
  g_viol.clear();
  g_jk_launches = g_quartets = g_other = 0;
  st().mallocs = 0;
}

// ---------------------------------------------------------------------------
// runtime
// A call that real HIP refuses (or that would run outside the graph) while
// a stream captures in thread-local mode.
static void not_in_capture(const char *what) {
  if (st().capturing) violation(std::string(what) + " while a stream captures");
}

extern "C" hipError_t hipMalloc(void **ptr, size_t size) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  not_in_capture("hipMalloc");
  ++st().mallocs;
  if (st().fail_malloc_after >= 0 && st().fail_malloc_after-- == 0) {
    *ptr = nullptr;
    return hipErrorOutOfMemory;
  }
  // exactly n bytes (256-aligned), so a sanitizer's redzone starts at the
  // allocation's end
  const size_t n = size ? size : 8;
  void *p = nullptr;
  if (posix_memalign(&p, 256, n) != 0) return hipErrorOutOfMemory;
  std::memset(p, 0, n);
  g_alloc[reinterpret_cast<uintptr_t>(p)] = Alloc{n};
  *ptr = p;
  return hipSuccess;
}
extern "C" hipError_t hipFree(void *ptr) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  if (!ptr) return hipSuccess;
  not_in_capture("hipFree");
  auto it = g_alloc.find(reinterpret_cast<uintptr_t>(ptr));
  if (it == g_alloc.end()) {
    violation("hipFree of a pointer that is not a live allocation");
    return hipErrorInvalidValue;
  }
  g_alloc.erase(it);
  std::free(ptr);
  return hipSuccess;
}
extern "C" hipError_t hipMemcpy(void *dst, const void *src, size_t n, hipMemcpyKind) {
  not_in_capture("hipMemcpy");
  if (n) std::memmove(dst, src, n);
  return hipSuccess;
}
extern "C" hipError_t hipMemcpyAsync(void *dst, const void *src, size_t n, hipMemcpyKind,
                                     hipStream_t) {
  not_in_capture("hipMemcpyAsync (a copy node of host memory the graph keeps reading)");
  if (n) std::memmove(dst, src, n);
  return hipSuccess;
}
extern "C" hipError_t hipMemset(void *dst, int v, size_t n) {
  not_in_capture("hipMemset");
  std::memset(dst, v, n);
  return hipSuccess;
}
extern "C" hipError_t hipMemsetAsync(void *dst, int v, size_t n, hipStream_t) {
  not_in_capture("hipMemsetAsync (a memset node)");
  std::memset(dst, v, n);
  return hipSuccess;
}
extern "C" hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned int) {
  static int dummy[64];
  static int k = 0;
  *s = reinterpret_cast<hipStream_t>(&dummy[(k++) % 64]);
  return hipSuccess;
}
extern "C" hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
extern "C" hipError_t hipStreamSynchronize(hipStream_t) {
  not_in_capture("hipStreamSynchronize");
  return hipSuccess;
}
extern "C" hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned int) {
  return hipSuccess;
}
extern "C" hipError_t hipEventCreate(hipEvent_t *e) {
  *e = reinterpret_cast<hipEvent_t>(new int(0));
  return hipSuccess;
}
extern "C" hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned int) {
  *e = reinterpret_cast<hipEvent_t>(new int(0));
  return hipSuccess;
}
extern "C" hipError_t hipEventDestroy(hipEvent_t e) {
  delete reinterpret_cast<int *>(e);
  return hipSuccess;
}
extern "C" hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
extern "C" hipError_t hipEventSynchronize(hipEvent_t) {
  not_in_capture("hipEventSynchronize");
  return hipSuccess;
}
extern "C" hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
  *ms = 0.0f;
  return hipSuccess;
}
extern "C" hipError_t hipGetLastError(void) {
  const hipError_t e = st().last;
  st().last = hipSuccess;
  return e;
}
extern "C" const char *hipGetErrorString(hipError_t e) {
  return e == hipErrorOutOfMemory     ? "out of memory (injected)"
         : e == hipErrorLaunchFailure ? "launch failure (injected)"
                                      : "replay";
}
extern "C" hipError_t hipGetDeviceCount(int *n) {
  *n = 1;
  return hipSuccess;
}
extern "C" hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
extern "C" hipError_t hipDeviceSynchronize(void) {
  not_in_capture("hipDeviceSynchronize");
  return hipSuccess;
}
extern "C" hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int) {
  switch (a) {
    case hipDeviceAttributeMultiprocessorCount: {
      // REPLAY_CU_COUNT: fewer CUs, so persistent grids stay small when the
      // kernels are executed on the host
      const char *e = std::getenv("REPLAY_CU_COUNT");
      *v = e ? std::atoi(e) : 256;
      break;
    }
    case hipDeviceAttributeMaxSharedMemoryPerBlock: *v = 160 * 1024; break;
    case hipDeviceAttributeMaxSharedMemoryPerMultiprocessor: *v = 160 * 1024; break;
    default: *v = 0;
  }
  return hipSuccess;
}
extern "C" hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, const void *, int,
                                                                    size_t) {
  const char *e = std::getenv("REPLAY_PER_CU");
  *n = e ? std::atoi(e) : 2;
  return hipSuccess;
}
extern "C" hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  *p = const_cast<void *>(sym);
  return hipSuccess;
}
// graphs: a capture records the launches on the capturing stream; the graph
// is that list, its executable the same list, and each graph launch runs
// every recorded launch against the memory as it is then
static void dispatch(const std::string &name, const Config &c, void **args);
extern "C" hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  if (st().capturing) violation("a capture begun inside another");
  st().capturing = true;
  st().capture_stream = s;
  st().rec = new std::vector<Recorded>();
  return hipSuccess;
}
extern "C" hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *graph) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  if (!st().capturing || s != st().capture_stream) violation("hipStreamEndCapture on a stream that does not capture");
  st().capturing = false;
  st().capture_stream = nullptr;
  *graph = reinterpret_cast<hipGraph_t>(st().rec);
  st().rec = nullptr;
  return hipSuccess;
}
extern "C" hipError_t hipGraphInstantiate(hipGraphExec_t *exec, hipGraph_t graph,
                                          hipGraphNode_t *, char *, size_t) {
  // the executable shares the graph's list: a graph destroyed while its
  // executable lives is reported at the executable's next launch
  *exec = reinterpret_cast<hipGraphExec_t>(graph);
  st().live_graphs.insert(graph);
  return hipSuccess;
}
extern "C" hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  not_in_capture("hipGraphLaunch");
  ++st().graph_launches;
  auto *list = reinterpret_cast<std::vector<Recorded> *>(exec);
  if (!st().live_graphs.count(reinterpret_cast<hipGraph_t>(exec))) {
    violation("hipGraphLaunch of an executable whose graph is destroyed");
    return hipErrorInvalidValue;
  }
  for (Recorded &r : *list) {
    if (r.run) {
      r.run(r.job);
      continue;
    }
    if (!r.known) {
      ++g_other;
      continue;
    }
    dispatch(r.name, r.c, r.args.data());
  }
  return hipSuccess;
}
extern "C" hipError_t hipGraphDestroy(hipGraph_t graph) {
  if (!graph) return hipSuccess;
  st().live_graphs.erase(graph);
  for (Recorded &r : *reinterpret_cast<std::vector<Recorded> *>(graph))
    if (r.del) r.del(r.job);
  delete reinterpret_cast<std::vector<Recorded> *>(graph);
  return hipSuccess;
}
extern "C" hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
extern "C" long long replay_graph_launches() { return st().graph_launches; }

// kernel registration and launch
extern "C" void **__hipRegisterFatBinary(const void *) {
  static void *handle = nullptr;
  return &handle;
}
extern "C" void __hipUnregisterFatBinary(void **) {}
extern "C" void __hipRegisterFunction(void **, const void *host, char *, const char *dev,
                                      unsigned int, uint3 *, uint3 *, dim3 *, dim3 *, int *) {
  g_kname[host] = demangle(dev);
}
extern "C" void __hipRegisterVar(void **, void *var, char *, char *, int, size_t size, int, int) {
  // a __device__ variable: its host shadow stands in for the device copy
  g_alloc[reinterpret_cast<uintptr_t>(var)] = Alloc{size};
}
extern "C" hipError_t __hipPushCallConfiguration(dim3 grid, dim3 block, size_t shm,
                                                 hipStream_t s) {
  g_cfg.push_back(Config{grid, block, shm, s});
  return hipSuccess;
}
extern "C" hipError_t __hipPopCallConfiguration(dim3 *grid, dim3 *block, size_t *shm,
                                                hipStream_t *s) {
  Config c = g_cfg.empty() ? Config{} : g_cfg.back();
  if (!g_cfg.empty()) g_cfg.pop_back();
  *grid = c.grid;
  *block = c.block;
  *shm = c.shm;
  *s = c.stream;
  return hipSuccess;
}
// Argument sizes of the kernels a graph launch re-runs (empty: unknown).
static std::vector<size_t> arg_sizes(const std::string &name) {
  auto has = [&](const char *k) { return name.find(k) != std::string::npos; };
  const size_t P = sizeof(void *), I = sizeof(int), D = sizeof(double);
  const size_t B = sizeof(BasisDev), Pr = sizeof(PairsDev), R = sizeof(RysDev),
               Q = sizeof(QuartetSet), Dd = sizeof(DensDev);
  if (has("fill_kernel")) return {P, sizeof(size_t), D};
  if (has("split_kernel")) return {I, P, I, I, P, P};
  if (has("to_cart_kernel") || has("block_max_kernel")) return {B, P, I, P};
  if (has("fold_copies_kernel")) return {P, sizeof(size_t), sizeof(size_t), I};
  if (has("jk_finish_kernel")) return {B, sizeof(DensDev), P, P};
  if (has("jk_rowacc_batch_kernel<"))
    return {P, P, R, P, P, P, I, P, sizeof(long long), I, P};
  if (has("jk_rowacc_kernel<") || has("jk_reg_kernel<") || has("jk_kernel<") ||
      has("jk_hrr_kernel<"))
    return {B, Pr, R, Q, Dd, D, P};
  if (has("jk_big_kernel")) return {I, I, I, I, I, B, Pr, R, Q, Dd, D, P, P, sizeof(size_t), P};
  if (has("jk_rt_kernel<")) return {I, I, I, I, I, I, B, Pr, R, Q, Dd, D, P, P};
  return {};
}

extern "C" hipError_t hipLaunchKernel(const void *f, dim3 grid, dim3 block, void **args,
                                      size_t shm, hipStream_t s) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  auto it = g_kname.find(f);
  const std::string name = it == g_kname.end() ? std::string("?") : it->second;
  const Config c{grid, block, shm, s};
  if (!st().fail_launch.empty() && name.find(st().fail_launch) != std::string::npos) {
    st().last = hipErrorLaunchFailure;
    return hipErrorLaunchFailure;
  }
  if (st().capturing) {
    if (s != st().capture_stream) {
      // thread-local capture lets this run at once, outside the graph
      violation(name.substr(0, name.find('(')) + ": launched on another stream while a stream captures");
      dispatch(name, c, args);
      return hipSuccess;
    }
    Recorded r;
    r.name = name;
    r.c = c;
    const std::vector<size_t> sz = arg_sizes(name);
    r.known = !sz.empty();
    for (size_t k = 0; k < sz.size(); ++k) {
      const char *a = static_cast<const char *>(args[k]);
      r.store.emplace_back(a, a + sz[k]);
    }
    st().rec->push_back(std::move(r));
    // the argument pointers into the copies (stable once the list is built)
    Recorded &last = st().rec->back();
    for (auto &v : last.store) last.args.push_back(v.data());
    return hipSuccess;
  }
  dispatch(name, c, args);
  return hipSuccess;
}

static void dispatch(const std::string &name, const Config &c, void **args) {
  const dim3 block = c.block;
  const size_t shm = c.shm;
  if (block.x * block.y * block.z > 1024) violation(name + ": more than 1024 lanes");
  if (shm > 160 * 1024) violation(name + ": more than 160 KB of dynamic LDS");
  auto has = [&](const char *k) { return name.find(k) != std::string::npos; };
  if (has("schwarz_big_kernel")) {
    run_schwarz(args, true);
  } else if (has("schwarz_kernel<")) {
    run_schwarz(args, false);
  } else if (has("xcc_probe_kernel")) {
    unsigned *hist = *static_cast<unsigned **>(args[0]);
    for (int k = 0; k < 8; ++k) hist[k] = 512;
  } else if (has("asym_flag_kernel")) {
    int *flag = *static_cast<int **>(args[3]);
    *flag = 0;
  } else if (has("fill_kernel")) {
    double *x = *static_cast<double **>(args[0]);
    const size_t n = *static_cast<size_t *>(args[1]);
    const double v = *static_cast<double *>(args[2]);
    if (!inside(x, n * sizeof(double))) violation("fill_kernel outside every live allocation");
    else for (size_t k = 0; k < n; ++k) x[k] = v;
  } else if (has("split_kernel") || has("to_cart_kernel") || has("block_max_kernel") ||
             has("fold_copies_kernel") || has("jk_finish_kernel")) {
    check_dens_kernel(name, args, c);
  } else if (has("jk_rowacc_kernel<") || has("jk_reg_kernel<") || has("jk_kernel<") ||
             has("jk_hrr_kernel<")) {
    int cls[4];
    std::vector<std::string> rest;
    if (!class_of(name, cls, rest))
      violation("unparsed J/K kernel name " + name);
    else
      check_jk(name, cls[0], cls[1], cls[2], cls[3], args, rest, c);
  } else if (has("jk_rowacc_batch_kernel<")) {
    check_batch(name, args, c);
    // the kernel's workgroups drain the queue: every chunk taken once
    unsigned *queue = *static_cast<unsigned **>(args[7]);
    if (inside(queue, sizeof(unsigned))) *queue += c.grid.x;
  } else if (has("jk_big_kernel") || has("jk_rt_kernel<")) {
    check_jk_runtime_l(name, args, c);
  } else if (has("eri_kernel<") || has("eri_big_kernel")) {
    // eri_kernel<C>(b, p, R, qs, EriOut); eri_big_kernel(la, lb, lc, ld, rb,
    // b, p, R, qs, EriOut, eoff, scratch, slot): views, quartets and the
    // output buffer (sized by its mode, at least live)
    int la, lb, lc, ld;
    void **v;
    if (has("eri_big_kernel")) {
      la = *static_cast<int *>(args[0]);
      lb = *static_cast<int *>(args[1]);
      lc = *static_cast<int *>(args[2]);
      ld = *static_cast<int *>(args[3]);
      v = args + 5;
      const double *slots = *static_cast<double **>(args[11]);
      const size_t slot = *static_cast<size_t *>(args[12]);
      need(name, "ERI block slots", slots, size_t(c.grid.x) * slot * sizeof(double));
    } else {
      int cls[4];
      std::vector<std::string> rest;
      if (!class_of(name, cls, rest)) {
        violation("unparsed ERI kernel name " + name);
        return;
      }
      la = cls[0], lb = cls[1], lc = cls[2], ld = cls[3];
      v = args;
    }
    const EriOut &eo = *static_cast<const EriOut *>(v[4]);
    const BasisDev &eb = *static_cast<const BasisDev *>(v[0]);
    const size_t nao = size_t(eb.nao);
    // the dense tensor spans nao^4 (the packed and DF modes at least one
    // element: their extents depend on the builder's ranges)
    need(name, "ERI output", eo.out,
         (eo.mode == kEriDense ? nao * nao * nao * nao : size_t(1)) * sizeof(double));
    check_jk(name, la, lb, lc, ld, v, {}, c, true, true);
  } else if (has("grad_lane_kernel<")) {
    int cls[4];
    std::vector<std::string> rest;
    if (!class_of(name, cls, rest))
      violation("unparsed gradient kernel name " + name);
    else
      check_jk(name, cls[0], cls[1], cls[2], cls[3], args, {}, c, true);
  } else if (has("grad_big_kernel")) {
    // grad_big_kernel(la, lb, lc, ld, rb, b, p, R, qs, gd, scratch, slot, counters):
    // one slot per workgroup (the contracted block of the raised tables)
    const int la = *static_cast<int *>(args[0]), lb = *static_cast<int *>(args[1]),
              lc = *static_cast<int *>(args[2]), ld = *static_cast<int *>(args[3]);
    const double *slots = *static_cast<double **>(args[10]);
    const size_t slot = *static_cast<size_t *>(args[11]);
    need(name, "gradient block slots", slots, size_t(c.grid.x) * slot * sizeof(double));
    check_jk(name, la, lb, lc, ld, args + 5, {}, c, true);
  } else {
    ++g_other;
  }
}

// Hooks of the host-executed kernels (tests/emu/emu_core.cpp): a launch while
// the stream captures is recorded as a job that every graph launch runs.
extern "C" int replay_emu_capturing(hipStream_t s) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  if (!st().capturing) return 0;
  if (s != st().capture_stream) {
    violation("a launch on another stream while a stream captures");
    return 0;
  }
  return 1;
}
extern "C" void replay_emu_record(void (*run)(void *), void (*del)(void *), void *job,
                                  const char *name) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  Recorded r;
  r.name = name;
  r.run = run;
  r.del = del;
  r.job = job;
  st().rec->push_back(std::move(r));
}
extern "C" void replay_emu_violation(const char *what) {
  std::lock_guard<std::recursive_mutex> g(g_mu);
  violation(what);
}
