// Fiber scheduler that runs the kernels' workgroups on the host: one
// workgroup at a time, its lanes as fibers of this thread.  Each lane runs
// until it reaches a wave operation, a workgroup barrier or its end; then
// the wave operations of every wavefront whose live lanes are all waiting
// are resolved (a subset when the wavefront diverged, as the hardware's
// exec mask would), and a barrier releases once every live lane is at it.
#include <sanitizer/common_interface_defs.h>
#include <sys/mman.h>

#include <cstdio>
#include <cstdlib>
#include <cxxabi.h>
#include <dlfcn.h>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "hip/hip_runtime.h"

emu::U3 threadIdx, blockIdx, blockDim, gridDim;

extern "C" void emu_switch(void **save, void *to);
asm(R"(
.text
.globl emu_switch
.type emu_switch, @function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size emu_switch, .-emu_switch
.section .note.GNU-stack,"",@progbits
.text
)");

// hooks of the host runtime stand-in (tests/replay/hip_replay.cpp)
extern "C" int replay_emu_capturing(hipStream_t s);
extern "C" void replay_emu_record(void (*run)(void *), void (*del)(void *), void *job,
                                  const char *name);
extern "C" void replay_emu_violation(const char *what);

namespace emu {
namespace {
enum State { kRun, kWave, kBar, kDone };
struct Lane {
  void *sp = nullptr;
  char *stack = nullptr;
  void *fake = nullptr;
  int state = kRun;
  int op = 0, arg = 0, width = 64;
  uint64_t val = 0, res = 0;
  unsigned tx = 0, ty = 0, tz = 0;
};
constexpr size_t kStack = 1 << 20;
std::vector<Lane> g_lanes;
Lane *g_cur = nullptr;
void *g_sched_sp = nullptr;
const void *g_sched_bottom = nullptr;
size_t g_sched_size = 0;
const std::function<void()> *g_body = nullptr;
char *g_lds = nullptr;
unsigned g_block_linear = 0;
std::map<std::string, long long> g_notes;
const Job *g_job = nullptr;

void note(const std::string &what) {
  const std::string k = (g_job ? g_job->name + ": " : std::string()) + what;
  if (g_notes[k]++ == 0) replay_emu_violation(k.c_str());
}

// lane -> scheduler
void to_sched() {
  Lane *L = g_cur;
  __sanitizer_start_switch_fiber(&L->fake, g_sched_bottom, g_sched_size);
  emu_switch(&L->sp, g_sched_sp);
  __sanitizer_finish_switch_fiber(L->fake, nullptr, nullptr);
}

void lane_entry() {
  __sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
  (*g_body)();
  g_cur->state = kDone;
  __sanitizer_start_switch_fiber(nullptr, g_sched_bottom, g_sched_size);
  emu_switch(&g_cur->sp, g_sched_sp);
  __builtin_unreachable();
}

void init_lane(Lane &L) {
  if (!L.stack) {
    void *m = mmap(nullptr, kStack, PROT_READ | PROT_WRITE,
                   MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (m == MAP_FAILED) {
      std::fprintf(stderr, "emu: fiber stack mmap failed\n");
      std::abort();
    }
    L.stack = static_cast<char *>(m);
  }
  uintptr_t top = (reinterpret_cast<uintptr_t>(L.stack) + kStack) & ~uintptr_t(15);
  void **sp = reinterpret_cast<void **>(top);
  *--sp = nullptr;                                   // return slot of lane_entry (alignment)
  *--sp = reinterpret_cast<void *>(&lane_entry);     // first ret lands here
  for (int i = 0; i < 6; ++i) *--sp = nullptr;       // rbp rbx r12 r13 r14 r15
  L.sp = sp;
  L.state = kRun;
  L.fake = nullptr;
}

void run_lane(Lane &L) {
  g_cur = &L;
  threadIdx = U3{L.tx, L.ty, L.tz};
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, L.stack, kStack);
  emu_switch(&g_sched_sp, L.sp);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
  g_cur = nullptr;
}

void resolve(std::vector<Lane *> &grp, int wave_base, int n) {
  // grp: lanes of one wavefront at the same operation (the active set)
  uint64_t vals[64];
  bool act[64] = {};
  for (Lane *L : grp) {
    const int li = int(L - g_lanes.data()) - wave_base;
    vals[li] = L->val;
    act[li] = true;
  }
  const int op = grp[0]->op;
  uint64_t mask = 0;
  bool any = false, all = true;
  for (int li = 0; li < 64; ++li)
    if (act[li]) {
      if (vals[li]) mask |= 1ull << li, any = true;
      else all = false;
    }
  for (Lane *L : grp) {
    const int self = int(L - g_lanes.data()) - wave_base;
    const int w = L->width;
    int src = self;
    switch (op) {
      case kShfl: src = (L->arg + (self & ~(w - 1))) & 63; break;
      case kDown: src = ((self & (w - 1)) + L->arg >= w) ? self : self + L->arg; break;
      case kXor: {
        const int idx = self ^ L->arg;
        src = idx >= ((self + w) & ~(w - 1)) ? self : idx;
        break;
      }
      case kBallot: L->res = mask; continue;
      case kAny: L->res = any; continue;
      case kAll: L->res = all; continue;
      default: L->res = 0; continue;
    }
    if (src < 0 || src >= 64 || !act[src] || wave_base + src >= n) {
      note("shuffle reads a lane outside the active set");
      L->res = 0;
    } else {
      L->res = vals[src];
    }
  }
}

// Advance wavefront w by one segment: its runnable lanes run to their next
// wave operation, barrier or end; then a wave operation every stopped live
// lane of it waits at is resolved (part of the wavefront when it diverged).
// Returns whether anything moved.
bool step_wave(int w, int n) {
  const int b0 = w * 64, b1 = std::min(n, b0 + 64);
  bool moved = false;
  for (int t = b0; t < b1; ++t)
    if (g_lanes[t].state == kRun) {
      run_lane(g_lanes[t]);
      moved = true;
    }
  int live = 0;
  std::vector<Lane *> grp;
  for (int t = b0; t < b1; ++t) {
    Lane &L = g_lanes[t];
    if (L.state != kDone) ++live;
    if (L.state == kWave) grp.push_back(&L);
  }
  if (grp.empty()) return moved;
  // one operation at a time: the first lane's; others wait
  std::vector<Lane *> same;
  for (Lane *L : grp)
    if (L->op == grp[0]->op && L->width == grp[0]->width) same.push_back(L);
  if (same.size() != grp.size()) note("different wave operations pending in one wavefront");
  if (int(same.size()) != live) note("wave operation with part of the wavefront (divergent)");
  resolve(same, b0, n);
  for (Lane *L : same) L->state = kRun;
  return true;
}

// EMU_SCHEDULE=random:SEED: each step advances one wavefront, drawn at random
// among those that can move, so wavefronts drift apart between barriers as
// on the hardware (default: every wavefront one segment per pass, in order).
std::mt19937_64 *schedule_rng() {
  static std::mt19937_64 *rng = [] {
    const char *e = std::getenv("EMU_SCHEDULE");
    if (!e || std::strncmp(e, "random:", 7) != 0) return (std::mt19937_64 *)nullptr;
    return new std::mt19937_64(std::strtoull(e + 7, nullptr, 10));
  }();
  return rng;
}

void run_block(const Job &J) {
  const int n = int(J.block.x * J.block.y * J.block.z);
  if (int(g_lanes.size()) < n) g_lanes.resize(n);
  for (int t = 0; t < n; ++t) {
    Lane &L = g_lanes[t];
    init_lane(L);
    L.tx = t % J.block.x;
    L.ty = (t / J.block.x) % J.block.y;
    L.tz = t / (J.block.x * J.block.y);
  }
  const int nw = (n + 63) / 64;
  std::mt19937_64 *rng = schedule_rng();
  std::vector<int> cand;
  for (;;) {
    bool moved = false;
    if (rng) {
      cand.clear();
      for (int w = 0; w < nw; ++w)
        for (int t = w * 64; t < std::min(n, w * 64 + 64); ++t)
          if (g_lanes[t].state == kRun || g_lanes[t].state == kWave) {
            cand.push_back(w);
            break;
          }
      if (!cand.empty()) moved = step_wave(cand[(*rng)() % cand.size()], n);
    } else {
      for (int w = 0; w < nw; ++w) moved = step_wave(w, n) || moved;
    }
    if (moved) continue;
    int live = 0, bar = 0;
    for (int t = 0; t < n; ++t) {
      if (g_lanes[t].state != kDone) ++live;
      if (g_lanes[t].state == kBar) ++bar;
    }
    if (live == 0) break;
    if (bar == live) {
      for (int t = 0; t < n; ++t)
        if (g_lanes[t].state == kBar) g_lanes[t].state = kRun;
      continue;
    }
    note("deadlock: lanes wait at different points");
    std::fprintf(stderr, "emu: deadlock in %s\n", J.name.c_str());
    std::abort();
  }
}

void run_job(void *p) {
  const Job &J = *static_cast<Job *>(p);
  g_job = &J;
  const unsigned long long lanes = 1ull * J.block.x * J.block.y * J.block.z;
  if (1ull * J.grid.x * J.grid.y * J.grid.z == 0 || lanes == 0) note("launch with an empty grid or block");
  if (lanes > 1024) note("more than 1024 lanes per workgroup");
  if (J.shm > 160 * 1024) note("more than 160 KB of dynamic LDS");
  g_body = &J.body;
  blockDim = U3{J.block.x, J.block.y, J.block.z};
  gridDim = U3{J.grid.x, J.grid.y, J.grid.z};
  for (unsigned z = 0; z < J.grid.z; ++z)
    for (unsigned y = 0; y < J.grid.y; ++y)
      for (unsigned x = 0; x < J.grid.x; ++x) {
        blockIdx = U3{x, y, z};
        g_block_linear = x + J.grid.x * (y + J.grid.y * z);
        // dynamic LDS: exactly the launch's bytes (so an access past them is
        // reported), filled with NaN bytes (LDS is not cleared on the GPU)
        g_lds = static_cast<char *>(std::malloc(J.shm ? J.shm : 1));
        std::memset(g_lds, 0xff, J.shm);
        run_block(J);
        std::free(g_lds);
        g_lds = nullptr;
      }
  g_job = nullptr;
}
void del_job(void *p) { delete static_cast<Job *>(p); }
}  // namespace

uint64_t wave_op(int op, uint64_t val, int arg, int width) {
  Lane *L = g_cur;
  L->state = kWave;
  L->op = op;
  L->val = val;
  L->arg = arg;
  L->width = width;
  to_sched();
  return L->res;
}
void block_barrier() {
  g_cur->state = kBar;
  to_sched();
}
void *dyn_lds() { return g_lds; }
// workgroups go round-robin over the 8 XCDs
unsigned xcc() { return g_block_linear % 8; }

void submit(Job *j, hipStream_t s) {
  // EMU_TRACE: each launch's kernel and shape on stderr
  static const bool trace = std::getenv("EMU_TRACE") != nullptr;
  if (trace)
    std::fprintf(stderr, "emu launch %s grid %u block %u lds %zu\n", j->name.c_str(), j->grid.x,
                 j->block.x, j->shm);
  if (replay_emu_capturing(s)) {
    replay_emu_record(&run_job, &del_job, j, j->name.c_str());
    return;
  }
  run_job(j);
  del_job(j);
}
}  // namespace emu
