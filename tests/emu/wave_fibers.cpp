// Fiber scheduler for the wavefront emulation of tests/emu/hip/hip_runtime.h
// (x86-64 System V): 64 lane fibers on one thread, switched by a minimal
// register-save routine; AddressSanitizer is told about every stack switch.
// Test infrastructure only.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>

extern "C" void emu_switch(void **save_sp, void *new_sp);
asm(R"(
.text
.globl emu_switch
.type emu_switch,@function
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
)");

namespace emu {
constexpr int kMaxLanes = 16 * 64; /* up to 16 wavefronts per workgroup */
int lane_cur = 0, wave_cur = 0;
static int fib_cur = 0;
static uint64_t ra_all[kMaxLanes], rb_all[kMaxLanes];
static uint64_t pa[kMaxLanes], pb[kMaxLanes];
static int pop[kMaxLanes];
static bool done[kMaxLanes];
static void *lane_sp[kMaxLanes];
static void *sched_sp;
static char *stacks[kMaxLanes];
static const size_t kStack = 1 << 20;
static std::function<void()> body;
static const void *sched_bottom = nullptr;
static size_t sched_size = 0;

const uint64_t *wave_ra() { return ra_all + 64 * wave_cur; }
const uint64_t *wave_rb() { return rb_all + 64 * wave_cur; }

static void to_sched() {
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, sched_bottom, sched_size);
  emu_switch(&lane_sp[fib_cur], sched_sp);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
}

void park(uint64_t a, uint64_t b, int op) {
  pa[fib_cur] = a;
  pb[fib_cur] = b;
  pop[fib_cur] = op;
  to_sched();
}

extern "C" void emu_fiber_entry() {
  __sanitizer_finish_switch_fiber(nullptr, &sched_bottom, &sched_size);
  body();
  done[fib_cur] = true;
  pop[fib_cur] = -1;
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(nullptr, sched_bottom, sched_size);
  emu_switch(&lane_sp[fib_cur], sched_sp);
  std::abort();
}

static void resume(int f) {
  fib_cur = f;
  wave_cur = f / 64;
  lane_cur = f % 64;
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, stacks[f], kStack);
  emu_switch(&sched_sp, lane_sp[f]);
  __sanitizer_finish_switch_fiber(fake, &sched_bottom, &sched_size);
}

/* Runs fn() as nwaves wavefronts of 64 lanes: the lanes of a wavefront in lock
 * step at cross-lane operations, all wavefronts at workgroup barriers. */
void run_block(int nwaves, const std::function<void()> &fn) {
  if (nwaves < 1 || nwaves * 64 > kMaxLanes) std::abort();
  body = fn;
  const int nf = nwaves * 64;
  for (int f = 0; f < nf; ++f) {
    if (!stacks[f]) stacks[f] = static_cast<char *>(std::aligned_alloc(64, kStack));
    done[f] = false;
    char *top = stacks[f] + kStack;
    top = reinterpret_cast<char *>(reinterpret_cast<uintptr_t>(top) & ~(uintptr_t)15);
    void **sp = reinterpret_cast<void **>(top);
    *--sp = nullptr;                                       /* alignment slot */
    *--sp = reinterpret_cast<void *>(&emu_fiber_entry);    /* popped by ret */
    for (int k = 0; k < 6; ++k) *--sp = nullptr;           /* rbp rbx r12-r15 */
    lane_sp[f] = sp;
  }
  bool finished[16] = {false}, at_barrier[16] = {false};
  for (;;) {
    for (int w = 0; w < nwaves; ++w) {
      if (finished[w] || at_barrier[w]) continue;
      for (int l = 0; l < 64; ++l)
        if (!done[64 * w + l]) resume(64 * w + l);
      int live = 0, op = -2;
      bool mixed = false;
      for (int l = 0; l < 64; ++l) {
        const int f = 64 * w + l;
        if (done[f]) continue;
        ++live;
        if (op == -2) op = pop[f];
        else if (pop[f] != op) mixed = true;
      }
      if (live == 0) {
        finished[w] = true;
        continue;
      }
      if (mixed || live != 64) {
        std::fprintf(stderr, "emu: lanes of wave %d diverged at a cross-lane operation (live %d)\n", w, live);
        std::abort();
      }
      if (op == 4) {
        at_barrier[w] = true;
        continue;
      }
      std::memcpy(ra_all + 64 * w, pa + 64 * w, 64 * sizeof(uint64_t));
      std::memcpy(rb_all + 64 * w, pb + 64 * w, 64 * sizeof(uint64_t));
    }
    int n_fin = 0, n_bar = 0;
    for (int w = 0; w < nwaves; ++w) {
      n_fin += finished[w];
      n_bar += at_barrier[w];
    }
    if (n_fin == nwaves) break;
    if (n_bar > 0 && n_bar + n_fin == nwaves) {
      if (n_fin) {
        std::fprintf(stderr, "emu: %d wave(s) wait at a barrier that finished waves never reach\n", n_bar);
        std::abort();
      }
      for (int w = 0; w < nwaves; ++w) at_barrier[w] = false;
    }
  }
}

/* Runs fn() as 64 lanes in lock step at cross-lane operations. */
void run_wave(const std::function<void()> &fn) { run_block(1, fn); }
}  // namespace emu
