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
int lane_cur = 0;
uint64_t ra[64], rb[64];
static uint64_t pa[64], pb[64];
static int pop[64];
static bool done[64];
static void *lane_sp[64];
static void *sched_sp;
static char *stacks[64];
static const size_t kStack = 1 << 20;
static std::function<void()> body;
static const void *sched_bottom = nullptr;
static size_t sched_size = 0;

static void to_sched() {
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, sched_bottom, sched_size);
  emu_switch(&lane_sp[lane_cur], sched_sp);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
}

void park(uint64_t a, uint64_t b, int op) {
  pa[lane_cur] = a;
  pb[lane_cur] = b;
  pop[lane_cur] = op;
  to_sched();
}

extern "C" void emu_fiber_entry() {
  __sanitizer_finish_switch_fiber(nullptr, &sched_bottom, &sched_size);
  body();
  done[lane_cur] = true;
  pop[lane_cur] = -1;
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(nullptr, sched_bottom, sched_size);
  emu_switch(&lane_sp[lane_cur], sched_sp);
  std::abort();
}

static void resume(int l) {
  lane_cur = l;
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, stacks[l], kStack);
  emu_switch(&sched_sp, lane_sp[l]);
  __sanitizer_finish_switch_fiber(fake, &sched_bottom, &sched_size);
}

/* Runs fn() as 64 lanes in lock step at cross-lane operations. */
void run_wave(const std::function<void()> &fn) {
  body = fn;
  for (int l = 0; l < 64; ++l) {
    if (!stacks[l]) stacks[l] = static_cast<char *>(std::aligned_alloc(64, kStack));
    done[l] = false;
    char *top = stacks[l] + kStack;
    top = reinterpret_cast<char *>(reinterpret_cast<uintptr_t>(top) & ~(uintptr_t)15);
    void **sp = reinterpret_cast<void **>(top);
    *--sp = nullptr;                                       /* alignment slot */
    *--sp = reinterpret_cast<void *>(&emu_fiber_entry);    /* popped by ret */
    for (int k = 0; k < 6; ++k) *--sp = nullptr;           /* rbp rbx r12-r15 */
    lane_sp[l] = sp;
  }
  for (;;) {
    for (int l = 0; l < 64; ++l)
      if (!done[l]) resume(l);
    int live = 0, op = -2;
    bool mixed = false;
    for (int l = 0; l < 64; ++l) {
      if (done[l]) continue;
      ++live;
      if (op == -2) op = pop[l];
      else if (pop[l] != op) mixed = true;
    }
    if (live == 0) break;
    if (mixed || live != 64) {
      std::fprintf(stderr, "emu: lanes diverged at a cross-lane operation (live %d)\n", live);
      std::abort();
    }
    std::memcpy(ra, pa, sizeof ra);
    std::memcpy(rb, pb, sizeof rb);
  }
}
}  // namespace emu
