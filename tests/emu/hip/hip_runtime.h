// Host emulation of the HIP device features used by the tiled kernel
// (fastjet_gpu_amd/csrc/tiled.hip), for the CPU test suite: one 64-lane
// wavefront is 64 fibers (tests/emu/wave_fibers.cpp) that run the kernel source
// and meet at every cross-lane operation (readlane, readfirstlane, ballot, DPP,
// fences), where the operation's result is computed from all lanes' operands.
// Between those points the lanes run one after another, which matches the
// hardware for code whose lanes exchange data only through those operations
// and wavefront-ordered memory.  Test infrastructure only.
#pragma once
#include <algorithm>
#include <atomic>
#include <barrier>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __constant__

using std::max;
using std::min;

struct alignas(16) int4 { int x, y, z, w; };
struct alignas(16) uint4 { unsigned x, y, z, w; };
struct alignas(16) double2 { double x, y; };
struct alignas(32) double4 { double x, y, z, w; };
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }
inline double2 make_double2(double a, double b) { return double2{a, b}; }
inline double4 make_double4(double a, double b, double c, double d) { return double4{a, b, c, d}; }

namespace emu {
/* Lanes are fibers on one OS thread, 64 per wavefront, one or more wavefronts
 * per workgroup.  A lane runs until it reaches a cross-lane operation (or the
 * end of the kernel), then parks; when every live lane of its wavefront is
 * parked at the same operation, the results are computed from that
 * wavefront's operands and its lanes resume in order 0..63.  A workgroup
 * barrier waits for every wavefront. */
extern int lane_cur; /* lane within the wavefront */
extern int wave_cur; /* wavefront within the workgroup */
void park(uint64_t a, uint64_t b, int op); /* stores the lane's operands, yields */
const uint64_t *wave_ra();                   /* the current wavefront's operands at its last operation */
const uint64_t *wave_rb();
inline void sync() { park(0, 0, 1); }
inline void all(uint64_t v, uint64_t *out) {
  park(v, 0, 2);
  std::memcpy(out, wave_ra(), 64 * sizeof(uint64_t));
}
inline void all2(uint64_t v, uint64_t w, uint64_t *o1, uint64_t *o2) {
  park(v, w, 3);
  std::memcpy(o1, wave_ra(), 64 * sizeof(uint64_t));
  std::memcpy(o2, wave_rb(), 64 * sizeof(uint64_t));
}
inline void fence() { sync(); }
inline void barrier() { park(0, 0, 4); }
}  // namespace emu

struct EmuThreadIdx {
  int x, y, z;
};
inline EmuThreadIdx emu_tid() { return EmuThreadIdx{emu::wave_cur * 64 + emu::lane_cur, 0, 0}; }
#define threadIdx emu_tid()

inline int __builtin_amdgcn_readfirstlane(int v) {
  uint64_t o[64];
  emu::all((uint32_t)v, o);
  return (int)(uint32_t)o[0];
}
inline int __builtin_amdgcn_readlane(int v, int l) {
  uint64_t o[64];
  emu::all((uint32_t)v, o);
  if (l < 0 || l > 63) { std::fprintf(stderr, "readlane lane %d\n", l); std::abort(); }
  return (int)(uint32_t)o[l];
}
inline unsigned long long __ballot(int p) {
  uint64_t o[64];
  emu::all(p ? 1 : 0, o);
  unsigned long long m = 0;
  for (int i = 0; i < 64; ++i) m |= (o[i] ? 1ull : 0ull) << i;
  return m;
}
inline int __builtin_amdgcn_ds_bpermute(int addr, int src) { /* src of lane (addr / 4) mod 64 */
  uint64_t oa[64], os[64];
  emu::all2((uint32_t)addr, (uint32_t)src, oa, os);
  const int s = (int)(((uint32_t)oa[emu::lane_cur] >> 2) & 63u);
  return (int)(uint32_t)os[s];
}
inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned m, unsigned add) {
  const int l = emu::lane_cur;
  return add + (unsigned)__builtin_popcount(l < 32 ? (m & ((1u << l) - 1u)) : m);
}
inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned m, unsigned add) {
  const int l = emu::lane_cur;
  return add + (l < 32 ? 0u : (unsigned)__builtin_popcount(m & ((l - 32 == 0) ? 0u : ((1u << (l - 32)) - 1u))));
}
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound) {
  uint64_t os[64], ss[64];
  emu::all2((uint32_t)old, (uint32_t)src, os, ss);
  const int i = emu::lane_cur;
  const int row = i >> 4, bank = (i & 15) >> 2;
  if (!((row_mask >> row) & 1) || !((bank_mask >> bank) & 1)) return old;
  int s = -1;
  if (ctrl >= 0 && ctrl <= 0xFF) {  // quad_perm
    const int sel = (ctrl >> (2 * (i & 3))) & 3;
    s = (i & ~3) | sel;
  } else if (ctrl >= 0x111 && ctrl <= 0x11F) {  // row_shr
    const int n = ctrl - 0x110;
    s = ((i & 15) >= n) ? i - n : -1;
  } else if (ctrl == 0x141) {  // row_half_mirror
    s = (i & ~7) | (7 - (i & 7));
  } else if (ctrl == 0x140) {  // row_mirror
    s = (i & ~15) | (15 - (i & 15));
  } else if (ctrl == 0x142) {  // row_bcast15
    s = row >= 1 ? 16 * row - 1 : -1;
  } else if (ctrl == 0x143) {  // row_bcast31
    s = row >= 2 ? 31 : -1;
  } else {
    std::fprintf(stderr, "dpp ctrl 0x%x not emulated\n", ctrl);
    std::abort();
  }
  if (s < 0) return bound ? 0 : old;
  return (int)(uint32_t)ss[s];
}
#define __builtin_amdgcn_s_sleep(n) std::this_thread::yield()
#define __builtin_amdgcn_fence(o, s) emu::fence()
inline void __syncthreads() { emu::barrier(); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_system() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline unsigned long long wall_clock64() {
  return (unsigned long long)std::chrono::steady_clock::now().time_since_epoch().count() / 10;
}
inline unsigned long long clock64() { return wall_clock64(); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline double __hiloint2double(int hi, int lo) {
  const uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}
inline int __double2hiint(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return (int)(uint32_t)(u >> 32);
}
inline int __double2loint(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return (int)(uint32_t)u;
}
inline double __longlong_as_double(long long x) {
  double d;
  std::memcpy(&d, &x, 8);
  return d;
}
inline long long __double_as_longlong(double d) {
  long long x;
  std::memcpy(&x, &d, 8);
  return x;
}
template <class T>
inline T atomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicSub(T *p, T v) { return __atomic_fetch_sub(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicOr(T *p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicExch(T *p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicCAS(T *p, T cmp, T v) {
  __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST);
  return cmp;
}
template <class T>
inline T atomicMin(T *p, T v) {
  T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return cur;
}
template <class T>
inline T atomicMax(T *p, T v) {
  T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v > cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
  return cur;
}
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __HIP_MEMORY_SCOPE_SYSTEM 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 0
#define __hip_atomic_load(p, o, s) __atomic_load_n(p, __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, o, s) __atomic_store_n(p, v, __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_max(p, v, o, s) atomicMax(p, v)
