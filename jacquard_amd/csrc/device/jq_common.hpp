// Shared device helpers for the gfx950 kernels: grid barrier, error state,
// script-format constants.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "jacquard_hip.h"

namespace jqd {

constexpr int kThreads = 256;      // words per 8192-bit boomerang layer
constexpr u32 kSramWords = 8192;   // one AIGPDK SRAM (src/aigpdk.rs:22)

// Each kernel source owns its barrier slots and a sticky error word as
// file-scope device arrays.  Every launch takes its own slot; generations
// only grow and counters return to zero after every barrier, so no
// per-launch reset is needed.

// Host side: records the message returned by jq_last_error() for this
// thread; returns -1 (or 0 for an empty message).
int set_error(const char *msg);

// Host side: reads (and optionally clears) the batched kernels' error word.
int batched_error_flag(u32 *out, int clear);

// Host side: sets the batched kernels' wait bound (polls); 0 or -1.
int batched_set_wait_limit(u32 polls);

// Agent-coherent access to memory that other workgroups of the launch read
// or write (relaxed agent-scope atomics: on gfx950 loads and stores with the
// device-scope cache policy, which bypass the non-coherent cached copies).
__device__ __forceinline__ u32 ld_coherent(const u32 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent(u32 *p, u32 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until this wave's memory operations have completed.
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Grid-wide barrier among `n` co-resident workgroups whose inter-workgroup
// data moves only through ld_coherent / st_coherent.  Those accesses are
// coherent at agent scope by themselves, so the barrier needs ordering, not
// cache maintenance: every wave waits for its stores to complete, then the
// workgroup arrives; after the generation changes, the workgroup's coherent
// loads are issued and see every store made before any arrival.  An
// agent-scope release / acquire fence pair per workgroup (an L2 writeback
// and an L2 invalidate on gfx950) serialises in each XCD's L2 and cost about
// 2 us per barrier at 156 workgroups.
//
// Arrivals and polling are relaxed agent-scope atomics: an acquire on every
// poll would invalidate the XCD's L2 once per poll.  Waiters poll a
// generation word on its own 128-byte line, apart from the counters' lines.
// The wait is bounded (a limit in device memory, set by jq_wait_limit, read
// once per wait after the arrival).  A wait that gives up raises `err` and
// sets the slot's abort word; every other wait of the launch reads that
// word every 256 polls (not on entry: that would add a round trip to every
// barrier), returns once it is set and
// reports the abort, and the kernels then leave their cycle loops, so a
// launch whose workgroups are not all resident drains after one time-out
// instead of timing out at every barrier.  The word stays set until the
// host clears the error flag, which zeroes every slot of the kind
// (jq_device_error_flag).
//
// A slot is kCohSlotWords u32: the generation at [32], the abort word at
// [kAbortWord] and two counter sets
// at 256 and 2304, used by alternate barriers, so the last arrivers' resets
// need no wait (a set is reused two barriers later, after the resetting
// workgroup's next arrival has waited for its stores).  Words 64..87 are
// left to the co-simulation launch's flash state buffers.  A set holds a top
// counter at [0] and, for n >= 128, eight sub-counters at [128 * (1 + x)],
// x = member % 8 (per XCD when members are consecutive workgroups, which
// are dispatched round-robin over the 8 XCDs): one contended atomic per
// arrival costs ~10 ns at the coherence point, so 156 flat arrivals take
// ~1.5 us, two levels ~0.3 us.  Measured empty major stage at 156
// workgroups: 5.2 us with the fenced single-counter barrier, 2.0 us with
// this one.
constexpr int kCohSlots = 64;
constexpr int kCohSlotWords = 4096;
constexpr int kAbortWord = 36;

// Default bound of a wait, in polls (about 1-2 us each).
constexpr u32 kDefaultWaitPolls = 1u << 23;

// Marks the launch that owns `bar` aborted and raises `err`.
__device__ __forceinline__ void abort_launch(u32 *bar, u32 *err) {
  atomicOr(err, 1u);
  st_coherent(&bar[kAbortWord], 1u);
}

// One poll's bound check of a wait that started polling with limit `lim`
// (read from device memory once, after the arrival): true when the wait
// has taken more than lim polls (it then aborts the launch) or, checked
// every 256 polls, when the launch has aborted.
__device__ __forceinline__ bool wait_gives_up(u32 *bar, u32 spins, u32 lim, u32 *err) {
  if (spins > lim) {
    abort_launch(bar, err);
    return true;
  }
  return (spins & 255u) == 0 && ld_coherent(&bar[kAbortWord]) != 0;
}

// The arrival and wait of grid_sync_coherent, for callers that have already
// waited for their stores (wait_vmem) in every wave and may have loads in
// flight that the barrier must not wait for.  `member` is the workgroup's
// index among the n (0..n-1).  When the launch has aborted (this wait or
// another gave up), thread 0 sets the caller's LDS word `aborted` to 1
// before the closing workgroup barrier; it is never cleared here.
__device__ __forceinline__ void grid_arrive_wait(u32 *bar, u32 n, u32 member, u32 &gen,
                                                 u32 *err, const u32 *limit, u32 &aborted) {
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 *set = bar + 256 + ((gen + 1u) & 1u) * 2048;
    bool last = false;
    if (n >= 128) {
      const u32 x = member & 7u;
      const u32 nx = (n - x + 7u) >> 3;  // members with index % 8 == x
      u32 a = __hip_atomic_fetch_add(&set[128 * (1 + x)], 1u, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
      if (a == nx - 1u) {
        __hip_atomic_store(&set[128 * (1 + x)], 0u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        u32 top = __hip_atomic_fetch_add(&set[0], 1u, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        last = top == 7u;
      }
    } else {
      u32 a = __hip_atomic_fetch_add(&set[0], 1u, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
      last = a == n - 1u;
    }
    if (last) {
      __hip_atomic_store(&set[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(&bar[32], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const u32 lim = *limit;
      u32 spins = 0;
      while (__hip_atomic_load(&bar[32], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(1);
        if (wait_gives_up(bar, ++spins, lim, err)) {
          aborted = 1;
          break;
        }
      }
    }
    gen += 1;
  }
  asm volatile("" ::: "memory");
  __syncthreads();
}

__device__ __forceinline__ void grid_sync_coherent(u32 *bar, u32 n, u32 member, u32 &gen,
                                                   u32 *err, const u32 *limit, u32 &aborted) {
  wait_vmem();
  grid_arrive_wait(bar, n, member, gen, err, limit, aborted);
}

__device__ __forceinline__ u32 read_gen(const u32 *bar) {
  return __hip_atomic_load(&bar[32], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}

// One AND-with-invert word of the boomerang tree
// (csrc/kernel_v1_impl.cuh:209-211).
__device__ __forceinline__ u32 gate(u32 a, u32 b, u32 xa, u32 xb, u32 ob) {
  return (a ^ xa) & ((b ^ xb) | ob);
}

// Its X lane (kernel_v1_impl.cuh:214-219).
__device__ __forceinline__ u32 gate_x(u32 a, u32 ax, u32 b, u32 bx, u32 xa,
                                      u32 xb, u32 ob) {
  u32 ae = a ^ xa, be = (b ^ xb) | ob, bex = bx & ~ob;
  return (ax | bex) & (ae | ax) & (be | bex);
}

}  // namespace jqd
