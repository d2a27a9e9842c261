// Plain drop-in interpreter for gfx950 (wave64, 4 waves per workgroup).
//
// One workgroup per script block, persistent over all cycles and major
// stages, a grid barrier after every major stage (the contract of
// csrc/kernel_v1_impl.cuh:565-626) or, in value-only launches, a wait for
// just the blocks whose words the next stage reads.  Inside a partition the
// work is mapped to CDNA4 as follows:
//   - script sections are streamed with 16-byte per-thread loads (each
//     section is one fully coalesced 4 KiB request), the next boomerang
//     stage's permutation words are in flight while the current tree runs,
//     and a partition's prologue loads are issued before the barrier that
//     precedes it;
//   - snapshot words move between workgroups only through agent-coherent
//     loads and stores, so the grid barrier carries no cache maintenance;
//   - the 8192-bit shuffle is 32 LDS bit-gathers per thread on all 4 waves;
//   - the whole AND tree (hier 0..7 and the in-word tail) runs in wave 0
//     with permlane / DPP cross-lane moves, so a boomerang stage costs two
//     workgroup barriers instead of one per tree level;
//   - X masks and u16 arrivals ride in the same wave-0 pass when enabled.
// The computed values are bit-identical to the reference CPU interpreter
// (src/sim/cpu_reference.rs:17-674) and to the CUDA kernel's X, arrival and
// setup/hold rules, with the writeout arrival reset at every partition.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "jq_common.hpp"
#include "jq_cosim.hpp"

namespace jqd {

__device__ u32 g_barrier[kCohSlots * kCohSlotWords];
__device__ u32 g_error;
__device__ u32 g_wait_polls = kDefaultWaitPolls;  // bound of every wait here
// progress words of the dataflow-synchronised launches: per barrier slot,
// 256 words one 128-byte line apart (packed into 8 lines, every block's
// polls and the producers' stores met on the same lines, and the 1M-gate
// launch ran 20% slower than with grid barriers)
constexpr u32 kProgStride = 32;
__device__ u32 g_prog[kCohSlots * 256 * kProgStride];

namespace {

struct EvHdr {
  u32 count, overflow, r0, r1;
};
struct Ev {
  u32 type, msg, cycle, r;
  u32 data[4];
};

struct ScanArgs {
  u32 num_blocks, num_major_stages;
  const usize *blocks_start;
  const u32 *blocks_data;
  u32 *sram_data, *sram_xmask;
  u32 num_cycles;
  usize state_size;
  u32 *states;
  const u32 *timing;       // [clock_ps, per-word constraints...] or null
  const u32 *constraints;  // timing + 1, set on device
  u32 clock_ps;            // timing[0], set on device
  EvHdr *events;
  u32 ev_cap;
  i32 arrival_state_offset;
  u32 *bar;
  // dataflow synchronisation (value-only launches): the producer table of
  // build_flow_tables, the flag its builders raise when a script does not
  // fit it, and the blocks' progress words; dep == null: grid barriers
  const u32 *dep;
  const u32 *dep_bad;
  u32 *prog;
  // profiling
  const u32 *part_index;
  u32 num_parts;
  u32 *samples;
  u32 sample_cycles;
};

__device__ __forceinline__ uint4 ld4(const u32 *p) {
  return *reinterpret_cast<const uint4 *>(p);
}

// 32 LDS bit-gathers: word bit 2k (2k+1) <- src bit lo16 (hi16) of perm
// word k (src/flatten.rs:1027-1034).
__device__ __forceinline__ u32 gather_bits(const u32 *src, uint4 q0, uint4 q1,
                                           uint4 q2, uint4 q3) {
  u32 r = 0;
  const u32 w[16] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w,
                     q2.x, q2.y, q2.z, q2.w, q3.x, q3.y, q3.z, q3.w};
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    u32 i1 = w[k] & 0xffffu, i2 = w[k] >> 16;
    r |= ((src[i1 >> 5] >> (i1 & 31)) & 1u) << (2 * k);
    r |= ((src[i2 >> 5] >> (i2 & 31)) & 1u) << (2 * k + 1);
  }
  return r;
}

__device__ __forceinline__ u16 arr_step(u16 a, u16 b, u32 ob, u32 dly) {
  u16 base = (ob == 0xFFFFFFFFu) ? a : (a > b ? a : b);
  return (u16)(base + (u16)(dly & 0xffffu));
}

__device__ __forceinline__ void push_event(EvHdr *ev, u32 cap, u32 type,
                                           u32 word, u32 cycle, int slack,
                                           u32 arrival, u32 lim) {
  u32 k = atomicAdd(&ev->count, 1u);
  Ev *e = reinterpret_cast<Ev *>(ev + 1);
  if (k < cap) {
    e[k].type = type;
    e[k].msg = word;
    e[k].cycle = cycle;
    e[k].data[0] = word;
    e[k].data[1] = (u32)slack;
    e[k].data[2] = arrival;
    e[k].data[3] = lim;
  } else {
    atomicExch(&ev->overflow, 1u);
  }
}

// Header of one partition (src/flatten.rs:198-258) and this thread's
// writeout hook; `base` is the header's word offset in blocks_data.
struct PartHdr {
  size_t base;
  u32 num_stages, is_last, num_ios, io_offset, num_srams, sram_offset, rounds,
      num_dups, xoff, hook;
  bool is_x;
};

__device__ __forceinline__ PartHdr load_hdr(const u32 *__restrict__ bd, size_t off) {
  PartHdr h;
  const u32 *s = bd + off;
  h.base = off;
  h.num_stages = s[0];
  h.is_last = s[1];
  h.num_ios = s[2];
  h.io_offset = s[3];
  h.num_srams = s[4];
  h.sram_offset = s[5];
  h.rounds = s[6];
  h.num_dups = s[7];
  h.is_x = s[8] != 0;
  h.xoff = s[9];
  const u32 hook = s[128 + (threadIdx.x >> 1)];
  h.hook = (threadIdx.x & 1) ? (hook >> 16) : (hook & 0xffffu);
  return h;
}

__device__ __forceinline__ PartHdr no_hdr() {
  PartHdr h{};
  return h;  // num_stages == 0: nothing to run
}

// Words of one partition's script section (src/flatten.rs:873-1127).
__device__ __forceinline__ size_t part_words(const PartHdr &h) {
  return 256 + (size_t)512 * h.rounds + (size_t)5120 * (h.num_stages + 2);
}

// Everything a partition loads before its state gather: the global-read
// entries (up to kPreRounds rounds), the stage-0 permutation and wave 0's
// stage-0 tree flags, both writeout permutation sections and the commit's
// old values.  Script words are read-only and the old values are the
// partition's own writeouts of the previous cycle, so these loads are
// issued before the barrier that precedes the partition, and their latency
// hides behind the barrier wait (or the previous partition of the block).
constexpr u32 kPreRounds = 16;
struct PartPre {
  uint2 e[kPreRounds];
  uint4 q0, q1, q2, q3, fA, fB, fC, fD;
  uint4 sd0, sd1, sd2, sd3, sdf, ck0, ck1, ck2, ck3, ckf;
  u32 old_v, old_x;
};

__device__ __forceinline__ void load_pre(PartPre &P, const u32 *__restrict__ bd,
                                         const PartHdr &h, const u32 *in) {
  const int t = threadIdx.x;
  const int lane = t & 63;
  const size_t pi = h.base + 256;
  const uint2 *gr = reinterpret_cast<const uint2 *>(bd + pi);
#pragma unroll
  for (u32 k = 0; k < kPreRounds; ++k)
    P.e[k] = (k < h.rounds) ? gr[256 * k + t] : make_uint2(0, 0);
  const size_t p_st = pi + (size_t)512 * h.rounds;
  const size_t p_sd = p_st + (size_t)5120 * h.num_stages;
  const size_t p_ck = p_sd + 5120;
  P.q0 = ld4(bd + p_st + 4 * t);
  P.q1 = ld4(bd + p_st + 1024 + 4 * t);
  P.q2 = ld4(bd + p_st + 2048 + 4 * t);
  P.q3 = ld4(bd + p_st + 3072 + 4 * t);
  // wave 0's tree flags of stage 0 (words 0..255 of the flag block)
  P.fA = P.fB = P.fC = P.fD = make_uint4(0, 0, 0, 0);
  if (t < 64) {
    const u32 *fl0 = bd + p_st + 4096;
    P.fA = ld4(fl0 + 4 * lane);
    P.fB = ld4(fl0 + 4 * (64 + lane));
    P.fC = ld4(fl0 + 4 * (128 + lane));
    P.fD = ld4(fl0 + 4 * (192 + lane));
  }
  P.sd0 = ld4(bd + p_sd + 4 * t);
  P.sd1 = ld4(bd + p_sd + 1024 + 4 * t);
  P.sd2 = ld4(bd + p_sd + 2048 + 4 * t);
  P.sd3 = ld4(bd + p_sd + 3072 + 4 * t);
  P.sdf = ld4(bd + p_sd + 4096 + 4 * t);
  P.ck0 = ld4(bd + p_ck + 4 * t);
  P.ck1 = ld4(bd + p_ck + 1024 + 4 * t);
  P.ck2 = ld4(bd + p_ck + 2048 + 4 * t);
  P.ck3 = ld4(bd + p_ck + 3072 + 4 * t);
  P.ckf = ld4(bd + p_ck + 4096 + 4 * t);
  P.old_v = P.old_x = 0;
  if ((u32)t < h.num_ios) {
    P.old_v = ld_coherent(in + h.io_offset + (u32)t);
    if (h.is_x) P.old_x = ld_coherent(in + h.xoff + h.io_offset + (u32)t);
  }
}

// Value of lane L + HW on lane L (for the lanes where that stays inside
// the wave's 64 lanes, rows of 16 for HW < 16): v_permlane32_swap moves the
// upper half down, v_permlane16_swap odd rows to even rows, DPP row_shl
// shifts within a row.
template <int HW>
__device__ __forceinline__ u32 from_lane_up(u32 v) {
  if constexpr (HW == 32) {
    return __builtin_amdgcn_permlane32_swap(v, v, false, false)[1];
  } else if constexpr (HW == 16) {
    return __builtin_amdgcn_permlane16_swap(v, v, false, false)[1];
  } else {
    return __builtin_amdgcn_update_dpp(0u, v, 0x100 + HW, 0xF, 0xF, false);
  }
}

// One hier level of wave 0's tree (kernel_v1_impl.cuh:236-293): on entry
// lane i holds word 2HW + i (lanes i < 2HW); on exit lane i < HW holds word
// HW + i, also written to the layer.  Word w's flags are fA on lane w.
template <int HW, bool TIMING>
__device__ __forceinline__ void tree_level(u32 &cur, u32 &curx, u16 &cura, const uint4 &fA,
                                           bool is_x, int lane, u32 *st_v, u32 *st_x,
                                           u16 *st_arr) {
  const u32 b = from_lane_up<HW>(cur);
  const u32 xa = from_lane_up<HW>(fA.x), xb = from_lane_up<HW>(fA.y);
  const u32 ob = from_lane_up<HW>(fA.z);
  const u32 n = gate(cur, b, xa, xb, ob);
  u32 nx = 0;
  if (is_x) nx = gate_x(cur, curx, b, from_lane_up<HW>(curx), xa, xb, ob);
  u16 na = 0;
  if (TIMING) {
    const u32 dly = from_lane_up<HW>(fA.w);
    na = arr_step(cura, (u16)from_lane_up<HW>((u32)cura), ob, dly);
  }
  if (lane < HW) {
    st_v[HW + lane] = n;
    if (is_x) st_x[HW + lane] = nx;
    if (TIMING) st_arr[HW + lane] = na;
  }
  cur = n;
  curx = nx;
  cura = na;
}

// One partition (src/sim/cpu_reference.rs:25-262 per partition;
// kernel_v1_impl.cuh:62-558), its prologue loads already in P.
template <bool TIMING, bool PROFILE>
__device__ void run_partition(const ScanArgs &A, const u32 *__restrict__ bd,
                              const PartHdr &h, PartPre &P,
                              const u32 *__restrict__ in, u32 *__restrict__ out,
                              u32 cycle, u32 part, u32 *st_v, u32 *st_x,
                              u32 *hin_v, u32 *hin_x, u32 *wo_v, u32 *wo_x,
                              u16 *st_arr, u16 *wo_arr) {
  const int t = threadIdx.x;
  const int lane = t & 63;
  const bool w0 = t < 64;
  long long t_start = 0;
  if (PROFILE) t_start = clock64();
  const u32 num_stages = h.num_stages;
  const u32 num_ios = h.num_ios;
  const u32 io_offset = h.io_offset;
  const u32 num_srams = h.num_srams;
  const u32 sram_offset = h.sram_offset;
  const u32 rounds = h.rounds;
  const u32 num_dups = h.num_dups;
  const bool is_x = h.is_x;
  const u32 xoff = h.xoff;
  const u32 hook = h.hook;
  const u32 *script = bd + h.base;
  size_t pi = 256;
  const uint2 *gr = reinterpret_cast<const uint2 *>(script + pi);
  const uint2 *e = P.e;
  uint4 q0 = P.q0, q1 = P.q1, q2 = P.q2, q3 = P.q3;
  uint4 fA = P.fA, fB = P.fB, fC = P.fC, fD = P.fD;
  const uint4 sd0 = P.sd0, sd1 = P.sd1, sd2 = P.sd2, sd3 = P.sd3, sdf = P.sdf;
  const uint4 ck0 = P.ck0, ck1 = P.ck1, ck2 = P.ck2, ck3 = P.ck3, ckf = P.ckf;
  const u32 old_v = P.old_v, old_x = P.old_x;
  const size_t p_st = pi + (size_t)512 * rounds;

    // ---- global read (pext-gather), kernel_v1_impl.cuh:87-151: the state
    // words are gathered now, after the barrier that published them
    u32 acc_v = 0, acc_x = 0;
    {
      u32 v[kPreRounds], x[kPreRounds];
#pragma unroll
      for (u32 k = 0; k < kPreRounds; ++k) {
        v[k] = 0;
        x[k] = 0;
        if (e[k].y) {
          const u32 *base = (e[k].x >> 31) ? out : in;
          u32 w = e[k].x & 0x7fffffffu;
          v[k] = ld_coherent(base + w);
          if (is_x) x[k] = ld_coherent(base + xoff + w);
        }
      }
      // pext, lowest mask bit first into the accumulator (kernel_v1_impl.cuh:
      // 112-119): one ctz, one bit-field extract and one shift-or per bit;
      // the X lane's work only in X-capable partitions
      if (is_x) {
#pragma unroll
        for (u32 k = 0; k < kPreRounds; ++k)
          for (u32 m = e[k].y; m; m &= m - 1) {
            const u32 p = __builtin_ctz(m);
            acc_v = (acc_v << 1) | ((v[k] >> p) & 1u);
            acc_x = (acc_x << 1) | ((x[k] >> p) & 1u);
          }
      } else {
#pragma unroll
        for (u32 k = 0; k < kPreRounds; ++k)
          for (u32 m = e[k].y; m; m &= m - 1)
            acc_v = (acc_v << 1) | ((v[k] >> __builtin_ctz(m)) & 1u);
      }
    }
    for (u32 r = kPreRounds; r < rounds; ++r) {  // rare: very wide reads
      uint2 er = gr[256 * r + t];
      if (!er.y) continue;
      const u32 *base = (er.x >> 31) ? out : in;
      u32 w = er.x & 0x7fffffffu;
      u32 v = ld_coherent(base + w), x = is_x ? ld_coherent(base + xoff + w) : 0u;
      for (u32 m = er.y; m; m &= m - 1) {
        const u32 p = __builtin_ctz(m);
        acc_v = (acc_v << 1) | ((v >> p) & 1u);
        acc_x = (acc_x << 1) | ((x >> p) & 1u);
      }
    }
    pi = p_st;
    st_v[t] = acc_v;
    st_x[t] = acc_x;
    if (TIMING) wo_arr[t] = 0;  // pinned rule: reset per partition
    __syncthreads();

    // ---- boomerang stages
    for (u32 bs = 0; bs < num_stages; ++bs) {
      u32 hv = gather_bits(st_v, q0, q1, q2, q3);
      u32 hx = is_x ? gather_bits(st_x, q0, q1, q2, q3) : 0u;
      pi += 5120;
      // next stage's permutation and tree flags are loaded now, a stage
      // ahead of their use; this stage's came with the previous one
      uint4 nA = fA, nB = fB, nC = fC, nD = fD;
      if (bs + 1 < num_stages) {
        q0 = ld4(script + pi + 4 * t);
        q1 = ld4(script + pi + 1024 + 4 * t);
        q2 = ld4(script + pi + 2048 + 4 * t);
        q3 = ld4(script + pi + 3072 + 4 * t);
        if (w0) {
          const u32 *fn = script + pi + 4096;
          nA = ld4(fn + 4 * lane);
          nB = ld4(fn + 4 * (64 + lane));
          nC = ld4(fn + 4 * (128 + lane));
          nD = ld4(fn + 4 * (192 + lane));
        }
      }
      hin_v[t] = hv;
      hin_x[t] = hx;
      __syncthreads();
      if (w0) {
        // hier0: words 128+L, 192+L; hier1: word 64+L
        u32 s0 = hin_v[lane], s1 = hin_v[64 + lane];
        u32 s2 = hin_v[128 + lane], s3 = hin_v[192 + lane];
        u32 o2 = gate(s0, s2, fC.x, fC.y, fC.z);
        u32 o3 = gate(s1, s3, fD.x, fD.y, fD.z);
        u32 o1 = gate(o2, o3, fB.x, fB.y, fB.z);
        u32 o2x = 0, o3x = 0, o1x = 0;
        if (is_x) {
          u32 x0 = hin_x[lane], x1 = hin_x[64 + lane];
          u32 x2 = hin_x[128 + lane], x3 = hin_x[192 + lane];
          o2x = gate_x(s0, x0, s2, x2, fC.x, fC.y, fC.z);
          o3x = gate_x(s1, x1, s3, x3, fD.x, fD.y, fD.z);
          o1x = gate_x(o2, o2x, o3, o3x, fB.x, fB.y, fB.z);
        }
        u16 a2 = 0, a3 = 0, a1 = 0;
        if (TIMING) {
          a2 = arr_step(0, 0, fC.z, fC.w);
          a3 = arr_step(0, 0, fD.z, fD.w);
          a1 = arr_step(a2, a3, fB.z, fB.w);
        }
        st_v[128 + lane] = o2;
        st_v[192 + lane] = o3;
        st_v[64 + lane] = o1;
        if (is_x) {
          st_x[128 + lane] = o2x;
          st_x[192 + lane] = o3x;
          st_x[64 + lane] = o1x;
        }
        if (TIMING) {
          st_arr[128 + lane] = a2;
          st_arr[192 + lane] = a3;
          st_arr[64 + lane] = a1;
        }
        // hier2..7: the node of word hw + i (hw = 32, 16, ..., 1) is
        // computed on lane i from lanes i (word 2hw + i) and i + hw (word
        // 3hw + i); values and the node's flags (word L's on lane L in fA)
        // move up by hw with v_permlane32_swap, v_permlane16_swap and DPP
        // row shifts, a few cycles each, instead of LDS permutes
        u32 cur = o1, curx = o1x;
        u16 cura = a1;
        tree_level<32, TIMING>(cur, curx, cura, fA, is_x, lane, st_v, st_x, st_arr);
        tree_level<16, TIMING>(cur, curx, cura, fA, is_x, lane, st_v, st_x, st_arr);
        tree_level<8, TIMING>(cur, curx, cura, fA, is_x, lane, st_v, st_x, st_arr);
        tree_level<4, TIMING>(cur, curx, cura, fA, is_x, lane, st_v, st_x, st_arr);
        tree_level<2, TIMING>(cur, curx, cura, fA, is_x, lane, st_v, st_x, st_arr);
        tree_level<1, TIMING>(cur, curx, cura, fA, is_x, lane, st_v, st_x, st_arr);
        // in-word tail hier8..12 on lane 0 from word 1, now on lane 0
        // (kernel_v1_impl.cuh:294-346)
        if (lane == 0) {
          const u32 v1 = cur, v1x = curx;
          u32 xa = fA.x, xb = fA.y, ob = fA.z;
          u32 r8 = ((v1 << 16) ^ xa) & ((v1 ^ xb) | ob) & 0xffff0000u;
          u32 r9 = ((r8 >> 8) ^ xa) & (((r8 >> 16) ^ xb) | ob) & 0xff00u;
          u32 r10 = ((r9 >> 4) ^ xa) & (((r9 >> 8) ^ xb) | ob) & 0xf0u;
          u32 r11 = ((r10 >> 2) ^ xa) & (((r10 >> 4) ^ xb) | ob) & 0xcu;
          u32 r12 = ((r11 >> 1) ^ xa) & (((r11 >> 2) ^ xb) | ob) & 0x2u;
          st_v[0] = r8 | r9 | r10 | r11 | r12;
          if (is_x) {
            u32 x8 = gate_x(v1 << 16, v1x << 16, v1, v1x, xa, xb, ob) & 0xffff0000u;
            u32 x9 = gate_x(r8 >> 8, x8 >> 8, r8 >> 16, x8 >> 16, xa, xb, ob) & 0xff00u;
            u32 x10 = gate_x(r9 >> 4, x9 >> 4, r9 >> 8, x9 >> 8, xa, xb, ob) & 0xf0u;
            u32 x11 = gate_x(r10 >> 2, x10 >> 2, r10 >> 4, x10 >> 4, xa, xb, ob) & 0xcu;
            u32 x12 = gate_x(r11 >> 1, x11 >> 1, r11 >> 2, x11 >> 2, xa, xb, ob) & 0x2u;
            st_x[0] = x8 | x9 | x10 | x11 | x12;
          }
          if (TIMING) st_arr[0] = 0;
        }
      }
      __syncthreads();
      if ((hook >> 8) == bs) {
        wo_v[t] = st_v[hook & 255];
        if (is_x) wo_x[t] = st_x[hook & 255];
        if (TIMING) wo_arr[t] = st_arr[hook & 255];
      }
      fA = nA;
      fB = nB;
      fC = nC;
      fD = nD;
    }
    __syncthreads();

    // ---- SRAM + duplicate permutation, kernel_v1_impl.cuh:365-443
    const u32 nperm = num_srams * 4 + num_dups;
    pi += 5120;
    // partitions without SRAMs or duplicate words skip the phase (and
    // its two workgroup barriers); the condition is uniform
    if (nperm) {
      u32 pv = 0, px = 0;
      if ((u32)t < nperm) {
        pv = (gather_bits(wo_v, sd0, sd1, sd2, sd3) & ~sdf.y) ^ sdf.x;
        if (is_x) px = gather_bits(wo_x, sd0, sd1, sd2, sd3) & ~sdf.y;
      }
      const u32 en = __shfl(pv, (lane + 1) & 63);
      const u32 wdata = __shfl(pv, (lane + 2) & 63);
      const u32 wdata_x = __shfl(px, (lane + 2) & 63);
      __syncthreads();  // every gather of wo_* is done before it is written
      if ((u32)t < num_srams * 4) {
        if ((t & 3) == 0) {
          u32 s = (u32)t >> 2;
          size_t st = (size_t)sram_offset + (size_t)s * kSramWords;
          u32 ra = (pv & 0xffffu) & (kSramWords - 1);
          u32 wa = (pv >> 16) & (kSramWords - 1);
          u32 *ram = A.sram_data + st;
          u32 rv = ram[ra], w0v = ram[wa];
          ram[wa] = (w0v & ~en) | (wdata & en);
          wo_v[num_ios - num_srams + s] = rv;
          if (is_x) {
            u32 rx = 0;
            if (A.sram_xmask) {
              u32 *ramx = A.sram_xmask + st;
              rx = ramx[ra];
              u32 w0x = ramx[wa];
              ramx[wa] = (w0x & ~en) | (wdata_x & en);
            }
            wo_x[num_ios - num_srams + s] = rx;
          }
        }
      } else if ((u32)t < nperm) {
        u32 d = num_ios - num_srams - num_dups + ((u32)t - num_srams * 4);
        wo_v[d] = pv;
        if (is_x) wo_x[d] = px;
      }
      __syncthreads();
    }

    // ---- clock-enable permutation + commit, kernel_v1_impl.cuh:446-558
    {
      const uint4 f = ckf;
      if ((u32)t < num_ios) {
        // a word whose clock-enable bits are all forced (set0 all ones,
        // every combinational writeout) needs no gather: ck = inv, ckx = 0;
        // only words holding DFFs pay for the 32 LDS reads
        const bool gated = f.y != 0xFFFFFFFFu;
        u32 ck = f.x;
        if (gated) ck = (gather_bits(wo_v, ck0, ck1, ck2, ck3) & ~f.y) ^ f.x;
        u32 data = wo_v[t] ^ f.z;
        u32 w = io_offset + (u32)t;
        st_coherent(out + w, (old_v & ~ck) | (data & ck));
        if (is_x) {
          u32 ckx = 0;
          if (gated) ckx = gather_bits(wo_x, ck0, ck1, ck2, ck3) & ~f.y;
          st_coherent(out + xoff + w, (old_x & ~ck & ~ckx) | (wo_x[t] & ck & ~ckx) | ckx);
        }
        if (TIMING) {
          u16 arr = wo_arr[t];
          if (A.arrival_state_offset != 0)
            st_coherent(out + (u32)A.arrival_state_offset + w, (u32)arr);
          if (ck != 0 && A.constraints) {
            u32 cons = A.constraints[w];
            if (cons != 0) {
              u32 setup = cons >> 16, hold = cons & 0xffffu;
              if (arr > 0 && (u32)arr + setup > A.clock_ps)
                push_event(A.events, A.ev_cap, 4u, w, cycle,
                           (int)A.clock_ps - (int)arr - (int)setup, arr, setup);
              if ((u32)arr < hold)
                push_event(A.events, A.ev_cap, 5u, w, cycle,
                           (int)arr - (int)hold, arr, hold);
            }
          }
        }
      }
    }
    if (PROFILE) {
      __syncthreads();
      if (t == 0 && cycle < A.sample_cycles && part < A.num_parts)
        A.samples[(size_t)cycle * A.num_parts + part] = (u32)(clock64() - t_start);
    }
    __syncthreads();
}

// ---- dataflow synchronisation of value-only launches
//
// Partitions read snapshot words that other blocks wrote in an earlier
// major stage of the same cycle (`out`, bit 31 of a global-read entry) or
// in any stage of the previous cycle (`in`); every cycle has its own
// snapshot and SRAMs belong to one partition, so those reads are the only
// dependencies between blocks.  Instead of a grid barrier after every major
// stage, each block publishes its progress (c * M + m + 1 once it has
// finished major stage m of cycle c and its stores have completed) and,
// before a stage, waits only for the blocks it reads from, each up to the
// stage that wrote the words.  The producer table holds, per (block, major
// stage), one u32 per lane of wave 0: byte j of lane L is the producer
// block L + 64 j's offset d (its progress must reach c * M + d), or 0x80
// when the stage reads nothing of that block.  Results are those of the
// barrier schedule: every read still sees exactly the writes it saw there.
constexpr u32 kFlowMaxBlocks = 256;
constexpr u32 kFlowMaxStages = 120;
constexpr u32 kFlowNone = 0xFFFFFFFFu;
constexpr u32 kFlowMinCycles = 8;

struct FlowBuild {
  const usize *blocks_start;
  const u32 *bd;
  u32 nb, M;
  usize state_size;
  u32 *owner;  // state_size words: (block << 16 | major stage) of the writer
  u32 *dep;    // nb * M * 64 words
  u32 *bad;    // nonzero: the launch falls back to grid barriers
};

__global__ __launch_bounds__(256) void flow_fill_kernel(u32 *p, usize n, u32 v) {
  for (usize i = (usize)blockIdx.x * 256u + threadIdx.x; i < n; i += (usize)gridDim.x * 256u)
    p[i] = v;
}

// Workgroup (m, b) records the writer of every writeout word of block b's
// partitions in major stage m; a word written twice or out of range makes
// the table unusable.
__global__ __launch_bounds__(256) void flow_owner_kernel(FlowBuild F) {
  const u32 m = blockIdx.x / F.nb, b = blockIdx.x % F.nb;
  size_t off = F.blocks_start[(size_t)m * F.nb + b];
  for (;;) {
    const u32 *s = F.bd + off;
    const u32 ns = s[0], is_last = s[1], nio = s[2], io = s[3], rounds = s[6];
    if (ns == 0) break;
    if (threadIdx.x < nio) {
      const usize w = (usize)io + threadIdx.x;
      if (w >= F.state_size)
        atomicOr(F.bad, 1u);
      else if (atomicExch(&F.owner[w], (b << 16) | m) != kFlowNone)
        atomicOr(F.bad, 2u);
    }
    if (is_last) break;
    off += 256 + (size_t)512 * rounds + (size_t)5120 * (ns + 2);
  }
}

// Workgroup (m, b) reduces the global-read entries of block b's partitions
// in major stage m to the latest progress it needs of every other block.
__global__ __launch_bounds__(256) void flow_dep_kernel(FlowBuild F) {
  __shared__ u32 r[kFlowMaxBlocks];
  const u32 m = blockIdx.x / F.nb, b = blockIdx.x % F.nb;
  const u32 t = threadIdx.x;
  r[t] = 0;
  __syncthreads();
  size_t off = F.blocks_start[(size_t)m * F.nb + b];
  for (;;) {
    const u32 *s = F.bd + off;
    const u32 ns = s[0], is_last = s[1], rounds = s[6];
    if (ns == 0) break;
    const uint2 *gr = reinterpret_cast<const uint2 *>(s + 256);
    for (u32 k = 0; k < rounds; ++k) {
      const uint2 e = gr[256 * k + t];
      if (!e.y) continue;
      const u32 w = e.x & 0x7fffffffu;
      const bool from_out = (e.x >> 31) != 0;
      if (w >= F.state_size) {
        atomicOr(F.bad, 4u);
        continue;
      }
      const u32 o = F.owner[w];
      if (o == kFlowNone) continue;  // a primary input: written by the host
      const u32 q = o >> 16, sq = o & 0xffffu;
      if (q == b) continue;  // the block's own earlier work
      if (from_out && sq >= m) {  // not written before this stage
        atomicOr(F.bad, 8u);
        continue;
      }
      const int d = from_out ? (int)sq + 1 : (int)sq + 1 - (int)F.M;
      atomicMax(&r[q], (u32)(d + 128));
    }
    if (is_last) break;
    off += 256 + (size_t)512 * rounds + (size_t)5120 * (ns + 2);
  }
  __syncthreads();
  if (t < 64) {
    u32 word = 0;
#pragma unroll
    for (u32 j = 0; j < 4; ++j) {
      const u32 v = r[t + 64 * j];
      word |= (v ? ((v - 128u) & 0xffu) : 0x80u) << (8 * j);
    }
    F.dep[((size_t)b * F.M + m) * 64 + t] = word;
  }
}

// Publishes `done` as block b's progress once every wave has waited for its
// stores, then wave 0 polls the progress of the producers in `row` until
// each reaches base + its offset.  The wait is bounded like the grid
// barrier's (the poll limit, and the launch's abort word in `bar` read every
// 256 polls); an abort sets the caller's LDS word s_aborted to 1.
__device__ __forceinline__ void flow_publish_wait(u32 *prog, u32 b, u32 done, u32 row,
                                                  u32 base, u32 *bar, u32 *err,
                                                  u32 &s_aborted) {
  __syncthreads();
  if (threadIdx.x == 0) st_coherent(&prog[b * kProgStride], done);
  bool aborted = false;
  if (threadIdx.x < 64 && row != 0x80808080u) {
    const u32 lane = threadIdx.x;
    u32 tg[4];
    bool need[4];
#pragma unroll
    for (u32 j = 0; j < 4; ++j) {
      const u32 d = (row >> (8 * j)) & 0xffu;
      need[j] = d != 0x80u;
      tg[j] = base + (u32)(int)(signed char)d;
    }
    const u32 lim = g_wait_polls;
    u32 spins = 0;
    for (;;) {
      u32 p[4];
#pragma unroll
      for (u32 j = 0; j < 4; ++j) p[j] = need[j] ? ld_coherent(&prog[(lane + 64 * j) * kProgStride]) : 0u;
      bool pend = false;
#pragma unroll
      for (u32 j = 0; j < 4; ++j)
        if (need[j]) {
          if ((int)(p[j] - tg[j]) >= 0)
            need[j] = false;
          else
            pend = true;
        }
      if (!__any(pend)) break;
      __builtin_amdgcn_s_sleep(1);
      // the outcome is wave-uniform (every lane counts the same polls and
      // reads the same word); one lane raises the abort
      if (++spins > lim) {
        if (lane == 0) abort_launch(bar, err);
        aborted = true;
        break;
      }
      if ((spins & 255u) == 0 && ld_coherent(&bar[kAbortWord]) != 0) {
        aborted = true;
        break;
      }
    }
  }
  if (threadIdx.x == 0 && aborted) s_aborted = 1;
  asm volatile("" ::: "memory");
  __syncthreads();
}

template <bool TIMING, bool PROFILE>
__global__ __launch_bounds__(256) void scan_kernel(ScanArgs A) {
  if (A.timing) {
    A.clock_ps = A.timing[0];
    A.constraints = A.timing + 1;
  }
  __shared__ u32 st_v[256], st_x[256], hin_v[256], hin_x[256];
  __shared__ u32 wo_v[256], wo_x[256];
  __shared__ u16 st_arr[256], wo_arr[256];
  __shared__ u32 s_aborted;  // set by a wait when the launch has aborted
  const int t = threadIdx.x;
  if (t == 0) s_aborted = 0;
  st_x[t] = 0;
  hin_x[t] = 0;
  wo_x[t] = 0;
  wo_arr[t] = 0;
  st_arr[t] = 0;
  u32 gen = read_gen(A.bar);
  __syncthreads();
  const u32 b = blockIdx.x;
  const u32 nb = A.num_blocks, M = A.num_major_stages;
  const u32 *bd = A.blocks_data;
  // the block's partitions in execution order: cycle, major stage, then
  // position in the stage's section of blocks_data (kernel_v1_impl.cuh:
  // 565-626); h / P always describe the next partition to run
  // value-only launches whose producer table was built synchronise by
  // dataflow; the flag is the same word for every block
  const bool flow = !TIMING && !PROFILE && A.dep != nullptr && *A.dep_bad == 0;
  if (flow) {
    if (t == 0) st_coherent(&A.prog[b * kProgStride], 0u);
    grid_sync_coherent(A.bar, nb, b, gen, &g_error, &g_wait_polls, s_aborted);
  }
  PartHdr h = M ? load_hdr(bd, A.blocks_start[b]) : no_hdr();
  PartPre P;
  if (h.num_stages) load_pre(P, bd, h, A.states);
  for (u32 c = 0; c < A.num_cycles; ++c) {
    // an aborted launch (a bounded wait gave up; the waits that follow end
    // within 256 polls each) leaves at the next cycle
    if (s_aborted) return;
    const u32 *in = A.states + (size_t)c * A.state_size;
    u32 *out = A.states + (size_t)(c + 1) * A.state_size;
    for (u32 m = 0; m < M; ++m) {
      u32 part = PROFILE ? A.part_index[(size_t)m * nb + b] : 0u;
      const bool wrap = m + 1 == M;
      const u32 nc = wrap ? c + 1 : c, nm = wrap ? 0u : m + 1;
      // the next stage's producer row is read a stage ahead of its use
      u32 row = 0x80808080u;
      if (flow && t < 64 && nc < A.num_cycles)
        row = A.dep[((size_t)b * M + nm) * 64 + t];
      for (;;) {
        const bool last = h.num_stages == 0 || h.is_last != 0;
        // the successor's header is requested before this partition runs
        PartHdr hn = no_hdr();
        if (!last)
          hn = load_hdr(bd, h.base + part_words(h));
        else if (nc < A.num_cycles)
          hn = load_hdr(bd, A.blocks_start[(size_t)nm * nb + b]);
        if (h.num_stages) {
          run_partition<TIMING, PROFILE>(A, bd, h, P, in, out, c, part, st_v, st_x,
                                         hin_v, hin_x, wo_v, wo_x, st_arr, wo_arr);
          ++part;
        }
        if (last) {
          // this wave's stores complete before the next stage's prologue
          // loads are issued, so the barrier need not wait for those loads
          wait_vmem();
          if (hn.num_stages) load_pre(P, bd, hn, A.states + (size_t)nc * A.state_size);
          h = hn;
          break;
        }
        if (hn.num_stages) load_pre(P, bd, hn, in);
        h = hn;
      }
      if (flow)
        flow_publish_wait(A.prog, b, c * M + m + 1, row, nc * M, A.bar, &g_error,
                          s_aborted);
      else
        grid_arrive_wait(A.bar, nb, b, gen, &g_error, &g_wait_polls, s_aborted);
    }
  }
}
// Persistent co-simulation: C.num_edges edges of one cycle each over the
// two snapshots of C.states (timing events carry cycle 0, as every edge of
// the reference's co-simulation dispatch does, src/sim/cosim_metal.rs:
// 591-600), every workgroup doing its share of the
// between-edges work of jq_cosim.hpp ahead of every cycle (and once after
// the last).  A cycle
// walks the partitions as scan_kernel does, except that the first
// partition of a cycle is loaded after the barrier that follows the
// preparation, which may change any input word.
template <bool TIMING>
__global__ __launch_bounds__(256) void cosim_kernel(ScanArgs A, cosim::CosimArgs C) {
  if (A.timing) {
    A.clock_ps = A.timing[0];
    A.constraints = A.timing + 1;
  }
  __shared__ u32 st_v[256], st_x[256], hin_v[256], hin_x[256];
  __shared__ u32 wo_v[256], wo_x[256];
  __shared__ u16 st_arr[256], wo_arr[256];
  __shared__ u32 s_aborted;  // set by a wait when the launch has aborted
  const int t = threadIdx.x;
  if (t == 0) s_aborted = 0;
  st_x[t] = 0;
  hin_x[t] = 0;
  wo_x[t] = 0;
  wo_arr[t] = 0;
  st_arr[t] = 0;
  u32 gen = read_gen(A.bar);
  __syncthreads();
  const u32 b = blockIdx.x;
  const u32 nb = A.num_blocks, M = A.num_major_stages;
  const u32 *bd = A.blocks_data;
  const u32 *in = A.states;
  u32 *out = A.states + A.state_size;
  // co-simulation snapshots are 2-state: a partition flagged X-capable runs
  // its value lane only
  auto hdr2 = [&](size_t off) {
    PartHdr x = load_hdr(bd, off);
    x.is_x = false;
    return x;
  };
  for (u32 e = 0; e <= C.num_edges; ++e) {
    // an aborted launch (a bounded wait gave up) leaves at the next edge
    if (s_aborted) return;
    cosim::between_phase(C, (u32)t, b, nb, e);
    if (e == C.num_edges) break;
    grid_sync_coherent(A.bar, nb, b, gen, &g_error, &g_wait_polls, s_aborted);
    PartHdr h = M ? hdr2(A.blocks_start[b]) : no_hdr();
    PartPre P;
    if (h.num_stages) load_pre(P, bd, h, in);
    for (u32 m = 0; m < M; ++m) {
      const bool wrap = m + 1 == M;
      for (;;) {
        const bool last = h.num_stages == 0 || h.is_last != 0;
        PartHdr hn = no_hdr();
        if (!last)
          hn = hdr2(h.base + part_words(h));
        else if (!wrap)
          hn = hdr2(A.blocks_start[(size_t)(m + 1) * nb + b]);
        if (h.num_stages)
          run_partition<TIMING, false>(A, bd, h, P, in, out, 0u, 0u, st_v, st_x, hin_v,
                                       hin_x, wo_v, wo_x, st_arr, wo_arr);
        if (last) {
          wait_vmem();
          if (hn.num_stages) load_pre(P, bd, hn, in);
          h = hn;
          break;
        }
        if (hn.num_stages) load_pre(P, bd, hn, in);
        h = hn;
      }
      grid_arrive_wait(A.bar, nb, b, gen, &g_error, &g_wait_polls, s_aborted);
    }
  }
}

thread_local std::string t_err;
u32 g_slot = 0;

int fail(const std::string &m) {
  t_err = m;
  return -1;
}

}  // namespace

int set_error(const char *msg) {
  t_err = msg;
  return msg[0] ? -1 : 0;
}

namespace {

// 1: value-only launches synchronise by dataflow; 0: by grid barriers.
int g_scan_sync = 1;

template <bool TIMING, bool PROFILE>
int launch_scan(ScanArgs &A, hipStream_t stream) {
  static_assert(sizeof(EvHdr) == 16 && sizeof(Ev) == 32, "event layout");
  u32 *bar_base = nullptr;
  if (hipGetSymbolAddress((void **)&bar_base, HIP_SYMBOL(g_barrier)) != hipSuccess)
    return fail("hipGetSymbolAddress(g_barrier) failed");
  A.bar = bar_base + kCohSlotWords * (__atomic_fetch_add(&g_slot, 1u, __ATOMIC_RELAXED) % kCohSlots);
  if (A.dep) {
    u32 *pb = nullptr;
    if (hipGetSymbolAddress((void **)&pb, HIP_SYMBOL(g_prog)) != hipSuccess)
      return fail("hipGetSymbolAddress(g_prog) failed");
    A.prog = pb + (size_t)256 * kProgStride * ((A.bar - bar_base) / kCohSlotWords);
  }
  void *args[] = {&A};
  const void *kern = reinterpret_cast<const void *>(&scan_kernel<TIMING, PROFILE>);
  // Under stream capture (a hipGraph of the launch) the launch is an
  // ordinary kernel node: co-residency of the workgroups is checked here
  // from the occupancy query, and the bounded waits turn a shortfall at a
  // replay (other work holding CUs) into the error flag instead of a hang.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cap) != hipSuccess)
    return fail("hipStreamIsCapturing failed");
  hipError_t e;
  if (cap == hipStreamCaptureStatusActive) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 256, 0) != hipSuccess)
      return fail("occupancy query failed");
    if ((long long)per * cus < (long long)A.num_blocks)
      return fail("scan launch: more blocks than workgroups that can be resident at once");
    e = hipLaunchKernel(kern, dim3(A.num_blocks), dim3(256), args, 0, stream);
  } else {
    e = hipLaunchCooperativeKernel(kern, dim3(A.num_blocks), dim3(256), args, 0, stream);
  }
  if (e != hipSuccess)
    return fail(std::string("scan launch failed: ") + hipGetErrorString(e));
  return 0;
}

// Builds the producer table of a value-only launch on `stream` ahead of it
// and returns its stream-ordered allocation (freed after the launch), or
// null when the launch runs with grid barriers: the mode is off, the
// script is outside the table's limits, the launch is shorter than
// kFlowMinCycles (the table's ~30 us build is then comparable to the
// barriers it saves), the stream is being captured into a graph (no
// allocation or builder kernels go into the caller's graph), or the
// allocation failed.  A
// script that the builders find unusable (a writeout word written twice or
// out of range, a same-cycle read of a word not written in an earlier
// stage) raises the flag that the kernel reads, and the launch then uses
// grid barriers as well.
void *build_flow_tables(ScanArgs &A, hipStream_t stream) {
  A.dep = nullptr;
  A.dep_bad = nullptr;
  const u32 nb = A.num_blocks, M = A.num_major_stages;
  if (!__atomic_load_n(&g_scan_sync, __ATOMIC_RELAXED) || nb > kFlowMaxBlocks || M == 0 ||
      M > kFlowMaxStages || A.num_cycles < kFlowMinCycles || A.state_size == 0 ||
      A.state_size > (usize(1) << 28))
    return nullptr;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cap) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (cap != hipStreamCaptureStatusNone) return nullptr;
  const size_t own_words = ((size_t)A.state_size + 63) & ~(size_t)63;
  const size_t dep_words = (size_t)nb * M * 64;
  void *buf = nullptr;
  if (hipMallocAsync(&buf, 4 * (own_words + dep_words + 64), stream) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  FlowBuild F{};
  F.blocks_start = A.blocks_start;
  F.bd = A.blocks_data;
  F.nb = nb;
  F.M = M;
  F.state_size = A.state_size;
  F.owner = static_cast<u32 *>(buf);
  F.dep = F.owner + own_words;
  F.bad = F.dep + dep_words;
  const u32 fill_blocks = (u32)std::min<size_t>(1024, (own_words + 255) / 256);
  flow_fill_kernel<<<fill_blocks, 256, 0, stream>>>(F.owner, A.state_size, kFlowNone);
  flow_fill_kernel<<<1, 256, 0, stream>>>(F.bad, 1, 0u);
  flow_owner_kernel<<<nb * M, 256, 0, stream>>>(F);
  flow_dep_kernel<<<nb * M, 256, 0, stream>>>(F);
  if (hipGetLastError() != hipSuccess) {
    (void)hipFreeAsync(buf, stream);
    return nullptr;
  }
  A.dep = F.dep;
  A.dep_bad = F.bad;
  return buf;
}

int scan_common(usize num_blocks, usize num_major_stages, const usize *blocks_start,
                const u32 *blocks_data, u32 *sram_data, u32 *sram_xmask,
                usize num_cycles, usize state_size, u32 *states,
                const u32 *timing_constraints, u8 *event_buffer, u32 ev_cap,
                i32 arrival_state_offset, hipStream_t stream,
                const u32 *part_index, u32 num_parts, u32 *samples,
                u32 sample_cycles) {
  t_err.clear();
  if (num_blocks == 0 || num_blocks > 65535) return fail("num_blocks out of range");
  if (num_cycles == 0) return 0;
  if (num_cycles > 0xFFFFFFFFull) return fail("num_cycles out of range");
  ScanArgs A{};
  A.num_blocks = (u32)num_blocks;
  A.num_major_stages = (u32)num_major_stages;
  A.blocks_start = blocks_start;
  A.blocks_data = blocks_data;
  A.sram_data = sram_data;
  A.sram_xmask = sram_xmask;
  A.num_cycles = (u32)num_cycles;
  A.state_size = state_size;
  A.states = states;
  A.events = reinterpret_cast<EvHdr *>(event_buffer);
  A.ev_cap = ev_cap;
  A.arrival_state_offset = arrival_state_offset;
  A.timing = timing_constraints;  // [clock_ps, words...], read on device
  if (A.timing && !A.events) return fail("timing constraints need an event buffer");
  A.part_index = part_index;
  A.num_parts = num_parts;
  A.samples = samples;
  A.sample_cycles = sample_cycles;
  const bool timing = A.timing != nullptr || arrival_state_offset != 0;
  if (samples) return launch_scan<false, true>(A, stream);
  if (timing) return launch_scan<true, false>(A, stream);
  void *flow_buf = build_flow_tables(A, stream);
  const int rc = launch_scan<false, false>(A, stream);
  if (flow_buf) (void)hipFreeAsync(flow_buf, stream);
  return rc;
}

void die_on(int rc) {
  if (rc) {
    printf("HIP error in jacquard backend: %s\n", t_err.c_str());
    exit(EXIT_FAILURE);
  }
}

}  // namespace
}  // namespace jqd

using namespace jqd;

extern "C" void simulate_v1_noninteractive_simple_scan_hip(
    usize num_blocks, usize num_major_stages, const usize *blocks_start,
    const u32 *blocks_data, u32 *sram_data, u32 *sram_xmask, usize num_cycles,
    usize state_size, u32 *states_noninteractive, i32 arrival_state_offset) {
  die_on(scan_common(num_blocks, num_major_stages, blocks_start, blocks_data,
                     sram_data, sram_xmask, num_cycles, state_size,
                     states_noninteractive, nullptr, nullptr, 0,
                     arrival_state_offset, (hipStream_t)0, nullptr, 0, nullptr, 0));
}

extern "C" void simulate_v1_noninteractive_timed_hip(
    usize num_blocks, usize num_major_stages, const usize *blocks_start,
    const u32 *blocks_data, u32 *sram_data, u32 *sram_xmask, usize num_cycles,
    usize state_size, u32 *states_noninteractive, const u32 *timing_constraints,
    u8 *event_buffer, i32 arrival_state_offset) {
  die_on(scan_common(num_blocks, num_major_stages, blocks_start, blocks_data,
                     sram_data, sram_xmask, num_cycles, state_size,
                     states_noninteractive, timing_constraints, event_buffer,
                     1024u, arrival_state_offset, (hipStream_t)0, nullptr, 0,
                     nullptr, 0));
}

extern "C" int jq_simulate_scan(usize num_blocks, usize num_major_stages,
                                const usize *blocks_start, const u32 *blocks_data,
                                u32 *sram_data, u32 *sram_xmask, usize num_cycles,
                                usize state_size, u32 *states,
                                const u32 *timing_constraints, u8 *event_buffer,
                                u32 event_capacity, i32 arrival_state_offset,
                                void *stream) {
  return scan_common(num_blocks, num_major_stages, blocks_start, blocks_data,
                     sram_data, sram_xmask, num_cycles, state_size, states,
                     timing_constraints, event_buffer, event_capacity,
                     arrival_state_offset, (hipStream_t)stream, nullptr, 0,
                     nullptr, 0);
}

extern "C" int jq_simulate_scan_profiled(
    usize num_blocks, usize num_major_stages, const usize *blocks_start,
    const u32 *blocks_data, u32 *sram_data, u32 *sram_xmask, usize num_cycles,
    usize state_size, u32 *states, const u32 *part_index, u32 num_parts,
    u32 *samples, u32 sample_cycles, void *stream) {
  if (!part_index || !samples) return fail("profiling buffers missing");
  return scan_common(num_blocks, num_major_stages, blocks_start, blocks_data,
                     sram_data, sram_xmask, num_cycles, state_size, states,
                     nullptr, nullptr, 0, 0, (hipStream_t)stream, part_index,
                     num_parts, samples, sample_cycles);
}

extern "C" const char *jq_last_error(void) { return t_err.c_str(); }

namespace jqd {
// occupancy of the batched kernel, defined next to it
int batched_blocks_per_cu(int *out);
}  // namespace jqd

extern "C" int jq_max_resident_blocks(int kind, int *out_blocks) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail("hipGetDevice failed");
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return fail("CU count query failed");
  if (kind == 0) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per, reinterpret_cast<const void *>(&scan_kernel<true, false>), 256, 0) !=
        hipSuccess)
      return fail("occupancy query failed");
  } else {
    if (batched_blocks_per_cu(&per)) return -1;
  }
  *out_blocks = per * cus;
  return 0;
}

extern "C" usize jq_wait_limit(usize polls) {
  u32 prev = 0;
  if (hipMemcpyFromSymbol(&prev, HIP_SYMBOL(g_wait_polls), 4) != hipSuccess) return 0;
  if (polls) {
    const u32 v = polls > 0xFFFFFFFFull ? 0xFFFFFFFFu : (u32)polls;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_wait_polls), &v, 4) != hipSuccess ||
        batched_set_wait_limit(v))
      return 0;
  }
  return prev;
}

extern "C" int jq_scan_sync_mode(int mode) {
  if (mode < 0) return __atomic_load_n(&g_scan_sync, __ATOMIC_RELAXED);
  return __atomic_exchange_n(&g_scan_sync, mode ? 1 : 0, __ATOMIC_RELAXED);
}

extern "C" int jq_device_error_flag(u32 *out, int clear) {
  u32 v = 0, vb = 0;
  if (hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_error), 4) != hipSuccess) return -1;
  if (clear) {
    u32 z = 0;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_error), &z, 4) != hipSuccess) return -1;
    // a timed-out wait leaves partial arrival counts behind in its slot:
    // start every slot over (no scan launch may be running)
    if (v) {
      void *bar = nullptr;
      if (hipGetSymbolAddress(&bar, HIP_SYMBOL(g_barrier)) != hipSuccess ||
          hipMemset(bar, 0, sizeof(u32) * kCohSlots * kCohSlotWords) != hipSuccess ||
          hipDeviceSynchronize() != hipSuccess)
        return -1;
    }
  }
  if (batched_error_flag(&vb, clear)) return -1;
  *out = v | vb;
  return 0;
}

extern "C" size_t jq_cosim_flash_state_bytes(void) { return sizeof(cosim::FlashState); }

extern "C" int jq_cosim_edges(usize num_blocks, usize num_major_stages,
                              const usize *blocks_start, const u32 *blocks_data,
                              u32 *sram_data, usize state_size, u32 *states,
                              const u32 *edge_ops, const u32 *edge_op_start,
                              u32 num_sched, u32 sched_offset, u32 num_edges,
                              void *flash_state, const u32 *flash_pins,
                              const u8 *flash_data, u32 flash_data_size,
                              void *uart_state, const void *uart_params,
                              void *uart_channels, void *bus_trace,
                              const void *bus_trace_params,
                              const u32 *timing_constraints, u8 *event_buffer,
                              u32 event_capacity, i32 arrival_state_offset, u32 *trace,
                              void *stream) {
  t_err.clear();
  if (num_edges == 0) return 0;
  if (uart_state && (!uart_params || !uart_channels))
    return fail("cosim: UART state without its parameters or channels");
  if (bus_trace && !bus_trace_params) return fail("cosim: bus trace without its parameters");
  if (num_blocks == 0 || num_blocks > 65535) return fail("cosim: num_blocks out of range");
  if (!states || state_size == 0 || state_size > 0xFFFFFFFFull / 2)
    return fail("cosim: state buffer missing or state_size out of range");
  if (num_sched == 0 || !edge_op_start || !edge_ops) return fail("cosim: empty edge schedule");
  if (flash_state && !flash_pins) return fail("cosim: flash state without flash pins");
  if (flash_data_size && !flash_data) return fail("cosim: flash image missing");
  ScanArgs A{};
  A.num_blocks = (u32)num_blocks;
  A.num_major_stages = (u32)num_major_stages;
  A.blocks_start = blocks_start;
  A.blocks_data = blocks_data;
  A.sram_data = sram_data;
  A.sram_xmask = nullptr;  // 2-state: no SRAM X shadow
  A.num_cycles = 1;
  A.state_size = state_size;
  A.states = states;
  A.timing = timing_constraints;
  A.events = reinterpret_cast<EvHdr *>(event_buffer);
  A.ev_cap = event_capacity;
  A.arrival_state_offset = arrival_state_offset;
  if (A.timing && !A.events) return fail("cosim: timing constraints need an event buffer");
  const bool timing = A.timing != nullptr || arrival_state_offset != 0;
  cosim::CosimArgs C{};
  C.states = states;
  C.ss = (u32)state_size;
  C.ops = edge_ops;
  C.op_start = edge_op_start;
  C.num_sched = num_sched;
  C.sched_offset = sched_offset;
  C.num_edges = num_edges;
  C.fs = reinterpret_cast<cosim::FlashState *>(flash_state);
  C.pins = flash_pins;
  C.flash_data = flash_data;
  C.flash_size = flash_data_size;
  C.trace = trace;
  C.us = reinterpret_cast<cosim::UartState *>(uart_state);
  C.up = reinterpret_cast<const cosim::UartParams *>(uart_params);
  C.uc = reinterpret_cast<cosim::UartChannel *>(uart_channels);
  C.wbc = reinterpret_cast<cosim::WbTraceChannel *>(bus_trace);
  C.wbp = reinterpret_cast<const cosim::WbTraceParams *>(bus_trace_params);
  u32 *bar_base = nullptr;
  if (hipGetSymbolAddress((void **)&bar_base, HIP_SYMBOL(g_barrier)) != hipSuccess)
    return fail("hipGetSymbolAddress(g_barrier) failed");
  A.bar = bar_base + kCohSlotWords * (__atomic_fetch_add(&g_slot, 1u, __ATOMIC_RELAXED) % kCohSlots);
  // the flash state's double buffer: words 64..87 of the launch's barrier
  // slot, which the barrier does not use (generation at 32, abort word at
  // 36, counters from 256)
  C.fsbuf = A.bar + 64;
  void *args[] = {&A, &C};
  const void *k = timing ? reinterpret_cast<const void *>(&cosim_kernel<true>)
                        : reinterpret_cast<const void *>(&cosim_kernel<false>);
  hipError_t e = hipLaunchCooperativeKernel(k, dim3(A.num_blocks), dim3(256), args, 0,
                                            (hipStream_t)stream);
  if (e != hipSuccess) return fail(std::string("cosim launch failed: ") + hipGetErrorString(e));
  return 0;
}
