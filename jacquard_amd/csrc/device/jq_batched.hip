// Batched lane-sliced interpreter for gfx950.
//
// The same script runs for 32 independent stimulus lanes at once: every
// bit position of the reference layout (state bits, the 8192-bit boomerang
// layers, writeout bits) becomes one u32 "lane-word" whose bit l belongs to
// lane l.  The AND tree is bitwise, so evaluating it position by position on
// lane-words gives every lane exactly the value the packed interpreter
// computes for it (src/sim/cpu_reference.rs:17-289), while the 8192 bit
// gathers per boomerang stage become 8192 word gathers shared by 32 lanes.
//
// Work mapping inside a workgroup (256 threads, 4 x wave64):
//   - global read: a per-partition gather table (built once per script from
//     its (idx, mask) rounds) gives the source lane-word of each of the 8192
//     first-layer positions, so each thread issues 32 independent loads
//     instead of a chain of dependent pext steps;
//   - every 5120-word script section (boomerang stage, SRAM/dup, clock
//     enable) is copied into LDS with coalesced 16-byte loads, one section
//     ahead of its use (double buffer), so the per-position permutation and
//     flag reads are LDS reads;
//   - thread t owns level-1 outputs t + 256k (k < 16), so levels 1..5 of its
//     subtree stay in registers; wave 0 finishes levels 6..13 with
//     cross-lane moves: two workgroup barriers per boomerang stage;
//   - the three 8192-word LDS arrays (two after-layouts, the writeout area)
//     are row-swizzled: word w of 32-word row r lives at w ^ (r & 31), so
//     accesses where each lane touches the same word of a different row
//     spread over all banks;
//   - one persistent workgroup per (lane group, script block); each lane
//     group has its own barrier among its blocks, so groups with different
//     cycle counts run independently;
//   - stimulus is generated on the device with the same counter hash the
//     host uses, observed state bits are captured every cycle.
#include <hip/hip_runtime.h>

#include <string>

#include "jq_common.hpp"

namespace jqd {

constexpr int kBatchedSlots = 256;  // lane groups in flight
__device__ u32 gb_barrier[kBatchedSlots * kCohSlotWords];
__device__ u32 gb_wait_polls = kDefaultWaitPolls;  // bound of every wait here
__device__ u32 gb_error;

namespace {

constexpr int kLayerWords = 8192;
constexpr int kSecWords = 5120;  // 4096 permutation words + 256 x 4 flags
// LDS: the layer, the writeout area, two script-section buffers and the
// commit's old writeout words (136 KB of the 160 KB), then 4 words whose
// first is the launch's abort flag (kept in the dynamic region: a static
// __shared__ word would move the region's start off its 256-byte boundary)
constexpr int kLdsWords = 3 * kLayerWords + 2 * kSecWords + 4;
constexpr int kAbortLds = 3 * kLayerWords + 2 * kSecWords;
constexpr u32 kNoSource = 0xFFFFFFFFu;

struct BArgs {
  u32 G, W, num_major;
  const usize *bs;
  const u32 *bd;
  const u32 *gsrc;   // gather table: 8192 u32 per partition (blocks_data order)
  const u32 *gidx;   // first partition index of each (major stage, block)
  u32 rio, sram_size;
  u32 *sram;
  u32 *states;
  const u32 *data_bits;
  u32 n_data;
  const u32 *flag_bits;
  u32 n_flag;
  // lane queue (lanes in queue order)
  u32 n_lanes;
  const u32 *lane_cycles, *lane_seed, *lane_thresh;
  u32 *lane_info;    // per lane: group*32 + slot, start cycle of the group
  u32 *slots;        // per group 32 slots: lane or kNoLane
  u32 *queue;        // next lane to hand out
  u32 *group_cycles; // in: group cycle to resume from; out: cycles each group ran
  u32 budget;        // group cycles per launch (0: run every lane to the end)
  // packed input words per lane (null: hashed stimulus): snapshot k of the
  // lane in queue order q is stim[stim_off[q] + k*in_words ...], and the
  // state bits in_bits[0..n_in) take their values from it
  const u32 *stim;
  const usize *stim_off;
  const u32 *in_bits;
  u32 n_in, in_words;
  const u32 *obs_bits;
  u32 n_obs, obs_cap;
  u32 *obs_out;
  u32 *bar;
  // optional per-partition clock samples of group 0: prof[c*n_parts + part]
  u32 *prof;
  u32 prof_cycles, n_parts;
};

constexpr u32 kNoLane = 0xFFFFFFFFu;

__device__ __forceinline__ u32 mix32(u32 x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Row swizzle of the 8192-word LDS arrays.
__device__ __forceinline__ u32 sw(u32 p) { return p ^ ((p >> 5) & 31u); }

// u16 permutation entry of bit position p of a 4x1024-word section
// (src/flatten.rs:1027-1034).
__device__ __forceinline__ u32 perm_at(const u32 *sec, u32 p) {
  u32 i = p >> 5, b = p & 31;
  u32 w = sec[1024 * (b >> 3) + 4 * i + ((b >> 1) & 3)];
  return (b & 1) ? (w >> 16) : (w & 0xffffu);
}

// bit b of a flag word as an all-lanes mask
__device__ __forceinline__ u32 bmask(u32 word, int b) { return 0u - ((word >> b) & 1u); }

// flag bit of position p as an all-lanes mask
__device__ __forceinline__ u32 fbit(u32 word, u32 p) {
  return 0u - ((word >> (p & 31)) & 1u);
}

struct Flags {
  u32 xa, xb, ob;
};

__device__ __forceinline__ Flags flags_at(const u32 *fl, u32 p) {
  uint4 f = *reinterpret_cast<const uint4 *>(fl + 4 * (p >> 5));
  return Flags{fbit(f.x, p), fbit(f.y, p), fbit(f.z, p)};
}

__device__ __forceinline__ u32 lgate(u32 a, u32 b, Flags f) {
  return (a ^ f.xa) & ((b ^ f.xb) | f.ob);
}

// Copy one 5120-word script section from global memory straight into LDS
// (global_load_lds_dwordx4: no registers, completes asynchronously).  Lane
// l of wave w moves the 16 bytes at word 4*(t + 256j), t = 64w + l, to the
// same word offset in LDS; the LDS base of each instruction is wave-uniform.
__device__ __forceinline__ void sec_fetch(const u32 *g, u32 *l) {
  const u32 t = threadIdx.x, w = t >> 6;
#pragma unroll
  for (int j = 0; j < 5; ++j)
    __builtin_amdgcn_global_load_lds(g + 4 * (t + 256 * j), l + 4 * (64 * w + 256 * j),
                                     16, 0, 0);
}

// Copy the n (a multiple of 32) words at g into LDS at l, 16 bytes per
// lane, with the agent-coherent cache policy of ld_coherent (sc1): the
// words may have been written by other workgroups of the lane group.
__device__ __forceinline__ void coherent_fetch(const u32 *g, u32 *l, u32 n) {
  const u32 t = threadIdx.x, w = t >> 6;
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (4 * (t + 256 * j) < n)
      __builtin_amdgcn_global_load_lds(g + 4 * (t + 256 * j), l + 4 * (64 * w + 256 * j),
                                       16, 0, 16);
}

// Wait until this wave's outstanding global loads (including the LDS
// copies above) have landed: s_waitcnt vmcnt(0).  Every wave runs it before
// the workgroup barrier that precedes the first read of a fetched section.
__device__ __forceinline__ void sec_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Header of one partition (src/flatten.rs:198-258) and this thread's
// writeout hook.
struct BHdr {
  u32 S, is_last, num_ios, io, ns, so, R, nd, hook;
};

__device__ __forceinline__ BHdr load_bhdr(const u32 *p) {
  BHdr h;
  h.S = p[0];
  h.is_last = p[1];
  h.num_ios = p[2];
  h.io = p[3];
  h.ns = p[4];
  h.so = p[5];
  h.R = p[6];
  h.nd = p[7];
  const u32 hk = p[128 + (threadIdx.x >> 1)];
  h.hook = (threadIdx.x & 1) ? (hk >> 16) : (hk & 0xffffu);
  return h;
}

// Runs the partitions of one (major stage, block).  On entry h and src hold
// the first partition's header and gather-table entries (src only when the
// script is not empty); on return they hold those of the first partition of
// the block's next major stage (script nscript, table ngsrc), loaded while
// this stage ran.
__device__ __forceinline__ void run_block_lanes(const BArgs &A, const u32 *__restrict__ script,
                                const u32 *__restrict__ gsrc,
                                const u32 *__restrict__ in, u32 *__restrict__ out,
                                u32 *ram_group, u32 *lds, u32 *prof, BHdr &h,
                                u32 (&src)[32], const u32 *nscript, const u32 *ngsrc) {
  const u32 t = threadIdx.x;
  const u32 lane = t & 63;
  u32 *layer = lds, *WO = lds + kLayerWords;
  u32 *OLD = lds + 2 * kLayerWords + 2 * kSecWords;
  // section buffer k & 1 (pointer arithmetic, not an array: no scratch)
  auto secb = [lds](u32 k) { return lds + 2 * kLayerWords + (k & 1u) * kSecWords; };
  size_t pi = 0;
  // state lane-words of the next partition's first layer, loaded during
  // the current partition when both are in this major stage (their sources
  // are earlier major stages and the previous snapshot, complete since the
  // barrier before this stage)
  u32 vnext[32];
  bool have_next = false;
  for (;; gsrc += kLayerWords) {
    if (h.S == 0) {  // no partition in this stage: fetch the next stage's first
      h = load_bhdr(nscript);
      if (h.S != 0) {
#pragma unroll
        for (int j = 0; j < 32; ++j) src[j] = ngsrc[t + 256 * j];
      }
      break;
    }
    const unsigned long long tick0 = prof ? clock64() : 0;
    const u32 S = h.S, is_last = h.is_last, num_ios = h.num_ios, io = h.io;
    const u32 ns = h.ns, so = h.so, R = h.R, nd = h.nd, hook = h.hook;
    // sections: stage 0..S-1, SRAM/dup, clock enable, contiguous
    const u32 *secg = script + pi + 256 + (size_t)512 * R;
    pi += 256 + (size_t)512 * R + (size_t)kSecWords * (S + 2);
    // the successor's header (the next partition here, or the next stage's
    // first) is requested now, a partition ahead of its use
    h = is_last ? load_bhdr(nscript) : load_bhdr(script + pi);
    const u32 next_s = h.S;

    sec_fetch(secg, secb(0));
    // the commit's old values (this partition's writeouts in the previous
    // snapshot, which nothing changes during the cycle) go to LDS now and
    // land while the partition runs
    coherent_fetch(in + (size_t)io * 32, OLD, num_ios * 32);
    // ---- global read through the gather table: 32 independent loads of
    // state words at entries prefetched one partition ahead (or the words
    // themselves, when the previous partition of this stage loaded them)
    {
      u32 v[32];
      if (have_next) {
#pragma unroll
        for (int j = 0; j < 32; ++j) v[j] = vnext[j];
      } else {
#pragma unroll
        for (int j = 0; j < 32; ++j) {
          u32 e = src[j];
          v[j] = e != kNoSource ? ld_coherent(((e >> 31) ? out : in) + (e & 0x7fffffffu)) : 0u;
        }
      }
      // the successor's entries load while this partition runs (the table
      // is read-only)
      if (next_s != 0) {
        const u32 *nx = is_last ? ngsrc : gsrc + kLayerWords;
#pragma unroll
        for (int j = 0; j < 32; ++j) src[j] = nx[t + 256 * j];
      }
#pragma unroll
      for (int j = 0; j < 32; ++j) layer[sw(t + 256 * j)] = v[j];
    }
    sec_wait();
    __syncthreads();

    const u32 nperm = ns * 4 + nd;  // SRAM port and duplicate words
    // ---- boomerang stages, in place on the one layer: every thread first
    // gathers its 32 level-1 operands into registers, then (after a
    // barrier) the layer is overwritten with this stage's after-layout
    for (u32 s = 0; s < S; ++s) {
      const uint4 *S4 = reinterpret_cast<const uint4 *>(secb(s));
      const uint4 P0 = S4[t], P1 = S4[256 + t], P2 = S4[512 + t], P3 = S4[768 + t];
      const uint4 F = S4[1024 + t];
      // the next section goes into the other buffer, last read by the
      // previous stage, which finished before the last barrier; after the
      // last stage that is the SRAM/dup section, or the clock-enable
      // section when the partition has no SRAM or duplicate words
      const u32 nxt = (s + 1 == S && nperm == 0) ? S + 1 : s + 1;
      sec_fetch(secg + (size_t)kSecWords * nxt, secb(s + 1));
      const u32 pw[16] = {P0.x, P0.y, P0.z, P0.w, P1.x, P1.y, P1.z, P1.w,
                          P2.x, P2.y, P2.z, P2.w, P3.x, P3.y, P3.z, P3.w};
      u32 g[32];
#pragma unroll
      for (int m = 0; m < 32; ++m) g[m] = layer[(pw[m >> 1] >> (16 * (m & 1))) & 0xffffu];
      __syncthreads();  // all operands read before the layer is rewritten
      // levels 1..5 of the subtree rooted at position 256 + t; gate g's
      // flags are bit g of F.x / F.y / F.z
      u32 L1[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        L1[k] = (g[k] ^ bmask(F.x, k)) & ((g[k + 16] ^ bmask(F.y, k)) | bmask(F.z, k));
        layer[sw(4096 + t + 256u * k)] = L1[k];
      }
      u32 L2[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        L2[k] = (L1[k] ^ bmask(F.x, 16 + k)) & ((L1[k + 8] ^ bmask(F.y, 16 + k)) |
                                                 bmask(F.z, 16 + k));
        layer[sw(2048 + t + 256u * k)] = L2[k];
      }
      u32 L3[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        L3[k] = (L2[k] ^ bmask(F.x, 24 + k)) & ((L2[k + 4] ^ bmask(F.y, 24 + k)) |
                                                 bmask(F.z, 24 + k));
        layer[sw(1024 + t + 256u * k)] = L3[k];
      }
      u32 L4[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        L4[k] = (L3[k] ^ bmask(F.x, 28 + k)) & ((L3[k + 2] ^ bmask(F.y, 28 + k)) |
                                                 bmask(F.z, 28 + k));
        layer[sw(512 + t + 256u * k)] = L4[k];
      }
      layer[sw(256 + t)] = (L4[0] ^ bmask(F.x, 30)) & ((L4[1] ^ bmask(F.y, 30)) | bmask(F.z, 30));
      __syncthreads();
      if (t < 64) {
        // wave 0 lane L: flags of positions 128+L, 192+L, 64+L, L in F.w
        const u32 fw = F.w;
        auto wg = [fw](u32 a, u32 b, int i) {
          return (a ^ bmask(fw, 3 * i)) & ((b ^ bmask(fw, 3 * i + 1)) | bmask(fw, 3 * i + 2));
        };
        // level 6: positions 128+L and 192+L; level 7: 64+L
        u32 o6a = wg(layer[sw(256 + lane)], layer[sw(384 + lane)], 0);
        u32 o6b = wg(layer[sw(320 + lane)], layer[sw(448 + lane)], 1);
        u32 o7 = wg(o6a, o6b, 2);
        // level 8: lanes 32..63 -> position L from L+32 (lane L-32) and
        // L+64 (own level-7 word)
        u32 c = 0;
        u32 a8 = __shfl(o7, (lane + 32) & 63);
        if (lane >= 32) c = wg(a8, o7, 3);
#pragma unroll
        for (int len = 16; len >= 1; len >>= 1) {
          u32 a = __shfl(c, (lane + len) & 63);
          u32 b = __shfl(c, (lane + 2 * len) & 63);
          if (lane >= (u32)len && lane < 2u * len) c = wg(a, b, 3);
        }
        layer[sw(128 + lane)] = o6a;
        layer[sw(192 + lane)] = o6b;
        layer[sw(64 + lane)] = o7;
        layer[sw(lane)] = c;  // lane 0 keeps position 0 at zero
      }
      sec_wait();  // section s+1 is read right after this barrier
      __syncthreads();
      if ((hook >> 8) == s) {
        const u32 src = 32u * (hook & 255u);
#pragma unroll
        for (u32 q = 0; q < 32; ++q) WO[sw(32 * t + q)] = layer[sw(src + q)];
      }
    }
    // the next partition of this stage: its first-layer words (its gather
    // entries arrived during this partition)
    have_next = !is_last && next_s != 0;
    if (have_next) {
#pragma unroll
      for (int j = 0; j < 32; ++j) {
        u32 e = src[j];
        vnext[j] = e != kNoSource ? ld_coherent(((e >> 31) ? out : in) + (e & 0x7fffffffu)) : 0u;
      }
    }
    __syncthreads();

    // ---- SRAM + duplicate permutation into the layer (free now); a
    // partition without SRAM or duplicate words skips the phase and its
    // barriers (the condition is uniform), its clock-enable section already
    // in place
    if (nperm) {
      {
        const u32 *sec = secb(S);
        const u32 *fl = sec + 4096;
        // the clock-enable section goes into the buffer of the last stage
        sec_fetch(secg + (size_t)kSecWords * (S + 1), secb(S + 1));
        for (u32 q = t; q < nperm * 32; q += 256) {
          uint4 f = *reinterpret_cast<const uint4 *>(fl + 4 * (q >> 5));
          u32 v = WO[sw(perm_at(sec, q))];
          layer[sw(q)] = (v & ~fbit(f.y, q)) ^ fbit(f.x, q);
        }
      }
      __syncthreads();
      // SRAM ports: thread per (sram, lane)
      for (u32 base = 0; base < ns * 32; base += 256) {
        u32 id = base + t;
        u32 s = id >> 5, l = id & 31;
        u32 r = 0;
        bool act = id < ns * 32;
        if (act) {
          const u32 P = 128 * s;
          u32 ra = 0, wa = 0, en = 0, data = 0;
          for (int k = 0; k < 13; ++k) {
            ra |= ((layer[sw(P + k)] >> l) & 1u) << k;
            wa |= ((layer[sw(P + 16 + k)] >> l) & 1u) << k;
          }
          for (int k = 0; k < 32; ++k) {
            en |= ((layer[sw(P + 32 + k)] >> l) & 1u) << k;
            data |= ((layer[sw(P + 64 + k)] >> l) & 1u) << k;
          }
          u32 *ram = ram_group + (size_t)l * A.sram_size + so + (size_t)s * kSramWords;
          r = ld_coherent(ram + ra);
          u32 w0 = ld_coherent(ram + wa);
          st_coherent(ram + wa, (w0 & ~en) | (data & en));
        }
        // transpose the 32 read words of each SRAM into lane-words
        u32 mine = 0;
        for (int k = 0; k < 32; ++k) {
          unsigned long long m = __ballot(act && ((r >> k) & 1u));
          u32 half = (u32)(m >> (32 * ((t >> 5) & 1)));
          if (l == (u32)k) mine = half;
        }
        if (act) WO[sw(32 * (num_ios - ns + s) + l)] = mine;
      }
      for (u32 q = t; q < nd * 32; q += 256)
        WO[sw(32 * (num_ios - ns - nd) + q)] = layer[sw(ns * 128 + q)];
      sec_wait();  // the clock-enable section is read after this barrier
      __syncthreads();
    }

    // ---- clock-enable permutation + commit (coalesced lane-words)
    {
      const u32 *sec = secb(nperm ? S + 1 : S);
      const u32 *fl = sec + 4096;
      u32 *o = out + (size_t)io * 32;
      for (u32 q = t; q < num_ios * 32; q += 256) {
        uint4 f = *reinterpret_cast<const uint4 *>(fl + 4 * (q >> 5));
        // a bit whose enable is forced (set0) needs no gather: ck = inv
        u32 ck = fbit(f.x, q);
        if (!((f.y >> (q & 31)) & 1u)) ck ^= WO[sw(perm_at(sec, q))];
        u32 d = WO[sw(q)] ^ fbit(f.z, q);
        st_coherent(o + q, (OLD[q] & ~ck) | (d & ck));
      }
    }
    __syncthreads();
    if (prof) {
      if (t == 0) *prof = (u32)(clock64() - tick0);
      ++prof;
    }
    if (is_last) break;
  }
}

__global__ __launch_bounds__(256) void batched_kernel(BArgs A) {
  extern __shared__ u32 lds[];
  u32 &s_aborted = lds[kAbortLds];  // set by a group wait when the launch aborts
  if (threadIdx.x == 0) s_aborted = 0;
  const u32 g = blockIdx.x / A.W, b = blockIdx.x % A.W;
  u32 *bar = A.bar + kCohSlotWords * g;
  u32 gen = read_gen(bar);
  const size_t snap = (size_t)A.rio * 32;
  u32 *st0 = A.states + (size_t)g * 2 * snap;
  u32 *st1 = st0 + snap;
  u32 *ram_group = A.sram + (size_t)g * 32 * A.sram_size;
  u32 *slots = A.slots + 32 * g;
  const u32 t = threadIdx.x, l = t & 31;
  const u32 par = b * 256 + t, npar = A.W * 256;  // group-wide thread index
  // every thread mirrors slot l = t & 31 of its group; a launch resumes
  // the group at its recorded cycle, each slot's lane at the group cycle it
  // started
  const u32 c0 = A.group_cycles[g];
  u32 my_lane = slots[l], my_start = 0, my_end = 0, my_seed = 0, my_thr = 0;
  if (my_lane != kNoLane) {
    my_start = A.lane_info[2 * my_lane + 1];
    my_end = my_start + A.lane_cycles[my_lane];
    if (!A.stim) {
      my_seed = A.lane_seed[my_lane];
      my_thr = A.lane_thresh[my_lane];
    }
  }
  // header and gather entries of the block's first partition of major
  // stage 0
  u32 src[32];
  BHdr h = load_bhdr(A.bd + A.bs[b]);
  if (h.S != 0) {
    const u32 *g0 = A.gsrc + (size_t)A.gidx[b] * kLayerWords;
#pragma unroll
    for (int j = 0; j < 32; ++j) src[j] = g0[t + 256 * j];
  }
  __syncthreads();
  u32 c = c0;
  for (;; ++c) {
    // an aborted launch (a bounded wait gave up; the waits that follow end
    // within 256 polls each) leaves at the next cycle
    if (s_aborted) break;
    u32 *in = (c & 1) ? st1 : st0;
    u32 *out = (c & 1) ? st0 : st1;
    // every group finishes within obs_cap cycles (queue bound); a launch
    // whose lane table breaks that bound stops here with an error instead
    // of looping (the same decision in every block of the group)
    if (c > A.obs_cap) {
      if (b == 0 && t == 0) atomicOr(&gb_error, 2u);
      break;
    }
    if (b == 0 && c > 0 && c - 1 < A.obs_cap) {
      u32 *o = A.obs_out + ((size_t)g * A.obs_cap + (c - 1)) * A.n_obs;
      for (u32 k = t; k < A.n_obs; k += 256) o[k] = ld_coherent(in + A.obs_bits[k]);
    }
    // a launch with a budget stops after that many group cycles, with the
    // observation of its last cycle written (the same decision in every
    // block of the group)
    if (A.budget && c - c0 == A.budget) break;
    // lanes that completed their cycles hand their slot to the next lane
    const bool fin = my_lane != kNoLane && c >= my_end;
    if (__syncthreads_or(fin)) {
      const u32 finmask = (u32)__ballot(fin);  // bit l: slot l finished
      if (b == 0 && t < 32 && fin) {
        u32 nl = atomicAdd(A.queue, 1u);
        if (nl < A.n_lanes) {
          st_coherent(slots + l, nl);
          A.lane_info[2 * nl] = g * 32 + l;
          A.lane_info[2 * nl + 1] = c;
        } else {
          st_coherent(slots + l, kNoLane);
        }
      }
      grid_sync_coherent(bar, A.W, b, gen, &gb_error, &gb_wait_polls, s_aborted);
      // reset the finished slots' lane bit: state of this cycle's input,
      // clock flags set, the lane's SRAM words zeroed
      for (size_t w = par; w < snap; w += npar) st_coherent(in + w, ld_coherent(in + w) & ~finmask);
      if (!A.stim)
        for (u32 k = par; k < A.n_flag; k += npar)
          st_coherent(in + A.flag_bits[k], ld_coherent(in + A.flag_bits[k]) | finmask);
      for (u32 m = finmask; m; m &= m - 1) {
        u32 *ram = ram_group + (size_t)(__ffs(m) - 1) * A.sram_size;
        for (u32 w = par; w < A.sram_size; w += npar) st_coherent(ram + w, 0u);
      }
      if (fin) {
        my_lane = __hip_atomic_load(&slots[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        my_start = c;
        if (my_lane != kNoLane) {
          my_end = c + A.lane_cycles[my_lane];
          if (!A.stim) {
            my_seed = A.lane_seed[my_lane];
            my_thr = A.lane_thresh[my_lane];
          }
        }
      }
      grid_sync_coherent(bar, A.W, b, gen, &gb_error, &gb_wait_polls, s_aborted);
      if (A.stim) {
        // packed stimulus: the new lanes' input bits of their snapshot 0
        for (u32 k0 = b * 8; k0 < A.n_in; k0 += A.W * 8) {
          const u32 k = k0 + (t >> 5);
          bool bit = false;
          if (k < A.n_in && fin && my_lane != kNoLane) {
            const u32 p = A.in_bits[k];
            bit = (A.stim[A.stim_off[my_lane] + (p >> 5)] >> (p & 31)) & 1u;
          }
          const unsigned long long m = __ballot(bit);
          const u32 set = (u32)(m >> (32 * ((t >> 5) & 1)));
          if (k < A.n_in && l == 0 && set) {
            const u32 p = A.in_bits[k];
            st_coherent(in + p, ld_coherent(in + p) | set);
          }
        }
        grid_sync_coherent(bar, A.W, b, gen, &gb_error, &gb_wait_polls, s_aborted);
      }
    }
    if (!__syncthreads_or(my_lane != kNoLane)) break;
    if (A.stim) {
      // stimulus of snapshot c+1 from the lanes' packed input words: input
      // bit k of slot l is bit in_bits[k] of the lane's snapshot
      // c - start + 1 (inactive slots read 0)
      for (u32 k0 = b * 8; k0 < A.n_in; k0 += A.W * 8) {
        const u32 k = k0 + (t >> 5);
        bool bit = false;
        if (k < A.n_in && my_lane != kNoLane) {
          const u32 p = A.in_bits[k];
          const size_t at = A.stim_off[my_lane] + (size_t)(c - my_start + 1) * A.in_words;
          bit = (A.stim[at + (p >> 5)] >> (p & 31)) & 1u;
        }
        const unsigned long long m = __ballot(bit);
        const u32 word = (u32)(m >> (32 * ((t >> 5) & 1)));
        if (k < A.n_in && l == 0) st_coherent(out + A.in_bits[k], word);
      }
    }
    // stimulus of snapshot c+1: data bit k toggles for the lane in slot l
    // when hash(seed, lane cycle + 1, k) < thresh
    for (u32 k0 = b * 8; !A.stim && k0 < A.n_data; k0 += A.W * 8) {
      u32 k = k0 + (t >> 5);
      bool act = k < A.n_data && my_lane != kNoLane;
      u32 bit = 0;
      if (act) {
        u32 h = mix32(mix32(my_seed + 0x9E3779B9u * k) ^ (c - my_start + 1));
        bit = h < my_thr;
      }
      unsigned long long m = __ballot(act && bit);
      u32 tog = (u32)(m >> (32 * ((t >> 5) & 1)));
      if (k < A.n_data && l == 0) {
        u32 p = A.data_bits[k];
        st_coherent(out + p, ld_coherent(in + p) ^ tog);
      }
    }
    for (u32 mj = 0; mj < A.num_major; ++mj) {
      const size_t k = (size_t)mj * A.W + b;
      u32 *prof = (A.prof && g == 0 && c < A.prof_cycles)
                      ? A.prof + (size_t)c * A.n_parts + A.gidx[k] : nullptr;
      // next major stage of this block (stage 0 of the next cycle after
      // the last)
      const size_t kn = mj + 1 < A.num_major ? k + A.W : (size_t)b;
      run_block_lanes(A, A.bd + A.bs[k], A.gsrc + (size_t)A.gidx[k] * kLayerWords, in,
                      out, ram_group, lds, prof, h, src, A.bd + A.bs[kn],
                      A.gsrc + (size_t)A.gidx[kn] * kLayerWords);
      grid_sync_coherent(bar, A.W, b, gen, &gb_error, &gb_wait_polls, s_aborted);
    }
  }
  if (b == 0 && t == 0) A.group_cycles[g] = c;
}

u32 b_slot = 0;

}  // namespace

int batched_set_wait_limit(u32 polls) {
  return hipMemcpyToSymbol(HIP_SYMBOL(gb_wait_polls), &polls, 4) == hipSuccess ? 0 : -1;
}

int batched_error_flag(u32 *out, int clear) {
  u32 v = 0;
  if (hipMemcpyFromSymbol(&v, HIP_SYMBOL(gb_error), 4) != hipSuccess) return -1;
  if (clear) {
    u32 z = 0;
    if (hipMemcpyToSymbol(HIP_SYMBOL(gb_error), &z, 4) != hipSuccess) return -1;
    // a timed-out wait leaves partial arrival counts behind in its slot:
    // start every slot over (no batched launch may be running)
    if (v) {
      void *bar = nullptr;
      if (hipGetSymbolAddress(&bar, HIP_SYMBOL(gb_barrier)) != hipSuccess ||
          hipMemset(bar, 0, sizeof(u32) * kBatchedSlots * kCohSlotWords) != hipSuccess ||
          hipDeviceSynchronize() != hipSuccess)
        return -1;
    }
  }
  *out = v;
  return 0;
}

int batched_blocks_per_cu(int *out) {
  static bool attr = false;
  if (!attr) {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(&batched_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            kLdsWords * 4) != hipSuccess)
      return -1;
    attr = true;
  }
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per, reinterpret_cast<const void *>(&batched_kernel), 256,
          kLdsWords * 4) != hipSuccess)
    return -1;
  *out = per;
  return 0;
}

}  // namespace jqd

using namespace jqd;

namespace {

int batched_impl(usize num_groups, usize blocks_per_group, usize num_major_stages,
                 const usize *blocks_start, const u32 *blocks_data,
                 const u32 *gather_table, const u32 *gather_start,
                 u32 reg_io_state_size, u32 sram_storage_size, u32 *sram_data,
                 u32 *states, const u32 *data_bits, u32 n_data, const u32 *flag_bits,
                 u32 n_flag, u32 n_lanes, const u32 *lane_cycles, const u32 *lane_seed,
                 const u32 *lane_thresh, u32 *lane_info, u32 *slots, u32 *queue,
                 u32 *group_cycles, const u32 *obs_bits, u32 num_obs,
                 u32 obs_capacity, u32 *obs_out, u32 num_parts, u32 *samples,
                 u32 sample_cycles, u32 budget, const u32 *stim, const usize *stim_off,
                 const u32 *in_bits, u32 n_in, u32 in_words, void *stream) {
  set_error("");
  if (stim) {
    if (!stim_off || (n_in && !in_bits) || in_words == 0 ||
        (usize)in_words > (usize)reg_io_state_size)
      return set_error("batched launch: packed stimulus needs lane offsets, input bits "
                       "and 0 < in_words <= reg_io_state_size");
  } else if (!lane_seed || !lane_thresh) {
    return set_error("batched launch: hashed stimulus needs lane seeds and thresholds");
  }
  if (num_groups == 0 || blocks_per_group == 0) return 0;
  if (num_groups > (usize)kBatchedSlots) {
    return set_error("too many lane groups");
  }
  if (!gather_table || !gather_start || !slots || !queue || !lane_info ||
      !group_cycles) {
    return set_error("batched launch: missing gather table or lane-queue buffers");
  }
  int per = 0, dev = 0, cus = 0;
  if (batched_blocks_per_cu(&per) || hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    return set_error("batched occupancy query failed");
  }
  if (num_groups * blocks_per_group > (usize)per * (usize)cus) {
    return set_error("batched launch: groups x blocks exceed the co-resident workgroups");
  }
  BArgs A{};
  A.G = (u32)num_groups;
  A.W = (u32)blocks_per_group;
  A.num_major = (u32)num_major_stages;
  A.bs = blocks_start;
  A.bd = blocks_data;
  A.gsrc = gather_table;
  A.gidx = gather_start;
  A.rio = reg_io_state_size;
  A.sram_size = sram_storage_size;
  A.sram = sram_data;
  A.states = states;
  A.data_bits = data_bits;
  A.n_data = n_data;
  A.flag_bits = flag_bits;
  A.n_flag = n_flag;
  A.n_lanes = n_lanes;
  A.lane_cycles = lane_cycles;
  A.lane_seed = lane_seed;
  A.lane_thresh = lane_thresh;
  A.lane_info = lane_info;
  A.slots = slots;
  A.queue = queue;
  A.group_cycles = group_cycles;
  A.obs_bits = obs_bits;
  A.n_obs = num_obs;
  A.obs_cap = obs_capacity;
  A.obs_out = obs_out;
  A.prof = samples;
  A.prof_cycles = samples ? sample_cycles : 0;
  A.n_parts = num_parts;
  A.budget = budget;
  A.stim = stim;
  A.stim_off = stim_off;
  A.in_bits = in_bits;
  A.n_in = stim ? n_in : 0;
  A.in_words = in_words;
  u32 *bar_base = nullptr;
  if (hipGetSymbolAddress((void **)&bar_base, HIP_SYMBOL(gb_barrier)) != hipSuccess) {
    return set_error("hipGetSymbolAddress(gb_barrier) failed");
  }
  // the launch's groups take consecutive slots from the rotation; a run that
  // would pass the end starts at slot 0, and the rotation continues after it
  u32 cur = __atomic_load_n(&b_slot, __ATOMIC_RELAXED), first = 0;
  do {
    first = cur % kBatchedSlots;
    if (first + num_groups > (usize)kBatchedSlots) first = 0;
  } while (!__atomic_compare_exchange_n(&b_slot, &cur, first + (u32)num_groups, true,
                                        __ATOMIC_RELAXED, __ATOMIC_RELAXED));
  A.bar = bar_base + kCohSlotWords * first;
  void *args[] = {&A};
  // Under stream capture (a hipGraph of the run) the launch is an ordinary
  // kernel node: co-residency of all workgroups is established above from
  // the occupancy query, and the group barriers' bounded wait turns any
  // shortfall into an error flag instead of a hang.
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing((hipStream_t)stream, &cap) != hipSuccess) {
    return set_error("hipStreamIsCapturing failed");
  }
  hipError_t e;
  if (cap == hipStreamCaptureStatusActive) {
    e = hipLaunchKernel(reinterpret_cast<const void *>(&batched_kernel),
                        dim3((u32)(num_groups * blocks_per_group)), dim3(256), args,
                        kLdsWords * 4, (hipStream_t)stream);
  } else {
    e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&batched_kernel),
                                   dim3((u32)(num_groups * blocks_per_group)), dim3(256),
                                   args, kLdsWords * 4, (hipStream_t)stream);
  }
  if (e != hipSuccess) {
    return set_error((std::string("batched launch failed: ") + hipGetErrorString(e)).c_str());
  }
  return 0;
}

}  // namespace

extern "C" int jq_simulate_batched(
    usize num_groups, usize blocks_per_group, usize num_major_stages,
    const usize *blocks_start, const u32 *blocks_data, const u32 *gather_table,
    const u32 *gather_start, u32 reg_io_state_size, u32 sram_storage_size,
    u32 *sram_data, u32 *states, const u32 *data_bits, u32 n_data,
    const u32 *flag_bits, u32 n_flag, u32 n_lanes, const u32 *lane_cycles,
    const u32 *lane_seed, const u32 *lane_thresh, u32 *lane_info, u32 *slots,
    u32 *queue, u32 *group_cycles, const u32 *obs_bits, u32 num_obs,
    u32 obs_capacity, u32 *obs_out, void *stream) {
  return batched_impl(num_groups, blocks_per_group, num_major_stages, blocks_start,
                      blocks_data, gather_table, gather_start, reg_io_state_size,
                      sram_storage_size, sram_data, states, data_bits, n_data, flag_bits,
                      n_flag, n_lanes, lane_cycles, lane_seed, lane_thresh, lane_info,
                      slots, queue, group_cycles, obs_bits, num_obs, obs_capacity, obs_out,
                      0, nullptr, 0, 0, nullptr, nullptr, nullptr, 0, 0, stream);
}

extern "C" int jq_simulate_batched_chunk(
    usize num_groups, usize blocks_per_group, usize num_major_stages,
    const usize *blocks_start, const u32 *blocks_data, const u32 *gather_table,
    const u32 *gather_start, u32 reg_io_state_size, u32 sram_storage_size,
    u32 *sram_data, u32 *states, const u32 *data_bits, u32 n_data,
    const u32 *flag_bits, u32 n_flag, u32 n_lanes, const u32 *lane_cycles,
    const u32 *lane_seed, const u32 *lane_thresh, u32 *lane_info, u32 *slots,
    u32 *queue, u32 *group_cycles, u32 max_group_cycles, const u32 *obs_bits,
    u32 num_obs, u32 obs_capacity, u32 *obs_out, void *stream) {
  return batched_impl(num_groups, blocks_per_group, num_major_stages, blocks_start,
                      blocks_data, gather_table, gather_start, reg_io_state_size,
                      sram_storage_size, sram_data, states, data_bits, n_data, flag_bits,
                      n_flag, n_lanes, lane_cycles, lane_seed, lane_thresh, lane_info,
                      slots, queue, group_cycles, obs_bits, num_obs, obs_capacity, obs_out,
                      0, nullptr, 0, max_group_cycles, nullptr, nullptr, nullptr, 0, 0,
                      stream);
}

extern "C" int jq_simulate_batched_stimulus(
    usize num_groups, usize blocks_per_group, usize num_major_stages,
    const usize *blocks_start, const u32 *blocks_data, const u32 *gather_table,
    const u32 *gather_start, u32 reg_io_state_size, u32 sram_storage_size,
    u32 *sram_data, u32 *states, const u32 *in_bits, u32 n_in, u32 in_words,
    const u32 *stimulus, const usize *stimulus_offset, u32 n_lanes,
    const u32 *lane_cycles, u32 *lane_info, u32 *slots, u32 *queue, u32 *group_cycles,
    u32 max_group_cycles, const u32 *obs_bits, u32 num_obs, u32 obs_capacity,
    u32 *obs_out, void *stream) {
  if (!stimulus) return set_error("batched stimulus launch: stimulus missing");
  return batched_impl(num_groups, blocks_per_group, num_major_stages, blocks_start,
                      blocks_data, gather_table, gather_start, reg_io_state_size,
                      sram_storage_size, sram_data, states, nullptr, 0, nullptr, 0,
                      n_lanes, lane_cycles, nullptr, nullptr, lane_info, slots, queue,
                      group_cycles, obs_bits, num_obs, obs_capacity, obs_out, 0, nullptr,
                      0, max_group_cycles, stimulus, stimulus_offset, in_bits, n_in,
                      in_words, stream);
}

extern "C" int jq_simulate_batched_profiled(
    usize num_groups, usize blocks_per_group, usize num_major_stages,
    const usize *blocks_start, const u32 *blocks_data, const u32 *gather_table,
    const u32 *gather_start, u32 reg_io_state_size, u32 sram_storage_size,
    u32 *sram_data, u32 *states, const u32 *data_bits, u32 n_data,
    const u32 *flag_bits, u32 n_flag, u32 n_lanes, const u32 *lane_cycles,
    const u32 *lane_seed, const u32 *lane_thresh, u32 *lane_info, u32 *slots,
    u32 *queue, u32 *group_cycles, const u32 *obs_bits, u32 num_obs,
    u32 obs_capacity, u32 *obs_out, u32 num_parts, u32 *samples, u32 sample_cycles,
    void *stream) {
  if (!samples || num_parts == 0) return set_error("profiled batched launch needs samples");
  return batched_impl(num_groups, blocks_per_group, num_major_stages, blocks_start,
                      blocks_data, gather_table, gather_start, reg_io_state_size,
                      sram_storage_size, sram_data, states, data_bits, n_data, flag_bits,
                      n_flag, n_lanes, lane_cycles, lane_seed, lane_thresh, lane_info,
                      slots, queue, group_cycles, obs_bits, num_obs, obs_capacity, obs_out,
                      num_parts, samples, sample_cycles, 0, nullptr, nullptr, nullptr, 0, 0,
                      stream);
}
