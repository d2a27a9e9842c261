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
//   - thread t owns level-1 outputs t + 256k (k < 16), so levels 1..5 of its
//     subtree stay in registers; wave 0 finishes levels 6..13 with
//     cross-lane moves: two workgroup barriers per boomerang stage;
//   - LDS (dynamic, 96 KiB): two after-layouts (ping-pong) and the 256-word
//     writeout area, each 8192 lane-words;
//   - one persistent workgroup per (lane group, script block); each lane
//     group has its own barrier among its blocks, so groups with different
//     cycle counts run independently;
//   - stimulus is generated on the device with the same counter hash the
//     host uses, observed state bits are captured every cycle.
#include <hip/hip_runtime.h>

#include <string>

#include "jq_common.hpp"

namespace jqd {

__device__ u32 gb_barrier[kBarrierSlots * kSlotWords];
__device__ u32 gb_error;

namespace {

constexpr int kLdsWords = 3 * 8192;

struct BArgs {
  u32 G, W, num_major;
  const usize *bs;
  const u32 *bd;
  u32 rio, sram_size;
  u32 *sram;
  const u32 *gcycles;
  u32 max_cycles;
  u32 *states;
  u32 n_pi_words;
  const u32 *data_bits;
  u32 n_data;
  const u32 *lane_seed, *lane_thresh;
  const u32 *obs_bits;
  u32 n_obs;
  u32 *obs_out;
  u32 *bar;
};

__device__ __forceinline__ u32 mix32(u32 x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// u16 permutation entry of bit position p of a 4x1024-word section
// (src/flatten.rs:1027-1034).
__device__ __forceinline__ u32 perm_at(const u32 *sec, u32 p) {
  u32 i = p >> 5, b = p & 31;
  u32 w = sec[1024 * (b >> 3) + 4 * i + ((b >> 1) & 3)];
  return (b & 1) ? (w >> 16) : (w & 0xffffu);
}

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

__device__ void run_block_lanes(const BArgs &A, const u32 *__restrict__ script,
                                const u32 *__restrict__ in, u32 *__restrict__ out,
                                u32 *ram_group, u32 *lds) {
  const u32 t = threadIdx.x;
  const u32 lane = t & 63;
  u32 *bufA = lds, *bufB = lds + 8192, *WO = lds + 16384;
  size_t pi = 0;
  for (;;) {
    const u32 S = script[pi];
    if (S == 0) break;
    const u32 is_last = script[pi + 1];
    const u32 num_ios = script[pi + 2];
    const u32 io = script[pi + 3];
    const u32 ns = script[pi + 4];
    const u32 so = script[pi + 5];
    const u32 R = script[pi + 6];
    const u32 nd = script[pi + 7];
    u32 hook = script[pi + 128 + (t >> 1)];
    hook = (t & 1) ? (hook >> 16) : (hook & 0xffffu);
    pi += 256;

    // ---- global read: thread t fills positions 32t .. 32t+31 in the
    // packed order (first extracted bit lands highest), zero above
    {
      const uint2 *gr = reinterpret_cast<const uint2 *>(script + pi);
      u32 B = 0;
      for (u32 r = 0; r < R; ++r) B += __popc(gr[256 * r + t].y);
      u32 pos = B;
      for (u32 r = 0; r < R; ++r) {
        uint2 e = gr[256 * r + t];
        if (!e.y) continue;
        const u32 *base = (e.x >> 31) ? out : in;
        const u32 *wbase = base + (size_t)(e.x & 0x7fffffffu) * 32;
        u32 m = e.y;
        while (m) {
          u32 k = __ffs(m) - 1;
          m &= m - 1;
          --pos;
          bufA[32 * t + pos] = wbase[k];
        }
      }
      for (u32 q = B; q < 32; ++q) bufA[32 * t + q] = 0;
      pi += (size_t)512 * R;
    }
    __syncthreads();

    u32 *prev = bufA, *cur = bufB;
    for (u32 s = 0; s < S; ++s) {
      const u32 *sec = script + pi;
      const u32 *fl = sec + 4096;
      pi += 5120;
      // levels 1..5 of the subtree rooted at position 256 + t
      u32 L1[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        u32 j = t + 256u * k;
        u32 a = prev[perm_at(sec, j)];
        u32 b = prev[perm_at(sec, j + 4096)];
        L1[k] = lgate(a, b, flags_at(fl, 4096 + j));
        cur[4096 + j] = L1[k];
      }
      u32 L2[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        u32 p = 2048 + t + 256u * k;
        L2[k] = lgate(L1[k], L1[k + 8], flags_at(fl, p));
        cur[p] = L2[k];
      }
      u32 L3[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        u32 p = 1024 + t + 256u * k;
        L3[k] = lgate(L2[k], L2[k + 4], flags_at(fl, p));
        cur[p] = L3[k];
      }
      u32 L4[2];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        u32 p = 512 + t + 256u * k;
        L4[k] = lgate(L3[k], L3[k + 2], flags_at(fl, p));
        cur[p] = L4[k];
      }
      cur[256 + t] = lgate(L4[0], L4[1], flags_at(fl, 256 + t));
      __syncthreads();
      if (t < 64) {
        // level 6: positions 128+L and 192+L; level 7: 64+L
        u32 o6a = lgate(cur[256 + lane], cur[384 + lane], flags_at(fl, 128 + lane));
        u32 o6b = lgate(cur[320 + lane], cur[448 + lane], flags_at(fl, 192 + lane));
        u32 o7 = lgate(o6a, o6b, flags_at(fl, 64 + lane));
        // level 8: lanes 32..63 -> position L from L+32 (lane L-32) and
        // L+64 (own level-7 word)
        u32 c = 0;
        u32 a8 = __shfl(o7, (lane + 32) & 63);
        if (lane >= 32) c = lgate(a8, o7, flags_at(fl, lane));
#pragma unroll
        for (int len = 16; len >= 1; len >>= 1) {
          u32 a = __shfl(c, (lane + len) & 63);
          u32 b = __shfl(c, (lane + 2 * len) & 63);
          if (lane >= (u32)len && lane < 2u * len) c = lgate(a, b, flags_at(fl, lane));
        }
        cur[128 + lane] = o6a;
        cur[192 + lane] = o6b;
        cur[64 + lane] = o7;
        cur[lane] = c;  // lane 0 keeps position 0 at zero
      }
      __syncthreads();
      if ((hook >> 8) == s) {
        const uint4 *src = reinterpret_cast<const uint4 *>(cur + 32 * (hook & 255));
        uint4 *dst = reinterpret_cast<uint4 *>(WO + 32 * t);
#pragma unroll
        for (int q = 0; q < 8; ++q) dst[q] = src[q];
      }
      u32 *tmp = prev;
      prev = cur;
      cur = tmp;
    }
    __syncthreads();

    // ---- SRAM + duplicate permutation into bufA (free now)
    const u32 nperm = ns * 4 + nd;
    {
      const u32 *sec = script + pi;
      const u32 *fl = sec + 4096;
      for (u32 q = t; q < nperm * 32; q += 256) {
        uint4 f = *reinterpret_cast<const uint4 *>(fl + 4 * (q >> 5));
        u32 v = WO[perm_at(sec, q)];
        bufA[q] = (v & ~fbit(f.y, q)) ^ fbit(f.x, q);
      }
      pi += 5120;
    }
    __syncthreads();
    // SRAM ports: thread per (sram, lane)
    for (u32 base = 0; base < ns * 32; base += 256) {
      u32 id = base + t;
      u32 s = id >> 5, l = id & 31;
      u32 r = 0;
      bool act = id < ns * 32;
      if (act) {
        const u32 *P = bufA + 128 * s;
        u32 ra = 0, wa = 0, en = 0, data = 0;
        for (int k = 0; k < 13; ++k) {
          ra |= ((P[k] >> l) & 1u) << k;
          wa |= ((P[16 + k] >> l) & 1u) << k;
        }
        for (int k = 0; k < 32; ++k) {
          en |= ((P[32 + k] >> l) & 1u) << k;
          data |= ((P[64 + k] >> l) & 1u) << k;
        }
        u32 *ram = ram_group + (size_t)l * A.sram_size + so + (size_t)s * kSramWords;
        r = ram[ra];
        u32 w0 = ram[wa];
        ram[wa] = (w0 & ~en) | (data & en);
      }
      // transpose the 32 read words of each SRAM into lane-words
      u32 mine = 0;
      for (int k = 0; k < 32; ++k) {
        unsigned long long m = __ballot(act && ((r >> k) & 1u));
        u32 half = (u32)(m >> (32 * ((t >> 5) & 1)));
        if (l == (u32)k) mine = half;
      }
      if (act) WO[32 * (num_ios - ns + s) + l] = mine;
    }
    for (u32 q = t; q < nd * 32; q += 256)
      WO[32 * (num_ios - ns - nd) + q] = bufA[ns * 128 + q];
    __syncthreads();

    // ---- clock-enable permutation + commit (coalesced lane-words)
    {
      const u32 *sec = script + pi;
      const u32 *fl = sec + 4096;
      u32 *o = out + (size_t)io * 32;
      const u32 *i0 = in + (size_t)io * 32;
      for (u32 q = t; q < num_ios * 32; q += 256) {
        uint4 f = *reinterpret_cast<const uint4 *>(fl + 4 * (q >> 5));
        u32 ck = (WO[perm_at(sec, q)] & ~fbit(f.y, q)) ^ fbit(f.x, q);
        u32 d = WO[q] ^ fbit(f.z, q);
        o[q] = (i0[q] & ~ck) | (d & ck);
      }
      pi += 5120;
    }
    __syncthreads();
    if (is_last) break;
  }
}

__global__ __launch_bounds__(256) void batched_kernel(BArgs A) {
  extern __shared__ u32 lds[];
  const u32 g = blockIdx.x / A.W, b = blockIdx.x % A.W;
  u32 *bar = A.bar + kSlotWords * g;
  u32 gen = read_gen(bar);
  const u32 ncyc = A.gcycles[g];
  const size_t snap = (size_t)A.rio * 32;
  u32 *st0 = A.states + (size_t)g * 2 * snap;
  u32 *st1 = st0 + snap;
  u32 *ram_group = A.sram + (size_t)g * 32 * A.sram_size;
  const u32 t = threadIdx.x;
  __syncthreads();
  for (u32 c = 0; c < ncyc; ++c) {
    const u32 *in = (c & 1) ? st1 : st0;
    u32 *out = (c & 1) ? st0 : st1;
    if (b == 0 && c > 0) {
      u32 *o = A.obs_out + ((size_t)g * A.max_cycles + (c - 1)) * A.n_obs;
      for (u32 k = t; k < A.n_obs; k += 256) o[k] = in[A.obs_bits[k]];
    }
    // stimulus of snapshot c+1: data bit k toggles for lane l when
    // hash(seed_l, c+1, k) < thresh_l
    for (u32 k0 = b * 8; k0 < A.n_data; k0 += A.W * 8) {
      u32 k = k0 + (t >> 5), l = t & 31;
      bool act = k < A.n_data;
      u32 bit = 0;
      if (act) {
        u32 gl = g * 32 + l;
        u32 h = mix32(mix32(A.lane_seed[gl] + 0x9E3779B9u * k) ^ (c + 1));
        bit = h < A.lane_thresh[gl];
      }
      unsigned long long m = __ballot(act && bit);
      u32 tog = (u32)(m >> (32 * ((t >> 5) & 1)));
      if (act && l == 0) {
        u32 p = A.data_bits[k];
        out[p] = in[p] ^ tog;
      }
    }
    for (u32 mj = 0; mj < A.num_major; ++mj) {
      usize s0 = A.bs[(size_t)mj * A.W + b];
      run_block_lanes(A, A.bd + s0, in, out, ram_group, lds);
      grid_sync(bar, A.W, gen, &gb_error);
    }
  }
  if (b == 0 && ncyc > 0) {
    const u32 *fin = (ncyc & 1) ? st1 : st0;
    u32 *o = A.obs_out + ((size_t)g * A.max_cycles + (ncyc - 1)) * A.n_obs;
    for (u32 k = t; k < A.n_obs; k += 256) o[k] = fin[A.obs_bits[k]];
  }
}

u32 b_slot = 0;

}  // namespace

int batched_error_flag(u32 *out, int clear) {
  u32 v = 0;
  if (hipMemcpyFromSymbol(&v, HIP_SYMBOL(gb_error), 4) != hipSuccess) return -1;
  if (clear) {
    u32 z = 0;
    if (hipMemcpyToSymbol(HIP_SYMBOL(gb_error), &z, 4) != hipSuccess) return -1;
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

extern "C" int jq_simulate_batched(
    usize num_groups, usize blocks_per_group, usize num_major_stages,
    const usize *blocks_start, const u32 *blocks_data, u32 reg_io_state_size,
    u32 sram_storage_size, u32 *sram_data, const u32 *group_cycles,
    u32 max_cycles, u32 *states, u32 n_pi_words, const u32 *data_bits,
    u32 n_data, const u32 *lane_seed, const u32 *lane_thresh,
    const u32 *obs_bits, u32 num_obs, u32 *obs_out, void *stream) {
  set_error("");
  if (num_groups == 0 || blocks_per_group == 0) return 0;
  if (num_groups > (usize)kBarrierSlots) {
    return set_error("too many lane groups");
  }
  int per = 0;
  if (batched_blocks_per_cu(&per)) {
    return set_error("batched occupancy query failed");
  }
  BArgs A{};
  A.G = (u32)num_groups;
  A.W = (u32)blocks_per_group;
  A.num_major = (u32)num_major_stages;
  A.bs = blocks_start;
  A.bd = blocks_data;
  A.rio = reg_io_state_size;
  A.sram_size = sram_storage_size;
  A.sram = sram_data;
  A.gcycles = group_cycles;
  A.max_cycles = max_cycles;
  A.states = states;
  A.n_pi_words = n_pi_words;
  A.data_bits = data_bits;
  A.n_data = n_data;
  A.lane_seed = lane_seed;
  A.lane_thresh = lane_thresh;
  A.obs_bits = obs_bits;
  A.n_obs = num_obs;
  A.obs_out = obs_out;
  u32 *bar_base = nullptr;
  if (hipGetSymbolAddress((void **)&bar_base, HIP_SYMBOL(gb_barrier)) != hipSuccess) {
    return set_error("hipGetSymbolAddress(gb_barrier) failed");
  }
  u32 first = __atomic_fetch_add(&b_slot, (u32)num_groups, __ATOMIC_RELAXED) % kBarrierSlots;
  if (first + num_groups > (usize)kBarrierSlots) first = 0;
  A.bar = bar_base + kSlotWords * first;
  void *args[] = {&A};
  hipError_t e = hipLaunchCooperativeKernel(
      reinterpret_cast<const void *>(&batched_kernel),
      dim3((u32)(num_groups * blocks_per_group)), dim3(256), args, kLdsWords * 4,
      (hipStream_t)stream);
  if (e != hipSuccess) {
    return set_error((std::string("batched launch failed: ") + hipGetErrorString(e)).c_str());
  }
  return 0;
}
