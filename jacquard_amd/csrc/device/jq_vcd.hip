// Output change extraction for the VCD writer, on the device.
//
// The reference host copies every snapshot back and scans all output bits
// on the CPU for changes (write_output_vcd_impl, src/sim/vcd_io.rs:758-820).
// Here one workgroup per emitted snapshot compares each output's value (0,
// 1, or X from the X mask) with the previous emitted snapshot's and writes
// the changed ones, compacted in output order, as (output << 2 | value)
// records.  Two passes: the first counts changes per snapshot, the caller
// turns the counts into exclusive offsets, the second writes the records.
// Only the records cross PCIe.
#include <hip/hip_runtime.h>

#include <string>

#include "jq_common.hpp"

namespace jqd {
namespace {

constexpr u32 kConst0 = 0xFFFFFFFFu;  // output tied to 0
constexpr u32 kConst1 = 0xFFFFFFFEu;  // output tied to 1
constexpr u32 kNone = 2u;             // "no previous value": every output is emitted

__device__ __forceinline__ u32 out_value(const u32 *snap, u32 bit, u32 xoff) {
  if (bit == kConst0) return 0u;
  if (bit == kConst1) return 1u;
  const u32 w = bit >> 5, s = bit & 31;
  if (xoff && ((snap[xoff + w] >> s) & 1u)) return 3u;
  return (snap[w] >> s) & 1u;
}

__global__ __launch_bounds__(256) void changes_kernel(
    const u32 *__restrict__ states, const usize *__restrict__ snap_off, u32 n_snap,
    const u32 *__restrict__ bits, u32 n_out, u32 xoff, u32 *__restrict__ counts,
    u32 *__restrict__ records, const usize *__restrict__ rec_off) {
  __shared__ u32 wave_sum[4];
  const u32 k = blockIdx.x;
  if (k >= n_snap) return;
  const u32 t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const u32 *cur = states + snap_off[k];
  const u32 *prv = k ? states + snap_off[k - 1] : nullptr;
  u32 base = 0;
  for (u32 i0 = 0; i0 < n_out; i0 += 256) {
    const u32 i = i0 + t;
    u32 v = 0;
    bool ch = false;
    if (i < n_out) {
      const u32 b = bits[i];
      v = out_value(cur, b, xoff);
      ch = v != (prv ? out_value(prv, b, xoff) : kNone);
    }
    const unsigned long long m = __ballot(ch);
    const u32 below = (u32)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_sum[wave] = (u32)__popcll(m);
    __syncthreads();
    u32 before = 0, total = 0;
    for (u32 w = 0; w < 4; ++w) {
      before += w < wave ? wave_sum[w] : 0u;
      total += wave_sum[w];
    }
    if (ch && records) records[rec_off[k] + base + before + below] = (i << 2) | v;
    base += total;
    __syncthreads();
  }
  if (t == 0 && counts) counts[k] = base;
}

}  // namespace
}  // namespace jqd

using namespace jqd;

extern "C" int jq_output_changes(const u32 *states, const usize *snapshot_offsets,
                                 u32 num_snapshots, const u32 *output_bits,
                                 u32 num_outputs, u32 xmask_offset, u32 *counts,
                                 u32 *records, const usize *record_offsets,
                                 void *stream) {
  set_error("");
  if (num_snapshots == 0) return 0;
  if (!states || !snapshot_offsets || (num_outputs && !output_bits))
    return set_error("jq_output_changes: missing buffer");
  if (records && !record_offsets)
    return set_error("jq_output_changes: records need record_offsets");
  if (num_outputs >= (1u << 30)) return set_error("jq_output_changes: too many outputs");
  hipLaunchKernelGGL(changes_kernel, dim3(num_snapshots), dim3(256), 0,
                     (hipStream_t)stream, states, snapshot_offsets, num_snapshots,
                     output_bits, num_outputs, xmask_offset, counts, records,
                     record_offsets);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error((std::string("output change launch failed: ") + hipGetErrorString(e)).c_str());
  return 0;
}
