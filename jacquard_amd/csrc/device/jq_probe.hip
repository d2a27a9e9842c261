// LDS read-bandwidth probe: the measured peak the LDS roofline of the
// interpreter kernels is quoted against.  Every workgroup fills 64 KB of LDS
// and streams it back with 16-byte reads, 64 consecutive lanes reading 1 KB
// of consecutive banks, eight independent reads in flight per iteration.
#include <hip/hip_runtime.h>

#include <string>

#include "jq_common.hpp"

namespace jqd {
namespace {

constexpr u32 kProbeVecs = 4096;  // 64 KB of uint4

__global__ __launch_bounds__(256) void lds_read_kernel(u32 iters, u32 *__restrict__ sink) {
  __shared__ uint4 buf[kProbeVecs];
  const u32 t = threadIdx.x;
  for (u32 i = t; i < kProbeVecs; i += 256)
    buf[i] = make_uint4(i, i * 3u + blockIdx.x, i * 5u, i * 7u);
  __syncthreads();
  uint4 acc = make_uint4(0u, 0u, 0u, 0u);
  for (u32 i = 0; i < iters; ++i) {
#pragma unroll
    for (u32 k = 0; k < 8; ++k) {
      // row (8i + k) mod 16 of 256 vectors: lanes hit consecutive banks
      const uint4 v = buf[(t + ((i * 8u + k) & 15u) * 256u) & (kProbeVecs - 1u)];
      acc.x ^= v.x;
      acc.y += v.y;
      acc.z ^= v.z;
      acc.w += v.w;
    }
  }
  sink[(size_t)blockIdx.x * 256u + t] = acc.x ^ acc.y ^ acc.z ^ acc.w;
}

}  // namespace
}  // namespace jqd

using namespace jqd;

extern "C" int jq_lds_read_probe(u32 num_blocks, u32 iters, u32 *sink, void *stream) {
  set_error("");
  if (num_blocks == 0 || num_blocks > 65535u || iters == 0 || !sink)
    return set_error("jq_lds_read_probe: bad arguments");
  hipLaunchKernelGGL(lds_read_kernel, dim3(num_blocks), dim3(256), 0,
                     (hipStream_t)stream, iters, sink);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error((std::string("LDS probe launch failed: ") + hipGetErrorString(e)).c_str());
  return 0;
}
