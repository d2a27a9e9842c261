// Reactive co-simulation on the device: the work between two simulated clock
// edges of the reference's Metal co-simulation (state preparation, SPI flash
// data-in and the SPI flash model, csrc/kernel_v1.metal:654-1025), driven
// edge by edge as src/sim/cosim_metal.rs:547-651 does:
//
//   per edge e:  prepare(e) -> flash data-in -> simulate one cycle
//                -> flash model step (sees the outputs of edge e)
//
// The state buffer holds two snapshots: the input of the edge at [0, ss) and
// its output at [ss, 2ss).  The flash step of edge e and the preparation of
// edge e+1 run in one single-workgroup launch (`between_kernel`), so an edge
// costs two launches: that kernel and one cycle of the scan kernel.  Nothing
// returns to the host between edges.
#include <hip/hip_runtime.h>

#include <string>

#include "jq_common.hpp"

namespace jqd {
namespace {

// The reference's FlashState (csrc/kernel_v1.metal:731-749,
// src/sim/cosim_metal.rs:90-111), 48 bytes, persistent across edges.
struct FlashState {
  int32_t bit_count, byte_count;
  u32 data_width, addr;
  u8 curr_byte, command, out_buffer, pad1;
  u32 prev_clk, prev_csn;
  u8 d_i, pad2[3];
  u8 prev_d_out, pad3[3];
  u32 in_reset, last_error_cmd, model_prev_csn;
};
static_assert(sizeof(FlashState) == 48, "FlashState layout");

// Pin words of the flash (a device array of 11 u32): state bits of the
// four data inputs d0..d3 (0xFFFFFFFF: not connected), of the SPI clock and
// chip-select outputs and of the four data outputs, then has_flash.

struct BetweenArgs {
  u32 *states;
  u32 ss;
  const u32 *ops;        // (position, value) pairs of every schedule edge
  const u32 *op_start;   // CSR over schedule edges
  u32 prep_edge;         // schedule edge to prepare, ~0u: none
  u32 flash_step;        // 1: run the flash step on the current outputs first
  FlashState *fs;        // null: no flash
  const u32 *pins;       // 11 pin words
  const u8 *flash_data;
  u32 flash_size;
  u32 *trace;            // null, or 2*ss words of this edge's snapshots
};

// Bit access inside one snapshot of ss words; positions past its end read
// as 0 and are not written.
__device__ __forceinline__ u32 bit_at(const u32 *s, u32 ss, u32 pos) {
  return (pos >> 5) < ss ? (s[pos >> 5] >> (pos & 31u)) & 1u : 0u;
}

__device__ __forceinline__ void put_bit(u32 *s, u32 ss, u32 pos, u32 v) {
  if ((pos >> 5) >= ss) return;
  const u32 m = 1u << (pos & 31u);
  if (v)
    s[pos >> 5] |= m;
  else
    s[pos >> 5] &= ~m;
}

// A byte completed on the SPI bus (csrc/spiflash_model.cc:40-101; the Metal
// port csrc/kernel_v1.metal:767-831): command decode, address bytes, data
// bytes from the image (0xFF past its end) and the read-ID bytes.
__device__ void flash_byte(FlashState &f, const u8 *data, u32 size) {
  f.out_buffer = 0;
  if (f.byte_count == 0) {
    f.addr = 0;
    f.data_width = 1;
    f.command = f.curr_byte;
    switch (f.command) {
      case 0xab: case 0x03: case 0x9f: case 0xff: case 0x35: case 0x31:
      case 0x50: case 0x05: case 0x01: case 0x06:
        break;
      case 0xeb:  // quad read
        f.data_width = 4;
        break;
      default:
        f.last_error_cmd = f.command;
    }
  } else if (f.command == 0x03 || f.command == 0xeb) {
    if (f.byte_count <= 3) f.addr |= (u32)f.curr_byte << ((3 - f.byte_count) * 8);
    // single read: data from the fourth byte on; quad read after a mode
    // byte and two dummy bytes
    if (f.byte_count >= (f.command == 0x03 ? 3 : 6)) {
      const u32 a = f.addr & 0xFFFFFFu;
      f.out_buffer = a < size ? data[a] : (u8)0xFF;
      f.addr = (a + 1u) & 0xFFFFFFu;
    }
  }
  if (f.command == 0x9f) {
    const u32 id = 0xFFA77CCAu;  // bytes CA 7C A7 FF, low byte first
    f.out_buffer = (u8)(id >> (8 * (f.byte_count & 3)));
  }
}

// One evaluation of the flash at the given pins: edges against the model's
// previous clock and chip select (csrc/spiflash_model.cc:117-160;
// csrc/kernel_v1.metal:835-890).  d_i changes only on a falling clock.
__device__ void flash_eval(FlashState &f, u32 clk, u32 csn, u32 d_out, u32 &model_prev_csn,
                           const u8 *data, u32 size) {
  const bool pos_clk = clk && !f.prev_clk;
  const bool neg_clk = !clk && f.prev_clk;
  const bool pos_csn = csn && !model_prev_csn;
  if (pos_csn) {
    f.bit_count = 0;
    f.byte_count = 0;
    f.data_width = 1;
  } else if (pos_clk && !csn) {
    const u32 w = f.data_width;
    f.curr_byte = (u8)(w == 4 ? (f.curr_byte << 4) | (d_out & 0xFu)
                              : (f.curr_byte << 1) | (d_out & 1u));
    f.out_buffer = (u8)(f.out_buffer << w);
    f.bit_count += (int32_t)w;
    if (f.bit_count >= 8) {
      flash_byte(f, data, size);
      ++f.byte_count;
      f.bit_count = 0;
    }
  } else if (neg_clk && !csn) {
    f.d_i = (u8)(f.data_width == 4 ? (f.out_buffer >> 4) & 0xFu
                                   : ((f.out_buffer >> 7) & 1u) << 1);
  }
  f.prev_clk = clk;
  model_prev_csn = csn;
}

// The flash model step after an edge (csrc/kernel_v1.metal:929-1025): the
// outputs of the edge are read from the output snapshot; the model sees the
// clock of this edge with the chip select and data of the previous step (a
// setup delay: clock, chip select and data flip-flops change on the same
// system clock edge).  The reference evaluates twice, with the previous and
// then the current data; the second evaluation sees no clock or chip-select
// edge and changes nothing, so one evaluation is made here.
__device__ void flash_step(const BetweenArgs &A) {
  FlashState f = *A.fs;
  const u32 *out = A.states + A.ss;
  const u32 clk = bit_at(out, A.ss, A.pins[4]);
  const u32 csn = bit_at(out, A.ss, A.pins[5]);
  u32 d_out = 0;
  for (u32 i = 0; i < 4; ++i) d_out |= bit_at(out, A.ss, A.pins[6 + i]) << i;
  if (f.in_reset) {
    f.d_i = 0x0F;
    f.prev_clk = clk;
    f.model_prev_csn = csn;
  } else {
    u32 mpc = f.model_prev_csn;
    flash_eval(f, clk, f.prev_csn, f.prev_d_out, mpc, A.flash_data, A.flash_size);
    f.model_prev_csn = mpc;
  }
  f.prev_csn = csn;
  f.prev_d_out = (u8)d_out;
  *A.fs = f;
}

__global__ __launch_bounds__(256) void between_kernel(BetweenArgs A) {
  const u32 t = threadIdx.x;
  const bool flash = A.fs != nullptr && A.pins[10] != 0;
  if (A.flash_step && flash && t == 0) flash_step(A);
  if (A.trace) {
    for (u32 i = t; i < 2 * A.ss; i += 256) A.trace[i] = A.states[i];
  }
  if (A.prep_edge == ~0u) return;
  __syncthreads();
  // state preparation (csrc/kernel_v1.metal:678-710): the output snapshot
  // becomes the next input, then the edge's bit operations in order
  for (u32 i = t; i < A.ss; i += 256) A.states[i] = A.states[A.ss + i];
  __syncthreads();
  if (t == 0) {
    const u32 b = A.op_start[A.prep_edge], e = A.op_start[A.prep_edge + 1];
    for (u32 k = b; k < e; ++k) put_bit(A.states, A.ss, A.ops[2 * k], A.ops[2 * k + 1]);
    // flash data-in (csrc/kernel_v1.metal:896-918): the model's d_i
    if (flash) {
      const u32 d_i = A.fs->d_i;
      for (u32 i = 0; i < 4; ++i) put_bit(A.states, A.ss, A.pins[i], (d_i >> i) & 1u);
    }
  }
}

int launch_between(const BetweenArgs &A, hipStream_t stream) {
  hipLaunchKernelGGL(between_kernel, dim3(1), dim3(256), 0, stream, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error((std::string("cosim launch failed: ") + hipGetErrorString(e)).c_str());
  return 0;
}

}  // namespace
}  // namespace jqd

using namespace jqd;

extern "C" size_t jq_cosim_flash_state_bytes(void) { return sizeof(FlashState); }

extern "C" int jq_cosim_edges(usize num_blocks, usize num_major_stages,
                              const usize *blocks_start, const u32 *blocks_data,
                              u32 *sram_data, usize state_size, u32 *states,
                              const u32 *edge_ops, const u32 *edge_op_start,
                              u32 num_sched, u32 sched_offset, u32 num_edges,
                              void *flash_state, const u32 *flash_pins,
                              const u8 *flash_data, u32 flash_data_size, u32 *trace,
                              void *stream) {
  set_error("");
  if (num_edges == 0) return 0;
  if (!states || state_size == 0 || state_size > 0xFFFFFFFFull / 2)
    return set_error("cosim: state buffer missing or state_size out of range");
  if (num_sched == 0 || !edge_op_start || !edge_ops)
    return set_error("cosim: empty edge schedule");
  if (flash_state && !flash_pins) return set_error("cosim: flash state without flash pins");
  if (flash_data_size && !flash_data) return set_error("cosim: flash image missing");
  hipStream_t s = (hipStream_t)stream;
  BetweenArgs A{};
  A.states = states;
  A.ss = (u32)state_size;
  A.ops = edge_ops;
  A.op_start = edge_op_start;
  A.fs = reinterpret_cast<FlashState *>(flash_state);
  A.pins = flash_pins;
  A.flash_data = flash_data;
  A.flash_size = flash_data_size;
  for (u32 e = 0; e < num_edges; ++e) {
    A.flash_step = e > 0;
    A.trace = (e > 0 && trace) ? trace + (size_t)(e - 1) * 2 * state_size : nullptr;
    A.prep_edge = (u32)(((usize)sched_offset + e) % num_sched);
    if (launch_between(A, s)) return -1;
    if (jq_simulate_scan(num_blocks, num_major_stages, blocks_start, blocks_data, sram_data,
                         sram_data, 1, state_size, states, nullptr, nullptr, 0, 0, stream))
      return -1;
  }
  A.flash_step = 1;
  A.trace = trace ? trace + (size_t)(num_edges - 1) * 2 * state_size : nullptr;
  A.prep_edge = ~0u;
  return launch_between(A, s);
}
