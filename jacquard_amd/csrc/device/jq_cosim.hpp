// Device side of the reactive co-simulation: the work between two simulated
// clock edges of the reference's Metal co-simulation (state preparation,
// SPI flash data-in and the SPI flash model, csrc/kernel_v1.metal:654-1025),
// in the edge order of src/sim/cosim_metal.rs:574-650:
//
//   per edge e:  prepare(e) -> flash data-in -> simulate one cycle
//                -> flash model step (sees the outputs of edge e)
//
// The state buffer holds two snapshots: the input of the edge at [0, ss) and
// its output at [ss, 2ss).  `between_phase` runs in one workgroup between
// two cycles of the persistent co-simulation kernel (jq_scan.hip); other
// workgroups read and write the snapshots, so every snapshot access here is
// agent-coherent.
#pragma once

#include "jq_common.hpp"

namespace jqd {
namespace cosim {

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

struct CosimArgs {
  u32 *states;
  u32 ss;
  const u32 *ops;        // (position, value) pairs of every schedule edge
  const u32 *op_start;   // CSR over schedule edges
  u32 num_sched, sched_offset, num_edges;
  FlashState *fs;        // null: no flash
  const u32 *pins;       // 11 pin words
  const u8 *flash_data;
  u32 flash_size;
  u32 *trace;            // null, or num_edges x 2*ss words
};

// Bit access inside one snapshot of ss words; positions past its end read
// as 0 and are not written.
__device__ __forceinline__ u32 bit_at(const u32 *s, u32 ss, u32 pos) {
  return (pos >> 5) < ss ? (ld_coherent(s + (pos >> 5)) >> (pos & 31u)) & 1u : 0u;
}

__device__ __forceinline__ void put_bit(u32 *s, u32 ss, u32 pos, u32 v) {
  if ((pos >> 5) >= ss) return;
  const u32 m = 1u << (pos & 31u);
  const u32 w = ld_coherent(s + (pos >> 5));
  st_coherent(s + (pos >> 5), v ? (w | m) : (w & ~m));
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
__device__ void flash_step(const CosimArgs &A) {
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

// dst[i] = src[i] for i < n over one workgroup, src read agent-coherently
// (dst written so when `coherent`), with kCopyBatch loads in flight per
// thread before their stores.
constexpr u32 kCopyBatch = 8;
__device__ __forceinline__ void copy_words(u32 *dst, const u32 *src, u32 n, u32 t,
                                           bool coherent) {
  for (u32 base = t; base < n; base += 256 * kCopyBatch) {
    u32 v[kCopyBatch];
#pragma unroll
    for (u32 k = 0; k < kCopyBatch; ++k) {
      const u32 i = base + 256 * k;
      v[k] = i < n ? ld_coherent(src + i) : 0u;
    }
#pragma unroll
    for (u32 k = 0; k < kCopyBatch; ++k) {
      const u32 i = base + 256 * k;
      if (i < n) {
        if (coherent)
          st_coherent(dst + i, v[k]);
        else
          dst[i] = v[k];
      }
    }
  }
}

// Between cycle e-1 and cycle e, in one workgroup (t = thread index):
// the flash step on the outputs of edge e-1 and its trace copy (when
// flash_step), then the preparation of edge e (prep_edge != ~0u).  The
// caller waits for this workgroup's stores before the grid barrier.
__device__ void between_phase(const CosimArgs &A, u32 t, bool step, u32 prep_edge,
                              u32 *trace) {
  const bool flash = A.fs != nullptr && A.pins[10] != 0;
  if (step && flash && t == 0) flash_step(A);
  if (trace) copy_words(trace, A.states, 2 * A.ss, t, false);
  if (prep_edge == ~0u) return;
  __syncthreads();
  // state preparation (csrc/kernel_v1.metal:678-710): the output snapshot
  // becomes the next input, then the edge's bit operations in order
  copy_words(A.states, A.states + A.ss, A.ss, t, true);
  wait_vmem();
  __syncthreads();
  if (t == 0) {
    const u32 b = A.op_start[prep_edge], e = A.op_start[prep_edge + 1];
    for (u32 k = b; k < e; ++k) put_bit(A.states, A.ss, A.ops[2 * k], A.ops[2 * k + 1]);
    // flash data-in (csrc/kernel_v1.metal:896-918): the model's d_i
    if (flash) {
      const u32 d_i = A.fs->d_i;
      for (u32 i = 0; i < 4; ++i) put_bit(A.states, A.ss, A.pins[i], (d_i >> i) & 1u);
    }
  }
}

}  // namespace cosim
}  // namespace jqd
