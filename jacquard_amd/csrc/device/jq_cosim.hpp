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

// UART decoders and the Wishbone bus trace (csrc/kernel_v1.metal:1027-1101),
// the same layouts: up to 4 UARTs, 4096-byte channels, 30 address and 32
// data bits.
constexpr u32 kMaxUarts = 4;
constexpr u32 kUartChannelCap = 4096;
constexpr u32 kWbAdrBits = 30;
constexpr u32 kWbDatBits = 32;

struct UartState {
  u32 state;  // 0 idle, 1 start bit, 2 data bits, 3 stop bit
  u32 last_tx, start_cycle, bits_received, value, current_cycle;
};
struct UartParams {
  u32 state_size, n_uarts, pad[2];
  struct {
    u32 tx_out_pos, cycles_per_bit;
  } ch[kMaxUarts];
};
struct UartChannel {
  u32 write_head, capacity, pad[2];
  u8 data[kUartChannelCap];
};
struct WbTraceParams {
  u32 ibus_cyc, ibus_stb, ibus_adr[kWbAdrBits], ibus_rdata[kWbDatBits];
  u32 dbus_cyc, dbus_stb, dbus_we, dbus_adr[kWbAdrBits];
  u32 spiflash_ack, sram_ack, csr_ack, has_trace;
};
struct WbTraceEntry {
  u32 tick, flags, ibus_adr, ibus_rdata, dbus_adr;
};
struct WbTraceChannel {
  u32 write_head, capacity, current_tick, prev_flags;  // entries follow
};
static_assert(sizeof(UartState) == 24 && sizeof(UartParams) == 48, "UART layouts");
static_assert(sizeof(UartChannel) == 16 + kUartChannelCap, "UART channel layout");
static_assert(sizeof(WbTraceParams) == 4 * (2 + 30 + 32 + 3 + 30 + 4), "trace layout");
static_assert(sizeof(WbTraceEntry) == 20 && sizeof(WbTraceChannel) == 16, "trace layout");

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
  u32 *fsbuf;            // 2 x 12 words: flash states after even / odd steps
  UartState *us;         // null: no UART decoding
  const UartParams *up;
  UartChannel *uc;
  WbTraceChannel *wbc;   // null: no bus trace
  const WbTraceParams *wbp;
};

// Bit access inside one snapshot of ss words; positions past its end read
// as 0.
__device__ __forceinline__ u32 bit_at(const u32 *s, u32 ss, u32 pos) {
  return (pos >> 5) < ss ? (ld_coherent(s + (pos >> 5)) >> (pos & 31u)) & 1u : 0u;
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
__device__ FlashState flash_step(const CosimArgs &A, FlashState f) {
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
  return f;
}

// FlashState words moved agent-coherently (the double buffer is read by
// workgroups on other XCDs).
__device__ __forceinline__ FlashState load_fs(const u32 *p) {
  FlashState f;
  u32 *w = reinterpret_cast<u32 *>(&f);
  for (u32 i = 0; i < sizeof(FlashState) / 4; ++i) w[i] = ld_coherent(p + i);
  return f;
}
__device__ __forceinline__ void store_fs(u32 *p, const FlashState &f) {
  const u32 *w = reinterpret_cast<const u32 *>(&f);
  for (u32 i = 0; i < sizeof(FlashState) / 4; ++i) st_coherent(p + i, w[i]);
}

// The UART decoders and the bus trace after an edge (gpu_io_step,
// csrc/kernel_v1.metal:1103-1231): each UART samples its TX output once per
// edge (a start bit confirmed at half a bit, data bits at bit centres, LSB
// first, the byte kept when the stop bit reads 1); the trace records the
// bus signals when a bus cycle is active or any flag changed.
__device__ void io_step(const CosimArgs &A) {
  const u32 *out = A.states + A.ss;
  if (A.us && A.up && A.uc) {
    const u32 n = A.up->n_uarts < kMaxUarts ? A.up->n_uarts : kMaxUarts;
    for (u32 ui = 0; ui < n; ++ui) {
      const u32 cpb = A.up->ch[ui].cycles_per_bit;
      const u32 tx = bit_at(out, A.ss, A.up->ch[ui].tx_out_pos);
      UartState u = A.us[ui];
      const u32 cyc = u.current_cycle;
      if (u.state == 0) {
        if (u.last_tx == 1 && tx == 0) {
          u.state = 1;
          u.start_cycle = cyc;
        }
      } else if (u.state == 1) {
        if (cyc >= u.start_cycle + cpb / 2) {
          if (tx == 0) {
            u.state = 2;
            u.start_cycle += cpb;
            u.bits_received = 0;
            u.value = 0;
          } else {
            u.state = 0;
          }
        }
      } else if (u.state == 2) {
        if (cyc >= u.start_cycle + u.bits_received * cpb + cpb / 2) {
          u.value |= tx << u.bits_received;
          if (u.bits_received >= 7) {
            u.state = 3;
            u.start_cycle += 8 * cpb;
          } else {
            ++u.bits_received;
          }
        }
      } else if (u.state == 3) {
        if (cyc >= u.start_cycle + cpb / 2) {
          if (tx == 1) {
            UartChannel &c = A.uc[ui];
            const u32 cap = c.capacity - 1u < kUartChannelCap ? c.capacity : kUartChannelCap;
            c.data[c.write_head % cap] = (u8)(u.value & 0xFFu);
            c.write_head += 1;
          }
          u.state = 0;
        }
      }
      u.last_tx = tx;
      u.current_cycle = cyc + 1;
      A.us[ui] = u;
    }
  }
  if (A.wbc && A.wbp && A.wbp->has_trace) {
    const WbTraceParams &P = *A.wbp;
    const u32 ic = bit_at(out, A.ss, P.ibus_cyc), is = bit_at(out, A.ss, P.ibus_stb);
    const u32 dc = bit_at(out, A.ss, P.dbus_cyc), ds = bit_at(out, A.ss, P.dbus_stb);
    const u32 flags = ic | is << 1 | dc << 2 | ds << 3 | bit_at(out, A.ss, P.dbus_we) << 4 |
                      bit_at(out, A.ss, P.spiflash_ack) << 5 |
                      bit_at(out, A.ss, P.sram_ack) << 6 | bit_at(out, A.ss, P.csr_ack) << 7;
    WbTraceChannel &c = *A.wbc;
    if ((ic && is) || (dc && ds) || flags != c.prev_flags) {
      u32 ia = 0, id = 0, da = 0;
      for (u32 i = 0; i < kWbAdrBits; ++i) ia |= bit_at(out, A.ss, P.ibus_adr[i]) << i;
      for (u32 i = 0; i < kWbDatBits; ++i) id |= bit_at(out, A.ss, P.ibus_rdata[i]) << i;
      for (u32 i = 0; i < kWbAdrBits; ++i) da |= bit_at(out, A.ss, P.dbus_adr[i]) << i;
      if (c.write_head < c.capacity) {
        WbTraceEntry *e = reinterpret_cast<WbTraceEntry *>(&c + 1) + c.write_head;
        *e = WbTraceEntry{c.current_tick, flags, ia, id, da};
        c.write_head += 1;
      }
    }
    c.prev_flags = flags;
    c.current_tick += 1;
  }
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

// Between cycle e-1 and cycle e (e = 0: before the first cycle; e =
// num_edges: after the last), in every workgroup b of nb.  The snapshot
// words are split in nb ranges; workgroup b, for its range only:
//   * copies both snapshots of edge e-1 to the trace;
//   * copies the output snapshot over the input snapshot and applies the
//     bit operations of edge e that fall in the range, in order, then the
//     flash's data-in bits that fall in it.
// The flash model is a pure function of its previous state and the edge's
// outputs, so every workgroup that owns a data-in word steps it itself from
// the state of the previous step; workgroup 0 keeps the states, alternately
// in fsbuf[0..11] and fsbuf[12..23] (state k in buffer k & 1, state 0 = the
// caller's), so no workgroup reads a buffer that is being written.
// Workgroup 0 also runs the UART decoders and the bus trace.  The caller
// waits for the stores and meets the grid barrier afterwards.
__device__ void between_phase(const CosimArgs &A, u32 t, u32 b, u32 nb, u32 e) {
  const bool flash = A.fs != nullptr && A.pins[10] != 0;
  const u32 chunk = (A.ss + nb - 1) / nb;
  const u32 lo = b * chunk < A.ss ? b * chunk : A.ss;
  const u32 hi = lo + chunk < A.ss ? lo + chunk : A.ss;
  const bool last = e == A.num_edges;
  auto owns = [&](u32 pos) { return (pos >> 5) >= lo && (pos >> 5) < hi; };
  bool need_fs = false;
  u32 din = 0;  // the flash's d_i for this edge's input (thread 0)
  if (flash)
    for (u32 i = 0; i < 4; ++i) need_fs |= owns(A.pins[i]);
  if (t == 0) {
    if (b == 0 && e > 0) io_step(A);
    if (flash && (b == 0 || need_fs)) {
      FlashState f;
      if (e == 0) {
        f = *A.fs;
        if (b == 0) store_fs(A.fsbuf, f);
      } else {
        f = flash_step(A, load_fs(A.fsbuf + 12 * ((e - 1) & 1)));
        if (b == 0) {
          store_fs(A.fsbuf + 12 * (e & 1), f);
          if (last) *A.fs = f;
        }
      }
      din = f.d_i;
    }
  }
  if (e > 0 && A.trace) {
    u32 *tr = A.trace + (size_t)(e - 1) * 2 * A.ss;
    copy_words(tr + lo, A.states + lo, hi - lo, t, false);
    copy_words(tr + A.ss + lo, A.states + A.ss + lo, hi - lo, t, false);
  }
  if (last) return;
  __shared__ u32 s_din;
  if (t == 0) s_din = din;
  __syncthreads();
  // state preparation (csrc/kernel_v1.metal:678-710, 896-918): the output
  // snapshot becomes the next input, with the edge's bit operations in order
  // and then the flash's data-in bits applied to each word on its way
  // (kCopyBatch agent-coherent loads in flight per thread)
  const u32 pe = (u32)(((usize)A.sched_offset + e) % A.num_sched);
  const u32 ob = A.op_start[pe], oe = A.op_start[pe + 1];
  const u32 dbits = s_din;
  for (u32 base = lo + t; base < hi; base += 256 * kCopyBatch) {
    u32 v[kCopyBatch];
#pragma unroll
    for (u32 k = 0; k < kCopyBatch; ++k) {
      const u32 i = base + 256 * k;
      v[k] = i < hi ? ld_coherent(A.states + A.ss + i) : 0u;
    }
#pragma unroll
    for (u32 k = 0; k < kCopyBatch; ++k) {
      const u32 i = base + 256 * k;
      if (i >= hi) continue;
      u32 w = v[k];
      for (u32 j = ob; j < oe; ++j) {
        const u32 pos = A.ops[2 * j];
        if ((pos >> 5) == i) {
          const u32 m = 1u << (pos & 31u);
          w = A.ops[2 * j + 1] ? (w | m) : (w & ~m);
        }
      }
      if (flash)
        for (u32 q = 0; q < 4; ++q)
          if ((A.pins[q] >> 5) == i) {
            const u32 m = 1u << (A.pins[q] & 31u);
            w = (dbits >> q) & 1u ? (w | m) : (w & ~m);
          }
      st_coherent(A.states + i, w);
    }
  }
}

}  // namespace cosim
}  // namespace jqd
