"""CPU restatement of the reactive co-simulation edge loop.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
co-simulation path (jq_cosim_edges), never by the product package.

Follows:
  * the SPI flash model, csrc/spiflash_model.cc:17-160 (its eval/commit
    step), with the device state of the Metal port,
    csrc/kernel_v1.metal:731-890 (FlashState, last_error_cmd for unknown
    commands);
  * the per-edge flash step with the setup delay and the reset override,
    csrc/kernel_v1.metal:929-1025;
  * state preparation and flash data-in, csrc/kernel_v1.metal:678-710 and
    896-918;
  * the UART decoders and the Wishbone bus trace, csrc/kernel_v1.metal:
    1103-1231;
  * the edge order of src/sim/cosim_metal.rs:574-650 (prepare, data-in,
    simulate, flash step, UART and bus trace, snapshot copy).

`Flash.step_model` is checked against the reference's own model compiled
from csrc/spiflash_model.cc (oracle/_ref/libspiflash_ref.so) where that
library is present.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

FLASH_DTYPE = np.dtype([
    ("bit_count", "<i4"), ("byte_count", "<i4"), ("data_width", "<u4"),
    ("addr", "<u4"), ("curr_byte", "u1"), ("command", "u1"),
    ("out_buffer", "u1"), ("pad1", "u1"), ("prev_clk", "<u4"),
    ("prev_csn", "<u4"), ("d_i", "u1"), ("pad2", "u1", (3,)),
    ("prev_d_out", "u1"), ("pad3", "u1", (3,)), ("in_reset", "<u4"),
    ("last_error_cmd", "<u4"), ("model_prev_csn", "<u4")])
assert FLASH_DTYPE.itemsize == 48

KNOWN_COMMANDS = (0xab, 0x03, 0x9f, 0xff, 0x35, 0x31, 0x50, 0x05, 0x01, 0x06)
READ_ID = (0xCA, 0x7C, 0xA7, 0xFF)


class Flash:
    """The SPI flash model's state and steps."""

    def __init__(self, data, cosim_init=True):
        self.data = bytes(data)
        self.bit_count = 0
        self.byte_count = 0
        self.data_width = 1
        self.addr = 0
        self.curr_byte = 0
        self.command = 0
        self.out_buffer = 0
        self.prev_clk = 0
        self.last_error_cmd = 0
        if cosim_init:
            # the co-simulation host's initial FlashState
            # (src/sim/cosim_metal.rs:2727-2734)
            self.prev_csn = 1
            self.model_prev_csn = 1
            self.d_i = 0x0F
            self.in_reset = 1
        else:
            # the C++ model's member initialisers (spiflash_model.cc:29-37)
            self.prev_csn = 1
            self.model_prev_csn = 1
            self.d_i = 0
            self.in_reset = 0
        self.prev_d_out = 0

    def _byte(self):
        # spiflash_model.cc:40-101
        self.out_buffer = 0
        if self.byte_count == 0:
            self.addr = 0
            self.data_width = 1
            self.command = self.curr_byte
            if self.command == 0xeb:
                self.data_width = 4
            elif self.command not in KNOWN_COMMANDS:
                self.last_error_cmd = self.command
        elif self.command in (0x03, 0xeb):
            if self.byte_count <= 3:
                self.addr |= self.curr_byte << ((3 - self.byte_count) * 8)
            first = 3 if self.command == 0x03 else 6
            if self.byte_count >= first:
                a = self.addr & 0xFFFFFF
                self.out_buffer = self.data[a] if a < len(self.data) else 0xFF
                self.addr = (a + 1) & 0xFFFFFF
        if self.command == 0x9f:
            self.out_buffer = READ_ID[self.byte_count % 4]

    def step_model(self, clk, csn, d_o):
        """One eval + commit of the C++ model (spiflash_model.cc:117-177)
        with the model's previous chip select in model_prev_csn.  Returns
        d_i."""
        clk, csn = int(bool(clk)), int(bool(csn))
        pos_clk = clk and not self.prev_clk
        neg_clk = (not clk) and self.prev_clk
        pos_csn = csn and not self.model_prev_csn
        if pos_csn:
            self.bit_count = 0
            self.byte_count = 0
            self.data_width = 1
        elif pos_clk and not csn:
            w = self.data_width
            if w == 4:
                self.curr_byte = ((self.curr_byte << 4) | (d_o & 0xF)) & 0xFF
            else:
                self.curr_byte = ((self.curr_byte << 1) | (d_o & 1)) & 0xFF
            self.out_buffer = (self.out_buffer << w) & 0xFF
            self.bit_count += w
            if self.bit_count >= 8:
                self._byte()
                self.byte_count += 1
                self.bit_count = 0
        elif neg_clk and not csn:
            if self.data_width == 4:
                self.d_i = (self.out_buffer >> 4) & 0xF
            else:
                self.d_i = ((self.out_buffer >> 7) & 1) << 1
        self.prev_clk = clk
        self.model_prev_csn = csn
        return self.d_i

    def cosim_step(self, clk, csn, d_out):
        """The per-edge flash step (csrc/kernel_v1.metal:929-1025): the
        model sees this edge's clock with the previous step's chip select
        and data; in reset it holds d_i at 0xF."""
        if self.in_reset:
            self.d_i = 0x0F
            self.prev_clk = int(bool(clk))
            self.model_prev_csn = int(bool(csn))
        else:
            self.step_model(clk, self.prev_csn, self.prev_d_out)
        self.prev_csn = int(bool(csn))
        self.prev_d_out = d_out & 0xFF

    def to_record(self):
        r = np.zeros(1, FLASH_DTYPE)
        for name in FLASH_DTYPE.names:
            if not name.startswith("pad"):
                r[name] = getattr(self, name)
        return r


class Uart:
    """One UART TX decoder, sampled once per edge (csrc/kernel_v1.metal:
    1120-1178), with its byte channel."""

    def __init__(self, tx_pos, cycles_per_bit):
        self.tx_pos = tx_pos
        self.cpb = cycles_per_bit
        self.state = 0
        self.last_tx = 1
        self.start_cycle = 0
        self.bits_received = 0
        self.value = 0
        self.current_cycle = 0
        self.data = []

    def step(self, tx):
        cyc, cpb = self.current_cycle, self.cpb
        if self.state == 0:
            if self.last_tx == 1 and tx == 0:
                self.state, self.start_cycle = 1, cyc
        elif self.state == 1:
            if cyc >= self.start_cycle + cpb // 2:
                if tx == 0:
                    self.state = 2
                    self.start_cycle += cpb
                    self.bits_received = 0
                    self.value = 0
                else:
                    self.state = 0
        elif self.state == 2:
            if cyc >= self.start_cycle + self.bits_received * cpb + cpb // 2:
                self.value |= tx << self.bits_received
                if self.bits_received >= 7:
                    self.state = 3
                    self.start_cycle += 8 * cpb
                else:
                    self.bits_received += 1
        elif self.state == 3:
            if cyc >= self.start_cycle + cpb // 2:
                if tx == 1:
                    self.data.append(self.value & 0xFF)
                self.state = 0
        self.last_tx = tx
        self.current_cycle = cyc + 1

    def record(self):
        return [self.state, self.last_tx, self.start_cycle, self.bits_received,
                self.value, self.current_cycle]


class BusTrace:
    """The Wishbone bus trace (csrc/kernel_v1.metal:1181-1229)."""

    def __init__(self, sig, capacity):
        self.sig = sig
        self.capacity = capacity
        self.entries = []
        self.write_head = 0
        self.current_tick = 0
        self.prev_flags = 0

    def step(self, out):
        def b(name):
            return _bit(out, self.sig[name]) if name in self.sig else 0

        def bus(name, width):
            bits = list(self.sig.get(name, ()))[:width]
            return sum(_bit(out, p) << i for i, p in enumerate(bits))

        names = ("ibus_cyc", "ibus_stb", "dbus_cyc", "dbus_stb", "dbus_we",
                 "spiflash_ack", "sram_ack", "csr_ack")
        flags = sum(b(n) << i for i, n in enumerate(names))
        active = (flags & 3) == 3 or (flags & 12) == 12
        if active or flags != self.prev_flags:
            if self.write_head < self.capacity:
                self.entries.append((self.current_tick, flags, bus("ibus_adr", 30),
                                     bus("ibus_rdata", 32), bus("dbus_adr", 30)))
                self.write_head += 1
        self.prev_flags = flags
        self.current_tick += 1


def _bit(words, pos):
    return int((int(words[pos >> 5]) >> (pos & 31)) & 1) if (pos >> 5) < len(words) else 0


def _put(words, pos, v):
    if (pos >> 5) >= len(words):
        return
    m = np.uint32(1 << (pos & 31))
    if v:
        words[pos >> 5] |= m
    else:
        words[pos >> 5] &= ~m


def run_edges(script, states, ops, op_start, sched_offset, num_edges, flash=None,
              pins=None, uarts=(), bus=None, constraints=None, events=None,
              arrival_state_offset=0):
    """The edge loop over a two-snapshot buffer (uint32[2*ss], updated in
    place).  ops: uint32[n, 2] (state bit, value); op_start: CSR over the
    schedule edges.  pins: 11 ints as jq_cosim_edges takes them.  With
    timing, each edge's setup/hold events (cycle 0) are appended to the
    list `events`.  Returns the trace uint32[num_edges, 2*ss] (both
    snapshots after each edge)."""
    import oracle
    ss = states.size // 2
    num_sched = len(op_start) - 1
    trace = np.zeros((num_edges, 2 * ss), np.uint32)
    use_flash = flash is not None and pins is not None and pins[10] != 0
    for e in range(num_edges):
        k = (sched_offset + e) % num_sched
        inp = states[:ss]
        inp[:] = states[ss:]
        for pos, val in ops[op_start[k]:op_start[k + 1]]:
            _put(inp, int(pos), int(val))
        if use_flash:
            for i in range(4):
                _put(inp, int(pins[i]), (flash.d_i >> i) & 1)
        if constraints is not None or arrival_state_offset:
            ev, cnt = oracle.run(script, states, 1, constraints=constraints,
                                 arrival_state_offset=arrival_state_offset,
                                 event_capacity=1 << 16)
            if events is not None:
                assert cnt <= 1 << 16
                events.append(ev)
        else:
            oracle.run(script, states, 1)
        if use_flash:
            out = states[ss:]
            d_out = sum(_bit(out, int(pins[6 + i])) << i for i in range(4))
            flash.cosim_step(_bit(out, int(pins[4])), _bit(out, int(pins[5])), d_out)
        out = states[ss:]
        for u in uarts:
            u.step(_bit(out, u.tx_pos))
        if bus is not None:
            bus.step(out)
        trace[e] = states
    return trace


class RefFlash:
    """The reference's SPI flash model (csrc/spiflash_model.cc) compiled into
    oracle/_ref; None where the library is absent."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            path = os.path.join(HERE, "_ref", "libspiflash_ref.so")
            if not os.path.exists(path):
                return None
            L = ctypes.CDLL(path)
            L.spiflash_new.restype = ctypes.c_void_p
            L.spiflash_new.argtypes = [ctypes.c_size_t]
            L.spiflash_free.argtypes = [ctypes.c_void_p]
            L.spiflash_load.restype = ctypes.c_int
            L.spiflash_load.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t,
                                        ctypes.c_size_t]
            L.spiflash_step.restype = ctypes.c_uint8
            L.spiflash_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_uint8]
            L.spiflash_get_command.restype = ctypes.c_uint8
            L.spiflash_get_command.argtypes = [ctypes.c_void_p]
            L.spiflash_get_byte_count.restype = ctypes.c_uint32
            L.spiflash_get_byte_count.argtypes = [ctypes.c_void_p]
            cls._lib = L
        return cls._lib

    def __init__(self, data):
        L = self.lib()
        self.h = L.spiflash_new(len(data))
        L.spiflash_load(self.h, bytes(data), len(data), 0)

    def step(self, clk, csn, d_o):
        return self.lib().spiflash_step(self.h, int(clk), int(csn), int(d_o))

    def command(self):
        return self.lib().spiflash_get_command(self.h)

    def byte_count(self):
        return self.lib().spiflash_get_byte_count(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib().spiflash_free(self.h)
            self.h = None
