"""Reactive co-simulation on the device: clock edges with host-free state
preparation and an SPI flash model between them (the GPU edge pipeline of
the reference's Metal co-simulation, src/sim/cosim_metal.rs, on the HIP scan
kernel through jq_cosim_edges).

  * `MultiClockScheduler` restates src/sim/cosim_metal.rs:1433-1508: the
    edge schedule of several clock domains over one least common period at
    the greatest common half-period resolution;
  * `MultiClockScheduler.edge_ops` restates :1510-1611: the bit operations
    of one schedule edge (reset, clock inputs, edge flags, constant inputs,
    placeholders for model-driven inputs);
  * `Cosim` owns the device buffers (two snapshots, the edge operations,
    the flash state and image) and runs edges in batches that do not cross
    the end of reset, updating the reset operation and the flash's reset
    flag between batches (:3470-3497).
"""
import ctypes
from dataclasses import dataclass, field
from math import gcd

import numpy as np

from . import _native as N

FLASH_STATE_BYTES = 48
# UART and bus-trace layouts (csrc/kernel_v1.metal:1027-1101)
UART_MAX = 4
UART_CHANNEL_CAP = 4096
UART_PARAMS_WORDS = 12
UART_CHANNEL_WORDS = 4 + UART_CHANNEL_CAP // 4
WB_ADR_BITS = 30
WB_DAT_BITS = 32
BUS_SIGNALS = ("ibus_cyc", "ibus_stb", "ibus_adr", "ibus_rdata", "dbus_cyc", "dbus_stb",
               "dbus_we", "dbus_adr", "spiflash_ack", "sram_ack", "csr_ack")


def bus_trace_params(sig):
    """WbTraceParams words from {signal: output state bit (lists for the
    address and data buses)}; missing signals are not connected."""
    words = []
    for name in BUS_SIGNALS:
        width = {"ibus_adr": WB_ADR_BITS, "dbus_adr": WB_ADR_BITS,
                 "ibus_rdata": WB_DAT_BITS}.get(name)
        if width is None:
            words.append(sig.get(name, NC))
        else:
            bits = list(sig.get(name, ()))[:width]
            words += bits + [NC] * (width - len(bits))
    words.append(1)  # has_trace
    return np.asarray(words, np.uint32)
_IN_RESET_OFFSET = 36  # FlashState.in_reset (csrc/kernel_v1.metal:731-749)
NC = 0xFFFFFFFF


def _lcm(a, b):
    return a // gcd(a, b) * b


@dataclass
class ClockTiming:
    half_period_ps: int
    phase_offset_ps: int = 0
    domain_index: int = 0


@dataclass
class ClockDomain:
    """State bits of one clock domain: its clock input (None: not an input)
    and its posedge / negedge flag bits."""
    clock_input_pos: int = None
    posedge_flag_bits: list = field(default_factory=list)
    negedge_flag_bits: list = field(default_factory=list)


class MultiClockScheduler:
    """Schedule of (falling, rising) per domain for every tick of one common
    period (src/sim/cosim_metal.rs:1433-1508)."""

    MAX_LEN = 1_000_000

    def __init__(self, timings):
        if not timings:
            raise ValueError("at least one clock domain is required")
        g = timings[0].half_period_ps
        for t in timings[1:]:
            g = gcd(g, t.half_period_ps)
        for t in timings:
            if t.phase_offset_ps > 0:
                g = gcd(g, t.phase_offset_ps)
        period = timings[0].half_period_ps * 2
        for t in timings[1:]:
            period = _lcm(period, t.half_period_ps * 2)
        n = period // g
        if n > self.MAX_LEN:
            raise ValueError(f"multi-clock schedule too large ({n} ticks)")
        ndom = max(t.domain_index + 1 for t in timings)
        self.gcd_ps = g
        self.schedule = []
        for tick in range(n):
            tps = tick * g
            edges = [[False, False] for _ in range(ndom)]
            for t in timings:
                if tps < t.phase_offset_ps:
                    continue
                adj = tps - t.phase_offset_ps
                if adj % t.half_period_ps:
                    continue
                if (adj // t.half_period_ps) % 2 == 0:
                    edges[t.domain_index][0] = True   # falling
                else:
                    edges[t.domain_index][1] = True   # rising
            self.schedule.append([tuple(e) for e in edges])

    def __len__(self):
        return len(self.schedule)

    def edge_ops(self, tick, domains, reset_pos=None, reset_val=0, constants=(),
                 model_driven=()):
        """[(state bit, value)] of schedule tick `tick`, in the reference's
        order: reset, per domain the clock input (falling: 0, rising: 1) and
        its flags (posedge = rising, negedge = falling), constant inputs,
        model-driven placeholders at 0."""
        edges = self.schedule[tick % len(self.schedule)]
        ops = []
        if reset_pos is not None:
            ops.append((reset_pos, int(reset_val)))
        for i, d in enumerate(domains):
            fall, rise = edges[i] if i < len(edges) else (False, False)
            if d.clock_input_pos is not None:
                if fall:
                    ops.append((d.clock_input_pos, 0))
                if rise:
                    ops.append((d.clock_input_pos, 1))
            ops += [(p, int(rise)) for p in d.posedge_flag_bits]
            ops += [(p, int(fall)) for p in d.negedge_flag_bits]
        ops += [(p, int(v)) for p, v in constants]
        ops += [(p, 0) for p in model_driven]
        return ops


def clock_domains_of(script):
    """One domain per clock input port of the netlist, with its posedge flag
    (the netlist's DFFs are positive-edge)."""
    ins, _ = script.netlist.ports()
    out = []
    for name, pin, flag in ins:
        if flag:
            out.append((name, ClockDomain(script.input_map[pin],
                                          [script.input_map[flag]], [])))
    return out


@dataclass
class FlashPins:
    """State bits of the flash's signals (NC: not connected)."""
    clk_out: int
    csn_out: int
    d_out: tuple = (NC, NC, NC, NC)
    d_in: tuple = (NC, NC, NC, NC)

    def words(self):
        """The 11 pin words of jq_cosim_edges (up to 4 data bits each way,
        missing ones not connected)."""
        d_in, d_out = list(self.d_in)[:4], list(self.d_out)[:4]
        d_in += [NC] * (4 - len(d_in))
        d_out += [NC] * (4 - len(d_out))
        return np.asarray(d_in + [self.clk_out, self.csn_out] + d_out + [1], np.uint32)


class Cosim:
    """Edges of one script on the device: `run(num_edges)` advances the
    simulation; `trace` collects both snapshots after every edge."""

    def __init__(self, script, scheduler, domains, *, reset_pos=None, reset_active=1,
                 reset_edges=0, constants=(), model_driven=(), flash_pins=None,
                 flash_image=b"", uarts=(), bus_trace=None, constraints=None,
                 event_capacity=1024, stream=None):
        import torch
        self.torch = torch
        self.script = script
        if script.xprop_enabled:
            raise ValueError("co-simulation runs the 2-state script (arrivals allowed)")
        self.ss = int(script.effective_state_size())
        script.validate(self.ss)
        if constraints is not None and len(constraints) != 1 + script.reg_io_state_size:
            raise ValueError("constraints must be [clock_ps] + one word per state word "
                             f"({1 + script.reg_io_state_size} words)")
        self.sched = scheduler
        self.domains = list(domains)
        self.reset_pos = reset_pos
        self.reset_active = int(reset_active)
        self.reset_edges = int(reset_edges)
        self.constants = list(constants)
        self.model_driven = list(model_driven)
        self.edge = 0
        self.stream = stream or torch.cuda.current_stream()
        dev = "cuda"
        self.d_bs = torch.from_numpy(np.ascontiguousarray(script.blocks_start, np.uint64)
                                     .view(np.int64)).to(dev)
        self.d_bd = torch.from_numpy(np.ascontiguousarray(script.blocks_data, np.uint32)
                                     .view(np.int32)).to(dev)
        self.d_sram = torch.zeros(max(script.sram_storage_size, 1), dtype=torch.int32,
                                  device=dev)
        self.states = torch.zeros(2 * self.ss, dtype=torch.int32, device=dev)
        self.flash = flash_pins is not None
        if self.flash:
            if int(N.hip().jq_cosim_flash_state_bytes()) != FLASH_STATE_BYTES:
                raise RuntimeError("FlashState size mismatch")
            fs = np.zeros(FLASH_STATE_BYTES // 4, np.uint32)
            b = fs.view(np.uint8)
            b[8:12] = np.frombuffer(np.uint32(1).tobytes(), np.uint8)   # data_width
            b[24:28] = np.frombuffer(np.uint32(1).tobytes(), np.uint8)  # prev_csn
            b[28] = 0x0F                                                 # d_i
            b[36:40] = np.frombuffer(np.uint32(1).tobytes(), np.uint8)  # in_reset
            b[44:48] = np.frombuffer(np.uint32(1).tobytes(), np.uint8)  # model_prev_csn
            self.d_flash = torch.from_numpy(fs.view(np.int32)).to(dev)
            self.d_pins = torch.from_numpy(flash_pins.words().view(np.int32)).to(dev)
            img = np.frombuffer(bytes(flash_image), np.uint8) if flash_image else np.zeros(1, np.uint8)
            self.d_image = torch.from_numpy(img.copy()).to(dev)
            self.image_size = len(flash_image)
        self._ops_reset_val = None
        # timing: arrivals after the values, setup/hold checks against
        # `constraints` ([clock_ps, per-word]) into an event buffer
        self.arrival_offset = int(script.arrival_state_offset) \
            if script.timing_arrivals_enabled else 0
        self.timed = constraints is not None
        if self.timed:
            self.d_cons = torch.from_numpy(np.ascontiguousarray(constraints, np.uint32)
                                           .view(np.int32)).to(dev)
            self.event_capacity = int(event_capacity)
            self.d_events = torch.zeros(4 + 8 * self.event_capacity, dtype=torch.int32,
                                        device=dev)
        # UART decoders: [(tx output bit, edges per bit)], at most 4
        self.uarts = list(uarts)
        if len(self.uarts) > UART_MAX:
            raise ValueError(f"at most {UART_MAX} UARTs")
        if self.uarts:
            prm = np.zeros(UART_PARAMS_WORDS, np.uint32)
            prm[0] = self.ss
            prm[1] = len(self.uarts)
            for i, (pos, cpb) in enumerate(self.uarts):
                prm[4 + 2 * i] = pos
                prm[5 + 2 * i] = cpb
            us = np.zeros((UART_MAX, 6), np.uint32)
            us[:, 1] = 1  # last_tx: the line idles high
            ch = np.zeros((UART_MAX, UART_CHANNEL_WORDS), np.uint32)
            ch[:, 1] = UART_CHANNEL_CAP
            self.d_uart_params = torch.from_numpy(prm.view(np.int32)).to(dev)
            self.d_uart_state = torch.from_numpy(us.view(np.int32)).to(dev)
            self.d_uart_ch = torch.from_numpy(ch.view(np.int32)).to(dev)
        # Wishbone bus trace: {signal: state bit or list of bits} and capacity
        self.bus = bus_trace is not None
        if self.bus:
            self.d_bus_params = torch.from_numpy(bus_trace_params(bus_trace).view(np.int32)).to(dev)
            cap = int(bus_trace.get("capacity", 16384))
            hdr = np.zeros(4 + 5 * cap, np.uint32)
            hdr[1] = cap
            self.d_bus = torch.from_numpy(hdr.view(np.int32)).to(dev)
        # every buffer above was filled on the current stream; the launches
        # go to self.stream, which must see those fills first
        self.stream.wait_stream(torch.cuda.current_stream())

    def _upload_ops(self, reset_val):
        if reset_val == self._ops_reset_val:
            return
        ops, start = [], [0]
        for k in range(len(self.sched)):
            ops += self.sched.edge_ops(k, self.domains, self.reset_pos, reset_val,
                                       self.constants, self.model_driven)
            start.append(len(ops))
        arr = np.asarray(ops if ops else [(NC, 0)], np.uint32).reshape(-1, 2)
        self.ops = arr
        self.op_start = np.asarray(start, np.uint32)
        self.d_ops = self.torch.from_numpy(arr.reshape(-1).view(np.int32)).to("cuda")
        self.d_start = self.torch.from_numpy(self.op_start.view(np.int32)).to("cuda")
        self._ops_reset_val = reset_val

    def run(self, num_edges, batch=1024, trace=False):
        """Run num_edges edges; returns the trace (uint32[num_edges, 2*ss])
        when asked."""
        torch = self.torch
        out = []
        done = 0
        while done < num_edges:
            n = min(batch, num_edges - done)
            in_reset = self.edge < self.reset_edges
            if in_reset and self.edge + n > self.reset_edges:
                n = self.reset_edges - self.edge
            reset_val = self.reset_active if in_reset else 1 - self.reset_active
            # buffer updates are ordered with the launches on the same stream
            with torch.cuda.stream(self.stream):
                self._upload_ops(reset_val)
                if self.flash:
                    self.d_flash.view(torch.uint8)[_IN_RESET_OFFSET] = 1 if in_reset else 0
                tr = (torch.zeros((n, 2 * self.ss), dtype=torch.int32, device="cuda")
                      if trace else None)
            N.check(N.hip().jq_cosim_edges(
                self.script.num_blocks, self.script.num_major_stages, self.d_bs.data_ptr(),
                self.d_bd.data_ptr(), self.d_sram.data_ptr(), self.ss, self.states.data_ptr(),
                self.d_ops.data_ptr(), self.d_start.data_ptr(), len(self.sched),
                self.edge % len(self.sched), n,
                self.d_flash.data_ptr() if self.flash else None,
                self.d_pins.data_ptr() if self.flash else None,
                self.d_image.data_ptr() if self.flash else None,
                self.image_size if self.flash else 0,
                self.d_uart_state.data_ptr() if self.uarts else None,
                self.d_uart_params.data_ptr() if self.uarts else None,
                self.d_uart_ch.data_ptr() if self.uarts else None,
                self.d_bus.data_ptr() if self.bus else None,
                self.d_bus_params.data_ptr() if self.bus else None,
                self.d_cons.data_ptr() if self.timed else None,
                self.d_events.data_ptr() if self.timed else None,
                self.event_capacity if self.timed else 0, self.arrival_offset,
                tr.data_ptr() if tr is not None else None,
                ctypes.c_void_p(self.stream.cuda_stream)))
            if tr is not None:
                out.append(tr)
            self.edge += n
            done += n
        self.stream.synchronize()
        from .sim import device_error
        if device_error():
            raise RuntimeError("device barrier timeout during co-simulation")
        if trace:
            return torch.cat(out).cpu().numpy().view(np.uint32) if out else \
                np.zeros((0, 2 * self.ss), np.uint32)
        return None

    def flash_state(self):
        """The device FlashState as 48 bytes."""
        return self.d_flash.cpu().numpy().view(np.uint8).copy() if self.flash else None

    def snapshots(self):
        return self.states.cpu().numpy().view(np.uint32).copy()

    def events(self):
        """(event records kept, total count, overflow flag) of every run so
        far; the flag is set once more events came than the capacity holds
        (the event buffer header's second word, csrc/event_buffer.h:27-33)."""
        if not self.timed:
            return None, 0, 0
        from .sim import EVENT_DTYPE
        self.stream.synchronize()
        raw = self.d_events.cpu().numpy().view(np.uint32)
        total = int(raw[0])
        kept = min(total, self.event_capacity)
        return raw[4:4 + 8 * kept].copy().view(EVENT_DTYPE), total, int(raw[1])

    def uart_bytes(self, i):
        """Bytes decoded by UART i so far (the last UART_CHANNEL_CAP when
        the channel wrapped)."""
        ch = self.d_uart_ch[i].cpu().numpy().view(np.uint32)
        head = int(ch[0])
        data = ch[4:].view(np.uint8)
        if head <= UART_CHANNEL_CAP:
            return bytes(data[:head])
        k = head % UART_CHANNEL_CAP
        return bytes(np.concatenate([data[k:], data[:k]]))

    def uart_states(self):
        return self.d_uart_state.cpu().numpy().view(np.uint32).copy()

    def bus_entries(self):
        """(header [write_head, capacity, current_tick, prev_flags],
        entries uint32[n, 5] = tick, flags, ibus_adr, ibus_rdata, dbus_adr)."""
        a = self.d_bus.cpu().numpy().view(np.uint32)
        n = min(int(a[0]), int(a[1]))
        return a[:4].copy(), a[4:4 + 5 * n].reshape(n, 5).copy()


# ---- synthetic co-simulation workload --------------------------------------
# A small SPI master in AIGPDK gate-level Verilog with a known answer: a
# 7-bit saturating counter drives a single-bit SPI read of two bytes at a
# fixed address (command 0x03, 24-bit address, MSB first), and a 16-bit
# shift register captures the flash's MISO bits.  Everything is a function of
# the counter, generated as multiplexer trees from truth tables.  Ports:
# clk, rst (active high), miso; flash_clk, flash_csn, flash_d0, rx_o0..15.
#
# Timeline in counter values c (one per system clock cycle):
#   c in [2, 98)      chip select low
#   c odd in window   SPI clock high
#   c in [2, 66)      MOSI = bit 31 - (c - 2) / 2 of (0x03 << 24 | address)
#   c odd, 67..97     the MISO input shifts into rx (16 samples)
# After the transaction rx (rx15..rx0) holds image[address] << 8 |
# image[address + 1].


CS_LO, CS_HI = 2, 98
CMD_END = 66
RX_FIRST, RX_LAST = 67, 97


class _Gates:
    def __init__(self):
        self.lines = []
        self.wires = []
        self.n = 0

    def _new(self):
        self.n += 1
        name = f"n{self.n}"
        self.wires.append(name)
        return name

    def and2(self, a, b, ia=0, ib=0, iy=0):
        y = self._new()
        self.lines.append(f"  AND2_{ia}{ib}_{iy} g{self.n} (.A({a}), .B({b}), .Y({y}));")
        return y

    def inv(self, a):
        y = self._new()
        self.lines.append(f"  INV g{self.n} (.A({a}), .Y({y}));")
        return y

    def mux(self, s, f1, f0):
        """s ? f1 : f0 with constants (Python bools) folded."""
        if isinstance(f1, bool) and isinstance(f0, bool):
            if f1 == f0:
                return f1
            return s if f1 else self.inv(s)
        if f1 is True:
            return self.and2(s, f0, 1, 1, 1)          # s | f0
        if f1 is False:
            return self.and2(s, f0, 1, 0, 0)          # ~s & f0
        if f0 is True:
            return self.and2(s, f1, 0, 1, 1)          # ~s | f1
        if f0 is False:
            return self.and2(s, f1)                   # s & f1
        if f1 == f0:
            return f0
        a = self.and2(s, f1)
        b = self.and2(s, f0, 1, 0, 0)
        return self.and2(a, b, 1, 1, 1)               # a | b

    def func(self, table, bits, memo):
        """Truth table over bits (bits[0] least significant) as gates."""
        key = (tuple(table), len(bits))
        if key in memo:
            return memo[key]
        if all(table) or not any(table):
            r = bool(table[0])
        else:
            half = len(table) // 2
            f0 = self.func(table[:half], bits[:-1], memo)
            f1 = self.func(table[half:], bits[:-1], memo)
            r = self.mux(bits[-1], f1, f0)
        memo[key] = r
        return r


def spi_master_verilog(address):
    word = (0x03 << 24) | (address & 0xFFFFFF)
    g = _Gates()
    cbits = [f"c{i}" for i in range(7)]
    rx = [f"rx{i}" for i in range(16)]
    memo = {}

    def f(pred):
        return g.func([bool(pred(c)) for c in range(128)], cbits, memo)

    csn = f(lambda c: not (CS_LO <= c < CS_HI))
    sclk = f(lambda c: CS_LO <= c < CS_HI and c & 1)
    mosi = f(lambda c: CS_LO <= c < CMD_END and (word >> (31 - (c - CS_LO) // 2)) & 1)
    shift = f(lambda c: RX_FIRST <= c <= RX_LAST and c & 1)
    dnext = []
    for i in range(7):
        nxt = f(lambda c, i=i: ((min(c + 1, 127)) >> i) & 1)
        dnext.append(nxt if isinstance(nxt, str) else None)
        if dnext[-1] is None:
            raise AssertionError("constant counter bit")
        dnext[-1] = g.and2(nxt, "rst", 0, 1, 0)  # held at 0 in reset
    rxn = [g.mux(shift, "miso" if i == 0 else rx[i - 1], rx[i]) for i in range(16)]

    def net(x):
        if isinstance(x, bool):
            return "1'b1" if x else "1'b0"
        return x

    outs = ["flash_clk", "flash_csn", "flash_d0"] + [f"rx_o{i}" for i in range(16)]
    ports = ["clk", "rst", "miso"] + outs
    L = [f"module spi_master({', '.join(ports)});"]
    L += ["  input clk;", "  input rst;", "  input miso;"]
    L += [f"  output {o};" for o in outs]
    wires = cbits + rx + [f"d{i}" for i in range(7)] + [f"rxd{i}" for i in range(16)] + g.wires
    L += [f"  wire {w};" for w in wires]
    L += g.lines
    for i in range(7):
        L.append(f"  assign d{i} = {dnext[i]};")
        L.append(f"  DFF q{i} (.CLK(clk), .D(d{i}), .Q(c{i}));")
    for i in range(16):
        L.append(f"  assign rxd{i} = {net(rxn[i])};")
        L.append(f"  DFF r{i} (.CLK(clk), .D(rxd{i}), .Q(rx{i}));")
    L.append(f"  assign flash_clk = {net(sclk)};")
    L.append(f"  assign flash_csn = {net(csn)};")
    L.append(f"  assign flash_d0 = {net(mosi)};")
    for i in range(16):
        L.append(f"  assign rx_o{i} = rx{i};")
    L.append("endmodule")
    return "\n".join(L) + "\n"


def uart_tx_verilog(data, first=4, cycles_per_bit=2, counter_bits=8):
    """A gate-level UART transmitter with a known answer: a saturating
    counter c (one per clock cycle) drives `tx` idle high, then for byte j
    of `data` a start bit, 8 data bits LSB first and a stop bit, each
    `cycles_per_bit` cycles long, from cycle `first` on.  Bus-shaped outputs
    for a trace: cyc (c % 8 in 1..3), stb (c % 8 in 2..3), we (bit 3 of c),
    ack (c % 8 == 3) and adr0..7 = c.  Ports: clk, rst; tx, cyc, stb, we,
    ack, adr0..7."""
    n = 1 << counter_bits
    frame = 10 * cycles_per_bit
    if first + frame * len(data) > n - 1:
        raise ValueError("too many bytes for the counter")

    def tx(c):
        k = c - first
        if k < 0 or k >= frame * len(data):
            return 1
        j, bit = divmod(k // cycles_per_bit, 10)
        if bit == 0:
            return 0
        if bit == 9:
            return 1
        return (data[j] >> (bit - 1)) & 1

    g = _Gates()
    cbits = [f"c{i}" for i in range(counter_bits)]
    memo = {}

    def f(pred):
        return g.func([bool(pred(c)) for c in range(n)], cbits, memo)

    def net(x):
        if isinstance(x, bool):
            return "1'b1" if x else "1'b0"
        return x

    outs = {"tx": f(tx), "cyc": f(lambda c: c % 8 in (1, 2, 3)),
            "stb": f(lambda c: c % 8 in (2, 3)), "we": f(lambda c: (c >> 3) & 1),
            "ack": f(lambda c: c % 8 == 3)}
    for i in range(counter_bits):
        outs[f"adr{i}"] = f"c{i}"
    dnext = [g.and2(f(lambda c, i=i: (min(c + 1, n - 1) >> i) & 1), "rst", 0, 1, 0)
             for i in range(counter_bits)]
    L = [f"module uart_tx(clk, rst, {', '.join(outs)});", "  input clk;", "  input rst;"]
    L += [f"  output {o};" for o in outs]
    L += [f"  wire {w};" for w in cbits + [f"d{i}" for i in range(counter_bits)] + g.wires]
    L += g.lines
    for i in range(counter_bits):
        L.append(f"  assign d{i} = {dnext[i]};")
        L.append(f"  DFF q{i} (.CLK(clk), .D(d{i}), .Q(c{i}));")
    for o, x in outs.items():
        L.append(f"  assign {o} = {net(x)};")
    L.append("endmodule")
    return "\n".join(L) + "\n"


# ---- pin mapping --------------------------------------------------------------
# Resolution of the co-simulation's GPIO indices and named ports to state bits
# (build_gpio_mapping, src/sim/cosim_metal.rs:955-1212).

def parse_gpio_index(pin_name, prefix):
    """N from "prefix[N]" or "prefix:N" anywhere in the name, else None
    (src/sim/cosim_metal.rs:1214-1233)."""
    k = pin_name.find(prefix + "[")
    if k >= 0:
        rest = pin_name[k + len(prefix) + 1:]
        end = rest.find("]")
        if end >= 0:
            try:
                return int(rest[:end])
            except ValueError:
                return None
    k = pin_name.find(prefix + ":")
    if k >= 0:
        rest = pin_name[k + len(prefix) + 1:]
        digits = ""
        for ch in rest:
            if not ch.isdigit():
                break
            digits += ch
        return int(digits) if digits else None
    return None


@dataclass
class GpioMapping:
    input_bits: dict            # gpio index -> input state bit
    output_bits: dict           # gpio index -> output state bit
    clock_domains: list         # [(clock port name, gpio index or None, ClockDomain)]
    named_input_bits: dict      # input port bit name -> state bit


def gpio_mapping(script, port_mapping=None):
    """GPIO indices of the netlist's ports: `port_mapping` = {"inputs":
    {"N": port bit name}, "outputs": {...}} first, else names of the form
    gpio_in[N] / gpio_out[N] (or gpio_in:N).  Outputs driven by a constant
    are skipped; an output whose literal is not in the output map falls back
    to its complement's position, as the reference does.  One clock domain
    per clock port, in port order, with its posedge flag."""
    pm = port_mapping or {}
    in_name = {v: int(k) for k, v in pm.get("inputs", {}).items() if str(k).isdigit()}
    out_name = {v: int(k) for k, v in pm.get("outputs", {}).items() if str(k).isdigit()}
    ins, outs = script.netlist.ports()
    input_bits, named, domains = {}, {}, []
    for name, pin, flag in ins:
        pos = script.input_map.get(pin)
        g = in_name.get(name, parse_gpio_index(name, "gpio_in"))
        if pos is not None:
            named[name] = pos
            if g is not None:
                input_bits[g] = pos
        if flag and flag in script.input_map:
            domains.append((name, g, ClockDomain(pos, [script.input_map[flag]], [])))
    output_bits = {}
    for name, lit in outs:
        g = out_name.get(name, parse_gpio_index(name, "gpio_out"))
        if g is None or lit <= 1:
            continue
        pos = script.output_map.get(lit, script.output_map.get(lit ^ 1))
        if pos is not None:
            output_bits[g] = pos
    return GpioMapping(input_bits, output_bits, domains, named)


# ---- stimulus VCD ---------------------------------------------------------------

def stimulus_vcd(script, trace, gcd_ps, first_edge=0, scope="top"):
    """The co-simulation's stimulus VCD (setup_stimulus_vcd,
    src/sim/vcd_io.rs:964-1017, and the per-edge writer of
    src/sim/cosim_metal.rs:3587-3619): every input port bit of the netlist
    with its value in each edge's input snapshot, at time edge * gcd_ps
    (timescale 1 ps), only changed values after the first edge.  `trace`
    is the uint32[n, 2*ss] trace of `Cosim.run(..., trace=True)`; edges are
    numbered from `first_edge`.  Returns the VCD text."""
    from .vcd import id_code
    ins, _ = script.netlist.ports()
    sig = []
    for name, pin, _ in ins:
        pos = script.input_map.get(pin)
        if pos is None:
            continue
        base, index = name, None
        if name.endswith("]") and "[" in name:
            base, idx = name[:-1].split("[", 1)
            if idx.isdigit():
                index = int(idx)
            else:
                base = name
        sig.append((pos, base, index))
    lines = ["$timescale 1 ps $end", f"$scope module {scope} $end"]
    for i, (_, base, index) in enumerate(sig):
        ref = base if index is None else f"{base} [{index}]"
        lines.append(f"$var wire 1 {id_code(i)} {ref} $end")
    lines += ["$upscope $end", "$enddefinitions $end", "$dumpvars"]
    lines += [f"0{id_code(i)}" for i in range(len(sig))]
    lines.append("$end")
    trace = np.asarray(trace, np.uint32)
    prev = [None] * len(sig)
    pos = np.asarray([p for p, _, _ in sig], np.int64)
    for k in range(trace.shape[0]):
        lines.append(f"#{(first_edge + k) * gcd_ps}")
        vals = (trace[k, pos >> 5] >> (pos & 31).astype(np.uint32)) & 1 if len(sig) else []
        for i, v in enumerate(vals):
            v = int(v)
            if v != prev[i]:
                lines.append(f"{v}{id_code(i)}")
                prev[i] = v
    return "\n".join(lines) + "\n"
