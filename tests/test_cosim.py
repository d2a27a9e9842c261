"""Co-simulation host logic and its CPU oracle (no GPU).

* The flash restatement (oracle/cosim_oracle.py) against the reference's own
  SPI flash model compiled from csrc/spiflash_model.cc into oracle/_ref,
  step by step over SPI transactions of every command class.
* The multi-clock scheduler (src/sim/cosim_metal.rs:1433-1508) against
  closed forms: alternating edges, edge counts per common period, phase
  offsets.
* The edge loop end to end: a gate-level SPI master reads two bytes of a
  random flash image through the loop; the bytes it captures must be the
  image's (a data-dependent check of the loop order, the setup delay and
  the data-in path).
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, HERE)

import cosim_oracle as co  # noqa: E402

from jacquard_amd import cosim  # noqa: E402


def _spi_sequence(rng, n_tx):
    """(clk, csn, d_o) steps of n_tx transactions: a command byte (one of
    every class), address and data bits, clock glitches while deselected,
    random lengths."""
    steps = []
    for _ in range(n_tx):
        cmd = int(rng.choice([0x03, 0xeb, 0x9f, 0xab, 0x05, 0x77, int(rng.integers(0, 256))]))
        steps.append((0, 1, int(rng.integers(0, 16))))
        steps.append((0, 0, 0))
        nbits = 8 + int(rng.integers(0, 80))
        byte = cmd
        for b in range(nbits):
            if b < 8:
                bit = (byte >> (7 - b)) & 1
                d = bit
            else:
                d = int(rng.integers(0, 16))
            steps.append((0, 0, d))
            steps.append((1, 0, d))
            if rng.random() < 0.05:  # a repeated level (no edge)
                steps.append((1, 0, d))
        steps.append((0, 0, 0))
        steps.append((0, 1, 0))
        if rng.random() < 0.3:  # clocks while deselected
            steps += [(1, 1, 0), (0, 1, 0)]
    return steps


@pytest.mark.parametrize("seed", range(6))
def test_flash_restatement_matches_reference_model(seed):
    if co.RefFlash.lib() is None:
        pytest.skip("oracle/_ref/libspiflash_ref.so not built (no reference tree here)")
    rng = np.random.default_rng(seed)
    size = int(rng.choice([64, 4096]))
    img = bytes(rng.integers(0, 256, size, dtype=np.uint8))
    ref = co.RefFlash(img)
    mine = co.Flash(img, cosim_init=False)
    seq = _spi_sequence(rng, 12)
    for i, (clk, csn, d) in enumerate(seq):
        got = mine.step_model(clk, csn, d)
        exp = ref.step(clk, csn, d)
        assert got == exp, f"step {i}: d_i {got:#x} != {exp:#x}"
        assert mine.command == ref.command() and mine.byte_count == ref.byte_count(), i


def test_flash_read_commands_return_image_bytes():
    # closed forms: a single read returns image[a], image[a+1], ...; past
    # the end 0xFF; read-ID returns CA 7C A7 FF
    img = bytes(range(256)) * 2
    f = co.Flash(img, cosim_init=False)

    def clock_byte(v, width=1):
        # the controller samples MISO before each rising clock: d_i changed
        # on the falling clock before it
        out = 0
        for k in range(8 // width):
            if width == 1:
                d = (v >> (7 - k)) & 1
                out = (out << 1) | ((f.d_i >> 1) & 1)
            else:
                d = (v >> (4 - 4 * k)) & 0xF
                out = (out << 4) | (f.d_i & 0xF)
            f.step_model(1, 0, d)
            f.step_model(0, 0, d)
        return out

    def read(cmd, addr, n, width=1, dummies=0):
        f.step_model(0, 1, 0)
        f.step_model(0, 0, 0)
        clock_byte(cmd)
        if addr is not None:
            for sh in (16, 8, 0):
                clock_byte((addr >> sh) & 0xFF, width)
        for _ in range(dummies):
            clock_byte(0, width)
        got = [clock_byte(0, width) for _ in range(n)]
        f.step_model(0, 1, 0)
        return got

    assert read(0x03, 0x10, 3) == [img[0x10], img[0x11], img[0x12]]
    assert read(0x03, 0x1FF, 2) == [img[0x1FF], 0xFF]
    # quad read: address in 4-bit mode, a mode byte and two dummy bytes
    assert read(0xeb, 0x20, 3, width=4, dummies=3) == [img[0x20], img[0x21], img[0x22]]
    assert read(0x9f, None, 5) == [0xCA, 0x7C, 0xA7, 0xFF, 0xCA]
    assert f.last_error_cmd == 0
    read(0x42, None, 1)
    assert f.last_error_cmd == 0x42


def test_scheduler_single_clock():
    s = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    assert s.gcd_ps == 5000
    assert s.schedule == [[(True, False)], [(False, True)]]


@pytest.mark.parametrize("hps,offsets", [((10, 15), (0, 0)), ((6, 4), (0, 2)),
                                         ((7, 7, 14), (0, 7, 0))])
def test_scheduler_multi_clock_closed_forms(hps, offsets):
    timings = [cosim.ClockTiming(h, o, i) for i, (h, o) in enumerate(zip(hps, offsets))]
    s = cosim.MultiClockScheduler(timings)
    g = 0
    for h in hps:
        g = np.gcd(g, h)
    for o in offsets:
        if o:
            g = np.gcd(g, o)
    period = int(np.lcm.reduce([2 * h for h in hps]))
    assert s.gcd_ps == g and len(s) == period // g
    for i, (h, o) in enumerate(zip(hps, offsets)):
        ticks = [k for k in range(len(s)) if any(s.schedule[k][i])]
        # edges every h ps from the offset, alternating falling / rising
        assert ticks == [k for k in range(len(s)) if k * g >= o and (k * g - o) % h == 0]
        kinds = ["f" if s.schedule[k][i][0] else "r" for k in ticks]
        assert kinds == ["f" if j % 2 == 0 else "r" for j in range(len(ticks))]
        assert not any(a and b for a, b in (s.schedule[k][i] for k in range(len(s))))


def test_scheduler_rejects_incommensurable_periods():
    with pytest.raises(ValueError):
        cosim.MultiClockScheduler([cosim.ClockTiming(1_000_003), cosim.ClockTiming(999_983)])
    with pytest.raises(ValueError):
        cosim.MultiClockScheduler([])


def test_edge_ops_order_and_values():
    s = cosim.MultiClockScheduler([cosim.ClockTiming(10, 0, 0), cosim.ClockTiming(5, 0, 1)])
    doms = [cosim.ClockDomain(3, [40], [41]), cosim.ClockDomain(None, [50], [])]
    # tick 0: both fall; tick 1: domain 1 rises
    assert s.edge_ops(0, doms, reset_pos=7, reset_val=1, constants=[(9, 1)],
                      model_driven=[11]) == [(7, 1), (3, 0), (40, 0), (41, 1), (50, 0),
                                             (9, 1), (11, 0)]
    assert s.edge_ops(1, doms) == [(40, 0), (41, 0), (50, 1)]
    assert s.edge_ops(2, doms) == [(3, 1), (40, 1), (41, 0), (50, 0)]


def spi_setup(address, image, reset_edges=8, blocks=1):
    """(script, flash pins, clock domains, input bits, output bits) of the
    SPI master, flattened for `blocks` workgroups."""
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.from_verilog(text=cosim.spi_master_verilog(address))
    s = Compiled(nl).flatten(blocks)
    ins, outs = nl.ports()
    inbit = {n: s.input_map[p] for n, p, _ in ins}
    outbit = {n: s.output_map[lit] for n, lit in outs if lit > 1}
    pins = cosim.FlashPins(clk_out=outbit["flash_clk"], csn_out=outbit["flash_csn"],
                           d_out=(outbit["flash_d0"], cosim.NC, cosim.NC, cosim.NC),
                           d_in=(cosim.NC, inbit["miso"], cosim.NC, cosim.NC))
    doms = [d for _, d in cosim.clock_domains_of(s)]
    return s, pins, doms, inbit, outbit


def oracle_edges(s, pins, doms, rst_pos, image, reset_edges, num_edges, sched):
    """The CPU edge loop with the reference's reset batching."""
    ss = s.effective_state_size()
    st = np.zeros(2 * ss, np.uint32)
    fl = co.Flash(image)
    traces = []
    done = 0
    for in_reset, n in ((1, min(reset_edges, num_edges)), (0, num_edges - min(reset_edges, num_edges))):
        if n == 0:
            continue
        ops, start = [], [0]
        for k in range(len(sched)):
            ops += sched.edge_ops(k, doms, rst_pos, in_reset)
            start.append(len(ops))
        fl.in_reset = in_reset
        traces.append(co.run_edges(s, st, np.asarray(ops, np.uint32).reshape(-1, 2),
                                   np.asarray(start, np.uint32), done % len(sched), n, fl,
                                   pins.words()))
        done += n
    return np.concatenate(traces), st, fl


@pytest.mark.parametrize("address", [0x000123, 0x000FFF, 0x000040])
def test_spi_master_reads_image_through_the_oracle_loop(address):
    rng = np.random.default_rng(address)
    image = bytes(rng.integers(0, 256, 4096, dtype=np.uint8))
    s, pins, doms, inbit, outbit = spi_setup(address, image)
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    _, st, fl = oracle_edges(s, pins, doms, inbit["rst"], image, 8, 300, sched)
    ss = s.effective_state_size()
    out = st[ss:]
    rx = sum(co._bit(out, outbit[f"rx_o{i}"]) << i for i in range(16))
    hi = image[address]
    lo = image[address + 1] if address + 1 < len(image) else 0xFF
    assert rx == (hi << 8 | lo)
    assert fl.command == 0x03 and fl.last_error_cmd == 0


def test_flash_held_in_reset_reads_nothing():
    # with the flash in reset throughout, MISO stays high: rx is all ones
    image = bytes(4096)
    s, pins, doms, inbit, outbit = spi_setup(0x10, image)
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    ss = s.effective_state_size()
    st = np.zeros(2 * ss, np.uint32)
    fl = co.Flash(image)
    ops, start = [], [0]
    for k in range(len(sched)):
        ops += sched.edge_ops(k, doms, inbit["rst"], 0)
        start.append(len(ops))
    co.run_edges(s, st, np.asarray(ops, np.uint32).reshape(-1, 2),
                 np.asarray(start, np.uint32), 0, 300, fl, pins.words())
    rx = sum(co._bit(st[ss:], outbit[f"rx_o{i}"]) << i for i in range(16))
    assert rx == 0xFFFF and fl.byte_count == 0


def test_uart_decoder_closed_forms():
    # a TX waveform sampled once per edge, 8 edges per bit: bytes framed
    # with start and stop bits decode LSB first; a frame whose stop bit
    # reads 0 is dropped; a start bit that is high again at half a bit is a
    # glitch and returns to idle
    cpb = 8

    def frame(v, stop=1):
        bits = [0] + [(v >> i) & 1 for i in range(8)] + [stop]
        return [b for b in bits for _ in range(cpb)]

    wave = [1] * 5 + frame(0x41) + [1] * 3 + frame(0xC3) + frame(0x55, stop=0) + [1] * 20
    wave += [0] * 3 + [1] * 30 + frame(0x7E) + [1] * 10
    u = co.Uart(0, cpb)
    for tx in wave:
        u.step(tx)
    assert bytes(u.data) == bytes([0x41, 0xC3, 0x7E])
    assert u.state == 0 and u.current_cycle == len(wave)


def test_bus_trace_closed_forms():
    # flags of (ibus_cyc, ibus_stb, dbus_cyc, dbus_stb, ...) at bits 0..7 of
    # word 0; the address at word 1: an entry on every active tick and on
    # every flag change, up to the capacity
    sig = {"ibus_cyc": 0, "ibus_stb": 1, "dbus_cyc": 2, "dbus_stb": 3, "dbus_we": 4,
           "ibus_adr": list(range(32, 62))}
    seq = [0b0000, 0b0001, 0b0011, 0b0011, 0b0000, 0b0000, 0b0100, 0b1100, 0b11100,
           0b0000]
    bt = co.BusTrace(sig, capacity=6)
    for k, fl in enumerate(seq):
        out = np.asarray([fl, 100 + k], np.uint32)
        bt.step(out)
    # ticks: 1 changed, 2 active+changed, 3 active, 4 changed, 6 changed,
    # 7 active (dbus), 8 active, 9 changed: capacity 6 keeps the first six
    assert [e[0] for e in bt.entries] == [1, 2, 3, 4, 6, 7]
    assert [e[2] for e in bt.entries] == [101, 102, 103, 104, 106, 107]
    assert bt.write_head == 6 and bt.current_tick == len(seq) and bt.prev_flags == 0


def uart_setup(data, blocks=1):
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.from_verilog(text=cosim.uart_tx_verilog(data))
    s = Compiled(nl).flatten(blocks)
    ins, outs = nl.ports()
    inbit = {n: s.input_map[p] for n, p, _ in ins}
    outbit = {n: s.output_map[lit] for n, lit in outs if lit > 1}
    sig = {"ibus_cyc": outbit["cyc"], "ibus_stb": outbit["stb"], "dbus_we": outbit["we"],
           "csr_ack": outbit["ack"], "ibus_adr": [outbit[f"adr{i}"] for i in range(8)],
           "dbus_adr": [outbit[f"adr{7 - i}"] for i in range(8)]}
    return s, inbit, outbit, sig


def test_uart_tx_decodes_through_the_oracle_loop():
    data = bytes([0x48, 0x00, 0xFF, 0xA5, 0x5A, 0x7F])
    s, inbit, outbit, sig = uart_setup(data)
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    doms = [d for _, d in cosim.clock_domains_of(s)]
    ops, start = [], [0]
    for k in range(len(sched)):
        ops += sched.edge_ops(k, doms, inbit["rst"], 0)
        start.append(len(ops))
    st = np.zeros(2 * s.effective_state_size(), np.uint32)
    u = co.Uart(outbit["tx"], 4)  # 2 clock cycles per bit = 4 edges
    bus = co.BusTrace(sig, 1000)
    co.run_edges(s, st, np.asarray(ops, np.uint32).reshape(-1, 2),
                 np.asarray(start, np.uint32), 0, 520, uarts=[u], bus=bus)
    assert bytes(u.data) == data
    # every entry holds the counter on the ibus address and its bit reverse
    # on the dbus address
    assert len(bus.entries) > 0 and all(e[2] == sum(((e[4] >> (7 - i)) & 1) << i
                                                    for i in range(8))
                                        for e in bus.entries)


GPIO_DESIGN = """module soc(clk, gpio_in, gpio_out, led);
  input clk;
  input [3:0] gpio_in;
  output [3:0] gpio_out;
  output led;
  wire q0, q1, n1;
  DFF r0 (.CLK(clk), .D(gpio_in[0]), .Q(q0));
  DFF r1 (.CLK(clk), .D(n1), .Q(q1));
  AND2_00_0 g1 (.A(gpio_in[1]), .B(gpio_in[2]), .Y(n1));
  INV g2 (.A(q1), .Y(gpio_out[1]));
  assign gpio_out[0] = q0;
  assign gpio_out[2] = 1'b1;
  assign gpio_out[3] = gpio_in[3];
  assign led = q1;
endmodule
"""


def test_gpio_index_parsing():
    assert cosim.parse_gpio_index("gpio_in[38]", "gpio_in") == 38
    assert cosim.parse_gpio_index("top.gpio_in:7_pad", "gpio_in") == 7
    assert cosim.parse_gpio_index("gpio_out[2]", "gpio_in") is None
    assert cosim.parse_gpio_index("gpio_in[x]", "gpio_in") is None


def test_gpio_mapping_names_ports_and_overrides():
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.from_verilog(text=GPIO_DESIGN)
    s = Compiled(nl).flatten(1)
    ins, outs = nl.ports()
    pin = {n: p for n, p, _ in ins}
    lit = dict(outs)
    m = cosim.gpio_mapping(s)
    assert m.input_bits == {i: s.input_map[pin[f"gpio_in[{i}]"]] for i in range(4)}
    # gpio_out[2] is a constant: skipped; the others resolve through the
    # output map (gpio_out[3] passes gpio_in[3] through)
    assert sorted(m.output_bits) == [0, 1, 3]
    for i in (0, 1, 3):
        assert m.output_bits[i] == s.output_map[lit[f"gpio_out[{i}]"]]
    assert [d[0] for d in m.clock_domains] == ["clk"]
    assert m.named_input_bits["clk"] == s.input_map[pin["clk"]]
    # an explicit port mapping names led as GPIO 9 and clk as input 5
    m2 = cosim.gpio_mapping(s, {"inputs": {"5": "clk"}, "outputs": {"9": "led"}})
    assert m2.input_bits[5] == s.input_map[pin["clk"]]
    assert m2.clock_domains[0][1] == 5
    assert m2.output_bits[9] == s.output_map[lit["led"]]


def test_cosim_entry_point_rejects_bad_arguments_before_any_device_work():
    # argument checks return an error (and a message) before any HIP call,
    # so they run without a GPU
    import ctypes
    from jacquard_amd import _native as N
    L = N.hip()
    dummy = ctypes.c_void_p(16)

    def call(num_blocks=1, states=dummy, ss=4, ops=dummy, start=dummy, nsched=2,
             edges=3, fs=None, pins=None, img=None, img_size=0, us=None, up=None, uc=None,
             bus=None, busp=None, cons=None, events=None):
        return L.jq_cosim_edges(num_blocks, 1, dummy, dummy, dummy, ss, states, ops, start,
                                nsched, 0, edges, fs, pins, img, img_size, us, up, uc, bus,
                                busp, cons, events, 0, 0, None, None)

    assert call(edges=0) == 0
    cases = [(dict(num_blocks=0), "num_blocks"), (dict(states=None), "state"),
             (dict(ss=0), "state"), (dict(nsched=0), "schedule"),
             (dict(start=None), "schedule"), (dict(fs=dummy), "flash pins"),
             (dict(img_size=16), "image"), (dict(us=dummy), "UART"),
             (dict(us=dummy, up=dummy), "UART"), (dict(bus=dummy), "bus trace"),
             (dict(cons=dummy), "event buffer")]
    for kw, word in cases:
        assert call(**kw) != 0, kw
        assert word in L.jq_last_error().decode(), (kw, L.jq_last_error())
    assert L.jq_cosim_flash_state_bytes() == 48


def test_cosim_runs_the_two_state_script_only():
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.from_verilog(text=GPIO_DESIGN)
    s = Compiled(nl).flatten(1, xprop=True)
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    with pytest.raises(ValueError):
        cosim.Cosim(s, sched, [])


def test_stimulus_vcd_replays_the_trace_inputs():
    # reading the stimulus VCD back (value changes per timestamp) gives every
    # input port bit of every edge's input snapshot
    image = bytes(range(256)) * 2
    s, pins, doms, inbit, outbit = spi_setup(0x10, image)
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    tr, _, _ = oracle_edges(s, pins, doms, inbit["rst"], image, 8, 60, sched)
    text = cosim.stimulus_vcd(s, tr, sched.gcd_ps, first_edge=100)
    head, body = text.split("$enddefinitions $end")
    ids = {}
    for line in head.splitlines():
        if line.startswith("$var"):
            f = line.split()
            ids[f[3]] = " ".join(f[4:-1])
    assert sorted(ids.values()) == ["clk", "miso", "rst"]
    name2bit = {"clk": inbit["clk"], "rst": inbit["rst"], "miso": inbit["miso"]}
    cur, t, seen = {}, None, 0
    for tok in body.split():
        if tok.startswith("#"):
            if t is not None:
                k = t // 5000 - 100
                for code, name in ids.items():
                    assert cur[code] == co._bit(tr[k], name2bit[name]), (k, name)
                seen += 1
            t = int(tok[1:])
        elif tok[0] in "01":
            cur[tok[1:]] = int(tok[0])
    k = t // 5000 - 100
    for code, name in ids.items():
        assert cur[code] == co._bit(tr[k], name2bit[name])
    assert seen + 1 == len(tr)


def test_flash_pin_words_layout():
    # d_in[4], clock, chip select, d_out[4], has_flash; short lists padded
    w = cosim.FlashPins(clk_out=10, csn_out=11, d_out=(20,), d_in=(30, 31)).words()
    NC = cosim.NC
    assert w.tolist() == [30, 31, NC, NC, 10, 11, 20, NC, NC, NC, 1]
