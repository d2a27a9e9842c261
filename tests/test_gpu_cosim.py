"""Device co-simulation edges (jq_cosim_edges) against the CPU edge loop of
oracle/cosim_oracle.py: every snapshot of every edge and the final flash
state, bit for bit."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, HERE)

import cosim_oracle as co  # noqa: E402
from helpers import CONFIG1, design  # noqa: E402
from jacquard_amd import cosim  # noqa: E402
from test_cosim import oracle_edges, spi_setup  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("address,batch,blocks", [(0x000123, 64, 1), (0x000FFF, 1000, 1),
                                                  (0x000321, 100, 4)])
def test_spi_master_cosim_matches_oracle(address, batch, blocks):
    # with 4 workgroups the snapshot words (9) are prepared in ranges of 3,
    # and the data-in word's owner steps the flash model itself
    rng = np.random.default_rng(address)
    image = bytes(rng.integers(0, 256, 4096, dtype=np.uint8))
    s, pins, doms, inbit, outbit = spi_setup(address, image, blocks=blocks)
    assert s.num_blocks == blocks
    if blocks > 1:
        # d_i bit 0 also drives a bit of the last-but-one state word, which
        # the last workgroup's range holds: that workgroup steps the model
        pins.d_in = (32 * (s.effective_state_size() - 2) + 5,) + tuple(pins.d_in[1:])
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    edges, reset = 300, 8
    print(f"cosim spi master: address {address:#x}, {edges} edges in batches of {batch}",
          flush=True)
    c = cosim.Cosim(s, sched, doms, reset_pos=inbit["rst"], reset_active=1,
                    reset_edges=reset, flash_pins=pins, flash_image=image)
    got = c.run(edges, batch=batch, trace=True)
    exp, st, fl = oracle_edges(s, pins, doms, inbit["rst"], image, reset, edges, sched)
    assert got.shape == exp.shape
    bad = np.nonzero((got != exp).any(axis=1))[0]
    assert bad.size == 0, f"first differing edge {bad[:1]}"
    assert np.array_equal(c.snapshots(), st)
    assert bytes(c.flash_state()) == fl.to_record().tobytes()
    if blocks == 1:
        ss = s.effective_state_size()
        rx = sum(co._bit(st[ss:], outbit[f"rx_o{i}"]) << i for i in range(16))
        lo = image[address + 1] if address + 1 < len(image) else 0xFF
        assert rx == (image[address] << 8 | lo)


def test_cosim_random_netlist_two_clock_domains():
    # a multi-block script, a second clock domain whose flags land on data
    # inputs, constant and model-driven inputs; no flash
    nl, comp = design("cfg1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    assert s.num_major_stages > 1
    ins, _ = nl.ports()
    data_bits = [s.input_map[p] for (_, p, f) in ins if not f and p in s.input_map]
    doms = [d for _, d in cosim.clock_domains_of(s)]
    doms.append(cosim.ClockDomain(data_bits[0], [data_bits[1]], [data_bits[2]]))
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000, 0, 0),
                                       cosim.ClockTiming(7500, 2500, 1)])
    consts = [(data_bits[3], 1), (data_bits[4], 0), (data_bits[5], 1)]
    driven = [data_bits[6]]
    edges = 96
    print(f"cosim random netlist: {s.num_blocks} blocks, {len(sched)} schedule edges, "
          f"{edges} edges", flush=True)
    c = cosim.Cosim(s, sched, doms, constants=consts, model_driven=driven)
    got = c.run(edges, batch=40, trace=True)
    ops, start = [], [0]
    for k in range(len(sched)):
        ops += sched.edge_ops(k, doms, None, 0, consts, driven)
        start.append(len(ops))
    st = np.zeros(2 * s.effective_state_size(), np.uint32)
    exp = co.run_edges(s, st, np.asarray(ops, np.uint32).reshape(-1, 2),
                       np.asarray(start, np.uint32), 0, edges)
    bad = np.nonzero((got != exp).any(axis=1))[0]
    assert bad.size == 0, f"first differing edge {bad[:1]}"
    # and without a trace: the same final snapshots
    c2 = cosim.Cosim(s, sched, doms, constants=consts, model_driven=driven)
    assert c2.run(edges, batch=edges) is None
    assert np.array_equal(c2.snapshots(), st)


@pytest.mark.parametrize("blocks", [1, 3])
def test_uart_and_bus_trace_match_oracle(blocks):
    from test_cosim import uart_setup
    data = bytes(range(0x41, 0x41 + 12))
    s, inbit, outbit, sig = uart_setup(data, blocks=blocks)
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    doms = [d for _, d in cosim.clock_domains_of(s)]
    uarts = [(outbit["tx"], 4), (outbit["we"], 6)]
    edges, cap = 600, 100
    print(f"cosim uart + bus trace: {edges} edges, trace capacity {cap}", flush=True)
    c = cosim.Cosim(s, sched, doms, reset_pos=inbit["rst"], reset_active=1, reset_edges=4,
                    uarts=uarts, bus_trace=dict(sig, capacity=cap))
    got = c.run(edges, batch=128, trace=True)
    ss = s.effective_state_size()
    st = np.zeros(2 * ss, np.uint32)
    us = [co.Uart(p, cpb) for p, cpb in uarts]
    bus = co.BusTrace(sig, cap)
    exp = []
    for in_reset, n, off in ((1, 4, 0), (0, edges - 4, 4)):
        ops, start = [], [0]
        for k in range(len(sched)):
            ops += sched.edge_ops(k, doms, inbit["rst"], in_reset)
            start.append(len(ops))
        exp.append(co.run_edges(s, st, np.asarray(ops, np.uint32).reshape(-1, 2),
                                np.asarray(start, np.uint32), off % len(sched), n,
                                uarts=us, bus=bus))
    assert np.array_equal(got, np.concatenate(exp))
    assert c.uart_bytes(0) == bytes(us[0].data) == data
    assert c.uart_bytes(1) == bytes(us[1].data)
    assert c.uart_states()[:2].tolist() == [u.record() for u in us]
    hdr, entries = c.bus_entries()
    assert hdr.tolist() == [bus.write_head, cap, bus.current_tick, bus.prev_flags]
    assert entries.tolist() == [list(e) for e in bus.entries]


def test_cosim_on_a_side_stream_matches_the_current_stream():
    # buffer updates between batches (reset operations, the flash's reset
    # flag) are ordered with the launches on the stream the run was given
    import torch
    image = bytes(range(256)) * 4
    s, pins, doms, inbit, outbit = spi_setup(0x40, image)
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    kw = dict(reset_pos=inbit["rst"], reset_active=1, reset_edges=10, flash_pins=pins,
              flash_image=image)
    print("cosim side stream: 300 edges in batches of 7", flush=True)
    a = cosim.Cosim(s, sched, doms, **kw).run(300, batch=7, trace=True)
    side = torch.cuda.Stream()
    c = cosim.Cosim(s, sched, doms, stream=side, **kw)
    b = c.run(300, batch=7, trace=True)
    assert np.array_equal(a, b)
    exp, _, fl = oracle_edges(s, pins, doms, inbit["rst"], image, 10, 300, sched)
    assert np.array_equal(b, exp)
    assert bytes(c.flash_state()) == fl.to_record().tobytes()


def timed_setup():
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.random(seed=5, gates=3000, depth=20, dffs=200, inputs=24, outputs=12,
                        delay=(20, 200))
    s = Compiled(nl, level_split=5).flatten(6)
    s.inject_timing()
    s.enable_timing_arrivals()
    cons = s.build_timing_constraint_buffer(600, 50, 20)  # setup and hold events
    doms = [d for _, d in cosim.clock_domains_of(s)]
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(750)])
    return s, cons, doms, sched


def timed_oracle(s, cons, doms, sched, edges):
    ops, start = [], [0]
    for k in range(len(sched)):
        ops += sched.edge_ops(k, doms)
        start.append(len(ops))
    st = np.zeros(2 * s.effective_state_size(), np.uint32)
    evs = []
    exp = co.run_edges(s, st, np.asarray(ops, np.uint32).reshape(-1, 2),
                       np.asarray(start, np.uint32), 0, edges, constraints=cons, events=evs,
                       arrival_state_offset=s.arrival_state_offset)
    return exp, (np.concatenate(evs) if evs else None)


@pytest.mark.parametrize("capacity", [1 << 16, 16])
@pytest.mark.parametrize("side_stream", [False, True])
def test_cosim_with_timing_matches_oracle(capacity, side_stream):
    # gate delays, arrivals in the state and setup/hold checks on every edge
    # (events carry cycle 0); every snapshot, the total event count, the
    # overflow flag and the records kept (all of them, or any `capacity` of
    # the oracle's multiset); on a side stream with no flash the buffers the
    # constructor fills must still reach the launches first
    import torch
    from jacquard_amd.sim import canonical_events
    s, cons, doms, sched = timed_setup()
    edges = 120
    print(f"cosim timing: {s.num_blocks} blocks, {s.num_major_stages} major stages, "
          f"{edges} edges, capacity {capacity}, side stream {side_stream}", flush=True)
    c = cosim.Cosim(s, sched, doms, constraints=cons, event_capacity=capacity,
                    stream=torch.cuda.Stream() if side_stream else None)
    got = c.run(edges, batch=50, trace=True)
    exp, want = timed_oracle(s, cons, doms, sched, edges)
    bad = np.nonzero((got != exp).any(axis=1))[0]
    assert bad.size == 0, f"first differing edge {bad[:1]}"
    ev, total, overflow = c.events()
    assert want is not None and total == len(want)
    assert set(want["event_type"].tolist()) == {4, 5}
    assert overflow == int(total > capacity)
    if total <= capacity:
        assert canonical_events(ev).tobytes() == canonical_events(want).tobytes()
    else:
        assert len(ev) == capacity
        pool = [r.tobytes() for r in want]
        for r in ev:
            pool.remove(r.tobytes())  # each kept record is one of the oracle's


def test_cosim_rejects_a_short_constraint_buffer():
    s, cons, doms, sched = timed_setup()
    with pytest.raises(ValueError):
        cosim.Cosim(s, sched, doms, constraints=cons[:-1])

