"""Parity of the plain drop-in kernel (through the C ABI) with the CPU oracle.

Every snapshot of every cycle must be bit-identical to the oracle's
kernel-semantics run; the final values are also checked against the direct
AIG simulation, which does not look at the script at all."""
import numpy as np
import pytest

import oracle
from helpers import CONFIG1, design, direct_check
from jacquard_amd import sim, stimulus

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


class _sync_mode:
    """Sets the value-only scan launches' synchronisation (1: dataflow, the
    default; 0: grid barriers) for the duration of a with-block."""

    def __init__(self, mode):
        from jacquard_amd import _native as N
        self.L, self.mode = N.hip(), mode

    def __enter__(self):
        self.prev = self.L.jq_scan_sync_mode(self.mode)

    def __exit__(self, *exc):
        self.L.jq_scan_sync_mode(self.prev)


@pytest.mark.parametrize("num_blocks,split", [(1, 0), (8, 0), (16, 6), (64, 4)])
def test_config1_2state(num_blocks, split):
    nl, comp = design("c1", level_split=split, **CONFIG1)
    s = comp.flatten(num_blocks)
    cycles = 300
    _say("config1", num_blocks, split, "parts", s.num_parts)
    st = stimulus.build_states(s, cycles, seed=11, p=0.5)
    ref = st.copy()
    oracle.run(s, ref, cycles)
    got = sim.run_scan(s, st.copy(), cycles)["states"]
    assert np.array_equal(got, ref)
    direct_check(s, got, cycles, 11, 0.5)


@pytest.mark.parametrize("num_blocks", [128, 200])
@pytest.mark.parametrize("sync", [0, 1])
def test_many_blocks_hierarchical_barrier(num_blocks, sync):
    """At 128 workgroups or more the grid barrier counts arrivals through
    eight sub-counters; most blocks here are empty and still take part
    (sync 0).  With dataflow synchronisation (sync 1) the empty blocks
    publish progress that nobody waits for."""
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.random(seed=5, gates=20000, depth=40, dffs=1024, inputs=64, outputs=32,
                        geometric_p=0.3, invert_prob=0.5)
    # small partitions: about ten per major stage, spread over the blocks
    s = Compiled(nl, level_split=4, max_part_nodes=500).flatten(num_blocks)
    cycles = 150
    _say("blocks", num_blocks, "major", s.num_major_stages, "parts", s.num_parts)
    st = stimulus.build_states(s, cycles, seed=21, p=0.3)
    ref = st.copy()
    oracle.run(s, ref, cycles)
    with _sync_mode(sync):
        got = sim.run_scan(s, st.copy(), cycles)["states"]
    assert np.array_equal(got, ref)


def test_reference_symbol_entry_point():
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(16)
    cycles = 100
    st = stimulus.build_states(s, cycles, seed=12, p=0.3)
    ref = st.copy()
    oracle.run(s, ref, cycles)
    got = sim.run_scan(s, st.copy(), cycles, reference_symbols=True)["states"]
    assert np.array_equal(got, ref)


def test_srams_and_major_stages():
    nl, comp = design("sram", level_split=8, seed=4, gates=30000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=4)
    s = comp.flatten(32)
    cycles = 120
    _say("sram", s.num_parts, s.num_major_stages, s.sram_storage_size)
    st = stimulus.build_states(s, cycles, seed=13, p=0.2)
    ref = st.copy()
    sram_ref = np.zeros(s.sram_storage_size, np.uint32)
    oracle.run(s, ref, cycles, sram=sram_ref)
    out = sim.run_scan(s, st.copy(), cycles)
    assert np.array_equal(out["states"], ref)
    assert np.array_equal(out["sram"][:s.sram_storage_size], sram_ref)
    direct_check(s, out["states"], cycles, 13, 0.2)


@pytest.mark.parametrize("frac", [0.0, 0.1])
def test_xprop(frac):
    nl, comp = design("xp", level_split=8, seed=5, gates=20000, depth=40,
                      dffs=2000, inputs=64, outputs=32, geometric_p=0.3, srams=2)
    s = comp.flatten(16, xprop=True)
    cycles = 80
    m, pick = stimulus.dff_initial_xmask(s, frac, seed=7)
    st = stimulus.build_states(s, cycles, seed=14, p=0.4, initial_xmask=m)
    ref = st.copy()
    sram = np.zeros(s.sram_storage_size, np.uint32)
    sram_x = np.full(s.sram_storage_size, 0xFFFFFFFF, np.uint32)
    oracle.run(s, ref, cycles, sram=sram, sram_x=sram_x)
    out = sim.run_scan(s, st.copy(), cycles)
    assert np.array_equal(out["states"], ref)
    assert np.array_equal(out["sram_x"][:s.sram_storage_size], sram_x)


@pytest.mark.parametrize("xprop", [False, True])
def test_dataflow_and_barrier_sync_agree(xprop):
    """Value-only launches wait only for the blocks whose words the next
    major stage reads (each up to the stage that wrote them) instead of
    taking a grid barrier.  Both schedules must give the oracle's bits:
    every snapshot, the SRAM contents and their X shadow."""
    from jacquard_amd import _native as N
    assert N.hip().jq_scan_sync_mode(-1) == 1, "dataflow is the default"
    nl, comp = design("sync", level_split=4, seed=9, gates=40000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=3)
    s = comp.flatten(48, xprop=xprop)
    cycles = 100
    _say("sync modes", xprop, s.num_blocks, s.num_major_stages, s.num_parts)
    m = None
    if xprop:
        m, _pick = stimulus.dff_initial_xmask(s, 0.1, seed=8)
    st = stimulus.build_states(s, cycles, seed=15, p=0.3, initial_xmask=m)
    ref = st.copy()
    sram = np.zeros(s.sram_storage_size, np.uint32)
    sram_x = np.full(s.sram_storage_size, 0xFFFFFFFF, np.uint32)
    oracle.run(s, ref, cycles, sram=sram, sram_x=sram_x if xprop else None)
    for mode in (1, 0):
        with _sync_mode(mode):
            out = sim.run_scan(s, st.copy(), cycles)
        assert np.array_equal(out["states"], ref), mode
        assert np.array_equal(out["sram"][:s.sram_storage_size], sram), mode
        if xprop:
            assert np.array_equal(out["sram_x"][:s.sram_storage_size], sram_x), mode


@pytest.mark.parametrize("cycles", [1, 7, 8, 9])
def test_launch_lengths_around_the_dataflow_threshold(cycles):
    """Launches shorter than 8 cycles keep grid barriers; from 8 cycles on
    the producer table is built.  Both sides of the threshold, on a
    multi-block script with SRAMs, give the oracle's snapshots and SRAMs."""
    nl, comp = design("sync", level_split=4, seed=9, gates=40000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=3)
    s = comp.flatten(48)
    st = stimulus.build_states(s, cycles, seed=16, p=0.3)
    ref = st.copy()
    sram = np.zeros(s.sram_storage_size, np.uint32)
    oracle.run(s, ref, cycles, sram=sram)
    out = sim.run_scan(s, st.copy(), cycles)
    assert np.array_equal(out["states"], ref)
    assert np.array_equal(out["sram"][:s.sram_storage_size], sram)


def test_more_major_stages_than_the_dataflow_table_holds():
    """One logic level per major stage on a depth-140 netlist gives more
    than 120 major stages: the launch keeps grid barriers and stays exact."""
    nl, comp = design("deep", level_split=1, seed=12, gates=6000, depth=140,
                      dffs=256, inputs=32, outputs=16, geometric_p=0.3)
    s = comp.flatten(8)
    assert s.num_major_stages > 120
    cycles = 20
    _say("deep", s.num_major_stages, s.num_parts)
    st = stimulus.build_states(s, cycles, seed=17, p=0.3)
    ref = st.copy()
    oracle.run(s, ref, cycles)
    got = sim.run_scan(s, st.copy(), cycles)["states"]
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("sync", [0, 1])
def test_wait_time_out_aborts_the_launch_and_the_next_launch_is_exact(sync):
    """The time-out path, forced from the host: with a one-poll bound nearly
    every wait gives up, the first one marks the launch aborted, every
    workgroup leaves at its next wait, and the launch ends promptly with the
    device flag raised (run_scan raises and clears it).  With the bound
    restored the next launch of the same script matches the oracle."""
    import time
    from jacquard_amd import _native as N
    L = N.hip()
    nl, comp = design("sync", level_split=4, seed=9, gates=40000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=3)
    s = comp.flatten(48)
    cycles = 200
    st = stimulus.build_states(s, cycles, seed=18, p=0.3)
    ref = st.copy()
    oracle.run(s, ref, cycles)
    with _sync_mode(sync):
        prev = L.jq_wait_limit(1)
        try:
            t = time.time()
            with pytest.raises(RuntimeError, match="timed out"):
                sim.run_scan(s, st.copy(), cycles)
            _say("aborted launch ended in", time.time() - t, "s")
            assert time.time() - t < 30
        finally:
            assert L.jq_wait_limit(prev) == 1
        got = sim.run_scan(s, st.copy(), cycles)["states"]
    assert np.array_equal(got, ref)


def test_captured_launch_replays_match_the_oracle():
    """A value-only launch through jq_simulate_scan captured into a hipGraph
    on a side stream: under capture it is an ordinary kernel node with grid
    barriers (no dataflow table is built into the graph).  Before each of
    three replays the snapshot buffer gets its input back, with a junk
    pattern in every word the launch writes (a replay that skipped a word
    would show it), the SRAMs are cleared, and an eager launch of another
    script runs; every replay reproduces the oracle's snapshots and SRAM
    contents, and no wait times out."""
    import ctypes
    import torch
    from jacquard_amd import _native as N
    L = N.hip()
    nl, comp = design("cap", level_split=4, seed=9, gates=40000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=3)
    s = comp.flatten(48)
    cycles = 60
    ss = s.effective_state_size()
    st = stimulus.build_states(s, cycles, seed=19, p=0.3)
    ref = st.copy()
    sram_ref = np.zeros(s.sram_storage_size, np.uint32)
    oracle.run(s, ref, cycles, sram=sram_ref)
    # the launch's input: snapshot 0 and the pre-filled primary inputs, and
    # a junk pattern in every writeout word of snapshots 1.. (the words
    # past the inputs' n // 32 + 1), which each cycle rewrites in full
    ws = len(s.input_layout) // 32 + 1
    inp = st.copy().reshape(cycles + 1, ss)
    inp[1:, ws:] = 0xA5C3F00D
    inp = inp.reshape(-1)
    d_bs = sim.to_device(np.ascontiguousarray(s.blocks_start, np.uint64))
    d_bd = sim.to_device(np.ascontiguousarray(s.blocks_data, np.uint32))
    d_in = sim.to_device(inp)
    d_st = d_in.clone()
    d_sram = torch.zeros(s.sram_storage_size, dtype=torch.int32, device="cuda")
    d_sx = torch.zeros(1, dtype=torch.int32, device="cuda")
    _, small = design("c1", level_split=0, **CONFIG1)
    s2 = small.flatten(8)
    st2 = stimulus.build_states(s2, 50, seed=3, p=0.5)
    ref2 = st2.copy()
    oracle.run(s2, ref2, 50)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            rc = L.jq_simulate_scan(
                s.num_blocks, s.num_major_stages, d_bs.data_ptr(), d_bd.data_ptr(),
                d_sram.data_ptr(), d_sx.data_ptr(), cycles, ss, d_st.data_ptr(),
                None, None, 0, 0, ctypes.c_void_p(side.cuda_stream))
    torch.cuda.current_stream().wait_stream(side)
    N.check(rc)
    for k in range(3):
        d_st.copy_(d_in)
        d_sram.zero_()
        assert np.array_equal(sim.run_scan(s2, st2.copy(), 50)["states"], ref2)
        _say("replay", k)
        g.replay()
        torch.cuda.synchronize()
        assert not sim.device_error(), f"replay {k}: a wait timed out"
        assert np.array_equal(sim.from_device(d_st, np.uint32), ref), f"replay {k}"
        assert np.array_equal(sim.from_device(d_sram, np.uint32), sram_ref), f"replay {k}"


def test_timing_events():
    nl, comp = design("tm", level_split=6, seed=6, gates=8000, depth=30,
                      dffs=800, inputs=32, outputs=16, geometric_p=0.3)
    d = np.zeros(nl.counts()["num_pins"] + 1, np.uint16)
    rng = np.random.default_rng(1)
    A = nl.arrays()
    ands = A["drv"] == 1
    d[ands] = rng.integers(20, 201, ands.sum())
    nl.set_delays(d)
    s = comp.flatten(16)
    s.inject_timing()
    s.enable_timing_arrivals()
    cons = s.build_timing_constraint_buffer(190, 50, 20)
    cycles = 60
    st = stimulus.build_states(s, cycles, seed=15, p=0.5)
    ref = st.copy()
    ev_ref, n_ref = oracle.run(s, ref, cycles, constraints=cons,
                               arrival_state_offset=s.arrival_state_offset,
                               event_capacity=1 << 20)
    out = sim.run_scan(s, st.copy(), cycles, constraints=cons,
                       event_capacity=1 << 20)
    assert np.array_equal(out["states"], ref)
    assert out["event_count"] == n_ref and n_ref > 0
    a = sim.canonical_events(out["events"])
    b = sim.canonical_events(ev_ref)
    assert np.array_equal(a, b)


def test_timed_reference_symbol_keeps_1024_events():
    """simulate_v1_noninteractive_timed_hip keeps the reference buffer's
    1024 records (csrc/event_buffer.h:99-111): the header counts every
    violation, sets the overflow flag, and the kept records are violations
    of the full list."""
    nl, comp = design("tm", level_split=6, seed=6, gates=8000, depth=30,
                      dffs=800, inputs=32, outputs=16, geometric_p=0.3)
    d = np.zeros(nl.counts()["num_pins"] + 1, np.uint16)
    rng = np.random.default_rng(1)
    A = nl.arrays()
    ands = A["drv"] == 1
    d[ands] = rng.integers(20, 201, ands.sum())
    nl.set_delays(d)
    s = comp.flatten(16)
    s.inject_timing()
    s.enable_timing_arrivals()
    cons = s.build_timing_constraint_buffer(190, 50, 20)
    cycles = 60
    st = stimulus.build_states(s, cycles, seed=15, p=0.5)
    ref = st.copy()
    ev_ref, n_ref = oracle.run(s, ref, cycles, constraints=cons,
                               arrival_state_offset=s.arrival_state_offset,
                               event_capacity=1 << 20)
    assert n_ref > 1024
    out = sim.run_scan(s, st.copy(), cycles, constraints=cons, reference_symbols=True)
    assert np.array_equal(out["states"], ref)
    assert out["event_count"] == n_ref and out["overflow"] == 1
    kept = sim.canonical_events(out["events"])
    assert len(kept) == 1024
    full = {bytes(e) for e in sim.canonical_events(ev_ref)}
    assert all(bytes(e) in full for e in kept)


def test_xprop_timing_srams_together():
    """X lane, arrivals, setup/hold and SRAMs in one run (the 4M-gate
    workload's feature mix at test size)."""
    nl, comp = design("xts", level_split=6, seed=9, gates=40000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=2,
                      delay=(20, 200))
    s = comp.flatten(24, xprop=True)
    s.inject_timing()
    s.enable_timing_arrivals()
    cons = s.build_timing_constraint_buffer(1500, 50, 20)
    cycles = 60
    m, _ = stimulus.dff_initial_xmask(s, 0.1, seed=8)
    st = stimulus.build_states(s, cycles, seed=16, p=0.3, initial_xmask=m)
    ref = st.copy()
    sram = np.zeros(s.sram_storage_size, np.uint32)
    sram_x = np.full(s.sram_storage_size, 0xFFFFFFFF, np.uint32)
    ev_ref, n_ref = oracle.run(s, ref, cycles, sram=sram, sram_x=sram_x, constraints=cons,
                               arrival_state_offset=s.arrival_state_offset,
                               event_capacity=1 << 20)
    out = sim.run_scan(s, st.copy(), cycles, constraints=cons, event_capacity=1 << 20)
    _say("xts", s.num_parts, s.num_major_stages, "events", n_ref)
    assert np.array_equal(out["states"], ref)
    assert np.array_equal(out["sram"][:s.sram_storage_size], sram)
    assert np.array_equal(out["sram_x"][:s.sram_storage_size], sram_x)
    assert out["event_count"] == n_ref and n_ref > 0
    assert np.array_equal(sim.canonical_events(out["events"]), sim.canonical_events(ev_ref))


@pytest.mark.parametrize("xprop", [False, True])
def test_device_output_changes_match_host_writer(xprop):
    """VCD text from device change records equals the host writer's text
    from the full snapshots (values and X), on a random design with many
    outputs and skipped timestamps."""
    from jacquard_amd import vcd
    nl, comp = design("xp", level_split=8, seed=5, gates=20000, depth=40,
                      dffs=2000, inputs=64, outputs=32, geometric_p=0.3, srams=2)
    s = comp.flatten(16, xprop=xprop)
    cycles = 70
    m = stimulus.dff_initial_xmask(s, 0.2, seed=3)[0] if xprop else None
    st = stimulus.build_states(s, cycles, seed=17, p=0.4, initial_xmask=m)
    out = sim.run_scan(s, st.copy(), cycles, return_device=True)
    rio, ess = s.reg_io_state_size, s.effective_state_size()
    offs = [((k + 1) * rio, vcd.U64_MAX if k % 7 == 3 else 10 * k) for k in range(cycles)]
    S = out["states"].reshape(-1, ess)
    values = np.ascontiguousarray(S[:, :rio]).reshape(-1)
    xm = np.ascontiguousarray(S[:, rio:2 * rio]).reshape(-1) if xprop else None
    host = vcd.write_output_vcd(s, offs, values, xmask_states=xm)
    cnt, rec = vcd.output_changes_device(s, out["device"]["states"], offs)
    dev = vcd.write_output_vcd_changes(s, offs, cnt, rec)
    assert dev == host
    assert len(rec) < S[1:].size  # compacted: far fewer words than snapshots


def test_tune_plain_chooses_a_fitting_width_that_stays_exact():
    """The measured partition-width choice only changes how the netlist is
    cut: whatever cap it keeps, the script fits the co-resident workgroups
    and simulates bit-exactly."""
    from jacquard_amd.netlist import Netlist
    nl = Netlist.random(seed=5, gates=20000, depth=40, dffs=1024, inputs=64, outputs=32,
                        geometric_p=0.3, invert_prob=0.5)
    comp, s, table = sim.tune_plain(nl, 4, caps=(2600, 400, 200), trial_cycles=20)
    _say("tune table", table)
    assert any(r["fits"] for r in table)
    assert s.num_blocks == max(comp.parts_per_stage())
    cycles = 60
    st = stimulus.build_states(s, cycles, seed=31, p=0.3)
    ref = st.copy()
    oracle.run(s, ref, cycles)
    got = sim.run_scan(s, st.copy(), cycles)["states"]
    assert np.array_equal(got, ref)


def test_scan_profiled_samples_every_partition_and_keeps_values():
    """jq_simulate_scan_profiled writes a positive tick count for every
    partition execution of the sampled cycles and simulates exactly like the
    unprofiled kernel."""
    from jacquard_amd import costmodel as CM
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    cycles, sampled = 40, 10
    st = stimulus.build_states(s, cycles, seed=41, p=0.4)
    ref = st.copy()
    oracle.run(s, ref, cycles)
    d = sim.to_device(st.copy())
    ticks = CM.profile_partitions(s, d, cycles, sampled)
    assert np.array_equal(d.cpu().numpy().view(np.uint32), ref)
    assert ticks.shape == (sampled, s.num_parts)
    assert (ticks > 0).all()


@pytest.mark.parametrize("xprop", [False, True])
def test_global_read_wider_than_prefetched_rounds(xprop):
    """Scripts whose global read needs more rounds than the kernel prefetches
    (up to 32 single-bit rounds per thread here): the extra rounds are read
    after the prefetched ones, with the X lane when X-capable."""
    from helpers import OneBitRounds
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.random(seed=5, gates=20000, depth=40, dffs=1024, inputs=64, outputs=32,
                        geometric_p=0.3, invert_prob=0.5)
    s = Compiled(nl, level_split=4, max_part_nodes=500).flatten(8, xprop=xprop)
    w = OneBitRounds(s)
    _say("xprop", xprop, "max rounds", w.max_rounds)
    assert w.max_rounds > 16
    cycles = 60
    st = stimulus.build_states(s, cycles, seed=33, p=0.3)
    ref = st.copy()
    oracle.run(s, ref, cycles)  # the original script: same values
    got = sim.run_scan(w, st.copy(), cycles)["states"]
    assert np.array_equal(got, ref)
