"""Parity of the batched lane-sliced kernel with per-lane CPU oracle runs:
every observed bit of every lane and cycle must equal the oracle's run of
that lane's stimulus through the plain (packed) semantics.  Runs with fewer
slots than lanes exercise the device lane queue: a slot that finishes a
lane is reset (state bit, clock flags, SRAM) and takes the next lane."""
import numpy as np
import pytest

import oracle
from helpers import CONFIG1, design
from jacquard_amd import stimulus
from jacquard_amd.batched import BatchedRun, plain_observations

pytestmark = pytest.mark.gpu


def _lanes(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(1, 2**31, n).astype(np.uint32),
            rng.uniform(0.02, 0.5, n), rng.integers(lo, hi + 1, n))


def _check(s, run, seeds, probs, cycles, lanes):
    got = run.run()
    for lane in lanes:
        c = int(cycles[lane])
        st = stimulus.build_states(s, c, seed=int(seeds[lane]), p=float(probs[lane]))
        sram = np.zeros(max(s.sram_storage_size, 1), np.uint32)
        oracle.run(s, st, c, sram=sram)
        ref = plain_observations(s, st, c, run.obs_lits)
        assert np.array_equal(got[lane], ref), f"lane {lane}"


@pytest.mark.parametrize("blocks,split", [(1, 0), (8, 6)])
def test_batched_config1(blocks, split):
    nl, comp = design("c1", level_split=split, **CONFIG1)
    s = comp.flatten(blocks)
    seeds, probs, cycles = _lanes(70, 20, 120, 1)
    run = BatchedRun(s, seeds, probs, cycles)
    print("groups", run.G, flush=True)
    _check(s, run, seeds, probs, cycles, range(70))
    assert run.slot_cycles_total() >= run.lane_cycles_total()


@pytest.mark.parametrize("groups", [1, 2])
def test_batched_lane_queue(groups):
    """70 lanes through 32 or 64 slots: most lanes start mid-run in a slot
    another lane used before."""
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    seeds, probs, cycles = _lanes(70, 5, 90, 3)
    run = BatchedRun(s, seeds, probs, cycles, groups=groups)
    _check(s, run, seeds, probs, cycles, range(70))
    gc = run.group_cycles()
    print("group cycles", gc.tolist(), "lane cycles", run.lane_cycles_total(), flush=True)
    # every slot stays busy until the queue runs dry
    assert gc.max() <= -(-run.lane_cycles_total() // (32 * groups)) + int(cycles.max())


def test_batched_srams_dups_major_stages():
    nl, comp = design("sram", level_split=8, seed=4, gates=30000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=4)
    s = comp.flatten(16)
    seeds, probs, cycles = _lanes(40, 10, 60, 2)
    run = BatchedRun(s, seeds, probs, cycles)
    _check(s, run, seeds, probs, cycles, [0, 1, 7, 19, 33, 39])


def test_batched_srams_through_queue():
    """SRAM lanes that reuse a slot start from zeroed SRAM contents."""
    nl, comp = design("sram", level_split=8, seed=4, gates=30000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=4)
    s = comp.flatten(16)
    seeds, probs, cycles = _lanes(48, 10, 40, 4)
    run = BatchedRun(s, seeds, probs, cycles, groups=1)
    _check(s, run, seeds, probs, cycles, range(48))


def test_batched_graph_replays():
    """The reset + launch sequence captured as one hipGraph: three replays,
    with eager work and synchronisation between them, each reproduce the
    eager run's outputs for every lane."""
    import torch
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    seeds, probs, cycles = _lanes(50, 5, 60, 7)
    run = BatchedRun(s, seeds, probs, cycles, groups=1)
    eager = run.run()
    run.capture_graph()
    for r in range(3):
        run.graph.replay()
        torch.cuda.synchronize()
        scratch = torch.ones(1 << 20, device="cuda").sum().item()  # eager work between
        assert scratch == float(1 << 20)
        got = run.lane_outputs()
        for lane in range(50):
            assert np.array_equal(got[lane], eager[lane]), (r, lane)


def test_batched_wide_group_hierarchical_barrier():
    """128 blocks per group: the group barrier counts arrivals through its
    eight sub-counters (the path taken at 128 members or more)."""
    nl, comp = design("c1", level_split=2, **CONFIG1)
    s = comp.flatten(128)
    seeds, probs, cycles = _lanes(40, 10, 60, 7)
    run = BatchedRun(s, seeds, probs, cycles, groups=1)
    print("blocks", s.num_blocks, "major", s.num_major_stages, flush=True)
    _check(s, run, seeds, probs, cycles, range(40))


def test_width_tuner_keeps_lanes_exact():
    """batched.tune_partition_width picks among cuts of the same netlist;
    the kept one runs the lanes bit-exactly."""
    from jacquard_amd.batched import tune_partition_width
    nl, comp = design("c1", level_split=6, **CONFIG1)
    comp2, table = tune_partition_width(comp, 6, 16, 1, caps=(60, 40), trial_cycles=16)
    print("width table", table, flush=True)
    assert table[0]["max_level_nodes"] is None and table[0]["tried"]
    assert [r["tried"] for r in table[1:]] == [True, False]  # widest 13 and 25 vs 24
    s = comp2.flatten(16)
    seeds, probs, cycles = _lanes(32, 10, 40, 9)
    run = BatchedRun(s, seeds, probs, cycles, groups=1)
    _check(s, run, seeds, probs, cycles, range(32))


def test_lane_vcd_matches_plain_path_vcd():
    """A batched lane's output VCD is the text the plain path writes for the
    same stimulus (whose writer is pinned byte for byte by the reference's
    golden VCDs)."""
    from jacquard_amd import vcd
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    seeds, probs, cycles = _lanes(40, 20, 80, 11)
    run = BatchedRun(s, seeds, probs, cycles)
    outs = run.run()
    rio = s.reg_io_state_size
    for lane in (0, 17, 39):
        c = int(cycles[lane])
        ts = [10 * (k + 1) for k in range(c)]
        got = vcd.write_lane_vcd(s, outs[lane], run.obs_lits, ts)
        st = stimulus.build_states(s, c, seed=int(seeds[lane]), p=float(probs[lane]))
        oracle.run(s, st, c)
        ref = vcd.write_output_vcd(s, [((k + 1) * rio, ts[k]) for k in range(c)], st)
        assert got == ref, f"lane {lane}"


def test_batched_chunks_equal_one_launch():
    """The batched job run as launches of a few group cycles each (every
    launch resumes the groups, slots and lanes where the previous one
    stopped) gives every lane the outputs of the single launch."""
    import torch
    nl, comp = design("sram", level_split=8, seed=4, gates=30000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=4)
    s = comp.flatten(16)
    seeds, probs, cycles = _lanes(48, 10, 50, 5)
    one = BatchedRun(s, seeds, probs, cycles, groups=1)
    ref = one.run()
    run = BatchedRun(s, seeds, probs, cycles, groups=1)
    run.prepare()
    torch.cuda.synchronize()
    launches = 0
    while not run.done():
        run.launch(budget=7)
        torch.cuda.synchronize()
        launches += 1
        assert launches < 200
    print("launches", launches, "group cycles", run.group_cycles().tolist(), flush=True)
    assert launches > 5
    got = run.lane_outputs()
    for lane in range(48):
        assert np.array_equal(got[lane], ref[lane]), f"lane {lane}"
    _check(s, one, seeds, probs, cycles, [0, 5, 47])


def random_input_vcd(rng, cycles, data_names, clock="clk", half=5):
    """An input VCD of `cycles` clock periods: the clock toggles every
    `half` time units from 0; each data input changes at a random time
    inside a period with probability 0.3 (sometimes at a clock edge's
    timestamp, which the reader applies before that edge)."""
    codes = {n: f"!{i}" if i < 90 else f"#{i}" for i, n in enumerate(data_names)}
    ck = "%c"
    lines = ["$timescale 1ps $end", "$scope module top $end",
             f"$var wire 1 {ck} {clock} $end"]
    for n in data_names:
        base, idx = n[:-1].split("[")
        lines.append(f"$var wire 1 {codes[n]} {base} [{idx}] $end")
    lines += ["$upscope $end", "$enddefinitions $end", "$dumpvars", f"0{ck}"]
    lines += [f"0{codes[n]}" for n in data_names] + ["$end"]
    events = {}
    for c in range(cycles):
        events.setdefault(2 * c * half, []).append(f"1{ck}")
        events.setdefault((2 * c + 1) * half, []).append(f"0{ck}")
        for n in data_names:
            if rng.random() < 0.3:
                t = 2 * c * half + int(rng.integers(0, 2 * half))
                events.setdefault(t, []).append(f"{int(rng.integers(0, 2))}{codes[n]}")
    for t in sorted(events):
        lines.append(f"#{t}")
        lines += events[t]
    lines.append(f"#{2 * cycles * half}")
    lines.append(f"1{ck}")
    return "\n".join(lines) + "\n"


def test_batched_lanes_from_input_vcds():
    """Lanes whose stimulus is read from input VCDs (parse_input_vcd's input
    words, clock flags included) run through the batched form and the lane
    queue; every lane's output VCD is the text the plain drop-in writes for
    the same VCD, and equals the CPU oracle's."""
    from jacquard_amd import sim, vcd
    from jacquard_amd.batched import stimulus_word_count
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    ins, _ = nl.ports()
    names = [n for (n, _, f) in ins if not f]
    rng = np.random.default_rng(17)
    nw = stimulus_word_count(s)
    rio = s.reg_io_state_size
    lanes = []
    for k in range(40):
        text = random_input_vcd(rng, int(rng.integers(8, 60)), names)
        states, offs, _ = vcd.parse_input_vcd(text, s)
        lanes.append((states, offs))
    stim = [st.reshape(-1, rio)[:, :nw] for st, _ in lanes]
    run = BatchedRun(s, None, None, None, stimulus=stim, groups=1)
    outs = run.run()
    for k in (0, 3, 17, 31, 32, 39):
        states, offs = lanes[k]
        c = len(offs)
        ref = states.copy()
        oracle.run(s, ref, c)
        plain = sim.run_scan(s, states.copy(), c)["states"]
        assert np.array_equal(plain, ref)
        want = vcd.write_output_vcd(s, offs, plain)
        rows = [i for i, (_, ts) in enumerate(offs) if ts != vcd.U64_MAX]
        got = vcd.write_lane_vcd(s, outs[k][rows], run.obs_lits, [offs[i][1] for i in rows])
        assert got == want, f"lane {k}"


def test_wait_time_out_aborts_the_batched_launch():
    """The batched kernel's time-out path, forced from the host with a
    one-poll bound: the launch ends promptly with the device flag raised
    (run raises and clears it), and with the bound restored a new run of
    the same lanes matches the oracle."""
    import time
    from jacquard_amd import _native as N
    L = N.hip()
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    seeds, probs, cycles = _lanes(40, 20, 80, 9)
    prev = L.jq_wait_limit(1)
    try:
        t = time.time()
        with pytest.raises(RuntimeError, match="timed out"):
            BatchedRun(s, seeds, probs, cycles, groups=1).run()
        print("aborted launch ended in", time.time() - t, "s", flush=True)
        assert time.time() - t < 30
    finally:
        assert L.jq_wait_limit(prev) == 1
    run = BatchedRun(s, seeds, probs, cycles, groups=1)
    _check(s, run, seeds, probs, cycles, range(40))
