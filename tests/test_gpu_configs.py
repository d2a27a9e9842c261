"""The benchmark workloads at their full sizes on the GPU, checked against
the CPU oracle and the direct AIG simulator in windows.

- 1M-gate plain drop-in (1M AND, depth 64, 64K DFF, 512 PI, 256 PO,
  geometric(0.3) fan-in; 100K cycles, toggle 0.1): every snapshot resident
  on the device; the first, middle and last 16-cycle windows against the
  oracle continued from the GPU's snapshot at the window start, and the
  same windows' outputs and registers against the direct AIG simulation
  started from that snapshot's registers.
- 4M-gate X + timing (depth 96, 256K DFF, 8 single-port 1024x32 SRAMs, 10%
  of DFFs initially X, delays U[20, 200] ps, setup/hold 50/20 ps at
  5000 ps; 50K cycles): first, middle and last windows against the oracle
  -- values, X masks, arrivals, SRAM contents and the canonical violation
  list of each window.
- 256 batched lanes of the 1M-gate netlist (seeds, cycles U[10K, 200K],
  toggle U[0.02, 0.5]; autotuned groups and width; one graph replay):
  eight lanes, including lanes that entered a reused slot, compared cycle
  by cycle with the plain drop-in over their full length, and each lane's
  first and last 16 cycles with the oracle.
- The 16M-gate structured netlist (64 16x16 multiplier tiles + 2M-gate
  control AIG, 1M DFF): 64 lanes x 20K cycles with the block assignment of
  the learned cost model (trained on >= 50K of the kernel's own partition
  timings); two lanes compared with the plain drop-in over all 20K cycles
  and their first and last 16 cycles with the oracle.

Sizes, seeds and level splits are those of bench.py's workloads.
"""
import ctypes

import numpy as np
import pytest

import oracle
from jacquard_amd import stimulus

V = 16  # window length in cycles
SEED = 2


def _say(*a):
    print(*a, flush=True)


def _torch():
    import torch
    return torch


def _max_resident():
    from jacquard_amd import _native as N
    r = ctypes.c_int(0)
    N.check(N.hip().jq_max_resident_blocks(0, ctypes.byref(r)))
    return r.value


def _netlist_1m():
    from jacquard_amd.netlist import Netlist
    return Netlist.random(seed=SEED, gates=1_000_000, depth=64, dffs=65536, inputs=512,
                          outputs=256, geometric_p=0.3, invert_prob=0.5)


def _netlist_4m():
    from jacquard_amd.netlist import Netlist
    return Netlist.random(seed=SEED, gates=4_000_000, depth=96, dffs=262_144, inputs=512,
                          outputs=256, geometric_p=0.3, invert_prob=0.5, srams=8,
                          sram_addr_bits=10, delay=(20, 200))


class PlainRun:
    """C cycles of the plain drop-in with every snapshot resident on the
    device, launched in segments split at every window boundary, with the
    SRAM contents and the events of every segment kept."""

    def __init__(self, s, C, seed, p, windows, *, xmask0=None, cons=None,
                 events_per_cycle=0):
        torch = _torch()
        from jacquard_amd import _native as N
        from jacquard_amd import sim
        self.s, self.C, self.seed, self.p = s, C, seed, p
        self.ss, self.rio = s.effective_state_size(), s.reg_io_state_size
        s.validate(self.ss)
        self.iw = stimulus.input_words(s, C, seed=seed, p=p)
        nw = self.iw.shape[1]
        st = torch.zeros((C + 1, self.ss), dtype=torch.int32, device="cuda")
        st[:, :nw] = torch.from_numpy(self.iw.view(np.int32)).to("cuda")
        if s.xprop_enabled:
            self.tmpl = s.xmask_template()
            st[:, self.rio:2 * self.rio] = torch.from_numpy(
                self.tmpl.view(np.int32)).to("cuda")[None, :]
            if xmask0 is not None:
                st[0, self.rio:2 * self.rio] = torch.from_numpy(
                    np.asarray(xmask0, np.uint32).view(np.int32)).to("cuda")
        nsram = max(s.sram_storage_size, 1)
        d_sram = torch.zeros(nsram, dtype=torch.int32, device="cuda")
        d_sx = (torch.full((nsram,), -1, dtype=torch.int32, device="cuda")
                if s.xprop_enabled else torch.zeros(1, dtype=torch.int32, device="cuda"))
        d_bs = sim.to_device(np.ascontiguousarray(s.blocks_start, np.uint64))
        d_bd = sim.to_device(np.ascontiguousarray(s.blocks_data, np.uint32))
        d_cons = sim.to_device(cons) if cons is not None else None
        bounds = sorted({0, C} | set(windows) | {w + V for w in windows})
        self.sram_at, self.events = {}, {}
        stream = torch.cuda.current_stream()
        for b0, b1 in zip(bounds[:-1], bounds[1:]):
            self.sram_at[b0] = (d_sram.clone(), d_sx.clone())
            ev, cap = None, 0
            if cons is not None:
                cap = events_per_cycle * (b1 - b0)
                ev = torch.zeros(16 + 32 * cap, dtype=torch.uint8, device="cuda")
                self.events[b0] = (ev, cap)
            N.check(N.hip().jq_simulate_scan(
                s.num_blocks, s.num_major_stages, d_bs.data_ptr(), d_bd.data_ptr(),
                d_sram.data_ptr(), d_sx.data_ptr(), b1 - b0, self.ss,
                st.data_ptr() + 4 * b0 * self.ss,
                None if d_cons is None else d_cons.data_ptr(),
                None if ev is None else ev.data_ptr(), cap, int(s.arrival_state_offset),
                ctypes.c_void_p(stream.cuda_stream)))
        self.sram_at[C] = (d_sram.clone(), d_sx.clone())
        torch.cuda.synchronize()
        if sim.device_error():
            raise RuntimeError("device barrier timed out")
        self.st = st

    def snapshots(self, k0, n):
        return self.st[k0:k0 + n].cpu().numpy().view(np.uint32).copy()

    def sram(self, k):
        a, b = self.sram_at[k]
        return a.cpu().numpy().view(np.uint32).copy(), b.cpu().numpy().view(np.uint32).copy()

    def window_events(self, k0):
        from jacquard_amd import sim
        ev, cap = self.events[k0]
        hdr = ev[:16].view(_torch().int32).cpu().numpy()
        cnt, ovf = int(hdr[0]), int(hdr[1])
        assert not ovf and cnt <= cap, "event capacity too small for the window"
        rec = ev[16:16 + 32 * cnt].cpu().numpy().view(sim.EVENT_DTYPE)
        return sim.canonical_events(rec)

    def oracle_window(self, k0, cons=None, xmask0=None):
        """Oracle run of [k0, k0+V) from the GPU's snapshot k0 (the
        caller's explicit X mask at k0 = 0) and SRAM contents at k0.
        Returns (snapshots, sram, sram_x, canonical events)."""
        from jacquard_amd import sim
        s, ss, rio = self.s, self.ss, self.rio
        ref = np.zeros((V + 1, ss), np.uint32)
        ref[0] = self.snapshots(k0, 1)[0]
        nw = self.iw.shape[1]
        ref[1:, :nw] = self.iw[k0 + 1:k0 + V + 1]
        if s.xprop_enabled:
            ref[1:, rio:2 * rio] = self.tmpl
        if k0 == 0:
            # the start snapshot is the caller's input, not a GPU result
            ref[0, :] = 0
            ref[0, :nw] = self.iw[0]
            if s.xprop_enabled:
                ref[0, rio:2 * rio] = self.tmpl if xmask0 is None else xmask0
        sram, sx = self.sram(k0)
        ref = ref.reshape(-1)
        ev, n = oracle.run(s, ref, V, sram=sram, sram_x=sx if s.xprop_enabled else None,
                           constraints=cons, arrival_state_offset=s.arrival_state_offset,
                           event_capacity=(1 << 22) if cons is not None else 0)
        evs = sim.canonical_events(ev[:n]) if cons is not None else None
        return ref.reshape(V + 1, ss), sram, sx, evs

    def free(self):
        torch = _torch()
        self.st = None
        self.sram_at = self.events = None
        torch.cuda.empty_cache()


def _direct_window(s, snaps, k0, seed, p):
    """Outputs and registers of [k0, k0+V) from the direct AIG simulation
    started from the registers of snapshot k0, against snapshots k0+1 ..
    k0+V."""
    A = s.netlist.arrays()
    data, fl = stimulus.input_roles(s)
    vals = stimulus.lane_inputs(len(data), k0 + V, seed, p)[k0:k0 + V]
    pos2pin = {v: k for k, v in s.input_map.items() if A["drv"][k] in (2, 3)}
    pins = [pos2pin[q] for q in data.tolist()] + [pos2pin[q] for q in fl.tolist()]
    bits = np.concatenate([vals.astype(np.uint8), np.ones((V, len(fl)), np.uint8)], axis=1)
    dpos = s.dff_data_pos
    valid = dpos != 0xFFFFFFFF
    bit_of = lambda S, q: ((S[:, q >> 5] >> (q & 31).astype(np.uint32)) & 1).astype(np.uint8)
    init = np.zeros(dpos.size, np.uint8)
    init[valid] = bit_of(snaps[:1], dpos[valid].astype(np.int64))[0]
    obs, regs = oracle.aig_sim(A, V, pins, bits, A["pos"], want_regs=True, dff_init=init)
    S = snaps[1:]
    lits = A["pos"].tolist()
    ks = [k for k, l in enumerate(lits) if l > 1]
    opos = np.array([s.output_map[lits[k]] for k in ks], np.int64)
    assert np.array_equal(bit_of(S, opos), obs[:, ks] & 1), f"outputs, window {k0}"
    assert np.array_equal(bit_of(S, dpos[valid].astype(np.int64)), regs[:, valid]), \
        f"registers, window {k0}"


@pytest.mark.gpu
def test_1m_gates_100k_cycles_plain_windows():
    from jacquard_amd.script import Compiled
    nl = _netlist_1m()
    comp = Compiled(nl, level_split=6)
    s = comp.flatten(min(max(comp.parts_per_stage()), _max_resident()))
    C = 100_000
    wins = [0, C // 2, C - V]
    _say("1M plain:", s.num_blocks, "blocks,", s.num_parts, "partitions,", C, "cycles")
    run = PlainRun(s, C, SEED, 0.1, wins)
    for k0 in wins:
        got = run.snapshots(k0, V + 1)
        ref, _, _, _ = run.oracle_window(k0)
        assert np.array_equal(got, ref), f"window at cycle {k0}"
        _direct_window(s, got, k0, SEED, 0.1)
        _say("  window", k0, "matches the oracle and the direct AIG simulation")
    run.free()


@pytest.mark.gpu
def test_4m_gates_xprop_timing_srams_50k_cycles_windows():
    from jacquard_amd.script import Compiled
    nl = _netlist_4m()
    comp = Compiled(nl, level_split=6)
    s = comp.flatten(min(max(comp.parts_per_stage()), _max_resident()), xprop=True)
    s.inject_timing()
    s.enable_timing_arrivals()
    cons = s.build_timing_constraint_buffer(5000, 50, 20)
    xm, _ = stimulus.dff_initial_xmask(s, 0.1, seed=SEED + 7)
    C = 50_000
    wins = [0, C // 2, C - V]
    _say("4M X+timing:", s.num_blocks, "blocks,", s.num_parts, "partitions,",
         s.effective_state_size(), "words per snapshot,", C, "cycles")
    run = PlainRun(s, C, SEED, 0.1, wins, xmask0=xm, cons=cons, events_per_cycle=1024)
    total = 0
    for k0 in wins:
        got = run.snapshots(k0, V + 1)
        ref, sram, sx, ev_ref = run.oracle_window(k0, cons=cons, xmask0=xm)
        assert np.array_equal(got, ref), f"snapshots, window at cycle {k0}"
        g_sram, g_sx = run.sram(k0 + V)
        assert np.array_equal(g_sram, sram) and np.array_equal(g_sx, sx), f"SRAM, window {k0}"
        ev = run.window_events(k0)
        assert len(ev) == len(ev_ref) and np.array_equal(ev, ev_ref), f"violations, window {k0}"
        total += len(ev)
        _say("  window", k0, "matches:", len(ev), "violations")
    assert total > 0
    run.free()


def _lane_plain(s, seed, p, cycles, obs_lits):
    """A lane through the plain drop-in in chunks of at most 2 GB of
    snapshots (each continuing from the previous chunk's last snapshot):
    observed outputs of every cycle, and snapshots cycles-V .. cycles."""
    torch = _torch()
    from jacquard_amd import _native as N
    from jacquard_amd import sim
    ss = s.reg_io_state_size
    chunk = max(64, min(8192, (2 << 30) // (4 * ss)))
    iw = stimulus.input_words(s, cycles, seed=seed, p=p)
    nw = iw.shape[1]
    d_bs = sim.to_device(np.ascontiguousarray(s.blocks_start, np.uint64))
    d_bd = sim.to_device(np.ascontiguousarray(s.blocks_data, np.uint32))
    d_sram = torch.zeros(max(s.sram_storage_size, 1), dtype=torch.int32, device="cuda")
    d_sx = torch.zeros(1, dtype=torch.int32, device="cuda")
    opos = torch.tensor([s.output_map[l] for l in obs_lits], dtype=torch.int64, device="cuda")
    buf = torch.zeros((chunk + 1, ss), dtype=torch.int32, device="cuda")
    buf[0, :nw] = torch.from_numpy(iw[0].view(np.int32)).to("cuda")
    obs = []
    stream = torch.cuda.current_stream()
    tail = None
    for c0 in range(0, cycles, chunk):
        n = min(chunk, cycles - c0)
        buf[1:].zero_()
        buf[1:n + 1, :nw] = torch.from_numpy(iw[c0 + 1:c0 + n + 1].view(np.int32)).to("cuda")
        N.check(N.hip().jq_simulate_scan(
            s.num_blocks, s.num_major_stages, d_bs.data_ptr(), d_bd.data_ptr(),
            d_sram.data_ptr(), d_sx.data_ptr(), n, ss, buf.data_ptr(), None, None, 0, 0,
            ctypes.c_void_p(stream.cuda_stream)))
        words = buf[1:n + 1].index_select(1, opos >> 5)
        obs.append(((words >> (opos & 31)) & 1).to(torch.uint8).cpu().numpy())
        # rows of this chunk among snapshots cycles-V .. cycles (row r is
        # snapshot c0 + r)
        r0 = max(cycles - V - c0, 0 if tail is None else 1)
        if r0 <= n:
            keep = buf[r0:n + 1].cpu().numpy().view(np.uint32).copy()
            tail = keep if tail is None else np.concatenate([tail, keep])
        buf[0].copy_(buf[n])
    torch.cuda.synchronize()
    if sim.device_error():
        raise RuntimeError("device barrier timed out")
    assert tail.shape[0] == V + 1
    return np.concatenate(obs), tail, iw


def _check_lanes(bs, ps, run, lanes, seeds, probs, cycles):
    """Batched lanes against the plain drop-in (every cycle) and the oracle
    (first and last V cycles).  `bs` is the batched script, `ps` a plain
    script of the same netlist."""
    outs = run.lane_outputs(lanes=lanes)
    for lane in lanes:
        c, sd, p = int(cycles[lane]), int(seeds[lane]), float(probs[lane])
        _say("  lane", lane, c, "cycles")
        obs, tail, iw = _lane_plain(ps, sd, p, c, run.obs_lits)
        assert np.array_equal(outs[lane], obs), f"lane {lane}: batched vs plain"
        # oracle: the first V cycles from reset (batched script), the last V
        # from the plain run's snapshot c - V (plain script)
        st = stimulus.build_states(bs, V, seed=sd, p=p)
        oracle.run(bs, st, V)
        from jacquard_amd.batched import plain_observations
        assert np.array_equal(outs[lane][:V], plain_observations(bs, st, V, run.obs_lits))
        ref = tail.copy()
        ref[1:, iw.shape[1]:] = 0
        ref = ref.reshape(-1)
        oracle.run(ps, ref, V)
        assert np.array_equal(ref, tail.reshape(-1)), f"lane {lane}: last window"


def _lanes_config4(n=256):
    rng = np.random.default_rng(SEED + 100)
    seeds = rng.integers(1, 2**31, n).astype(np.uint32)
    probs = rng.uniform(0.02, 0.5, n)
    cycles = rng.integers(10_000, 200_001, n)
    return seeds, probs, cycles


@pytest.mark.gpu
def test_256_batched_lanes_1m_gates_against_plain_and_oracle():
    torch = _torch()
    from jacquard_amd import batched as B
    from jacquard_amd.script import Compiled
    from jacquard_amd.sim import device_error
    nl = _netlist_1m()
    comp = Compiled(nl, level_split=4)
    seeds, probs, cycles = _lanes_config4()
    tune = B.autotune(comp, cycles)
    W, G = tune["blocks"], tune["groups"]
    comp_w, _ = B.tune_partition_width(comp, 4, W, G)
    bs = comp_w.flatten_shared(W)
    _say("batched 1M:", len(seeds), "lanes,", G, "groups x", W, "blocks,",
         int(cycles.sum()), "lane-cycles")
    run = B.BatchedRun(bs, seeds, probs, cycles, groups=G)
    bs.validate(bs.reg_io_state_size)
    run.prepare()
    torch.cuda.synchronize()
    run.capture_graph()
    run.graph.replay()
    torch.cuda.synchronize()
    assert not device_error()
    assert run.done()
    # lanes that entered a reused slot (queue position >= G*32: the
    # shortest), the last one the queue handed out, and two of the lanes
    # placed at the start
    order = run.order.tolist()
    reused = order[G * 32:]
    assert len(reused) >= 5
    lanes = sorted(set(reused[-5:] + [order[-1], reused[len(reused) // 2]] +
                       sorted(order[:G * 32], key=lambda l: cycles[l])[:2]))
    assert len(lanes) >= 8
    ps = Compiled(nl, level_split=6)
    ps = ps.flatten(min(max(ps.parts_per_stage()), _max_resident()))
    _check_lanes(bs, ps, run, lanes, seeds, probs, cycles)


@pytest.mark.gpu
def test_16m_gates_64_lanes_learned_costs():
    torch = _torch()
    from jacquard_amd import batched as B
    from jacquard_amd import costmodel as CM
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    from jacquard_amd.sim import device_error
    nl = Netlist.multiplier(seed=SEED)
    comp = Compiled(nl, level_split=5)  # the bench's split for this netlist
    lanes, C = 64, 20_000
    G = 2
    W = max(8, min(max(comp.parts_per_stage()), B.max_resident_batched() // G))
    costs, _model, info = CM.learn_costs(comp, W, G)
    _say("16M:", nl.counts()["num_and"], "AND;", G, "groups x", W, "blocks;",
         info["samples"], "timing samples")
    assert info["samples"] >= 50_000
    bs = comp.flatten(W, costs=costs)
    rng = np.random.default_rng(SEED + 200)
    seeds = rng.integers(1, 2**31, lanes).astype(np.uint32)
    probs = rng.uniform(0.02, 0.5, lanes)
    cycles = np.full(lanes, C)
    run = B.BatchedRun(bs, seeds, probs, cycles, groups=G)
    bs.validate(bs.reg_io_state_size)
    run.prepare()
    torch.cuda.synchronize()
    run.capture_graph()
    run.graph.replay()
    torch.cuda.synchronize()
    assert not device_error()
    assert run.done()
    ps = comp.flatten(min(max(comp.parts_per_stage()), _max_resident()))
    _check_lanes(bs, ps, run, [0, lanes - 1], seeds, probs, cycles)
