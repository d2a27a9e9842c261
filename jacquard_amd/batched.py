"""Batched lane-sliced form: many independent stimulus lanes of one netlist
in a single persistent launch (jq_simulate_batched).

Lanes run 32 to a group (one bit of every state lane-word each); every
group runs behind its own barrier and refills a slot from a device lane
queue (longest lanes first) whenever the slot's lane completes.  Each lane
is described by (seed, toggle probability, cycles) and sees exactly the
stimulus that ``stimulus.build_states(script, cycles, seed, p)`` gives the
plain path.  ``autotune`` picks the number of groups and the blocks per
group from short measured trial launches.
"""
import ctypes

import numpy as np

from . import _native as N
from . import stimulus as _stimulus


NO_SOURCE = 0xFFFFFFFF


def gather_table(script):
    """Per-partition gather table of the batched form (jqh_gather_table in
    libjacquard_host.so).

    For partition p (blocks_data order, the rows of ``script.parts``) the
    table holds 8192 u32: entry q is the state bit that lands at position q
    of the first boomerang layer after the global read, with bit 31 set when
    it is read from the current snapshot, or NO_SOURCE (position is zero).
    This is the global read of the script format
    (csrc/kernel_v1_impl.cuh:87-151, src/sim/cpu_reference.rs:47-71)
    resolved once per script: thread t's (idx, mask) entries of rounds
    0..R-1 are taken in order, each mask's set bits ascending, and the n-th
    of its B_t bits lands at position 32*t + B_t - 1 - n.

    Returns (table uint32[num_parts * 8192], start uint32[major * blocks]):
    start[m*num_blocks + b] = index of the first partition of block b in
    major stage m."""
    bd = np.ascontiguousarray(script.blocks_data, np.uint32)
    parts = np.ascontiguousarray(script.parts, np.uint64)
    n = len(parts)
    table = np.empty(max(n, 1) * 8192, np.uint32)
    if n == 0:
        table[:] = NO_SOURCE
    start = np.zeros(script.num_major_stages * script.num_blocks, np.uint32)
    if N.host().jqh_gather_table(bd.ctypes.data, bd.size, parts.ctypes.data, n,
                                 script.num_blocks, script.num_major_stages,
                                 table.ctypes.data, start.ctypes.data) != 0:
        raise RuntimeError(N.host_error())
    return table, start


def _lane_gate_positions():
    """Positions of thread t's 31 subtree gates, gate index g: level 1
    (g < 16) 4096 + t + 256g, level 2 2048 + t + 256(g-16), level 3
    1024 + ..., level 4 512 + ..., level 5 256 + t.  [31, 256]"""
    t = np.arange(256)
    rows = []
    for base, cnt in ((4096, 16), (2048, 8), (1024, 4), (512, 2), (256, 1)):
        for k in range(cnt):
            rows.append(base + t + 256 * k)
    return np.stack(rows)



def lane_blocks_data(script):
    """blocks_data with every boomerang-stage section in the batched
    kernel's lane layout (jqh_lane_blocks_data in libjacquard_host.so; the
    layout is documented with jq_simulate_batched in include/jacquard_hip.h).
    Headers, global-read rounds and the SRAM/dup and clock-enable sections
    are unchanged, so blocks_start applies as is."""
    bd = np.ascontiguousarray(script.blocks_data, np.uint32)
    parts = np.ascontiguousarray(script.parts, np.uint64)
    out = np.empty_like(bd)
    if N.host().jqh_lane_blocks_data(bd.ctypes.data, bd.size, parts.ctypes.data,
                                     len(parts), out.ctypes.data) != 0:
        raise RuntimeError(N.host_error())
    return out


def batched_program(script):
    """(lane blocks_data, gather table, gather start) of a script, built
    once and kept on the script object."""
    prog = getattr(script, "_batched_program", None)
    if prog is None:
        table, start = gather_table(script)
        prog = (lane_blocks_data(script), table, start)
        script._batched_program = prog
    return prog


class BatchedRun:
    """Lanes of one script in one persistent launch.

    Lanes enter a device queue longest first; the first groups*32 lanes are
    placed in the slots (interleaved across groups, so every group starts
    with a mix of long and short lanes) and each slot takes the next queued
    lane when its lane completes.  `groups` defaults to one group per 32
    lanes; fewer groups than that make the queue hand out the remaining
    lanes as slots free up.

    A lane's stimulus is either hashed on the device from (seed, toggle
    probability) or, with `stimulus`, recorded: one uint32[cycles+1,
    input_word_count] array per lane holding the input area of every
    snapshot, as ``vcd.parse_input_vcd`` writes it (clock flags included).
    ``launch(budget=k)`` runs k group cycles and returns; the next launch
    resumes where it stopped."""

    def __init__(self, script, seeds, probs, cycles, obs_lits=None, groups=None,
                 stimulus=None):
        self.script = script
        self.stim = None
        if stimulus is not None:
            nw = stimulus_word_count(script)
            self.stim = [np.ascontiguousarray(a, np.uint32).reshape(-1, nw) for a in stimulus]
            cycles = [len(a) - 1 for a in self.stim]
            seeds = np.zeros(len(self.stim), np.uint32)
            probs = np.zeros(len(self.stim))
        self.seeds = np.asarray(seeds, np.uint32)
        self.probs = np.asarray(probs, np.float64)
        self.cycles = np.asarray(cycles, np.int64)
        n = len(self.seeds)
        assert len(self.probs) == n and len(self.cycles) == n
        if n and int(self.cycles.min()) < 1:
            raise ValueError("every lane needs at least one cycle")
        if n and int(self.cycles.max()) >= 1 << 31:
            raise ValueError("lane cycle counts must fit 31 bits")
        if obs_lits is None:
            obs_lits = [l for l in script.netlist.arrays()["pos"].tolist() if l > 1]
        self.obs_lits = list(obs_lits)
        self.obs_bits = np.asarray([script.output_map[l] for l in self.obs_lits], np.uint32)
        self.order = np.argsort(-self.cycles, kind="stable")  # queue order
        self.G = groups or max(1, (n + 31) // 32)
        self.q_cycles = self.cycles[self.order].astype(np.uint32)
        self.q_seed = self.seeds[self.order]
        self.q_thr = np.asarray([_stimulus.threshold(self.probs[i]) for i in self.order.tolist()],
                                np.uint32)
        self.slots = np.full(self.G * 32, NO_SOURCE, np.uint32)
        self.lane_info = np.zeros(2 * max(n, 1), np.uint32)
        first = min(n, self.G * 32)
        for k in range(first):
            g, l = k % self.G, k // self.G
            self.slots[g * 32 + l] = k
            self.lane_info[2 * k] = g * 32 + l
        self.first = first
        total = int(self.cycles.sum())
        self.max_cycles = int(self.cycles.max()) if n else 0
        self.obs_cap = -(-total // (32 * self.G)) + self.max_cycles + 1
        self.data_bits, self.flag_bits = _stimulus.input_roles(script)

    def lane_cycles_total(self):
        return int(self.cycles.sum())

    def prepare(self):
        """Allocate and initialise the device buffers."""
        import torch
        from .sim import to_device
        s = self.script
        rio = s.reg_io_state_size
        st = np.zeros((self.G, 2, rio * 32), np.uint32)
        nz = lambda a: a if a.size else np.zeros(1, a.dtype)
        extra = {}
        if self.stim is None:
            st[:, :, self.flag_bits.astype(np.int64)] = 0xFFFFFFFF
        else:
            # lanes in queue order, concatenated; snapshot-0 input bits of
            # the lanes placed in slots go into the initial state
            in_bits = np.concatenate([self.data_bits, self.flag_bits]).astype(np.uint32)
            if in_bits.size and int(in_bits.max()) >= 32 * stimulus_word_count(s):
                raise ValueError("an input bit lies outside the recorded input words")
            q = [self.stim[i] for i in self.order.tolist()]
            off = np.cumsum([0] + [a.size for a in q[:-1]]).astype(np.uint64)
            for slot, k in enumerate(self.slots.tolist()):
                if k == NO_SOURCE:
                    continue
                g, l = divmod(slot, 32)
                w0 = q[k][0]
                bits = (w0[in_bits >> 5] >> (in_bits & 31)) & 1
                st[g, 0, in_bits.astype(np.int64)] |= bits.astype(np.uint32) << np.uint32(l)
            extra = dict(in_bits=to_device(nz(in_bits)),
                         stim=to_device(np.concatenate([a.reshape(-1) for a in q])),
                         stim_off=to_device(nz(off)))
            self.n_in = len(in_bits)
        lane_bd, table, start = batched_program(s)
        self.d = dict(extra,
            bs=to_device(np.ascontiguousarray(s.blocks_start, np.uint64)),
            bd=to_device(lane_bd),
            gtab=to_device(table),
            gstart=to_device(start),
            states=to_device(st.reshape(-1)),
            sram=torch.zeros(max(self.G * 32 * s.sram_storage_size, 1),
                             dtype=torch.int32, device="cuda"),
            data_bits=to_device(nz(self.data_bits)),
            flag_bits=to_device(nz(self.flag_bits)),
            lcyc=to_device(nz(self.q_cycles)),
            seed=to_device(nz(self.q_seed)),
            thr=to_device(nz(self.q_thr)),
            info=to_device(self.lane_info),
            slots=to_device(self.slots),
            queue=to_device(np.asarray([self.first], np.uint32)),
            gcyc=torch.zeros(self.G, dtype=torch.int32, device="cuda"),
            obs_bits=to_device(nz(self.obs_bits)),
            obs=torch.zeros(max(self.G * self.obs_cap * len(self.obs_bits), 1),
                            dtype=torch.int32, device="cuda"),
        )
        # pristine copies of everything the launch updates in place
        self.d0 = {k: self.d[k].clone() for k in ("states", "info", "slots", "queue")}
        return self.d

    def reset(self):
        """Restore the launch's in/out buffers to their initial contents
        with device copies and fill kernels (no memset nodes, so the
        sequence can be captured in a graph)."""
        for k, v in self.d0.items():
            self.d[k].copy_(v)
        self.d["sram"].zero_()
        self.d["gcyc"].zero_()
        self.d["obs"].zero_()

    def capture_graph(self):
        """A hipGraph of the whole batched run (reset + the persistent
        launch) on a side stream; replay() reruns every lane."""
        import torch
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.reset()
                self.launch(stream=s)
        torch.cuda.current_stream().wait_stream(s)
        self.graph = g
        return g

    def capture_chunk_graph(self, budget):
        """A hipGraph of one resumable launch of `budget` group cycles (no
        reset): every replay continues the job where the previous replay
        or launch left it, so the job runs as a sequence of replays.
        Returns the graph."""
        import torch
        if budget < 1:
            raise ValueError("a chunk needs a budget of at least one group cycle")
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.launch(stream=s, budget=budget)
        torch.cuda.current_stream().wait_stream(s)
        return g

    def predicted_group_cycles(self):
        """Group cycles each group runs when the groups advance in step:
        the device queue replayed on the host (lanes longest first, a freed
        slot takes the next lane at the cycle its lane ended).  On the
        device, groups that drift apart can take lanes in another order, so
        this is a close estimate, not a bound."""
        import heapq
        q = [int(c) for c in self.q_cycles.tolist()]
        heap = []
        for k in range(self.first):
            heapq.heappush(heap, (q[k], k % self.G, k // self.G))
        nxt, gend = self.first, [0] * self.G
        while heap:
            end, g, l = heapq.heappop(heap)
            gend[g] = max(gend[g], end)
            if nxt < len(q):
                heapq.heappush(heap, (end + q[nxt], g, l))
                nxt += 1
        return gend

    def _args(self):
        s = self.script
        d = self.d
        return [self.G, s.num_blocks, s.num_major_stages, d["bs"].data_ptr(),
                d["bd"].data_ptr(), d["gtab"].data_ptr(), d["gstart"].data_ptr(),
                s.reg_io_state_size, s.sram_storage_size, d["sram"].data_ptr(),
                d["states"].data_ptr(), d["data_bits"].data_ptr(), len(self.data_bits),
                d["flag_bits"].data_ptr(), len(self.flag_bits), len(self.seeds),
                d["lcyc"].data_ptr(), d["seed"].data_ptr(), d["thr"].data_ptr(),
                d["info"].data_ptr(), d["slots"].data_ptr(), d["queue"].data_ptr(),
                d["gcyc"].data_ptr(), d["obs_bits"].data_ptr(), len(self.obs_bits),
                self.obs_cap, d["obs"].data_ptr()]

    def launch(self, stream=None, budget=0):
        """One launch: every lane to its end, or `budget` group cycles per
        group (resuming where the previous launch stopped)."""
        import torch
        st = stream if stream is not None else torch.cuda.current_stream()
        sp = ctypes.c_void_p(st.cuda_stream)
        if self.stim is not None:
            s, d = self.script, self.d
            N.check(N.hip().jq_simulate_batched_stimulus(
                self.G, s.num_blocks, s.num_major_stages, d["bs"].data_ptr(),
                d["bd"].data_ptr(), d["gtab"].data_ptr(), d["gstart"].data_ptr(),
                s.reg_io_state_size, s.sram_storage_size, d["sram"].data_ptr(),
                d["states"].data_ptr(), d["in_bits"].data_ptr(), self.n_in,
                stimulus_word_count(s), d["stim"].data_ptr(), d["stim_off"].data_ptr(),
                len(self.seeds), d["lcyc"].data_ptr(), d["info"].data_ptr(),
                d["slots"].data_ptr(), d["queue"].data_ptr(), d["gcyc"].data_ptr(),
                int(budget), d["obs_bits"].data_ptr(), len(self.obs_bits), self.obs_cap,
                d["obs"].data_ptr(), sp))
        elif budget:
            a = self._args()
            N.check(N.hip().jq_simulate_batched_chunk(*a[:23], int(budget), *a[23:], sp))
        else:
            N.check(N.hip().jq_simulate_batched(*self._args(), sp))

    def done(self):
        """True once every lane has run all its cycles (after a synchronise)."""
        return bool((self.d["slots"].cpu().numpy().view(np.uint32) == NO_SOURCE).all()) and \
            int(self.d["queue"].cpu().numpy().view(np.uint32)[0]) >= len(self.seeds)

    def launch_profiled(self, sample_cycles, stream=None):
        """Launch with per-partition clock samples of group 0 for the first
        sample_cycles cycles; returns the device sample buffer
        (uint32[sample_cycles, num_parts] once synchronised)."""
        import torch
        n = self.script.num_parts
        samples = torch.zeros(max(sample_cycles * n, 1), dtype=torch.int32, device="cuda")
        st = stream if stream is not None else torch.cuda.current_stream()
        N.check(N.hip().jq_simulate_batched_profiled(
            *self._args(), n, samples.data_ptr(), sample_cycles,
            ctypes.c_void_p(st.cuda_stream)))
        return samples

    def run(self, validate=True):
        import torch
        from .sim import device_error
        if validate:
            self.script.validate(self.script.reg_io_state_size)
        self.prepare()
        torch.cuda.synchronize()
        self.launch()
        torch.cuda.synchronize()
        if device_error():
            raise RuntimeError("device barrier timed out in the batched launch")
        return self.lane_outputs()

    def group_cycles(self):
        return self.d["gcyc"].cpu().numpy().astype(np.int64)

    def slot_cycles_total(self):
        """Slot-cycles the groups ran (32 per group cycle)."""
        return int(self.group_cycles().sum()) * 32

    def lane_outputs(self, lanes=None, max_cycles=None):
        """{lane: uint8[cycles, n_obs]} observed values after every cycle,
        for the given lanes (all by default) and at most max_cycles cycles
        each; only those rows of the device buffer are copied."""
        n_obs = len(self.obs_bits)
        obs = self.d["obs"][:self.G * self.obs_cap * n_obs].view(self.G, self.obs_cap, n_obs)
        info = self.d["info"].cpu().numpy().view(np.uint32)
        where = {lane: k for k, lane in enumerate(self.order.tolist())}
        out = {}
        for lane in (self.order.tolist() if lanes is None else lanes):
            k = where[int(lane)]
            gs, start = int(info[2 * k]), int(info[2 * k + 1])
            g, l = divmod(gs, 32)
            c = int(self.cycles[lane])
            if max_cycles is not None:
                c = min(c, int(max_cycles))
            rows = obs[g, start:start + c, :].cpu().numpy().view(np.uint32)
            out[int(lane)] = ((rows >> np.uint32(l)) & 1).astype(np.uint8)
        return out


def stimulus_word_count(script):
    """Words of the snapshot's input area a recorded lane supplies per
    snapshot (primary inputs and clock flags)."""
    return _stimulus.input_word_count(script)


def plain_observations(script, states, cycles, obs_lits):
    """uint8[cycles, n_obs] of a finished plain snapshot buffer."""
    ss = script.effective_state_size()
    S = states.reshape(-1, ss)
    cols = []
    for lit in obs_lits:
        pos = script.output_map[lit]
        cols.append((S[1:cycles + 1, pos >> 5] >> np.uint32(pos & 31)) & 1)
    return np.stack(cols, 1).astype(np.uint8)


def max_resident_batched():
    """Co-resident workgroups of the batched kernel on this device."""
    r = ctypes.c_int(0)
    N.check(N.hip().jq_max_resident_blocks(1, ctypes.byref(r)))
    return r.value


def predicted_makespan_cycles(lane_cycles, groups):
    """Group cycles until the last lane completes with `groups`*32 slots fed
    longest-first: every slot is busy until the queue runs dry, so the
    makespan is bounded by max(total / slots, longest lane) (+ the tail of
    the last lanes, which the longest lane dominates)."""
    lane_cycles = np.asarray(lane_cycles, np.int64)
    return max(-(-int(lane_cycles.sum()) // (32 * groups)), int(lane_cycles.max()))


def autotune(compiled, lane_cycles, widths=(16, 24, 32, 48, 64), trial_cycles=24,
             log=None):
    """Choose (groups, blocks per group) for a lane set.

    For every candidate width W (blocks per group, i.e. the script is
    flattened for W blocks) and group count G that fits co-resident, one
    trial launch of G full groups runs `trial_cycles` cycles; its kernel time
    per cycle times the predicted makespan in cycles ranks the candidates.
    Returns dict(groups, blocks, script, us_per_cycle, trials)."""
    import torch
    resident = max_resident_batched()
    n = len(lane_cycles)
    g_full = max(1, -(-n // 32))
    trials = []
    best = None
    rng = np.random.default_rng(0)
    for W in widths:
        if W > resident:
            continue
        s = compiled.flatten_shared(W)
        gs = sorted({max(1, min(g_full, resident // W, g_full >> k)) for k in range(4)},
                    reverse=True)
        for G in gs:
            if G * W > resident:
                continue
            cyc = np.full(32 * G, trial_cycles)
            run = BatchedRun(s, rng.integers(1, 2**31, 32 * G), np.full(32 * G, 0.1), cyc,
                             obs_lits=[], groups=G)
            run.prepare()
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            run.launch()
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) * 1e3 / trial_cycles
            span = predicted_makespan_cycles(lane_cycles, G) * us
            trials.append(dict(blocks=W, groups=G, us_per_cycle=us, makespan_us=span))
            if log:
                log(f"autotune: W={W} G={G}: {us:.1f} us/cycle, predicted makespan "
                    f"{span / 1e6:.2f} s")
            if best is None or span < best[0]:
                best = (span, W, G, s, us)
            del run
    if best is None:
        raise RuntimeError("autotune: no candidate fits the device")
    return dict(groups=best[2], blocks=best[1], script=best[3], us_per_cycle=best[4],
                trials=trials)


def tune_partition_width(compiled, level_split, blocks, groups,
                         caps=(4000, 1800, 1300, 1100, 900), trial_cycles=48, log=None,
                         splits=()):
    """Partition tiling for the batched form at a chosen (blocks, groups).

    A shorter level split gives more major stages (more group barriers per
    cycle) with fewer, deeper partitions in each; capping the nodes per
    logic level of a partition (``max_level_nodes``) gives more, narrower
    partitions with fewer boomerang stages each.  Once a major stage holds
    more partitions than the group has blocks they serialise.  The given
    compilation, then every other level split in `splits`, then every cap
    at the best split, whose widest major stage is at most 1.5x the block
    count, are flattened for `blocks` and timed over `trial_cycles` cycles
    of `groups` full groups; the fastest wins.  Returns (Compiled, table)."""
    from .costmodel import batched_us_per_cycle
    from .script import Compiled
    table = []
    us = batched_us_per_cycle(compiled.flatten_shared(blocks), groups, cycles=trial_cycles)
    table.append({"level_split": level_split, "max_level_nodes": None,
                  "widest_stage": max(compiled.parts_per_stage()),
                  "tried": True, "us_per_cycle": us})
    best = (compiled, us)
    best_split = level_split
    if log:
        log(f"width tune: as compiled (level split {level_split}): {us:.1f} us/cycle")
    for sp in splits:
        if sp == level_split:
            continue
        comp = Compiled(compiled.netlist, level_split=sp)
        widest = max(comp.parts_per_stage())
        row = {"level_split": sp, "max_level_nodes": None, "widest_stage": widest}
        if widest > 1.5 * blocks:
            table.append(dict(row, tried=False))
            continue
        us = batched_us_per_cycle(comp.flatten_shared(blocks), groups, cycles=trial_cycles)
        table.append(dict(row, tried=True, us_per_cycle=us))
        if log:
            log(f"width tune: level split {sp}: widest stage {widest}, {us:.1f} us/cycle")
        if us < best[1]:
            best, best_split = (comp, us), sp
    for cap in caps:
        comp = Compiled(compiled.netlist, level_split=best_split, max_level_nodes=cap)
        widest = max(comp.parts_per_stage())
        row = {"level_split": best_split, "max_level_nodes": cap, "widest_stage": widest}
        if widest > 1.5 * blocks:
            table.append(dict(row, tried=False))
            continue
        us = batched_us_per_cycle(comp.flatten_shared(blocks), groups, cycles=trial_cycles)
        table.append(dict(row, tried=True, us_per_cycle=us))
        if log:
            log(f"width tune: level split {best_split}, max_level_nodes {cap}: widest "
                f"stage {widest}, {us:.1f} us/cycle")
        if us < best[1]:
            best = (comp, us)
    return best[0], table
