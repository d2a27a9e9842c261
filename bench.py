#!/usr/bin/env python3
"""Throughput of the MI355X Boolean-processor interpreter in simulated
lane-cycles per second: batched+learned against the plain GPU drop-in
against the CPU path, bit-exact.

Default workload (the headline): the 4M-gate netlist -- a seeded random AIG
of 4M AND gates, depth 96, 256K DFFs, 8 single-port 1024x32 SRAMs, 512
inputs, 256 outputs, fan-in level distance ~ Geometric(0.3) -- simulated as
256 independent value lanes (per-lane seeds, cycle counts U[10K, 200K],
toggle rates U[0.02, 0.5]) by the batched lane-sliced kernel with autotuned
groups and blocks and the learned cost model's block assignment.  `value`
is the whole lane job's lane-cycles over its wall time, with the lanes'
stimulus generated on the device; one "step" is one graph replay of a
resumable launch over an equal share of the job's group cycles.  Beside it:
the plain drop-in on the same netlist, the reference CPU path (the
oracle's literal restatement of sanity_check_cpu, one host core, a bounded
cycle sample), and under config.forms the 1M-gate plain drop-in over 100K
cycles, the 256 batched lanes of the 1M-gate netlist and the structured
16M-gate netlist's 64 lanes x 20K cycles with learned costs.  `roofline` is the
batched kernel's: algorithmic HBM bytes per lane-cycle over its kernel time
(HIP events on the launch stream), with the counter-measured traffic of the
committed FETCH_SIZE / WRITE_SIZE profile of this workload beside it.

Other workloads (--workload): plain (the 1M-gate drop-in line, with LDS
roofline and barrier time), xtiming (4M gates with X, timing and SRAMs),
mult16m (the 16M-gate structured netlist), cosim.

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "simulated lane-cycles/sec (batched+learned vs plain GPU vs CPU), bit-exact"
HBM_PEAK_GBPS = 8000.0
LDS_SPEC_PEAK_GBPS = 256 * 128 * 2.4  # 256 CUs x 128 B/clk x 2.4 GHz


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10,
                    help="north: graph replays the batched job is split into; plain: "
                         "timed launches of --cycles-per-step cycles")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cycles-per-step", type=int, default=10_000)
    ap.add_argument("--plain-launches", type=int, default=2,
                    help="north: launches of --cycles-per-step cycles of the plain "
                         "drop-in on the 4M-gate netlist")
    ap.add_argument("--no-forms", action="store_true",
                    help="north: skip the 1M- and 16M-gate forms")
    ap.add_argument("--no-16m", action="store_true",
                    help="north: skip the 16M-gate structured netlist's form")
    ap.add_argument("--workload", choices=["north", "plain", "xtiming", "lanes4m",
                                           "mult16m", "cosim"],
                    default="north",
                    help="north (the headline): batched+learned lanes vs the plain "
                         "drop-in on the 4M-gate netlist, the 1M-gate forms beside it; "
                         "plain: the 1M-gate 2-state drop-in; "
                         "xtiming: the 4M-gate netlist with 8 SRAMs, 10%% initial-X "
                         "DFFs, 20-200 ps gate delays and 50/20 ps setup/hold at 5 ns; "
                         "lanes4m: the headline without the 1M-gate forms; "
                         "mult16m: the 16M-gate structured "
                         "netlist (64 16x16 multiplier tiles + 2M-gate control AIG, 1M "
                         "DFF), --lanes lanes of --cycles cycles, learned vs heuristic; "
                         "cosim: reactive co-simulation edges (SPI master + flash model, "
                         "and the 1M-gate netlist clocked edge by edge)")
    ap.add_argument("--gates", type=int, default=None)
    ap.add_argument("--depth", type=int, default=None)
    ap.add_argument("--dffs", type=int, default=None)
    ap.add_argument("--cycles", type=int, default=50_000,
                    help="xtiming: cycles simulated (one run)")
    ap.add_argument("--verify-window", type=int, default=8,
                    help="xtiming: cycles checked against the oracle at the start "
                         "and at the end of the run")
    ap.add_argument("--inputs", type=int, default=512)
    ap.add_argument("--outputs", type=int, default=256)
    ap.add_argument("--toggle", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--split", type=int, default=6,
                    help="level-split period of the plain drop-in's script")
    ap.add_argument("--batched-split", type=int, default=4,
                    help="level-split period of the batched form's script (it runs "
                         "more partitions per block, so it prefers shorter major stages)")
    ap.add_argument("--blocks", type=int, default=0, help="plain blocks (0 = auto)")
    ap.add_argument("--no-plain-tune", action="store_true",
                    help="plain script with the compiler's default partition width "
                         "instead of the measured choice (sim.tune_plain)")
    ap.add_argument("--lanes", type=int, default=256)
    ap.add_argument("--batched-scale", type=float, default=1.0,
                    help="scale of the batched per-lane cycle counts "
                         "(1.0 = U[10K, 200K], the batched configuration)")
    ap.add_argument("--batched-blocks", type=int, default=0,
                    help="blocks per lane group (0 = autotune)")
    ap.add_argument("--batched-groups", type=int, default=0,
                    help="lane groups of 32 slots (0 = one per 32 lanes)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-cosim", action="store_true",
                    help="skip the co-simulation edges reported under config.forms.cosim")
    ap.add_argument("--no-width-tune", action="store_true",
                    help="batched form with the compiler's default partition width "
                         "instead of the measured choice (batched.tune_partition_width)")
    ap.add_argument("--tune-splits", default="2,3,4,5",
                    help="level splits the batched form's tiling tune also tries "
                         "(comma-separated; empty: only --batched-split)")
    ap.add_argument("--no-learned", action="store_true",
                    help="batched form with the heuristic block assignment only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--verify-cycles", type=int, default=4)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run an elementwise copy of MIB MiB three times "
                         "(the known-size kernel that scales the HBM counters)")
    a = ap.parse_args()
    big = a.workload in ("xtiming", "lanes4m", "north")
    if a.workload == "lanes4m":
        a.workload, a.no_forms = "north", True
    if a.gates is None:
        a.gates = 4_000_000 if big else 1_000_000
    if a.depth is None:
        a.depth = 96 if big else 64
    if a.dffs is None:
        a.dffs = 262_144 if big else 65_536
    return a


def copy_calibration(mib):
    """Three plain copies of `mib` MiB: each reads and writes exactly
    mib * 2^20 bytes, which calibrates FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)  # an elementwise kernel, not a DMA blit
    torch.cuda.synchronize()
    del src, dst
    log(f"copy calibration: 3 elementwise copies of {n * 4} bytes")


def build_design(a):
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    t = time.time()
    nl = Netlist.random(seed=a.seed, gates=a.gates, depth=a.depth, dffs=a.dffs,
                        inputs=a.inputs, outputs=a.outputs, geometric_p=0.3,
                        invert_prob=0.5)
    comp = Compiled(nl, level_split=a.split)
    log(f"netlist {nl.counts()} compiled in {time.time() - t:.1f}s")
    log(comp.log.strip().splitlines()[-1])
    return nl, comp


def bench_plain(a, script, torch):
    import ctypes
    from jacquard_amd import _native as N
    from jacquard_amd import sim, stimulus
    L = N.hip()
    ss = script.effective_state_size()
    cps = a.cycles_per_step
    total = (a.warmup + a.steps) * cps
    t = time.time()
    iw = stimulus.input_words(script, total, seed=a.seed, p=a.toggle)
    nw = iw.shape[1]
    st = torch.zeros((total + 1, ss), dtype=torch.int32, device="cuda")
    st[:, :nw] = torch.from_numpy(iw.view(np.int32)).to("cuda")
    d_bs = sim.to_device(np.ascontiguousarray(script.blocks_start, np.uint64))
    d_bd = sim.to_device(np.ascontiguousarray(script.blocks_data, np.uint32))
    d_sram = torch.zeros(max(script.sram_storage_size, 1), dtype=torch.int32, device="cuda")
    d_sx = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    log(f"plain: state {ss} words, {total} cycles staged in {time.time() - t:.1f}s")
    stream = torch.cuda.current_stream()
    base = st.data_ptr()

    def launch(i):
        N.check(L.jq_simulate_scan(
            script.num_blocks, script.num_major_stages, d_bs.data_ptr(),
            d_bd.data_ptr(), d_sram.data_ptr(), d_sx.data_ptr(), cps, ss,
            base + 4 * i * cps * ss, None, None, 0, 0,
            ctypes.c_void_p(stream.cuda_stream)))

    for i in range(a.warmup):
        launch(i)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        ev[k][0].record(stream)
        launch(a.warmup + k)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if sim.device_error():
        raise RuntimeError("device barrier timeout during the plain run")
    launch_ms = [s.elapsed_time(e) for s, e in ev]
    out = dict(wall=wall, launch_ms=launch_ms, cycles=a.steps * cps)
    # spot-check the first cycles against the CPU oracle
    if a.verify_cycles:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        v = a.verify_cycles
        got = st[:v + 1].cpu().numpy().view(np.uint32).reshape(-1)
        ref = np.zeros_like(got)
        ref.reshape(v + 1, ss)[:, :nw] = iw[:v + 1]
        oracle.run(script, ref, v)
        out["verified_cycles"] = v if np.array_equal(got, ref) else -1
        if out["verified_cycles"] < 0:
            raise AssertionError("plain GPU run differs from the CPU oracle")
        if not script.sram_storage_size:
            # and the last cycles of the run: the oracle continues from the
            # GPU's snapshot `total - v` (every word of it, including DFF
            # state) with the same inputs
            tail = st[total - v:total + 1].cpu().numpy().view(np.uint32).reshape(-1)
            ref = tail.copy()
            ref.reshape(v + 1, ss)[1:, nw:] = 0
            oracle.run(script, ref, v)
            if not np.array_equal(tail, ref):
                raise AssertionError("plain GPU run differs from the CPU oracle at its end")
        else:
            # SRAM contents at snapshot `total - v` are not kept, so both
            # continue for v more cycles from the final snapshot and SRAM
            sram_end = d_sram.cpu().numpy().view(np.uint32).copy()
            ext = np.zeros((v + 1, ss), np.uint32)
            ext[0] = st[total].cpu().numpy().view(np.uint32)
            ext[1:, :nw] = stimulus.input_words(script, v, seed=a.seed + 1, p=a.toggle)[1:]
            d_ext = torch.from_numpy(ext.reshape(-1).view(np.int32)).to("cuda")
            d_sram2 = d_sram.clone()
            N.check(L.jq_simulate_scan(
                script.num_blocks, script.num_major_stages, d_bs.data_ptr(),
                d_bd.data_ptr(), d_sram2.data_ptr(), d_sx.data_ptr(), v, ss,
                d_ext.data_ptr(), None, None, 0, 0, ctypes.c_void_p(stream.cuda_stream)))
            torch.cuda.synchronize()
            ref = ext.reshape(-1).copy()
            ref_sram = sram_end.copy()
            oracle.run(script, ref, v, sram=ref_sram)
            if not (np.array_equal(d_ext.cpu().numpy().view(np.uint32), ref) and
                    np.array_equal(d_sram2.cpu().numpy().view(np.uint32), ref_sram)):
                raise AssertionError("plain GPU run differs from the CPU oracle at its end")
        out["verified_cycles"] += v
    del st
    torch.cuda.empty_cache()
    return out


def bench_batched(a, comp, torch, cycles=None, split=None, steps=0, learned=None,
                  widths=None, merge=True, seeds_probs=None):
    """Batched form on the same netlist: --lanes lanes with per-lane seeds,
    cycle counts U[10K, 200K] x --batched-scale and toggle rates
    U[0.02, 0.5], in one persistent launch with the device lane queue.
    Groups and blocks per group come from the autotuner unless given; the
    block assignment uses the learned partition costs (trained here on the
    kernel's own partition timings) unless --no-learned (or learned=False).

    steps = 0: the reset and the whole job captured as one hipGraph and
    timed as one replay.  steps = K: the job is timed as K replays of a
    graph of one resumable launch of ceil(1.02 * predicted group cycles / K)
    group cycles each (after --warmup short launches and a reset), plus a
    final launch to the end should the groups need more cycles than
    predicted; the timed region covers the whole job either way."""
    from jacquard_amd import batched as B
    from jacquard_amd import costmodel as CM
    rng = np.random.default_rng(a.seed + 100)
    lo, hi = int(10_000 * a.batched_scale), int(200_000 * a.batched_scale)
    seeds = rng.integers(1, 2**31, a.lanes).astype(np.uint32)
    probs = rng.uniform(0.02, 0.5, a.lanes)
    if seeds_probs is not None:
        seeds, probs = seeds_probs
    if cycles is None:
        cycles = rng.integers(max(lo, 1), max(hi, 1) + 1, a.lanes)
    if learned is None:
        learned = not a.no_learned
    tune = None
    if a.batched_blocks:
        W = a.batched_blocks
        G = a.batched_groups or (a.lanes + 31) // 32
    else:
        t = time.time()
        tune = B.autotune(comp, cycles, log=log, **({"widths": widths} if widths else {}))
        W, G = tune["blocks"], tune["groups"]
        log(f"autotune: chose {G} groups x {W} blocks in {time.time() - t:.1f}s")
    width = None
    if not a.no_width_tune:
        t = time.time()
        comp, width = B.tune_partition_width(
            comp, a.batched_split if split is None else split, W, G, log=log,
            splits=tuple(int(x) for x in a.tune_splits.split(",") if x))
        log(f"width tune: {time.time() - t:.1f}s")
    lrn = None
    costs = None
    if learned:
        # the learned model sets the block assignment and decides the
        # partition merges; measured kernel time picks among the heuristic,
        # the learned assignment and the learned merge + assignment
        costs, model, info = CM.learn_costs(comp, W, G, log=log)
        variants = {"heuristic": None, "learned": costs}
        minfo = None
        if merge:
            merged, mcosts, minfo = CM.learned_merge(comp, model, W, groups=G, log=log)
            if merged is not comp:
                variants["learned+cut/merge"] = (merged, mcosts)
        spread = {}
        ab = CM.compare_assignments(comp, W, G, variants, repeats=3, cycles=32,
                                    spread=spread)
        best = min(ab, key=ab.get)
        lrn = dict(info, merge=minfo, us_per_cycle=ab, spread_us=spread,
                   gain=ab["heuristic"] / ab[best], used=best)
        log("cost model: us/cycle " + ", ".join(f"{k} {v:.1f}" for k, v in ab.items())
            + f"; using {best}")
        if best == "heuristic":
            costs = None  # the heuristic stays the fallback; measured time judges
        elif best == "learned+cut/merge":
            comp, costs = merged, mcosts
    script = comp.flatten_shared(W) if costs is None else comp.flatten(W, costs=costs)
    run = B.BatchedRun(script, seeds, probs, cycles, groups=G)
    script.validate(script.reg_io_state_size)
    run.prepare()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    from jacquard_amd.sim import device_error
    launch_ms, launches, drains, budget = [], 0, 0, 0
    if steps <= 0:
        run.capture_graph()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        run.graph.replay()
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        launch_ms.append(e0.elapsed_time(e1))
        launches = 1
    else:
        pred = max(run.predicted_group_cycles())
        budget = max(1, -(-int(1.02 * pred) // steps))
        graph = run.capture_chunk_graph(budget)
        for _ in range(a.warmup):
            run.launch(budget=8)
        torch.cuda.synchronize()
        run.reset()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps + 1)]
        t0 = time.perf_counter()
        for k in range(steps):
            ev[k][0].record(stream)
            graph.replay()
            ev[k][1].record(stream)
        torch.cuda.synchronize()
        if not run.done():
            ev[steps][0].record(stream)
            run.launch()
            ev[steps][1].record(stream)
            drains = 1
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        launches = steps + drains
        launch_ms = [s_.elapsed_time(e_) for s_, e_ in ev[:launches]]
        log(f"batched: job of {steps} graph replays of {budget} group cycles "
            f"(predicted {pred} group cycles){', plus a drain launch' if drains else ''}")
    if device_error():
        raise RuntimeError("device barrier timeout during the batched run")
    if not run.done():
        raise RuntimeError("batched job did not complete")
    lc = run.lane_cycles_total()
    log(f"batched: {a.lanes} lanes through {G} groups x {W} blocks, "
        f"{lc} lane-cycles in {wall:.2f}s")
    # spot-check the first cycles of the shortest lane and of the last lane
    # the queue handed out against the CPU oracle
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from jacquard_amd import stimulus
    spot = {int(np.argmin(cycles)), int(run.order[-1])}
    outs = run.lane_outputs(lanes=sorted(spot), max_cycles=a.verify_cycles or 4)
    checked = 0
    for lane in spot:
        v = min(a.verify_cycles or 4, int(cycles[lane]))
        st = stimulus.build_states(script, v, seed=int(seeds[lane]), p=float(probs[lane]))
        oracle.run(script, st, v)
        if not np.array_equal(outs[lane][:v], B.plain_observations(script, st, v, run.obs_lits)):
            raise AssertionError(f"batched GPU lane {lane} differs from the CPU oracle")
        checked += v
    return dict(wall=wall, kernel_ms=float(np.sum(launch_ms)), launch_ms=launch_ms,
                launches=launches, drains=drains, budget=budget, lane_cycles=lc,
                slot_cycles=run.slot_cycles_total(), groups=G, blocks=W,
                verified_cycles=checked, script=script, learned=lrn,
                tune=None if tune is None else tune["trials"], width=width)


def bench_xtiming(a, torch):
    """The 4M-gate X + timing run: a seeded random AIG with 8 single-port
    1024x32 SRAMs, a seeded 10% of DFFs starting X, per-gate delays
    U[20, 200] ps, setup/hold 50/20 ps at a 5000 ps clock.

    All (cycles+1) snapshots are resident in HBM.  The run is two launches,
    cycles [0, C-V) and [C-V, C), timed together; the first V and the last V
    cycles are checked against the oracle: every snapshot word (values, X
    masks, arrivals), the SRAM contents and the canonical violation list."""
    import ctypes
    from jacquard_amd import _native as N
    from jacquard_amd import sim, stimulus
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    L = N.hip()
    t = time.time()
    nl = Netlist.random(seed=a.seed, gates=a.gates, depth=a.depth, dffs=a.dffs,
                        inputs=a.inputs, outputs=a.outputs, geometric_p=0.3,
                        invert_prob=0.5, srams=8, sram_addr_bits=10, delay=(20, 200))
    comp = Compiled(nl, level_split=a.split)
    res = ctypes.c_int(0)
    N.check(L.jq_max_resident_blocks(0, ctypes.byref(res)))
    blocks = a.blocks or min(max(comp.parts_per_stage()), res.value)
    s = comp.flatten(blocks, xprop=True)
    s.inject_timing()
    s.enable_timing_arrivals()
    cons = s.build_timing_constraint_buffer(5000, 50, 20)
    xm, _ = stimulus.dff_initial_xmask(s, 0.1, seed=a.seed + 7)
    ss, rio = s.effective_state_size(), s.reg_io_state_size
    s.validate(ss)
    log(f"xtiming: {nl.counts()}; {s.num_blocks} blocks, {s.num_major_stages} major "
        f"stages, {s.num_parts} partitions, script {s.blocks_data.nbytes / 1e6:.1f} MB, "
        f"snapshot {ss} words; built in {time.time() - t:.1f}s")
    C, V = a.cycles, a.verify_window
    if not 0 < V < C:
        raise ValueError("need 0 < --verify-window < --cycles")
    t = time.time()
    iw = stimulus.input_words(s, C, seed=a.seed, p=a.toggle)
    nw = iw.shape[1]
    tmpl = s.xmask_template()
    st = torch.zeros((C + 1, ss), dtype=torch.int32, device="cuda")
    st[:, :nw] = torch.from_numpy(iw.view(np.int32)).to("cuda")
    st[:, rio:2 * rio] = torch.from_numpy(tmpl.view(np.int32)).to("cuda")[None, :]
    st[0, rio:2 * rio] = torch.from_numpy(xm.view(np.int32)).to("cuda")
    nsram = max(s.sram_storage_size, 1)
    d_sram = torch.zeros(nsram, dtype=torch.int32, device="cuda")
    d_sx = torch.full((nsram,), -1, dtype=torch.int32, device="cuda")
    d_bs = sim.to_device(np.ascontiguousarray(s.blocks_start, np.uint64))
    d_bd = sim.to_device(np.ascontiguousarray(s.blocks_data, np.uint32))
    d_cons = sim.to_device(cons)
    cap_a, cap_b = 1024 * (C - V), 1024 * V
    ev_a = torch.zeros(16 + 32 * cap_a, dtype=torch.uint8, device="cuda")
    ev_b = torch.zeros(16 + 32 * cap_b, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    log(f"xtiming: {C + 1} snapshots ({st.numel() * 4 / 1e9:.1f} GB) staged in "
        f"{time.time() - t:.1f}s")
    stream = torch.cuda.current_stream()

    def launch(c0, n, ev, cap):
        N.check(L.jq_simulate_scan(
            s.num_blocks, s.num_major_stages, d_bs.data_ptr(), d_bd.data_ptr(),
            d_sram.data_ptr(), d_sx.data_ptr(), n, ss, st.data_ptr() + 4 * c0 * ss,
            d_cons.data_ptr(), ev.data_ptr(), cap, int(s.arrival_state_offset),
            ctypes.c_void_p(stream.cuda_stream)))

    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    t0 = time.perf_counter()
    e[0].record(stream)
    launch(0, C - V, ev_a, cap_a)
    e[1].record(stream)
    sram_mid, sx_mid = d_sram.clone(), d_sx.clone()
    e[2].record(stream)
    launch(C - V, V, ev_b, cap_b)
    e[3].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if sim.device_error():
        raise RuntimeError("device barrier timeout during the xtiming run")
    kern_s = (e[0].elapsed_time(e[1]) + e[2].elapsed_time(e[3])) / 1e3

    def events_of(ev, cap, below=None):
        hdr = ev[:16].view(torch.int32).cpu().numpy()
        cnt, ovf = int(hdr[0]), int(hdr[1])
        rec = ev[16:16 + 32 * min(cnt, cap)].view(torch.int32).view(-1, 8)
        if below is not None:
            rec = rec[rec[:, 2] < below]
        out = rec.cpu().numpy().view(np.uint32).view(sim.EVENT_DTYPE).reshape(-1)
        return cnt, ovf, sim.canonical_events(out)

    cnt_a, ovf_a, ev_first = events_of(ev_a, cap_a, below=V)
    cnt_b, ovf_b, ev_last = events_of(ev_b, cap_b)
    if ovf_a or ovf_b:
        raise RuntimeError("event buffer overflow")
    # first window: from snapshot 0
    t = time.perf_counter()
    ref = stimulus.build_states(s, V, seed=a.seed, p=a.toggle, initial_xmask=xm)
    sram0 = np.zeros(nsram, np.uint32)
    sx0 = np.full(nsram, 0xFFFFFFFF, np.uint32)
    ev_ref, n_ref = oracle.run(s, ref, V, sram=sram0, sram_x=sx0, constraints=cons,
                               arrival_state_offset=s.arrival_state_offset,
                               event_capacity=1024 * V)
    cpu_s = time.perf_counter() - t
    got = st[:V + 1].cpu().numpy().view(np.uint32).reshape(-1)
    ok_first = (np.array_equal(got, ref) and len(ev_first) == n_ref
                and np.array_equal(ev_first, sim.canonical_events(ev_ref[:n_ref])))
    # last window: from the GPU's snapshot C-V and SRAM contents at that cycle
    ref = np.zeros((V + 1, ss), np.uint32)
    ref[0] = st[C - V].cpu().numpy().view(np.uint32)
    ref[1:, :nw] = iw[C - V + 1:]
    ref[1:, rio:2 * rio] = tmpl
    ref = ref.reshape(-1)
    sram1 = sram_mid.cpu().numpy().view(np.uint32).copy()
    sx1 = sx_mid.cpu().numpy().view(np.uint32).copy()
    ev_ref, n_ref = oracle.run(s, ref, V, sram=sram1, sram_x=sx1, constraints=cons,
                               arrival_state_offset=s.arrival_state_offset,
                               event_capacity=1024 * V)
    got = st[C - V:].cpu().numpy().view(np.uint32).reshape(-1)
    ok_last = (np.array_equal(got, ref)
               and np.array_equal(d_sram.cpu().numpy().view(np.uint32), sram1)
               and np.array_equal(d_sx.cpu().numpy().view(np.uint32), sx1)
               and cnt_b == n_ref
               and np.array_equal(ev_last, sim.canonical_events(ev_ref[:n_ref])))
    log(f"xtiming: {C} cycles in {kern_s:.2f}s kernel time; {cnt_a + cnt_b} violations; "
        f"oracle windows first={ok_first} last={ok_last}")
    if not (ok_first and ok_last):
        raise AssertionError("xtiming GPU run differs from the CPU oracle")
    bpc = s.algorithmic_bytes_per_cycle()
    return dict(script=s, cycles=C, wall=wall, kernel_s=kern_s, violations=cnt_a + cnt_b,
                bytes_per_cycle=bpc, cpu_cycles_per_s=V / cpu_s, cpu_cycles=V,
                cpu_s=cpu_s, verified_cycles=2 * V, counts=nl.counts())


def host_cpu():
    """(CPU model, online cores, cores this process may use) of the host."""
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def netlist_4m(a):
    from jacquard_amd.netlist import Netlist
    return Netlist.random(seed=a.seed, gates=a.gates, depth=a.depth, dffs=a.dffs,
                          inputs=a.inputs, outputs=a.outputs, geometric_p=0.3,
                          invert_prob=0.5, srams=8, sram_addr_bits=10, delay=(20, 200))


def bench_north(a, torch):
    """The headline: batched+learned lanes against the plain drop-in on the
    4M-gate netlist of the X/timing workload (same generator: 8 single-port
    1024x32 SRAMs; value lanes, 2-state), with the batched configuration's
    lanes (--lanes lanes, cycles U[10K, 200K] x --batched-scale, toggle
    U[0.02, 0.5]).  `value` is the batched job's lane-cycles over the wall
    time of the whole job, run as --steps graph replays (after --warmup
    short launches).  Beside it: the plain drop-in on the same netlist
    (--plain-launches x --cycles-per-step cycles), and under config.forms
    the 1M-gate plain drop-in (100K cycles), the 256 batched lanes of the
    1M-gate netlist (one graph replay) and the structured 16M-gate
    netlist's 64 lanes x 20K cycles with learned costs, unless --no-forms
    (--no-16m drops the last)."""
    from jacquard_amd.script import Compiled
    t = time.time()
    nl = netlist_4m(a)
    log(f"north: netlist {nl.counts()} generated in {time.time() - t:.1f}s")
    # the plain drop-in at its measured best width (sim.tune_plain): the
    # caps whose widest major stage fits the co-resident workgroups (at
    # level split 6 on this netlist only the compiled default), then every
    # partition held to one boomerang stage
    from jacquard_amd import sim
    _comp, script, plain_tune = sim.tune_plain(nl, a.split, caps=(2600, 1800), log=log)
    pa = argparse.Namespace(**vars(a))
    pa.steps, pa.warmup = a.plain_launches, 1
    p = bench_plain(pa, script, torch)
    plain = p["cycles"] / p["wall"]
    log(f"north: plain drop-in {plain:,.0f} cycles/s on the 4M-gate netlist")
    comp_b = Compiled(nl, level_split=a.batched_split)
    b = bench_batched(a, comp_b, torch, steps=a.steps, widths=(32, 48, 64))
    value = b["lane_cycles"] / b["wall"]
    log(f"north: batched {value:,.0f} lane-cycles/s, {value / plain:.1f}x the plain drop-in")
    bh = batched_hbm(b, BATCHED_4M_TRAFFIC_PROFILE)
    forms = {}
    if not a.no_forms:
        forms = forms_1m(a, torch)
        if not a.no_16m:
            forms.update(forms_16m(a, torch))
    cpu = None
    if not a.no_cpu:
        model, online, usable = host_cpu()
        c = cpu_baseline(a, script)
        cpu = {"value": c["value"], "unit": "lane-cycles/s", "cores": 1, "kind": "port",
               "sample": f"{c['cycles']} cycles of the 4M-gate netlist (2-state) through the "
                         f"oracle's literal restatement of sanity_check_cpu (gcc -O2, one "
                         f"thread) in {c['seconds']:.1f} s on {model}; {online} cores online, "
                         f"{usable} usable by this process",
               "cpu_model": model, "nproc": online, "usable_cores": usable}
        log(f"cpu: {c['value']:.2f} lane-cycles/s over {c['cycles']} cycles")
        # labelled apart from the one-thread port: the same restatement with
        # each major stage's blocks spread over the host's cores, and the
        # direct AIG cycle simulator on the same netlist
        cpu["variants"] = v = cpu_variants(a, script)
        log(f"cpu: OpenMP {v['port_openmp']['lane_cycles_per_s']:.2f} "
            f"({v['port_openmp']['threads']} threads), direct AIG "
            f"{v['direct_aig']['lane_cycles_per_s']:.2f} lane-cycles/s")
    bs = b["script"]
    line = {
        "metric": METRIC, "value": value, "unit": "lane-cycles/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": b["wall"] / a.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {
            "workload": f"batched+learned lanes vs plain drop-in: seeded random AIG, "
                        f"{a.gates} AND, depth {a.depth}, {a.dffs} DFF, 8 x 1024x32 SRAM, "
                        f"{a.inputs} PI, {a.outputs} PO, geometric(0.3) fan-in; {a.lanes} "
                        f"lanes, cycles U[{int(10_000 * a.batched_scale)}, "
                        f"{int(200_000 * a.batched_scale)}], toggle U[0.02, 0.5]",
            "batched": {"lanes": a.lanes, "groups": b["groups"],
                        "blocks_per_group": b["blocks"], "lane_cycles": b["lane_cycles"],
                        "slot_utilisation": b["lane_cycles"] / b["slot_cycles"],
                        "learned_cost_model": b["learned"], "autotune": b["tune"],
                        "partition_width": b["width"],
                        "launch": f"{b['launches']} launches: {a.steps} replays of one "
                                  f"hipGraph of a resumable persistent launch of "
                                  f"{b['budget']} group cycles"
                                  + (", then one launch to the end" if b["drains"] else ""),
                        "kernel_s": b["kernel_ms"] / 1e3,
                        "major_stages": bs.num_major_stages, "partitions": bs.num_parts,
                        "script_bytes": int(bs.blocks_data.nbytes),
                        "verified_cycles": b["verified_cycles"]},
            "plain": {"lane_cycles_per_s": plain, "blocks": script.num_blocks,
                      "partition_tune": plain_tune,
                      "cycles": p["cycles"], "major_stages": script.num_major_stages,
                      "partitions": script.num_parts,
                      "script_bytes": int(script.blocks_data.nbytes),
                      "verified_cycles": p.get("verified_cycles")},
            "vs_plain": value / plain,
            "forms": forms},
        "roofline": {"bound": "hbm", "achieved": bh["achieved"], "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": bh["frac"], "traffic": bh["traffic"],
                     "kernel": "batched_kernel",
                     "algorithmic_bytes_per_lane_cycle": bh["algorithmic_bytes_per_lane_cycle"],
                     "avg_launch_ms": b["kernel_ms"] / b["launches"],
                     "measured_gbps": bh["traffic_gbps"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def forms_1m(a, torch):
    """The 1M-gate workloads reported beside the headline: the plain
    drop-in over 100K cycles (10 launches of 10K cycles; partition width
    and blocks from sim.tune_plain) and the 256 batched lanes of the same
    netlist (autotuned, tuned width, heuristic assignment, one graph
    replay)."""
    from jacquard_amd import sim
    f = argparse.Namespace(**vars(a))
    f.gates, f.depth, f.dffs = 1_000_000, 64, 65_536
    f.steps, f.warmup, f.cycles_per_step = 10, 1, 10_000
    nl, comp = build_design(f)
    _comp, script, tune_table = sim.tune_plain(nl, f.split, log=log)
    p = bench_plain(f, script, torch)
    plain = p["cycles"] / p["wall"]
    avg_launch_s = float(np.mean(p["launch_ms"])) / 1e3
    bpc = script.algorithmic_bytes_per_cycle()
    out = {"plain_1m": {
        "workload": "plain GPU drop-in: seeded random AIG, 1000000 AND, depth 64, 65536 "
                    "DFF, 512 PI, 256 PO, geometric(0.3) fan-in; 100000 cycles, toggle 0.1",
        "lane_cycles_per_s": plain, "blocks": script.num_blocks,
        "major_stages": script.num_major_stages, "partitions": script.num_parts,
        "plain_partition_tune": tune_table, "verified_cycles": p.get("verified_cycles"),
        "roofline": {"kernel": "scan_kernel", "achieved": bpc * f.cycles_per_step /
                     avg_launch_s / 1e9, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "algorithmic_bytes_per_cycle": bpc, "avg_launch_ms": avg_launch_s * 1e3}}}
    out["plain_1m"]["roofline"]["frac"] = out["plain_1m"]["roofline"]["achieved"] / HBM_PEAK_GBPS
    out["plain_1m"]["sync_ab"] = ab = plain_sync_ab(script, torch)
    log(f"forms: 1M plain {plain:,.0f} cycles/s; sync A/B over {ab['cycles']} cycles: "
        f"dataflow {ab['dataflow_cycles_per_s']:,.0f}, barrier {ab['barrier_cycles_per_s']:,.0f}")
    from jacquard_amd.script import Compiled
    comp_b = Compiled(nl, level_split=a.batched_split)
    b = bench_batched(f, comp_b, torch, steps=0, learned=False)
    bh = batched_hbm(b)
    out["batched_1m"] = {
        "workload": "256 batched lanes of the 1M-gate netlist, cycles U[10K, 200K], "
                    "toggle U[0.02, 0.5]; one hipGraph replay of the reset and the "
                    "persistent launch",
        "lane_cycles_per_s": b["lane_cycles"] / b["wall"], "lanes": f.lanes,
        "groups": b["groups"], "blocks_per_group": b["blocks"],
        "lane_cycles": b["lane_cycles"], "partition_width": b["width"],
        "slot_utilisation": b["lane_cycles"] / b["slot_cycles"],
        "vs_plain_1m": b["lane_cycles"] / b["wall"] / plain, "hbm": bh,
        "verified_cycles": b["verified_cycles"]}
    log(f"forms: 1M batched {b['lane_cycles'] / b['wall']:,.0f} lane-cycles/s")
    return out


def bench_mult16m(a, torch):
    """The structured 16M-gate netlist: 64 replicated 16x16 AND-array
    multiplier tiles with pipeline DFFs plus a 2M-gate random control AIG,
    depth 128, 1M DFFs, at --split levels per major stage (5 is this
    netlist's fastest for the batched form).  --lanes lanes (64) of --cycles cycles
    (20K) run batched with the block assignment from the learned cost model
    (trained on >=50K samples of the kernel's own partition timings) and,
    for the makespan comparison, with the heuristic one; blocks per group
    default to the co-resident maximum for the lane groups (no width sweep:
    each candidate's compile and flattening of this script take about 6 s
    on 8 host cores).  The plain
    drop-in runs --steps x --cycles-per-step cycles beside it."""
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    from jacquard_amd import batched as B
    import ctypes
    from jacquard_amd import _native as N
    t = time.time()
    nl = Netlist.multiplier(seed=a.seed)
    t_gen = time.time() - t
    t = time.time()
    comp = Compiled(nl, level_split=a.split)
    t_comp = time.time() - t
    log(f"mult16m: netlist {nl.counts()} generated in {t_gen:.1f}s, compiled in "
        f"{t_comp:.1f}s; " + comp.log.strip().splitlines()[-1])
    groups = a.batched_groups or (a.lanes + 31) // 32
    if not a.batched_blocks:
        a.batched_blocks = max(8, min(max(comp.parts_per_stage()),
                                      B.max_resident_batched() // groups))
    a.batched_groups = groups
    cycles = np.full(a.lanes, a.cycles)
    b = bench_batched(a, comp, torch, cycles=cycles, split=a.split)
    value = b["lane_cycles"] / b["wall"]
    res = ctypes.c_int(0)
    N.check(N.hip().jq_max_resident_blocks(0, ctypes.byref(res)))
    script = comp.flatten(min(max(comp.parts_per_stage()), res.value))
    p = bench_plain(a, script, torch)
    plain = p["cycles"] / p["wall"]
    bs = b["script"]
    W_bytes = 4 * int(bs.blocks_start[-1])
    per_lane = W_bytes / 32 + (bs.algorithmic_bytes_per_cycle() - W_bytes)
    achieved = per_lane * b["lane_cycles"] / (b["kernel_ms"] / 1e3) / 1e9
    print(json.dumps({
        "metric": METRIC, "value": value, "unit": "lane-cycles/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": b["wall"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {
            "workload": f"batched+learned on the structured netlist: "
                        f"{nl.counts()['num_and']} AND (64 16x16 multiplier tiles + "
                        f"2M-gate control AIG), {nl.counts()['num_dffs']} DFF; "
                        f"{a.lanes} lanes x {a.cycles} cycles",
            "host_s": {"generate": t_gen, "compile": t_comp},
            "batched": {"lanes": a.lanes, "groups": b["groups"],
                        "blocks_per_group": b["blocks"], "lane_cycles": b["lane_cycles"],
                        "learned_cost_model": b["learned"],
                        "launch": "one hipGraph replay (reset + persistent kernel)",
                        "verified_cycles": b["verified_cycles"]},
            "plain": {"lane_cycles_per_s": plain, "blocks": script.num_blocks,
                      "verified_cycles": p.get("verified_cycles")},
            "vs_plain": value / plain,
            "major_stages": bs.num_major_stages, "partitions": bs.num_parts,
            "script_bytes": int(bs.blocks_data.nbytes)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "batched_kernel",
                     "algorithmic_bytes_per_lane_cycle": per_lane},
        "cpu_baseline": None,
    }), flush=True)


def forms_16m(a, torch):
    """The structured 16M-gate netlist reported beside the headline: 64 lanes
    x 20,000 cycles through 2 groups x 128 blocks at level split 5 with the
    learned block
    assignment (A/B against the heuristic; no merge and no tiling sweep,
    whose compile and flattening take about 6 s per candidate) in one
    graph replay, and the plain drop-in (compiled at level split 4) over
    2 x 1,000 cycles."""
    import ctypes
    from jacquard_amd import _native as N
    from jacquard_amd import batched as B
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    f = argparse.Namespace(**vars(a))
    f.lanes, f.batched_groups, f.no_width_tune = 64, 2, True
    f.steps, f.warmup, f.cycles_per_step = 2, 1, 1000
    t = time.time()
    nl = Netlist.multiplier(seed=a.seed)
    # 5 levels per major stage: 559 us per group cycle at 2 x 128 blocks
    # against 593 at split 4, 655 at 3, 664 at 6, 740 at 2 and 844 at 8
    # (24-cycle trials, heuristic assignment, one box)
    comp = Compiled(nl, level_split=5)
    t_host = time.time() - t
    log(f"forms: 16M netlist {nl.counts()} generated and compiled in {t_host:.1f}s")
    f.batched_blocks = max(8, min(max(comp.parts_per_stage()),
                                  B.max_resident_batched() // f.batched_groups))
    rng = np.random.default_rng(a.seed + 200)
    sp = (rng.integers(1, 2**31, 64).astype(np.uint32), rng.uniform(0.02, 0.5, 64))
    b = bench_batched(f, comp, torch, cycles=np.full(64, 20_000), split=5, steps=0,
                      learned=True, merge=False, seeds_probs=sp)
    value = b["lane_cycles"] / b["wall"]
    res = ctypes.c_int(0)
    N.check(N.hip().jq_max_resident_blocks(0, ctypes.byref(res)))
    # the plain drop-in at its own best level split: 4 levels per major
    # stage (32 stages) ran 2,870 cycles/s against 2,560 at split 8 and
    # 910 at split 12
    comp_p = Compiled(nl, level_split=4)
    script = comp_p.flatten(min(max(comp_p.parts_per_stage()), res.value))
    p = bench_plain(f, script, torch)
    plain = p["cycles"] / p["wall"]
    log(f"forms: 16M batched {value:,.0f} lane-cycles/s, plain {plain:,.0f} cycles/s")
    return {"batched_16m": {
        "workload": f"structured netlist: {nl.counts()['num_and']} AND (64 16x16 "
                    f"multiplier tiles with pipeline DFFs + 2M-gate control AIG), "
                    f"{nl.counts()['num_dffs']} DFF, level split 5; 64 lanes x 20000 "
                    f"cycles, toggle U[0.02, 0.5]; one hipGraph replay",
        "lane_cycles_per_s": value, "groups": b["groups"], "blocks_per_group": b["blocks"],
        "lane_cycles": b["lane_cycles"], "learned_cost_model": b["learned"],
        "partitions": b["script"].num_parts, "script_bytes": int(b["script"].blocks_data.nbytes),
        "host_generate_compile_s": t_host, "verified_cycles": b["verified_cycles"],
        "plain_lane_cycles_per_s": plain, "plain_blocks": script.num_blocks,
        "plain_level_split": 4, "plain_major_stages": script.num_major_stages,
        "plain_verified_cycles": p.get("verified_cycles"), "vs_plain": value / plain}}


def bench_cosim(a, torch, comp=None, emit=True):
    """Reactive co-simulation edges per second.  (1) The SPI master with the
    flash model: 20,000 edges in batches of 1024, the first 400 edges checked
    against the CPU edge loop and the two bytes it reads against the image.
    (2) The default 1M-gate netlist clocked edge by edge (one clock domain,
    no flash): 400 edges, the first 4 checked against the CPU loop.  Each edge
    is one between-edges launch and one scan-kernel cycle."""
    import numpy as np
    from jacquard_amd import cosim
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cosim_oracle as co
    res = {}
    # (1) SPI master + flash
    address = 0x000123
    rng = np.random.default_rng(a.seed)
    image = bytes(rng.integers(0, 256, 1 << 16, dtype=np.uint8))
    nl = Netlist.from_verilog(text=cosim.spi_master_verilog(address))
    s = Compiled(nl).flatten(1)
    ins, outs = nl.ports()
    inbit = {n: s.input_map[p] for n, p, _ in ins}
    outbit = {n: s.output_map[lit] for n, lit in outs if lit > 1}
    pins = cosim.FlashPins(clk_out=outbit["flash_clk"], csn_out=outbit["flash_csn"],
                           d_out=(outbit["flash_d0"], cosim.NC, cosim.NC, cosim.NC),
                           d_in=(cosim.NC, inbit["miso"], cosim.NC, cosim.NC))
    doms = [d for _, d in cosim.clock_domains_of(s)]
    sched = cosim.MultiClockScheduler([cosim.ClockTiming(5000)])
    reset, check, edges = 8, 400, 20_000

    def make():
        return cosim.Cosim(s, sched, doms, reset_pos=inbit["rst"], reset_active=1,
                           reset_edges=reset, flash_pins=pins, flash_image=image)
    c = make()
    got = c.run(check, batch=1024, trace=True)
    st = np.zeros(2 * s.effective_state_size(), np.uint32)
    fl = co.Flash(image)
    tr = []
    for in_reset, n, off in ((1, reset, 0), (0, check - reset, reset)):
        ops, start = [], [0]
        for k in range(len(sched)):
            ops += sched.edge_ops(k, doms, inbit["rst"], in_reset)
            start.append(len(ops))
        fl.in_reset = in_reset
        tr.append(co.run_edges(s, st, np.asarray(ops, np.uint32).reshape(-1, 2),
                               np.asarray(start, np.uint32), off % len(sched), n, fl,
                               pins.words()))
    if not np.array_equal(got, np.concatenate(tr)):
        raise AssertionError("co-simulation edges differ from the CPU edge loop")
    ss = s.effective_state_size()
    rx = sum(co._bit(st[ss:], outbit[f"rx_o{i}"]) << i for i in range(16))
    if rx != (image[address] << 8 | image[address + 1]):
        raise AssertionError("the SPI master read the wrong bytes")
    c = make()
    c.run(64, batch=64)  # warm-up
    c = make()
    torch.cuda.synchronize()
    t = time.perf_counter()
    c.run(edges, batch=1024)
    wall = time.perf_counter() - t
    t = time.perf_counter()
    fl_cpu = co.Flash(image)
    st = np.zeros(2 * ss, np.uint32)
    ops, start = [], [0]
    for k in range(len(sched)):
        ops += sched.edge_ops(k, doms, inbit["rst"], 0)
        start.append(len(ops))
    fl_cpu.in_reset = 0
    n_cpu = 2000
    co.run_edges(s, st, np.asarray(ops, np.uint32).reshape(-1, 2),
                 np.asarray(start, np.uint32), 0, n_cpu, fl_cpu, pins.words())
    cpu = n_cpu / (time.perf_counter() - t)
    res["spi_flash"] = {"edges": edges, "edges_per_s": edges / wall,
                        "us_per_edge": wall / edges * 1e6, "verified_edges": check,
                        "gates": nl.counts()["num_and"], "state_words": ss,
                        "cpu_edges_per_s": cpu, "cpu_edges": n_cpu}
    log(f"cosim spi+flash: {edges / wall:,.0f} edges/s ({wall / edges * 1e6:.1f} us per edge), "
        f"CPU loop {cpu:,.0f} edges/s")
    # (2) the 1M-gate netlist clocked edge by edge
    import ctypes
    from jacquard_amd import _native as N
    if comp is None:
        _, comp = build_design(a)
    r = ctypes.c_int(0)
    N.check(N.hip().jq_max_resident_blocks(0, ctypes.byref(r)))
    s2 = comp.flatten(min(max(comp.parts_per_stage()), r.value))
    doms2 = [d for _, d in cosim.clock_domains_of(s2)]
    c2 = cosim.Cosim(s2, sched, doms2)
    got = c2.run(4, batch=4, trace=True)
    ops, start = [], [0]
    for k in range(len(sched)):
        ops += sched.edge_ops(k, doms2)
        start.append(len(ops))
    st2 = np.zeros(2 * s2.effective_state_size(), np.uint32)
    t = time.perf_counter()
    exp = co.run_edges(s2, st2, np.asarray(ops, np.uint32).reshape(-1, 2),
                       np.asarray(start, np.uint32), 0, 4)
    cpu2 = 4 / (time.perf_counter() - t)
    if not np.array_equal(got, exp):
        raise AssertionError("1M-gate co-simulation edges differ from the CPU edge loop")
    e2 = 400
    torch.cuda.synchronize()
    t = time.perf_counter()
    c2.run(e2, batch=e2)
    wall2 = time.perf_counter() - t
    res["netlist_1m"] = {"edges": e2, "edges_per_s": e2 / wall2,
                         "us_per_edge": wall2 / e2 * 1e6, "verified_edges": 4,
                         "blocks": s2.num_blocks, "major_stages": s2.num_major_stages,
                         "cpu_edges_per_s": cpu2, "cpu_edges": 4}
    log(f"cosim 1M gates: {e2 / wall2:,.0f} edges/s ({wall2 / e2 * 1e6:.1f} us per edge)")
    if not emit:
        return res
    print(json.dumps({
        "metric": "co-simulation edges per second", "value": edges / wall,
        "unit": "edges/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"reactive co-simulation: gate-level SPI master "
                               f"({nl.counts()['num_and']} AND, 23 DFF) reading a 64 KiB "
                               f"flash image through the device SPI flash model, one clock "
                               f"domain, {edges} edges in batches of 1024; and the 1M-gate "
                               f"random AIG clocked edge by edge",
                   "forms": res},
        "cpu_baseline": {"value": cpu, "unit": "edges/s", "cores": 1, "kind": "port",
                         "sample": f"the CPU edge loop (oracle/cosim_oracle.py over the "
                                   f"C interpreter) for {n_cpu} edges of the SPI master"},
    }), flush=True)


def cpu_baseline(a, script):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from jacquard_amd import stimulus
    ss = script.reg_io_state_size
    n = 1
    spent = 0.0
    while True:
        st = stimulus.build_states(script, n, seed=a.seed, p=a.toggle)
        t = time.perf_counter()
        oracle.run_v1(script, st, n)
        spent = time.perf_counter() - t
        if spent >= a.cpu_seconds / 4 or n >= 4096:
            break
        n = min(4096, max(n * 2, int(n * (a.cpu_seconds / max(spent, 1e-3)))))
    return dict(value=n / spent, cycles=n, seconds=spent, state_words=ss)


def cpu_variants(a, script):
    """The two CPU figures reported beside the one-thread port: the same
    restatement with the blocks of each major stage spread over the host's
    cores (OpenMP), and the direct AIG cycle simulator (no script) on the
    same netlist.  Each on a bounded cycle sample of about a third of
    --cpu-seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from jacquard_amd import stimulus
    threads = max(1, min(16, os.cpu_count() or 1))
    budget = a.cpu_seconds / 3
    out = {}

    def bounded(run):
        n, spent = 8, 0.0
        while True:
            spent = run(n)
            if spent >= budget / 4 or n >= 4096:
                return n, spent
            n = min(4096, max(n * 2, int(n * (budget / max(spent, 1e-3)))))

    def omp(n):
        st = stimulus.build_states(script, n, seed=a.seed, p=a.toggle)
        t = time.perf_counter()
        oracle.run_v1(script, st, n, threads=threads)
        return time.perf_counter() - t

    n, s = bounded(omp)
    out["port_openmp"] = {"lane_cycles_per_s": n / s, "threads": threads, "cycles": n,
                          "seconds": s}
    A = script.netlist.arrays()
    data, fl = stimulus.input_roles(script)
    pos2pin = {v: k for k, v in script.input_map.items() if A["drv"][k] in (2, 3)}
    pins = [pos2pin[p] for p in data.tolist()] + [pos2pin[p] for p in fl.tolist()]

    def direct(n):
        vals = stimulus.lane_inputs(len(data), n, a.seed, a.toggle)
        bits = np.concatenate([vals[:n].astype(np.uint8),
                               np.ones((n, len(fl)), np.uint8)], axis=1)
        t = time.perf_counter()
        oracle.aig_sim(A, n, pins, bits, A["pos"])
        return time.perf_counter() - t

    n, s = bounded(direct)
    out["direct_aig"] = {"lane_cycles_per_s": n / s, "threads": 1, "cycles": n,
                         "seconds": s}
    return out


def lds_peak_gbps(torch):
    """Measured LDS read peak: the probe kernel (8 workgroups per CU, each
    streaming 64 KB of LDS with conflict-free 16-byte reads), best of 5
    launches timed with HIP events on the launch stream."""
    import ctypes
    from jacquard_amd import _native as N
    L = N.hip()
    blocks = torch.cuda.get_device_properties(0).multi_processor_count * 8
    iters = 2048
    sink = torch.empty(blocks * 256, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    N.check(L.jq_lds_read_probe(blocks, 64, sink.data_ptr(), sp))
    best = None
    for _ in range(5):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        N.check(L.jq_lds_read_probe(blocks, iters, sink.data_ptr(), sp))
        e.record(stream)
        e.synchronize()
        ms = s.elapsed_time(e)
        best = ms if best is None else min(best, ms)
    return blocks * 256 * 128 * iters / (best / 1e3) / 1e9


def barrier_us_per_stage(script, torch, cycles=2000):
    """Time of one major stage of the scan kernel when every block of the
    script is empty (a zero header): the grid barrier's latency at this
    block count, the limiter that is neither HBM nor LDS."""
    import ctypes
    from jacquard_amd import _native as N
    from jacquard_amd import sim
    L = N.hip()
    B, M = script.num_blocks, script.num_major_stages
    bs = np.arange(B * M + 1, dtype=np.uint64) * 256
    d_bs = sim.to_device(bs)
    d_bd = torch.zeros(B * M * 256 + 2048, dtype=torch.int32, device="cuda")
    st = torch.zeros(cycles + 1, dtype=torch.int32, device="cuda")
    one = torch.zeros(1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    def launch():
        N.check(L.jq_simulate_scan(B, M, d_bs.data_ptr(), d_bd.data_ptr(),
                                   one.data_ptr(), one.data_ptr(), cycles, 1,
                                   st.data_ptr(), None, None, 0, 0,
                                   ctypes.c_void_p(stream.cuda_stream)))

    # value-only launches of an empty script would wait for nobody under
    # dataflow synchronisation: time the grid barrier itself
    prev = L.jq_scan_sync_mode(0)
    try:
        launch()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(stream)
        for _ in range(3):
            launch()
        e.record(stream)
        e.synchronize()
    finally:
        L.jq_scan_sync_mode(prev)
    if sim.device_error():
        raise RuntimeError("device barrier timeout during the barrier measurement")
    return s.elapsed_time(e) / 3 / (cycles * M) * 1e3


def plain_sync_ab(script, torch, cycles=4000, seed=3, reps=2):
    """cycles/s of the plain drop-in under dataflow synchronisation (the
    default) and under grid barriers, alternating launches of the same
    stimulus; the two launches' snapshots must be bit-identical."""
    import ctypes
    from jacquard_amd import _native as N
    from jacquard_amd import sim, stimulus
    L = N.hip()
    ss = script.effective_state_size()
    st0 = sim.to_device(stimulus.build_states(script, cycles, seed=seed, p=0.1))
    d_bs = sim.to_device(np.ascontiguousarray(script.blocks_start, np.uint64))
    d_bd = sim.to_device(np.ascontiguousarray(script.blocks_data, np.uint32))
    stream = torch.cuda.current_stream()
    rates = {1: [], 0: []}
    finals = {}
    prev = L.jq_scan_sync_mode(-1)
    try:
        for _ in range(reps):
            for mode in (1, 0):
                L.jq_scan_sync_mode(mode)
                st = st0.clone()
                sram = torch.zeros(max(script.sram_storage_size, 1), dtype=torch.int32,
                                   device="cuda")
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                N.check(L.jq_simulate_scan(
                    script.num_blocks, script.num_major_stages, d_bs.data_ptr(),
                    d_bd.data_ptr(), sram.data_ptr(), sram.data_ptr(), cycles, ss,
                    st.data_ptr(), None, None, 0, 0, ctypes.c_void_p(stream.cuda_stream)))
                e1.record(stream)
                e1.synchronize()
                rates[mode].append(cycles / (e0.elapsed_time(e1) / 1e3))
                finals[mode] = st
    finally:
        L.jq_scan_sync_mode(prev)
    if sim.device_error():
        raise RuntimeError("device wait timed out during the sync A/B")
    same = bool(torch.equal(finals[0], finals[1]))
    if not same:
        raise RuntimeError("dataflow and barrier schedules disagree")
    return {"dataflow_cycles_per_s": float(np.mean(rates[1])),
            "barrier_cycles_per_s": float(np.mean(rates[0])),
            "cycles": cycles, "launches_each": reps, "identical": same}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "scan_kernel_1M_gates_hbm_traffic.json")
BATCHED_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "batched_kernel_1M_gates_hbm_traffic.json")
BATCHED_4M_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles",
                                          "batched_kernel_4M_gates_hbm_traffic.json")


def batched_hbm(b, profile=None):
    """HBM figures of the batched launch: algorithmic bytes per lane-cycle
    (a group cycle streams the script once for its 32 lanes, plus every
    lane's state traffic) over the kernel time, and the measured bytes per
    lane-cycle of the committed FETCH_SIZE / WRITE_SIZE profile of this
    workload at the same rate, when that profile is present."""
    bs = b["script"]
    w = 4 * int(bs.blocks_start[-1])
    per_lane = w / 32 + (bs.algorithmic_bytes_per_cycle() - w)
    rate = b["lane_cycles"] / (b["kernel_ms"] / 1e3)
    out = {"algorithmic_bytes_per_lane_cycle": per_lane, "achieved": per_lane * rate / 1e9,
           "peak": HBM_PEAK_GBPS, "unit": "GB/s", "traffic_gbps": None}
    out["frac"] = out["achieved"] / HBM_PEAK_GBPS
    out["traffic"] = None
    profile = profile or BATCHED_TRAFFIC_PROFILE
    if os.path.exists(profile):
        with open(profile) as f:
            prof = json.load(f)
        # the profile applies to this script, or to one within 2% of its
        # size (the tuned split and the learned merge can move a few
        # partitions between runs)
        pb = prof.get("script_bytes")
        if pb is None or abs(pb - int(bs.blocks_data.nbytes)) <= 0.02 * pb:
            out["traffic_gbps"] = prof["hbm_bytes_per_lane_cycle"] * rate / 1e9
            # HBM bytes per launch at this run's lane-cycles per launch
            out["traffic"] = prof["hbm_bytes_per_lane_cycle"] * b["lane_cycles"] / max(
                b.get("launches", 1), 1)
    return out


def traffic_from_profile(cycles_per_launch, script_bytes):
    """HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE
    profile of the same workload (already scaled by the copy-kernel
    calibration), or None when no profile of this script is present."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        p = json.load(f)
    if p.get("script_bytes") != script_bytes:
        return None
    per_cycle = p["hbm_bytes_per_launch"] / p["cycles_per_launch"]
    return per_cycle * cycles_per_launch


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(lanes would shard across GPUs with one script replica "
                          "each; only the one-GPU share is built and measured)"}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
    if a.workload == "north":
        return bench_north(a, torch)
    if a.workload == "mult16m":
        return bench_mult16m(a, torch)
    if a.workload == "cosim":
        return bench_cosim(a, torch)
    if a.workload == "xtiming":
        x = bench_xtiming(a, torch)
        s = x["script"]
        value = x["cycles"] / x["wall"]
        achieved = x["bytes_per_cycle"] * x["cycles"] / x["kernel_s"] / 1e9
        print(json.dumps({
            "metric": METRIC, "value": value, "unit": "lane-cycles/s", "n_gpus": 1,
            "steps": 1, "warmup": 0, "ms_per_step": x["wall"] * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32", "data": "synthetic",
            "config": {
                "workload": f"plain GPU drop-in with X-prop and timing: seeded random AIG, "
                            f"{a.gates} AND, depth {a.depth}, {a.dffs} DFF, 8 x 1024x32 "
                            f"SRAM, 10% DFFs initially X, delays U[20,200] ps, setup/hold "
                            f"50/20 ps at 5000 ps; {x['cycles']} cycles, toggle p={a.toggle}",
                "blocks": s.num_blocks, "major_stages": s.num_major_stages,
                "partitions": s.num_parts, "state_words": s.effective_state_size(),
                "script_bytes": int(s.blocks_data.nbytes),
                "violations": x["violations"], "verified_cycles": x["verified_cycles"]},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                         "kernel": "scan_kernel<true,false>",
                         "algorithmic_bytes_per_cycle": x["bytes_per_cycle"]},
            "cpu_baseline": {"value": x["cpu_cycles_per_s"], "unit": "lane-cycles/s",
                             "cores": 1, "kind": "port",
                             "sample": f"the oracle's kernel-semantics restatement (X lane, "
                                       f"arrivals, setup/hold; gcc -O2, one thread) over the "
                                       f"first {x['cpu_cycles']} cycles of the same run: "
                                       f"{x['cpu_s']:.1f} s"},
        }), flush=True)
        return
    from jacquard_amd import _native as N
    import ctypes
    nl, comp = build_design(a)
    res = ctypes.c_int(0)
    N.check(N.hip().jq_max_resident_blocks(0, ctypes.byref(res)))
    tune_table = None
    if a.blocks:
        script = comp.flatten(a.blocks)
    elif a.no_plain_tune:
        # one block per partition of the widest major stage: more blocks only
        # add grid-barrier arrivals, fewer serialise partitions
        script = comp.flatten(min(max(comp.parts_per_stage()), res.value))
    else:
        from jacquard_amd import sim
        t = time.time()
        comp, script, tune_table = sim.tune_plain(nl, a.split, log=log)
        log(f"plain tune: {time.time() - t:.1f}s")
    log(f"plain script: {script.num_blocks} blocks, {script.num_major_stages} major stages, "
        f"{script.num_parts} partitions, {script.blocks_data.nbytes / 1e6:.1f} MB "
        f"({script.blocks_data.nbytes} bytes)")
    p = bench_plain(a, script, torch)
    value = p["cycles"] / p["wall"]
    avg_launch_s = float(np.mean(p["launch_ms"])) / 1e3
    bpc = script.algorithmic_bytes_per_cycle()
    achieved = bpc * a.cycles_per_step / avg_launch_s / 1e9
    traffic = traffic_from_profile(a.cycles_per_step, int(script.blocks_data.nbytes))
    log(f"plain: {value:,.0f} lane-cycles/s, launch {avg_launch_s * 1e3:.1f} ms")
    forms = {"plain": {"lane_cycles_per_s": value, "verified_cycles": p.get("verified_cycles")}}
    lds_bpc = script.algorithmic_lds_bytes_per_cycle()
    lds_peak = lds_peak_gbps(torch)
    lds_achieved = lds_bpc * a.cycles_per_step / avg_launch_s / 1e9
    log(f"LDS: {lds_achieved:,.0f} GB/s of a measured {lds_peak:,.0f} GB/s peak")
    bar_us = barrier_us_per_stage(script, torch)
    stage_us = avg_launch_s * 1e6 / (a.cycles_per_step * script.num_major_stages)
    log(f"barrier: {bar_us:.2f} us per major stage with empty blocks, "
        f"{stage_us:.2f} us per major stage in the run")
    ab = plain_sync_ab(script, torch)
    log(f"sync A/B over {ab['cycles']} cycles: dataflow {ab['dataflow_cycles_per_s']:,.0f}, "
        f"barrier {ab['barrier_cycles_per_s']:,.0f} cycles/s")
    if not a.no_batched:
        from jacquard_amd.script import Compiled
        comp_b = comp if a.batched_split == a.split else Compiled(nl, level_split=a.batched_split)
        b = bench_batched(a, comp_b, torch)
        forms["batched"] = {"lane_cycles_per_s": b["lane_cycles"] / b["wall"],
                            "lanes": a.lanes, "groups": b["groups"],
                            "blocks_per_group": b["blocks"],
                            "lane_cycles": b["lane_cycles"],
                            "cycles_per_lane": f"U[{int(10_000 * a.batched_scale)}, "
                                               f"{int(200_000 * a.batched_scale)}]",
                            "slot_utilisation": b["lane_cycles"] / b["slot_cycles"],
                            "kernel_s": b["kernel_ms"] / 1e3,
                            "autotune": b["tune"],
                            "partition_width": b["width"],
                            "learned_cost_model": b["learned"],
                            "launch": "one hipGraph replay (reset + persistent kernel)",
                            "verified_cycles": b["verified_cycles"]}
        forms["batched"]["vs_plain"] = forms["batched"]["lane_cycles_per_s"] / value
        forms["batched"]["hbm"] = batched_hbm(b)
        log(f"batched: {forms['batched']['lane_cycles_per_s']:,.0f} lane-cycles/s "
            f"({forms['batched']['vs_plain']:.1f}x plain), "
            f"{forms['batched']['hbm']['achieved']:,.0f} GB/s algorithmic HBM traffic")
    if not a.no_cosim:
        # reactive co-simulation edges (the `--workload cosim` line's forms)
        forms["cosim"] = bench_cosim(a, torch, comp=comp, emit=False)
    cpu = None
    if not a.no_cpu:
        c = cpu_baseline(a, script)
        cpu = {"value": c["value"], "unit": "lane-cycles/s", "cores": 1,
               "kind": "port",
               "sample": f"{c['cycles']} cycles of the same 1M-gate workload through "
                         f"the oracle's literal restatement of sanity_check_cpu "
                         f"(gcc -O2, one thread) on {platform.processor() or 'host'}; "
                         f"{c['seconds']:.1f} s"}
        forms["cpu"] = {"lane_cycles_per_s": c["value"]}
        log(f"cpu: {c['value']:.2f} lane-cycles/s over {c['cycles']} cycles")
        forms["cpu"].update(cpu_variants(a, script))
        log(f"cpu: OpenMP {forms['cpu']['port_openmp']['lane_cycles_per_s']:.2f} "
            f"({forms['cpu']['port_openmp']['threads']} threads), direct AIG "
            f"{forms['cpu']['direct_aig']['lane_cycles_per_s']:.2f} lane-cycles/s")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "lane-cycles/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": p["wall"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "plain GPU drop-in: seeded random AIG, "
                        f"{a.gates} AND, depth {a.depth}, {a.dffs} DFF, {a.inputs} PI, "
                        f"{a.outputs} PO, geometric(0.3) fan-in; "
                        f"{a.steps * a.cycles_per_step} cycles, toggle p={a.toggle}",
            "cycles_per_step": a.cycles_per_step,
            "blocks": script.num_blocks,
            "major_stages": script.num_major_stages,
            "partitions": script.num_parts,
            "reg_io_state_words": script.reg_io_state_size,
            "plain_partition_tune": tune_table,
            "script_bytes": int(script.blocks_data.nbytes),
            "forms": forms,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "kernel": "scan_kernel",
            "algorithmic_bytes_per_cycle": bpc,
        },
        "lds_roofline": {
            "achieved": lds_achieved,
            "peak": lds_peak,
            "unit": "GB/s",
            "frac": lds_achieved / lds_peak,
            "peak_source": "measured: jq_lds_read_probe, 8 workgroups per CU",
            "spec_peak": LDS_SPEC_PEAK_GBPS,
            "algorithmic_bytes_per_cycle": lds_bpc,
        },
        "barrier": {
            "us_per_major_stage_empty": bar_us,
            "us_per_major_stage_run": stage_us,
            "blocks": script.num_blocks,
        },
        "sync_ab": ab,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
