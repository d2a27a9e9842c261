#!/usr/bin/env python3
"""Throughput of the MI355X Boolean-processor interpreter in simulated
lane-cycles per second.

Default workload (the plain GPU drop-in configuration): a seeded random AIG
of 1M AND gates, depth 64, 64K DFFs, 512 inputs, 256 outputs, fan-in level
distance ~ Geometric(0.3); 100K cycles of stimulus with every input toggling
with probability 0.1.  One "step" is one launch of the interpreter over
--cycles-per-step cycles of that stimulus, continuing from the previous
step's state, with the inputs already resident in HBM.

The same run also measures the batched lane-sliced form (256 lanes of the
same netlist, per-lane seeds, cycle counts U[10K, 200K] scaled by
--batched-scale, toggle rates U[0.02, 0.5]) and times the reference CPU
path (the oracle's literal restatement of sanity_check_cpu) on a bounded
cycle sample on one host core.

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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cycles-per-step", type=int, default=10_000)
    ap.add_argument("--gates", type=int, default=1_000_000)
    ap.add_argument("--depth", type=int, default=64)
    ap.add_argument("--dffs", type=int, default=65536)
    ap.add_argument("--inputs", type=int, default=512)
    ap.add_argument("--outputs", type=int, default=256)
    ap.add_argument("--toggle", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--split", type=int, default=4, help="level-split period")
    ap.add_argument("--blocks", type=int, default=0, help="plain blocks (0 = auto)")
    ap.add_argument("--lanes", type=int, default=256)
    ap.add_argument("--batched-scale", type=float, default=0.02,
                    help="scale of the batched per-lane cycle counts "
                         "(1.0 = U[10K, 200K])")
    ap.add_argument("--batched-blocks", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--verify-cycles", type=int, default=4)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run an elementwise copy of MIB MiB three times "
                         "(the known-size kernel that scales the HBM counters)")
    return ap.parse_args()


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
    del st
    torch.cuda.empty_cache()
    return out


def bench_batched(a, comp, torch):
    from jacquard_amd import _native as N
    from jacquard_amd.batched import BatchedRun
    import ctypes
    res = ctypes.c_int(0)
    N.check(N.hip().jq_max_resident_blocks(1, ctypes.byref(res)))
    groups = (a.lanes + 31) // 32
    W = a.batched_blocks or max(1, min(64, res.value // groups))
    script = comp.flatten(W)
    rng = np.random.default_rng(a.seed + 100)
    lo, hi = int(10_000 * a.batched_scale), int(200_000 * a.batched_scale)
    seeds = rng.integers(1, 2**31, a.lanes).astype(np.uint32)
    probs = rng.uniform(0.02, 0.5, a.lanes)
    cycles = rng.integers(max(lo, 1), max(hi, 1) + 1, a.lanes)
    run = BatchedRun(script, seeds, probs, cycles)
    script.validate(script.reg_io_state_size)
    run.prepare()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    run.launch()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    from jacquard_amd.sim import device_error
    if device_error():
        raise RuntimeError("device barrier timeout during the batched run")
    lc = run.lane_cycles_total()
    log(f"batched: {a.lanes} lanes x {groups} groups x {W} blocks, "
        f"{lc} lane-cycles in {wall:.2f}s")
    # spot-check one lane against the CPU oracle on its first cycles
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from jacquard_amd import stimulus
    from jacquard_amd.batched import plain_observations
    outs = run.lane_outputs()
    lane = int(np.argmin(cycles))
    v = min(a.verify_cycles or 4, int(cycles[lane]))
    st = stimulus.build_states(script, v, seed=int(seeds[lane]), p=float(probs[lane]))
    oracle.run(script, st, v)
    ok = np.array_equal(outs[lane][:v], plain_observations(script, st, v, run.obs_lits))
    if not ok:
        raise AssertionError("batched GPU lane differs from the CPU oracle")
    return dict(wall=wall, kernel_ms=e0.elapsed_time(e1), lane_cycles=lc,
                slot_cycles=run.slot_cycles_total(), groups=groups, blocks=W,
                verified_cycles=v)


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


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "scan_kernel_1M_gates_hbm_traffic.json")


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
                          "(lanes would shard across GPUs, see DESIGN.md)"}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
    from jacquard_amd import _native as N
    import ctypes
    nl, comp = build_design(a)
    res = ctypes.c_int(0)
    N.check(N.hip().jq_max_resident_blocks(0, ctypes.byref(res)))
    # one block per partition of the widest major stage: more blocks only
    # add grid-barrier arrivals, fewer serialise partitions
    blocks = a.blocks or min(max(comp.parts_per_stage()), res.value)
    script = comp.flatten(blocks)
    log(f"plain script: {script.num_blocks} blocks, {script.num_major_stages} major stages, "
        f"{script.num_parts} partitions, {script.blocks_data.nbytes / 1e6:.1f} MB")
    p = bench_plain(a, script, torch)
    value = p["cycles"] / p["wall"]
    avg_launch_s = float(np.mean(p["launch_ms"])) / 1e3
    bpc = script.algorithmic_bytes_per_cycle()
    achieved = bpc * a.cycles_per_step / avg_launch_s / 1e9
    traffic = traffic_from_profile(a.cycles_per_step, int(script.blocks_data.nbytes))
    log(f"plain: {value:,.0f} lane-cycles/s, launch {avg_launch_s * 1e3:.1f} ms")
    forms = {"plain": {"lane_cycles_per_s": value, "verified_cycles": p.get("verified_cycles")}}
    if not a.no_batched:
        b = bench_batched(a, comp, torch)
        forms["batched"] = {"lane_cycles_per_s": b["lane_cycles"] / b["wall"],
                            "lanes": a.lanes, "groups": b["groups"],
                            "blocks_per_group": b["blocks"],
                            "lane_cycles": b["lane_cycles"],
                            "slot_utilisation": b["lane_cycles"] / b["slot_cycles"],
                            "verified_cycles": b["verified_cycles"]}
        forms["batched"]["vs_plain"] = forms["batched"]["lane_cycles_per_s"] / value
        log(f"batched: {forms['batched']['lane_cycles_per_s']:,.0f} lane-cycles/s "
            f"({forms['batched']['vs_plain']:.1f}x plain)")
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
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
