"""Device runners behind the C ABI (the `jacquard sim` HIP path).

``run_scan`` mirrors ``sim_hip`` (src/bin/jacquard.rs:1259-1318): upload the
pre-filled snapshot buffer, zeroed SRAM (all-X SRAM shadow when X
propagation is on), launch the interpreter once for all cycles, read the
buffer back.  torch provides device memory and streams only; every launch
goes through libjacquard_hip.so and fails loudly when it is missing.
"""
import ctypes

import numpy as np

from . import _native as N

EVENT_DTYPE = np.dtype([("event_type", "<u4"), ("message_id", "<u4"),
                        ("cycle", "<u4"), ("reserved", "<u4"),
                        ("data", "<u4", (4,))])


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the jacquard_amd device path needs a MI355X")
    return torch


def to_device(a):
    torch = _torch()
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        t = torch.from_numpy(a.view(np.int32))
    elif a.dtype == np.uint64:
        t = torch.from_numpy(a.view(np.int64))
    elif a.dtype == np.uint8:
        t = torch.from_numpy(a)
    else:
        t = torch.from_numpy(a)
    return t.to("cuda", non_blocking=False)


def from_device(t, dtype):
    a = t.cpu().numpy()
    return a.view(dtype)


def device_error(clear=True):
    v = ctypes.c_uint32(0)
    N.check(N.hip().jq_device_error_flag(ctypes.byref(v), int(clear)))
    return v.value


def run_scan(script, states, num_cycles, *, constraints=None, event_capacity=0,
             sram_init=None, reference_symbols=False, validate=True,
             stream=None, return_device=False):
    """Run the plain drop-in over a host snapshot buffer.

    states      : uint32[(num_cycles+1) * effective_state_size], pre-filled
    constraints : uint32[1 + reg_io_state_size] ([clock, words...]) or None
    event_capacity : records kept (the reference symbol keeps 1024)
    Returns dict(states, events, event_count, overflow, sram, sram_x)."""
    torch = _torch()
    L = N.hip()
    ss = script.effective_state_size()
    assert states.size == (num_cycles + 1) * ss, "snapshot buffer size"
    if validate:
        script.validate(ss)
    d_bs = to_device(np.ascontiguousarray(script.blocks_start, np.uint64))
    d_bd = to_device(np.ascontiguousarray(script.blocks_data, np.uint32))
    d_st = to_device(states)
    nsram = max(script.sram_storage_size, 1)
    sram = np.zeros(nsram, np.uint32) if sram_init is None else sram_init
    d_sram = to_device(sram)
    d_sram_x = None
    if script.xprop_enabled:
        d_sram_x = to_device(np.full(nsram, 0xFFFFFFFF, np.uint32))
    else:
        d_sram_x = to_device(np.zeros(1, np.uint32))
    d_cons = to_device(constraints) if constraints is not None else None
    cap = event_capacity if constraints is not None else 0
    if reference_symbols and constraints is not None:
        cap = 1024
    d_ev = None
    if constraints is not None:
        d_ev = to_device(np.zeros(16 + 32 * max(cap, 1), np.uint8))
    ptr = lambda t: None if t is None else t.data_ptr()
    torch.cuda.synchronize()
    if reference_symbols:
        if constraints is None:
            L.simulate_v1_noninteractive_simple_scan_hip(
                script.num_blocks, script.num_major_stages, ptr(d_bs), ptr(d_bd),
                ptr(d_sram), ptr(d_sram_x), num_cycles, ss, ptr(d_st),
                int(script.arrival_state_offset))
        else:
            L.simulate_v1_noninteractive_timed_hip(
                script.num_blocks, script.num_major_stages, ptr(d_bs), ptr(d_bd),
                ptr(d_sram), ptr(d_sram_x), num_cycles, ss, ptr(d_st),
                ptr(d_cons), ptr(d_ev), int(script.arrival_state_offset))
    else:
        s = stream if stream is not None else torch.cuda.current_stream()
        N.check(L.jq_simulate_scan(
            script.num_blocks, script.num_major_stages, ptr(d_bs), ptr(d_bd),
            ptr(d_sram), ptr(d_sram_x), num_cycles, ss, ptr(d_st), ptr(d_cons),
            ptr(d_ev), cap, int(script.arrival_state_offset),
            ctypes.c_void_p(s.cuda_stream)))
    torch.cuda.synchronize()
    if device_error():
        raise RuntimeError("device barrier timed out (grid did not co-reside)")
    out = dict(states=from_device(d_st, np.uint32),
               sram=from_device(d_sram, np.uint32))
    if script.xprop_enabled:
        out["sram_x"] = from_device(d_sram_x, np.uint32)
    if d_ev is not None:
        raw = from_device(d_ev, np.uint8)
        hdr = raw[:16].view(np.uint32)
        cnt = int(hdr[0])
        out["event_count"] = cnt
        out["overflow"] = int(hdr[1])
        out["events"] = raw[16:16 + 32 * min(cnt, cap)].view(EVENT_DTYPE).copy()
    if return_device:
        out["device"] = dict(bs=d_bs, bd=d_bd, states=d_st, sram=d_sram)
    return out


def canonical_events(ev):
    """Sort event records by (cycle, word, kind): the GPU appends them with
    atomics in arbitrary order."""
    if ev is None or len(ev) == 0:
        return np.zeros(0, EVENT_DTYPE)
    order = np.lexsort((ev["event_type"], ev["message_id"], ev["cycle"]))
    return ev[order]


def tune_plain(netlist, level_split, caps=(2600, 1800, 1300, 1100, 900, 800, 700),
               trial_cycles=300, seed=1, log=None):
    """Partition width for the plain drop-in, chosen by measurement.

    The plain kernel runs one workgroup per partition of the widest major
    stage, and a major stage takes as long as its slowest partition, whose
    time grows with its boomerang-stage count.  Capping the nodes per logic
    level of a partition (``max_level_nodes``) gives more, narrower
    partitions with fewer boomerang stages each, up to the point where the
    widest major stage no longer fits the co-resident workgroups.  Every cap
    whose widest stage fits is compiled and timed over ``trial_cycles``
    cycles, then the best cap again with every partition held to one
    boomerang stage (``max_boomerang_stages=1``: partitions the mapper
    gives two stages are split by endpoints); the fastest wins.  Returns
    (Compiled, script, table)."""
    from .script import Compiled
    from . import stimulus
    torch = _torch()
    res = ctypes.c_int(0)
    N.check(N.hip().jq_max_resident_blocks(0, ctypes.byref(res)))
    stream = torch.cuda.current_stream()
    best, table = None, []
    # every cap as compiled, then the best cap with every partition held to
    # one boomerang stage (a stage's slowest partition sets its time, and
    # the two-stage ones are the slowest)
    variants = [(cap, 0) for cap in caps] + [(None, 1)]
    for cap, one_stage in variants:
        if cap is None:
            if best is None:
                break
            cap = best[3]
        comp = Compiled(netlist, level_split=level_split, max_level_nodes=cap,
                        max_boomerang_stages=one_stage)
        widest = max(comp.parts_per_stage())
        if widest > res.value:
            table.append({"max_level_nodes": cap, "max_boomerang_stages": one_stage,
                          "widest_stage": widest, "fits": False})
            continue
        s = comp.flatten(widest)
        ss = s.effective_state_size()
        st = to_device(stimulus.build_states(s, trial_cycles, seed=seed, p=0.1))
        d_bs = to_device(np.ascontiguousarray(s.blocks_start, np.uint64))
        d_bd = to_device(np.ascontiguousarray(s.blocks_data, np.uint32))
        one = torch.zeros(max(s.sram_storage_size, 1), dtype=torch.int32, device="cuda")

        def launch(n):
            N.check(N.hip().jq_simulate_scan(
                s.num_blocks, s.num_major_stages, d_bs.data_ptr(), d_bd.data_ptr(),
                one.data_ptr(), one.data_ptr(), n, ss, st.data_ptr(), None, None, 0, 0,
                ctypes.c_void_p(stream.cuda_stream)))

        launch(min(20, trial_cycles))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        launch(trial_cycles)
        e1.record(stream)
        e1.synchronize()
        if device_error():
            raise RuntimeError("device barrier timeout while tuning the plain script")
        rate = trial_cycles / (e0.elapsed_time(e1) / 1e3)
        table.append({"max_level_nodes": cap, "max_boomerang_stages": one_stage,
                      "widest_stage": widest, "fits": True,
                      "partitions": s.num_parts, "cycles_per_s": rate})
        if log:
            log(f"plain tune: max_level_nodes {cap}"
                + (", one boomerang stage per partition" if one_stage else "")
                + f": {s.num_parts} partitions, widest stage {widest}, {rate:,.0f} cycles/s")
        if best is None or rate > best[2]:
            best = (comp, s, rate, cap)
        del st, d_bs, d_bd, one
    if best is None:
        raise RuntimeError("no partition width fits the co-resident workgroups")
    return best[0], best[1], table
