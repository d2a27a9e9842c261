"""The reference's two launch symbols called from a compiled C program
(tests/abi_smoke.c, built by tests/Makefile with gcc against
include/jacquard_hip.h and libjacquard_hip.so), the way the reference's Rust
host calls them (src/bin/jacquard.rs:1270-1318).

CPU: the driver builds, and its only imports from the library are the two
reference symbols.  GPU: snapshots, SRAM contents and events of a run
through the driver against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from helpers import design
from jacquard_amd import stimulus

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "_bin", "abi_smoke")
MAGIC = 0x4A51534D4F4B4531
EV_BYTES = 16 + 1024 * 32


def _driver():
    subprocess.check_call(["make", "-s", "-C", HERE])
    return BIN


def test_c_driver_builds_and_imports_the_reference_symbols():
    path = _driver()
    out = subprocess.run(["nm", "-D", "--undefined-only", path], capture_output=True,
                         text=True, check=True).stdout
    imported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert "simulate_v1_noninteractive_simple_scan_hip" in imported
    assert "simulate_v1_noninteractive_timed_hip" in imported
    assert not [n for n in imported if n.startswith("jq_")]


def _job(path, s, states, cycles, sram, sram_x, cons):
    ss = s.effective_state_size()
    timed = cons is not None
    cons = np.zeros(0, np.uint32) if cons is None else np.asarray(cons, np.uint32)
    hdr = np.array([MAGIC, s.num_blocks, s.num_major_stages, cycles, ss,
                    len(s.blocks_start), len(s.blocks_data), len(sram), int(timed),
                    len(cons), np.int64(s.arrival_state_offset).view(np.uint64)],
                   np.uint64)
    with open(path, "wb") as f:
        for a in (hdr, np.asarray(s.blocks_start, np.uint64),
                  np.asarray(s.blocks_data, np.uint32), states, sram, sram_x, cons):
            f.write(np.ascontiguousarray(a).tobytes())


def _run(tmp_path, s, states, cycles, sram, sram_x, cons=None):
    job, res = str(tmp_path / "job.bin"), str(tmp_path / "result.bin")
    _job(job, s, states, cycles, sram, sram_x, cons)
    print("abi_smoke:", s.num_parts, "partitions,", cycles, "cycles,",
          "timed" if cons is not None else "scan", flush=True)
    subprocess.run([_driver(), job, res], check=True, timeout=120)
    raw = np.fromfile(res, np.uint8)
    n_st, n_sr = states.size * 4, sram.size * 4
    out = dict(states=raw[:n_st].view(np.uint32),
               sram=raw[n_st:n_st + n_sr].view(np.uint32),
               sram_x=raw[n_st + n_sr:n_st + 2 * n_sr].view(np.uint32))
    if cons is not None:
        ev = raw[n_st + 2 * n_sr:]
        assert ev.size == EV_BYTES
        out["count"], out["overflow"] = (int(v) for v in ev[:8].view(np.uint32))
        out["events"] = ev[16:].view(oracle.EVENT_DTYPE)
    return out


@pytest.mark.gpu
def test_simple_scan_symbol_from_c(tmp_path):
    nl, comp = design("abi-sram", level_split=8, seed=4, gates=30000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=4)
    s = comp.flatten(32, xprop=True)
    cycles = 80
    m, _ = stimulus.dff_initial_xmask(s, 0.1, seed=5)
    st = stimulus.build_states(s, cycles, seed=6, p=0.3, initial_xmask=m)
    n = s.sram_storage_size
    sram, sram_x = np.zeros(n, np.uint32), np.full(n, 0xFFFFFFFF, np.uint32)
    ref, ref_sram, ref_sx = st.copy(), sram.copy(), sram_x.copy()
    oracle.run(s, ref, cycles, sram=ref_sram, sram_x=ref_sx)
    out = _run(tmp_path, s, st, cycles, sram, sram_x)
    assert np.array_equal(out["states"], ref)
    assert np.array_equal(out["sram"], ref_sram)
    assert np.array_equal(out["sram_x"], ref_sx)


@pytest.mark.gpu
def test_timed_symbol_from_c(tmp_path):
    nl, comp = design("abi-timed", level_split=6, seed=9, gates=40000, depth=48,
                      dffs=3000, inputs=64, outputs=32, geometric_p=0.3, srams=2,
                      delay=(20, 200))
    s = comp.flatten(24)
    s.inject_timing()
    s.enable_timing_arrivals()
    cons = s.build_timing_constraint_buffer(700, 50, 20)  # 1,280 events
    cycles = 40
    st = stimulus.build_states(s, cycles, seed=16, p=0.3)
    n = s.sram_storage_size
    sram, sram_x = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    ref, ref_sram = st.copy(), sram.copy()
    ev_ref, n_ref = oracle.run(s, ref, cycles, sram=ref_sram, constraints=cons,
                               arrival_state_offset=s.arrival_state_offset,
                               event_capacity=1 << 20)
    out = _run(tmp_path, s, st, cycles, sram, sram_x, cons)
    assert np.array_equal(out["states"], ref)
    assert np.array_equal(out["sram"], ref_sram)
    # the reference buffer keeps 1024 records and counts them all
    assert n_ref > 1024
    assert out["count"] == n_ref and out["overflow"] == 1
    full = {r.tobytes() for r in ev_ref[:n_ref]}
    assert all(r.tobytes() in full for r in out["events"])
