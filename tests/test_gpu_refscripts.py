"""The reference's own script recipes, and the arrival probes, on the GPU.

- bench_script (benches/xprop.rs:11-80) at num_ios in {32, 128, 256} and
  1, 4 or 8 boomerang stages, 2-state and X-capable, and
  minimal_xprop_script (src/sim/cpu_reference.rs:994-1091), through both
  reference launch symbols: every word of every snapshot against the oracle.
- tree_probe_script: every tree word's arrival against the brute-force
  longest path of tests/test_timing_pins.py, and the setup/hold events
  against the closed form "setup fires iff arrival > 0 and arrival + setup
  > clock, hold iff arrival < hold" (csrc/kernel_v1_impl.cuh:529-558),
  through jq_simulate_scan and the timed reference symbol.
The CPU test at the top validates the same scripts and runs them on the
oracle, so the GPU runs only scripts the validator accepts.
"""
import numpy as np
import pytest

import oracle
from handscripts import (M, OneBlockScript, bench_script, minimal_xprop_script,
                         tree_probe_script)
from test_timing_pins import brute_arrivals, random_tree, wrap_chain

BENCH = [(x, ios, st) for x in (False, True) for ios in (32, 128, 256) for st in (1, 4, 8)]
CYCLES = 6


def _say(*a):
    print(*a, flush=True)


def ref_case(kind, is_x, ios, stages, seed=0):
    part = (bench_script(is_x, ios, stages) if kind == "bench"
            else minimal_xprop_script(is_x, ios))
    s = OneBlockScript(part, ios, xprop=is_x)
    ss = s.effective_state_size()
    r = np.random.default_rng(1000 * ios + 10 * stages + seed + int(is_x))
    st = np.zeros((CYCLES + 1) * ss, np.uint32)
    st[:ss] = r.integers(0, 1 << 32, ss, dtype=np.uint64).astype(np.uint32)
    if is_x:  # a sparse X mask in snapshot 0
        st[ios:ss] &= r.integers(0, 1 << 32, ios, dtype=np.uint64).astype(np.uint32)
        st[ios:ss] &= r.integers(0, 1 << 32, ios, dtype=np.uint64).astype(np.uint32)
    return s, st


def oracle_states(s, st):
    ref = st.copy()
    oracle.run(s, ref, CYCLES)
    return ref


CASES = [("bench",) + c for c in BENCH] + [("minimal", x, 4, 1) for x in (False, True)]


def test_reference_scripts_validate_and_run_on_the_oracle():
    for kind, is_x, ios, stages in CASES:
        s, st = ref_case(kind, is_x, ios, stages)
        s.validate()
        ref = oracle_states(s, st)
        ss = s.effective_state_size()
        # word 0 takes the tree's word 0 (every level a pass-through of its
        # a operand); every other word keeps its value (clock enable 0)
        assert np.array_equal(ref[ss:2 * ss][1:ios], st[1:ios])


@pytest.mark.gpu
@pytest.mark.parametrize("kind,is_x,ios,stages", CASES)
def test_reference_script_on_both_symbols(kind, is_x, ios, stages):
    from jacquard_amd import sim
    s, st = ref_case(kind, is_x, ios, stages)
    _say("refscript", kind, "x" if is_x else "2s", ios, stages)
    ref = oracle_states(s, st)
    got = sim.run_scan(s, st.copy(), CYCLES, reference_symbols=True)["states"]
    assert np.array_equal(got, ref)
    # the timed symbol with an all-zero constraint buffer: same values
    cons = np.zeros(1 + ios, np.uint32)
    cons[0] = 5000
    out = sim.run_scan(s, st.copy(), CYCLES, constraints=cons, reference_symbols=True)
    assert np.array_equal(out["states"], ref)
    assert out["event_count"] == 0


PROBES = [("random", k) for k in range(4)] + [("dense", 7), ("sparse", 8), ("wrap", 0)]


def probe(kind, seed):
    if kind == "wrap":
        fl, expect = wrap_chain()
        arr = np.zeros(256, np.int64)
        for w, e in expect.items():
            arr[w] = e
        known = sorted(expect)
    else:
        fl = random_tree(seed, pass_p={"random": 0.3, "dense": 0.0, "sparse": 0.8}[kind])
        arr = brute_arrivals(fl[2], fl[3])
        known = list(range(256))
    return tree_probe_script(*fl), arr, known


def closed_form_events(arr, known, clock, setup, hold, cycles):
    ev = []
    for c in range(cycles):
        for w in known:
            a = int(arr[w])
            if a > 0 and a + setup > clock:
                ev.append((c, w, 4, clock - a - setup, a))
            if a < hold:
                ev.append((c, w, 5, a - hold, a))
    return sorted(ev)


def as_tuples(ev, known):
    k = set(known)
    return sorted((int(e["cycle"]), int(e["message_id"]), int(e["event_type"]),
                   int(np.array(e["data"][1], np.uint32).view(np.int32)), int(e["data"][2]))
                  for e in ev if int(e["message_id"]) in k)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,seed", PROBES)
@pytest.mark.parametrize("symbol", [False, True])
def test_tree_arrivals_and_events_closed_form(kind, seed, symbol):
    from jacquard_amd import sim
    part, arr, known = probe(kind, seed)
    s = OneBlockScript(part, 256, arrivals=True)
    ss = s.effective_state_size()
    cycles = 3
    r = np.random.default_rng(seed + 77)
    st = np.zeros((cycles + 1) * ss, np.uint32)
    st[:256] = r.integers(0, 1 << 32, 256, dtype=np.uint64).astype(np.uint32)
    clock, setup, hold = 1500, 50, 20
    cons = np.zeros(1 + 256, np.uint32)
    cons[0] = clock
    cons[1:] = (setup << 16) | hold
    _say("probe", kind, seed, "reference symbol" if symbol else "jq_simulate_scan")
    ref = st.copy()
    ev_ref, n_ref = oracle.run(s, ref, cycles, constraints=cons,
                               arrival_state_offset=256, event_capacity=1 << 14)
    out = sim.run_scan(s, st.copy(), cycles, constraints=cons, event_capacity=1 << 14,
                       reference_symbols=symbol)
    got = out["states"].reshape(cycles + 1, ss)
    for c in range(1, cycles + 1):
        assert np.array_equal(got[c, 256 + np.array(known)], arr[known]), f"cycle {c}"
    assert np.array_equal(out["states"], ref)
    cf = closed_form_events(arr, known, clock, setup, hold, cycles)
    if symbol:  # the reference symbol keeps 1024 records, counts them all
        assert out["event_count"] == n_ref
        assert out["overflow"] == int(n_ref > 1024)
        if n_ref <= 1024:
            assert as_tuples(out["events"], known) == cf
    else:
        assert out["event_count"] == n_ref
        assert as_tuples(out["events"], known) == cf
    assert as_tuples(ev_ref, known) == cf
