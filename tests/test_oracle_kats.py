"""Known-answer tests pinning the CPU oracle to the reference's own tests.

Each test restates a reference test's inputs and expected outputs:
  - X-prop AND truth table, inversion and per-bit cases
    (src/sim/cpu_reference.rs:797-894), evaluated through a one-word script
    rather than by retyping the formula;
  - the minimal one-partition X script recipe and its three end-to-end
    checks (src/sim/cpu_reference.rs:994-1206);
  - setup/hold arithmetic with the exact slacks of src/flatten.rs:3068-3366,
    driven through a one-partition script whose writeout arrival is set by
    a delay slot, so the oracle's own event logic is what is checked.
"""
import numpy as np
import pytest

import oracle

M = 0xFFFFFFFF


def xprop_and(av, ax, bv, bx, xa, xb, ob):
    """One AND word through the oracle (hier0 of a one-stage script)."""
    sc = minimal_script(True, 4)
    # put a at hier word 0 and b at word 128 via the shuffle: word 0 takes
    # state word 0 bit k, word 128 takes state word 1 bit k
    pi = 256 + 512
    for ko in range(4):
        for t in (0, 128):
            for ki in range(4):
                k = ko * 4 + ki
                src = 0 if t == 0 else 1
                lo = src * 32 + 2 * k
                hi = src * 32 + 2 * k + 1
                sc[pi + ko * 1024 + t * 4 + ki] = lo | (hi << 16)
    fl = pi + 4096
    sc[fl + 128 * 4:fl + 128 * 4 + 3] = [xa, xb, ob]
    # global read: thread 0 <- state word 0 (all bits), thread 1 <- word 1
    g = 256
    sc[g:g + 4] = [0, M, 1, M]
    # hook writeout 0 <- stage 0 word 128
    sc[128] = (sc[128] & 0xFFFF0000) | 128
    inv = np.array([av, bv, 0, 0], np.uint32)
    ixm = np.array([ax, bx, 0, 0], np.uint32)
    # global read reverses bit order inside each thread's word
    rev = lambda w: int("{:032b}".format(w)[::-1], 2)
    inv[:2] = [rev(av), rev(bv)]
    ixm[:2] = [rev(ax), rev(bx)]
    out_v = np.zeros(4, np.uint32)
    out_x = np.zeros(4, np.uint32)
    rc = oracle.simulate_block_v1_xprop(sc, inv, out_v, ixm, out_x,
                                        np.zeros(1, np.uint32), np.zeros(1, np.uint32))
    assert rc == 0
    return int(out_v[0]), int(out_x[0])


def minimal_script(is_x, num_ios):
    """build_minimal_xprop_script (src/sim/cpu_reference.rs:994-1091)."""
    total = 256 + 512 + 5120 + 5120 + 5120 + 256
    s = np.zeros(total, np.uint32)
    s[0], s[1], s[2], s[3] = 1, 0, num_ios, 0
    s[4], s[5], s[6], s[7] = 0, 0, 1, 0
    s[8], s[9] = 1 if is_x else 0, num_ios
    s[128:256] = 0xFF00FF00
    s[128] = s[128] & 0xFFFF0000
    pi = 256
    s[pi], s[pi + 1] = 0, 1
    pi += 512 + 4096
    s[pi + 128 * 4 + 2] = M
    for hi in range(1, 8):
        hw = 1 << (7 - hi)
        s[pi + hw * 4 + 2] = M
    s[pi + 0 * 4 + 2] = M
    pi += 1024 + 5120 + 4096
    s[pi] = M
    pi += 1024
    s[pi], s[pi + 1] = 0, 1
    return s


@pytest.mark.parametrize("case,expect", [
    # (av, ax, bv, bx, xa, xb, ob) -> (value, X)
    ((0, 0, 0, M, 0, 0, 0), (0, 0)),             # 0 & X = 0
    ((M, 0, 0, M, 0, 0, 0), (0, M)),             # 1 & X = X
    ((0, M, 0, M, 0, 0, 0), (None, M)),          # X & X = X
    ((M, 0, M, 0, 0, 0, 0), (M, 0)),             # 1 & 1 = 1
    ((M, 0, 0, 0, 0, 0, 0), (0, 0)),             # 1 & 0 = 0
    ((0, M, 0, M, 0, 0, M), (None, M)),          # pass-through keeps a's X
    ((0xAAAAAAAA, 0, 0, M, 0, 0, M), (0xAAAAAAAA, 0)),
])
def test_xprop_and_truth_table(case, expect):
    v, x = xprop_and(*case)
    if expect[0] is not None:
        assert v == expect[0]
    assert x == expect[1]


def test_xprop_inversion_and_mixed_bits():
    _, x = xprop_and(0, 0, 0, 0xFFFF, 0xFFFF, 0, 0)
    assert x & 0xFFFF == 0xFFFF
    v, x = xprop_and(0b1010, 0b0100, 0b1000, 0b0011, 0, 0, 0)
    assert x & 1 == 0 and x & 2 == 2 and x & 4 == 0
    assert v & 8 == 8 and x & 8 == 0


def test_minimal_script_known_value_passthrough():
    s = minimal_script(True, 4)
    iv = np.array([1, 0, 0, 0], np.uint32)
    ov, ox = np.zeros(4, np.uint32), np.zeros(4, np.uint32)
    assert oracle.simulate_block_v1_xprop(s, iv, ov, np.zeros(4, np.uint32), ox,
                                          np.zeros(1, np.uint32), np.zeros(1, np.uint32)) == 0
    assert ox[0] & 1 == 0


def test_minimal_script_x_propagates():
    s = minimal_script(True, 4)
    ov, ox = np.zeros(4, np.uint32), np.zeros(4, np.uint32)
    oracle.simulate_block_v1_xprop(s, np.zeros(4, np.uint32), ov,
                                   np.array([1, 0, 0, 0], np.uint32), ox,
                                   np.zeros(1, np.uint32), np.zeros(1, np.uint32))
    assert ox[0] != 0


@pytest.mark.parametrize("quirk", [False, True])
def test_non_x_capable_equals_plain(quirk):
    s = minimal_script(False, 4)
    iv = np.array([0xAB, 0, 0, 0], np.uint32)
    a, b = np.zeros(4, np.uint32), np.zeros(4, np.uint32)
    oracle.simulate_block_v1_xprop(s, iv, a, np.zeros(4, np.uint32), np.zeros(4, np.uint32),
                                   np.zeros(1, np.uint32), np.zeros(1, np.uint32), quirk)
    assert oracle.simulate_block_v1(s, iv, b, np.zeros(1, np.uint32)) == 0
    assert np.array_equal(a, b)


# ---------------------------------------------------------------- timing
def timing_script(arrival, constraint, clock):
    """One partition, writeout 0 <- stage-0 word 128 (delay slot = arrival),
    clock enable all-ones; returns the events the oracle emits."""
    s = minimal_script(False, 1)
    s[128] = (s[128] & 0xFFFF0000) | 128
    fl = 256 + 512 + 4096
    s[fl + 128 * 4 + 3] = arrival
    s[fl + 128 * 4 + 2] = 0  # plain AND, arrival = 0 + delay
    state = np.zeros(2, np.uint32)
    out = np.zeros(2, np.uint32)
    cons = np.array([constraint], np.uint32)
    ev = np.zeros(8, oracle.EVENT_DTYPE)
    import ctypes
    cnt = ctypes.c_size_t(0)
    rc = oracle.lib().jqo_block_kernel(
        s.ctypes.data, s.size, state.ctypes.data, out.ctypes.data, None, None,
        cons.ctypes.data, clock, 0, 7, ev.ctypes.data, 8, ctypes.byref(cnt))
    assert rc == 0
    return ev[:cnt.value]


def _slack(ev, kind):
    e = [x for x in ev if x["event_type"] == kind]
    return None if not e else int(np.array(e[0]["data"][1], np.uint32).view(np.int32))


@pytest.mark.parametrize("arr,setup,clock,slack", [
    (900, 200, 1000, -100), (800, 200, 1000, None),
    (1323, 85, 1200, -208), (1323, 85, 1500, None), (1323, 85, 1400, -8),
    (850, 200, 1000, -50), (0, 200, 100, None), (1, 200, 100, -101),
    (65000, 1000, 60000, -6000), (65000, 1000, 70000, None),
    (65535, 65535, 131069, -1), (65535, 65535, 131070, None),
])
def test_setup_slack(arr, setup, clock, slack):
    ev = timing_script(arr, setup << 16, clock)
    assert _slack(ev, 4) == slack
    if slack is not None:
        e = [x for x in ev if x["event_type"] == 4][0]
        assert int(e["data"][0]) == 0 and int(e["data"][2]) == arr
        assert int(e["data"][3]) == setup and int(e["cycle"]) == 7


@pytest.mark.parametrize("arr,hold,slack", [
    (10, 50, -40), (50, 50, None), (0, 50, -50), (500, 0, None), (0, 0, None)])
def test_hold_slack(arr, hold, slack):
    ev = timing_script(arr, hold | (1 << 16), 1 << 30)
    assert _slack(ev, 5) == slack


def test_zero_constraint_never_fires():
    assert len(timing_script(0, 0, 1)) == 0
    assert len(timing_script(5000, 0, 1)) == 0


def test_aig_sim_resumes_from_register_values():
    """A direct AIG run started from the registers after cycle N-1 of a
    longer run (and the inputs from cycle N on) reproduces that run's
    outputs and registers from cycle N: the property the windowed checks of
    long GPU runs rely on."""
    from helpers import CONFIG1, design
    from jacquard_amd import stimulus
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    A = nl.arrays()
    data, fl = stimulus.input_roles(s)
    cycles, n = 60, 25
    vals = stimulus.lane_inputs(len(data), cycles, 4, 0.4)
    pos2pin = {v: k for k, v in s.input_map.items() if A["drv"][k] in (2, 3)}
    pins = [pos2pin[p] for p in data.tolist()] + [pos2pin[p] for p in fl.tolist()]
    bits = np.concatenate([vals[:cycles].astype(np.uint8),
                           np.ones((cycles, len(fl)), np.uint8)], axis=1)
    obs, regs = oracle.aig_sim(A, cycles, pins, bits, A["pos"], want_regs=True)
    assert regs[n - 1].any() and not regs[n - 1].all()
    obs2, regs2 = oracle.aig_sim(A, cycles - n, pins, bits[n:], A["pos"], want_regs=True,
                                 dff_init=regs[n - 1])
    assert np.array_equal(obs2, obs[n:]) and np.array_equal(regs2, regs[n:])
    # starting from the wrong registers changes the outputs (the start
    # state is used, not ignored)
    obs3, _ = oracle.aig_sim(A, cycles - n, pins, bits[n:], A["pos"], want_regs=True)
    assert not np.array_equal(obs3, obs[n:])
