"""Pins of the timing path that do not go through the oracle's own reading.

1. The script-side timing words, restating the reference's known-answer
   tests:
   - delay injection, src/flatten.rs:2868-2928 (raw picoseconds, max over
     the pins of a slot, zero stays zero, ordering);
   - the constraint buffer, src/flatten.rs:2969-3062 (empty, single DFF,
     per-word minimum, invalid and out-of-range positions skipped, clock
     saturation), and the clock-arrival fold of src/flatten.rs:2822-2823
     and :3290-3362.
2. The arrival recurrence of one boomerang stage
   (csrc/kernel_v1_impl.cuh:195-350), checked against a brute-force longest
   path: every leaf-to-word path through the tree is enumerated explicitly
   (a pass-through word, orb == ~0, admits only its a operand), and a word's
   arrival is the largest sum of delay slots along its paths.  In the u16
   wrap regime the closed form is a chain of pass-through words, whose
   arrival is the sum of its delays modulo 65536.
3. Planted mistakes: copies of the oracle with a swapped operand, a dropped
   pass-through delay, a missing tree level, a delay read one level off and
   a saturating add are compiled here, and each must fail (2).
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from handscripts import M, tree_probe_script
from jacquard_amd import _native as N

ORACLE_DIR = os.path.dirname(oracle.__file__)


# ------------------------------------------------------------ delay slots
def inject(delays, patch_map, size):
    """inject_timing_to_script over a blocks_data of `size` zero words;
    delays are (rise, fall) per pin, patch_map [(offset, [pins])]."""
    bd = np.zeros(size, np.uint32)
    off = np.array([o for o, _ in patch_map], np.uint64)
    start = np.cumsum([0] + [len(p) for _, p in patch_map]).astype(np.uint32)
    pins = np.array([q for _, p in patch_map for q in p], np.uint32)
    rise = np.array([d[0] for d in delays], np.uint16)
    fall = np.array([d[1] for d in delays], np.uint16)
    N.host().jqh_inject_delays(bd.ctypes.data, bd.size, off.ctypes.data,
                               start.ctypes.data, pins.ctypes.data, off.size,
                               rise.ctypes.data, fall.ctypes.data, rise.size)
    return bd


def test_basic_delay_injection():  # src/flatten.rs:2867-2879
    assert inject([(0, 0), (100, 100)], [(0, [1])], 1)[0] == 100


def test_max_of_thread_position():  # :2881-2895
    assert inject([(0, 0), (350, 350), (0, 0)], [(0, [1, 2])], 1)[0] == 350


def test_zero_stays_zero():  # :2897-2906
    assert inject([(0, 0), (0, 0)], [(0, [1])], 1)[0] == 0


def test_ordering_preserved():  # :2908-2927
    bd = inject([(0, 0), (5, 5), (15, 15)], [(0, [1]), (1, [2])], 2)
    assert (int(bd[0]), int(bd[1])) == (5, 15)


def test_delay_slot_takes_rise_fall_max_and_skips_unknown_pins():
    # PackedDelay::max_delay is max(rise, fall) (src/flatten.rs:55-57); pins
    # past the delay table are skipped, offsets past blocks_data untouched
    bd = inject([(0, 0), (7, 90), (40, 3)], [(0, [1, 2, 9]), (5, [2])], 2)
    assert list(bd) == [90, 0]


# ------------------------------------------------------- constraint words
def constraint_words(num_words, clock, cons):
    """build_timing_constraint_buffer over (setup, hold, arrival, pos)."""
    n = len(cons)
    su = np.array([c[0] for c in cons], np.uint16)
    ho = np.array([c[1] for c in cons], np.uint16)
    ar = np.array([c[2] for c in cons], np.int16)
    po = np.array([c[3] for c in cons], np.uint32)
    out = np.full(max(num_words, 1), 0xDEADBEEF, np.uint32)
    clk = N.host().jqh_constraint_words(num_words, clock, su.ctypes.data,
                                        ho.ctypes.data, ar.ctypes.data,
                                        po.ctypes.data, n, out.ctypes.data)
    return int(clk), out[:num_words]


def test_empty_constraints():  # src/flatten.rs:2969-2975
    clk, buf = constraint_words(4, 10000, [])
    assert clk == 10000 and len(buf) == 4 and not buf.any()


def test_single_dff_constraint():  # :2977-2997
    clk, buf = constraint_words(4, 25000, [(200, 50, 0, 35)])
    assert clk == 25000
    assert list(buf) == [0, (200 << 16) | 50, 0, 0]


def test_multiple_dffs_same_word_takes_min():  # :2999-3023
    _, buf = constraint_words(2, 10000, [(300, 100, 0, 5), (150, 200, 0, 10)])
    assert buf[0] == (150 << 16) | 100


def test_skip_invalid_data_state_pos():  # :3025-3040
    _, buf = constraint_words(2, 10000, [(200, 50, 0, M)])
    assert not buf.any()


def test_skip_out_of_range_word():  # :3042-3058
    _, buf = constraint_words(2, 10000, [(200, 50, 0, 100)])
    assert not buf.any()


def test_clock_period_saturation():  # :3060-3064
    clk, _ = constraint_words(1, (1 << 64) - 1, [])
    assert clk == 0xFFFFFFFF


def test_clock_arrival_folds_into_effective_setup_hold():  # :3290-3302
    _, buf = constraint_words(1, 10000, [(200, 50, 80, 0)])
    assert (buf[0] >> 16, buf[0] & 0xFFFF) == (120, 130)


def test_effective_setup_hold_clamps():  # :3306-3321
    _, buf = constraint_words(1, 10000, [(100, 65535 - 10, 500, 0)])
    assert (buf[0] >> 16, buf[0] & 0xFFFF) == (0, 65535)


def test_clock_arrival_fold_in_buffer_and_ir_values():
    # :3327-3344 (buffer keeps the per-DFF fold), :2822-2823 (setup 80 with
    # arrival 100 and setup 85 with arrival 230 clamp to 0, holds 100, 230)
    _, buf = constraint_words(2, 10000, [(200, 50, 80, 0), (80, 0, 100, 32)])
    assert buf[0] == (120 << 16) | 130 and buf[1] == 100
    _, buf = constraint_words(1, 10000, [(85, 0, 230, 3)])
    assert buf[0] == 230


def test_zero_clock_arrival_preserves_legacy_behaviour():  # :3349-3362
    _, buf = constraint_words(1, 10000, [(1234, 567, 0, 0)])
    assert buf[0] == (1234 << 16) | 567


def test_compiled_script_constraint_buffer_uses_the_same_rule():
    from helpers import CONFIG1, design
    nl, comp = design("cfg1", **CONFIG1)
    s = comp.flatten(4)
    buf = s.build_timing_constraint_buffer(5000, 50, 20)
    pos = [p for p in s.dff_data_pos.tolist()]
    clk, words = constraint_words(s.reg_io_state_size, 5000,
                                  [(50, 20, 0, p) for p in pos])
    assert int(buf[0]) == clk == 5000
    assert np.array_equal(buf[1:], words)
    assert words.any()


# --------------------------------------------------- arrival brute force
def children(n):
    """Operand words of tree word n (1 <= n < 256): the stage input layer for
    hier0 words 128..255, else words n + hw and n + 2hw at width hw."""
    if n >= 128:
        return None
    hw = 1 << (n.bit_length() - 1)
    return n + hw, n + 2 * hw


def all_paths(n, orb):
    """Every path from word n down to the stage inputs, as lists of words; a
    pass-through word only continues into its a operand."""
    ch = children(n)
    if ch is None:
        return [[n]]
    a, b = ch
    out = [[n] + p for p in all_paths(a, orb)]
    if orb[n] != M:
        out += [[n] + p for p in all_paths(b, orb)]
    return out


def brute_arrivals(orb, delay):
    """Longest-path arrival of every tree word (no u16 wrap in range)."""
    arr = np.zeros(256, np.int64)
    for n in range(1, 256):
        arr[n] = max(sum(int(delay[m]) for m in p) for p in all_paths(n, orb))
    return arr


def random_tree(seed, pass_p=0.3, dmax=300):
    r = np.random.default_rng(seed)
    xora = r.integers(0, 1 << 32, 256, dtype=np.uint64).astype(np.uint32)
    xorb = r.integers(0, 1 << 32, 256, dtype=np.uint64).astype(np.uint32)
    orb = r.integers(0, 1 << 32, 256, dtype=np.uint64).astype(np.uint32)
    orb[r.random(256) < pass_p] = M
    delay = r.integers(0, dmax + 1, 256).astype(np.uint32)
    return xora, xorb, orb, delay


def wrap_chain():
    """Pass-through chain 1, 2, 4, ..., 128 (each word's a operand) with
    60,000 ps per word: arrival of word 2^k = (8 - k) * 60000 mod 65536."""
    xora = np.zeros(256, np.uint32)
    xorb = np.zeros(256, np.uint32)
    orb = np.zeros(256, np.uint32)
    delay = np.full(256, 7, np.uint32)
    chain = [1, 2, 4, 8, 16, 32, 64, 128]
    for w in chain:
        orb[w] = M
        delay[w] = 60000
    expect = {w: ((8 - k) * 60000) % 65536 for k, w in enumerate(chain)}
    return (xora, xorb, orb, delay), expect


def run_kernel_semantics(lib, script, rio=256):
    """One cycle of the partition through jqo_block_kernel; returns the
    writeout arrivals (arrival words follow the rio value words)."""
    inp = np.zeros(2 * rio, np.uint32)
    inp[:rio] = np.arange(rio, dtype=np.uint32) * np.uint32(2654435761)
    out = inp.copy()
    cnt = ctypes.c_size_t(0)
    rc = lib.jqo_block_kernel(script.ctypes.data, script.size, inp.ctypes.data,
                              out.ctypes.data, None, None, None, 0, rio, 0,
                              None, 0, ctypes.byref(cnt))
    assert rc == 0
    return out[rio:]


CASES = [("random", s) for s in range(6)] + [("dense", 7), ("sparse", 8), ("wrap", 0)]


def case_script(kind, seed):
    if kind == "wrap":
        fl, expect = wrap_chain()
        return tree_probe_script(*fl), None, expect
    p = {"random": 0.3, "dense": 0.0, "sparse": 0.8}[kind]
    fl = random_tree(seed, pass_p=p)
    return tree_probe_script(*fl), brute_arrivals(fl[2], fl[3]), None


def check_case(lib, kind, seed):
    script, brute, expect = case_script(kind, seed)
    got = run_kernel_semantics(lib, script)
    if brute is not None:
        assert got[0] == 0, "word 0 (bit-level tail) keeps arrival 0"
        assert np.array_equal(got[1:].astype(np.int64), brute[1:]), \
            f"{kind}/{seed}: first mismatch at word " \
            f"{int(np.argmax(got[1:].astype(np.int64) != brute[1:])) + 1}"
    else:
        for w, e in expect.items():
            assert int(got[w]) == e, f"wrap chain word {w}"


@pytest.mark.parametrize("kind,seed", CASES)
def test_oracle_arrivals_equal_brute_force(kind, seed):
    check_case(oracle.lib(), kind, seed)


def test_brute_force_counts_paths():
    # with no pass-through word, word 1 has 2^7 paths through hier1..7 and
    # every one ends at a hier0 word; all pass-through: exactly one path
    orb = np.zeros(256, np.uint32)
    assert len(all_paths(1, orb)) == 128
    assert len(all_paths(1, np.full(256, M, np.uint32))) == 1
    assert all_paths(1, np.full(256, M, np.uint32))[0] == [1, 2, 4, 8, 16, 32, 64, 128]


# ------------------------------------------------------- planted mistakes
_ORIG_HIER = (
    "      u16 a = arr[t + hw], b = arr[t + 2 * hw];\n"
    "      u16 base = (ob[t] == 0xFFFFFFFFu) ? a : (a > b ? a : b);\n"
    "      arr[t] = (u16)(base + dly[t]);\n")
_LEVEL_LOOP = ("  for (int hi = 1; hi <= 7; ++hi) {\n"
               "    int hw = 1 << (7 - hi);\n"
               "    for (int t = hw; t < 2 * hw; ++t) {\n"
               "      u16 a = arr[t + hw]")
PLANTS = {
    "swapped_operand": (_ORIG_HIER, _ORIG_HIER.replace("? a :", "? b :")),
    "dropped_pass_delay": (_ORIG_HIER, _ORIG_HIER.replace(
        "(u16)(base + dly[t])", "(u16)(base + (ob[t] == 0xFFFFFFFFu ? 0 : dly[t]))")),
    "missing_level": (_LEVEL_LOOP, _LEVEL_LOOP.replace("hi <= 7", "hi <= 6")),
    "delay_one_level_off": (_ORIG_HIER, _ORIG_HIER.replace("dly[t])", "dly[t + hw])")),
    "saturating_add": (_ORIG_HIER, _ORIG_HIER.replace(
        "(u16)(base + dly[t])", "(u16)(base + dly[t] > 65535 ? 65535 : base + dly[t])")),
}


def build_planted(tmp_path, name):
    src = open(os.path.join(ORACLE_DIR, "jq_oracle.c")).read()
    old, new = PLANTS[name]
    assert src.count(old) == 1, "the oracle's arrival recurrence changed shape"
    d = tmp_path / name
    d.mkdir()
    shutil.copy(os.path.join(ORACLE_DIR, "jq_oracle.h"), d / "jq_oracle.h")
    (d / "jq_oracle.c").write_text(src.replace(old, new))
    so = d / "libplanted.so"
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-shared", "-std=c11", "-fopenmp",
                           "-o", str(so), str(d / "jq_oracle.c")])
    L = ctypes.CDLL(str(so))
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.jqo_block_kernel.argtypes = [vp, sz, vp, vp, vp, vp, vp, ctypes.c_uint32,
                                   ctypes.c_int32, ctypes.c_uint32, vp, sz, vp]
    L.jqo_block_kernel.restype = ctypes.c_int
    return L


@pytest.mark.parametrize("name", sorted(PLANTS))
def test_planted_arrival_mistake_is_caught(tmp_path, name):
    L = build_planted(tmp_path, name)
    caught = []
    for kind, seed in CASES:
        try:
            check_case(L, kind, seed)
        except AssertionError:
            caught.append((kind, seed))
    assert caught, f"planted mistake {name} passed every arrival check"
