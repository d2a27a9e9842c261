"""The native streaming reader of input VCDs (csrc/host/vcd.cpp) against the
pure-Python restatement in oracle/vcd_oracle.py, on the reference's own
input VCD fixtures and on random value-change sections that exercise every
rule of src/sim/vcd_io.rs:347-455: repeated timestamps, data versus clock
pins, deferred clock changes, posedge flags, partial vectors, x / z, real
values, comments, unknown codes and the clock-edge limit."""
import os
import random

import numpy as np
import pytest

import vcd_oracle
from jacquard_amd import vcd
from jacquard_amd.netlist import Netlist
from jacquard_amd.script import Compiled

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = [("timing_test", "dff_test_synth.gv", "dff_test.vcd"),
            ("assert_test", "assert_synth.gv", "assert_test.vcd")]


def _same(body, code_map, flag_bits, rio, limit=None):
    got, goff = vcd.parse_body(body, code_map, flag_bits, rio, limit)
    ref, roff = vcd_oracle.parse_body(body, code_map, flag_bits, rio, limit)
    assert goff == roff
    assert np.array_equal(got, ref)
    return got, goff


@pytest.mark.parametrize("case", FIXTURES, ids=[c[0] for c in FIXTURES])
@pytest.mark.parametrize("limit", [None, 1, 3])
def test_fixture_inputs(case, limit):
    d, gv, iv = case
    nl = Netlist.from_verilog(os.path.join(HERE, d, gv))
    s = Compiled(nl).flatten(1)
    with open(os.path.join(HERE, d, iv)) as f:
        text = f.read()
    _header, body, code_map, flag_bits, rio = vcd.input_plan(text, s)
    states, offs = _same(body, code_map, flag_bits, rio, limit)
    assert len(states) == rio * (len(offs) + 1)
    if limit:
        assert len(offs) <= limit


def _random_section(rng, codes, n_steps):
    out = []
    t = 0
    for _ in range(n_steps):
        r = rng.random()
        if r < 0.25:
            t += rng.choice([0, 0, 1, 5, 10])  # 0: a repeated timestamp
            out.append(f"#{t}")
        elif r < 0.3:
            out.append("$comment " + " ".join(rng.choice(["a", "#9", "1!", "b1 x"])
                                               for _ in range(3)) + " $end")
        elif r < 0.33:
            out.append(rng.choice(["$dumpvars", "$end", "$dumpall"]))
        elif r < 0.36:
            out.append(f"r{rng.random():.3f} {rng.choice(codes)}")
        else:
            code = rng.choice(codes + ["??"])
            width = rng.randint(1, 6)
            val = "".join(rng.choice("0101xz") for _ in range(width))
            if width == 1 and rng.random() < 0.5:
                out.append(f"{val}{code}")
            else:
                out.append(f"{rng.choice('bB')}{val} {code}")
    sep = rng.choice([" ", "\n", "\n\t ", "\r\n"])
    return sep.join(out) + "\n"


@pytest.mark.parametrize("seed", range(60))
def test_random_sections(seed):
    rng = random.Random(seed)
    rio = rng.randint(1, 4)
    nbits = 32 * rio
    codes = ["!", '"', "#a", "%%", "ab", "z9"][: rng.randint(2, 6)]
    free = list(range(nbits))
    rng.shuffle(free)
    flag_bits = [free.pop() for _ in range(rng.choice([0, 1, 1, 2, 2]))]
    code_map = {}
    for c in codes:
        m = {}
        for pos in range(rng.randint(1, 6)):
            if rng.random() < 0.2 or not free:
                continue  # a position with no input behind it
            flag = rng.choice(flag_bits) if flag_bits and rng.random() < 0.5 else None
            m[pos] = (free.pop(), flag)
        code_map[c] = m
    body = _random_section(rng, codes, rng.randint(0, 300))
    limit = rng.choice([None, None, 1, 2, 5])
    _same(body, code_map, flag_bits, rio, limit)


def test_empty_and_header_only_sections():
    code_map = {"!": {0: (3, None)}}
    for body in ["", "\n", "$dumpvars $end", "#0"]:
        states, offs = _same(body, code_map, [], 1)
        assert offs == [] and states.size == 1


def test_bad_timestamp_raises():
    with pytest.raises(ValueError):
        vcd.parse_body("#12x 1!", {"!": {0: (0, None)}}, [], 1)
    # a time that does not fit 64 bits is an error, not a wrapped value
    with pytest.raises(ValueError):
        vcd.parse_body("#18446744073709551616 1!", {"!": {0: (0, None)}}, [], 1)
    states, offs = vcd.parse_body("#0 1! #18446744073709551615 0!",
                                  {"!": {0: (0, None)}}, [], 1)[:2]
    assert states.tolist() == [1, 0] and len(offs) == 1
