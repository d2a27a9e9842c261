"""End-to-end golden VCDs produced by Jacquard itself.

tests/golden/ holds the reference's fixture files verbatim (netlist, input
VCD and the reference's output VCD; from tests/timing_test/ and
tests/assert_test/).  The host path (AIGPDK Verilog reader, compiler,
reference VCD reader, VCD writer) plus the CPU oracle must reproduce the
output files byte for byte; the GPU variant runs the same flow through the
HIP library.  These pin the PO lag convention, the data-immediate /
clock-deferred input semantics and the MSB-first partial-vector quirk.
"""
import os

import numpy as np
import pytest

import oracle
from jacquard_amd import sim, vcd
from jacquard_amd.netlist import Netlist
from jacquard_amd.script import Compiled

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = [
    ("timing_test", "dff_test_synth.gv", "dff_test.vcd", "gem_output_fixed.vcd"),
    ("assert_test", "assert_synth.gv", "assert_test.vcd", "assert_test_output.vcd"),
]


def _flow(case, blocks, runner):
    d, gv, iv, gold = case
    base = os.path.join(HERE, d)
    nl = Netlist.from_verilog(os.path.join(base, gv))
    s = Compiled(nl).flatten(blocks)
    with open(os.path.join(base, iv)) as f:
        states, offs, hdr = vcd.parse_input_vcd(f.read(), s)
    states = runner(s, states, len(offs))
    out = vcd.write_output_vcd(s, offs, states, hdr.timescale)
    with open(os.path.join(base, gold)) as f:
        return out, f.read()


def _oracle(s, states, n):
    oracle.run(s, states, n)
    return states


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("blocks", [1, 3])
def test_golden_vcd_cpu(case, blocks):
    out, gold = _flow(case, blocks, _oracle)
    assert out == gold


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_golden_vcd_gpu(case):
    out, gold = _flow(case, 2, lambda s, st, n: sim.run_scan(s, st, n)["states"])
    assert out == gold


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_golden_vcd_gpu_change_records(case):
    """The same flow with the output changes extracted on the device
    (jq_output_changes) instead of scanning host snapshots."""
    d, gv, iv, gold = case
    base = os.path.join(HERE, d)
    nl = Netlist.from_verilog(os.path.join(base, gv))
    s = Compiled(nl).flatten(2)
    with open(os.path.join(base, iv)) as f:
        states, offs, hdr = vcd.parse_input_vcd(f.read(), s)
    out = sim.run_scan(s, states, len(offs), return_device=True)
    cnt, rec = vcd.output_changes_device(s, out["device"]["states"], offs)
    text = vcd.write_output_vcd_changes(s, offs, cnt, rec, hdr.timescale)
    with open(os.path.join(base, gold)) as f:
        assert text == f.read()


def test_input_vcd_quirks():
    """Snapshot count and timestamps of the dff_test stimulus
    (src/sim/vcd_io.rs:347-455)."""
    base = os.path.join(HERE, "timing_test")
    nl = Netlist.from_verilog(os.path.join(base, "dff_test_synth.gv"))
    s = Compiled(nl).flatten(1)
    with open(os.path.join(base, "dff_test.vcd")) as f:
        st, offs, hdr = vcd.parse_input_vcd(f.read(), s)
    assert [t for _, t in offs] == [vcd.U64_MAX, 0, 10000, 40000, 70000, 100000]
    assert [o for o, _ in offs] == [(k + 1) * s.reg_io_state_size for k in range(6)]
    assert st.size == 7 * s.reg_io_state_size
    assert hdr.timescale == (1, "ps")
