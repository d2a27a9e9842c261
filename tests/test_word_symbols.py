"""Symbolic violation sites: WordSymbolMap (src/flatten.rs:125-178,
1848-1879).  The describe_word strings are the reference's own expected
values (src/flatten.rs:3525-3592); the map built from a script places each
DFF at the state bit its data input is written to."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

sys.path.insert(0, os.path.join(HERE, "..", "oracle"))

from jacquard_amd.events import WordSymbolMap, build_report  # noqa: E402


def test_describe_word_reference_cases():
    assert WordSymbolMap({}).describe_word(7, 4) == "<no DFF in word> [word=7]"
    assert (WordSymbolMap({3: [("top/cpu/state", 0)]}).describe_word(3, 4)
            == "top/cpu/state[bit 0] [word=3]")
    m = WordSymbolMap({5: [("top/regs[0]", 0), ("top/regs[1]", 1), ("top/regs[2]", 2)]})
    assert (m.describe_word(5, 4)
            == "top/regs[0][bit 0], top/regs[1][bit 1], top/regs[2][bit 2] [word=5]")
    m = WordSymbolMap({1: [(f"top/r{b}", b) for b in range(6)]})
    assert m.describe_word(1, 2) == "top/r0[bit 0], top/r1[bit 1], +4 more [word=1]"
    m = WordSymbolMap({0: [("a", 0), ("b", 1)]})
    assert m.describe_word(0, 0).startswith("a[bit 0], +1 more")
    m = WordSymbolMap({0: [("a", 0), ("b", 1)], 1: [("c", 0)]})
    assert m.num_words() == 2 and m.num_sites() == 3
    # sites are listed by bit whatever order they were given in
    assert WordSymbolMap({2: [("y", 9), ("x", 3)]}).describe_word(2) == \
        "x[bit 3], y[bit 9] [word=2]"


def test_map_from_script_names_dff_instances():
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    from test_cosim import GPIO_DESIGN
    nl = Netlist.from_verilog(text=GPIO_DESIGN)
    s = Compiled(nl).flatten(1)
    assert nl.dff_names() == ["r0", "r1"]
    d = nl.arrays()["dff_d"].tolist()
    m = WordSymbolMap.from_script(s)
    pos = [s.output_map[x] for x in d]
    for name, p in zip(["r0", "r1"], pos):
        assert (name, p % 32) in m.word_to_dffs[p // 32]
    # a report resolves its sites through the map
    import numpy as np
    from oracle import EVENT_DTYPE
    ev = np.zeros(1, EVENT_DTYPE)
    ev[0]["event_type"] = 4
    ev[0]["cycle"] = 3
    ev[0]["data"] = [pos[0] // 32, np.uint32(np.int32(-5).view(np.uint32)), 60, 50]
    rep = build_report(ev, 10, 100, describe=m.describe_word)
    assert rep["violations"][0]["site"] == m.describe_word(pos[0] // 32)
    assert "r0[bit" in rep["violations"][0]["site"]


def test_generated_netlists_fall_back_to_dff_numbers():
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.random(seed=1, gates=500, depth=8, dffs=40, inputs=8, outputs=4)
    s = Compiled(nl).flatten(1)
    m = WordSymbolMap.from_script(s)
    assert m.num_sites() == 40
    names = {n for v in m.word_to_dffs.values() for n, _ in v}
    assert names == {f"dff#{j}" for j in range(40)}
