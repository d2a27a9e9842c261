"""Compiler + oracle against the direct AIG simulation (closed form).

For any valid script the interpreted values must equal a direct cycle
simulation of the AIG, which never looks at the script; this pins the
compiler (levelization, level split, partitioning, mapping, flatten) and the
oracle together.  Also checks the literal 2-state restatement against the
kernel-semantics oracle, the script validator, and layout invariants of the
reference format.
"""
import numpy as np
import pytest

import oracle
from helpers import CONFIG1, design, direct_check
from jacquard_amd import stimulus
from jacquard_amd.script import SRAM_WORDS


@pytest.mark.parametrize("blocks,split", [(1, 0), (4, 0), (8, 6), (32, 3)])
def test_config1_vs_direct(blocks, split):
    nl, comp = design("c1", level_split=split, **CONFIG1)
    s = comp.flatten(blocks)
    s.validate()
    cycles = 120
    st = stimulus.build_states(s, cycles, seed=3, p=0.5)
    a, b = st.copy(), st.copy()
    oracle.run(s, a, cycles)
    oracle.run_v1(s, b, cycles)
    assert np.array_equal(a, b)
    direct_check(s, a, cycles, 3, 0.5)


@pytest.mark.parametrize("params", [
    dict(seed=21, gates=6000, depth=30, dffs=600, inputs=48, outputs=24,
         geometric_p=0.3, srams=2, level_split=5),
    dict(seed=22, gates=4000, depth=12, dffs=300, inputs=8, outputs=8,
         geometric_p=0.0, srams=1, level_split=0),
])
def test_srams_xprop_vs_direct(params):
    nl, comp = design("mix", **params)
    for xprop in (False, True):
        s = comp.flatten(8, xprop=xprop)
        s.validate()
        cycles = 60
        m, pick = (None, None)
        if xprop:
            m, pick = stimulus.dff_initial_xmask(s, 0.1, seed=4)
        st = stimulus.build_states(s, cycles, seed=5, p=0.3, initial_xmask=m)
        sram = np.zeros(s.sram_storage_size, np.uint32)
        sram_x = np.full(s.sram_storage_size, 0xFFFFFFFF, np.uint32) if xprop else None
        oracle.run(s, st, cycles, sram=sram, sram_x=sram_x)
        direct_check(s, st, cycles, 5, 0.3, xprop=xprop,
                     dff_init_x=None if pick is None else pick.astype(np.uint8))


def test_layout_invariants():
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    bs, bd = s.blocks_start, s.blocks_data
    assert bs.size == s.num_blocks * s.num_major_stages + 1
    assert bd.size == int(bs[-1]) + 2048          # padding, flatten.rs:1341
    nin = len(s.input_layout)
    start = (nin + 31) // 32 + (1 if nin % 32 == 0 else 0)
    # writeouts begin after the input words (flatten.rs:1163-1170)
    ios = sorted((int(bd[o + 3]), int(bd[o + 2])) for o in s.parts[:, 3])
    assert ios[0][0] == start
    for (a, n), (b, _) in zip(ios, ios[1:]):
        assert a + n == b                          # contiguous, block order
    assert ios[-1][0] + ios[-1][1] == s.reg_io_state_size
    for o in s.parts[:, 3]:
        S = int(bd[o])
        R = int(bd[o + 6])
        assert int(s.parts[list(s.parts[:, 3]).index(o), 4]) == 256 + 512 * R + 5120 * S + 10240


def test_validator_rejects_corruption():
    nl, comp = design("c1", level_split=0, **CONFIG1)
    s = comp.flatten(2)
    s.validate()
    o = int(s.parts[0, 3])
    saved = int(s.blocks_data[o + 3])
    s.blocks_data[o + 3] = s.reg_io_state_size  # writeout past the state
    with pytest.raises(ValueError):
        s.validate()
    s.blocks_data[o + 3] = saved
    g = o + 256
    saved = int(s.blocks_data[g + 1])
    s.blocks_data[g + 1] = 0xFFFFFFFF
    s.blocks_data[g] = s.reg_io_state_size + 5  # global read past the state
    with pytest.raises(ValueError):
        s.validate()


def test_sram_storage_layout():
    nl, comp = design("mix", seed=21, gates=6000, depth=30, dffs=600, inputs=48,
                      outputs=24, geometric_p=0.3, srams=2, level_split=5)
    s = comp.flatten(4)
    assert s.sram_storage_size == 2 * SRAM_WORDS
    assert sorted(s.sram_offsets.tolist()) == [0, SRAM_WORDS]


@pytest.mark.parametrize("xprop", [False, True])
def test_merged_partitions_vs_direct(xprop):
    """Partition merge (Compiled.merged): every stage's partitions merged in
    pairs where they fit one partition; the merged design has fewer
    partitions, passes the validator, and its interpreted values (and X)
    equal the direct AIG simulation."""
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.random(seed=21, gates=20000, depth=30, dffs=1500, inputs=48, outputs=24,
                        geometric_p=0.3, srams=2)
    comp = Compiled(nl, level_split=5, max_level_nodes=100)  # ~280 narrow partitions
    pps = comp.parts_per_stage()
    pairs = [(m, i, i + 1) for m, n in enumerate(pps) for i in range(0, n - 1, 2)]
    assert pairs
    mc, ok = comp.merged(pairs)
    assert ok.sum() > 100
    assert sum(mc.parts_per_stage()) == sum(pps) - int(ok.sum())
    # every endpoint is still produced by exactly one partition
    assert [sum(c) for c in mc.endpoint_counts()] == [sum(c) for c in comp.endpoint_counts()]
    s = mc.flatten(4, xprop=xprop)
    s.validate()
    cycles = 50
    m, pick = (None, None)
    if xprop:
        m, pick = stimulus.dff_initial_xmask(s, 0.1, seed=4)
    st = stimulus.build_states(s, cycles, seed=6, p=0.3, initial_xmask=m)
    sram = np.zeros(s.sram_storage_size, np.uint32)
    sram_x = np.full(s.sram_storage_size, 0xFFFFFFFF, np.uint32) if xprop else None
    oracle.run(s, st, cycles, sram=sram, sram_x=sram_x)
    direct_check(s, st, cycles, 6, 0.3, xprop=xprop,
                 dff_init_x=None if pick is None else pick.astype(np.uint8))


def test_merge_rejects_malformed_pairs():
    nl, comp = design("c1", level_split=6, **CONFIG1)
    n = comp.parts_per_stage()[0]
    with pytest.raises(RuntimeError):
        comp.merged([(0, 0, n)])
    with pytest.raises(RuntimeError):
        comp.merged([(len(comp.parts_per_stage()), 0, 1)])
    if n >= 3:
        with pytest.raises(RuntimeError):
            comp.merged([(0, 0, 1), (0, 1, 2)])


def test_one_boomerang_stage_per_partition_vs_direct():
    """max_boomerang_stages=1 splits every partition the mapper would give
    two or more boomerang stages, down to single endpoints (one endpoint
    whose cone is deeper than a stage keeps its stages); the values stay
    those of the direct AIG simulation."""
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.random(seed=23, gates=20000, depth=40, dffs=1500, inputs=48, outputs=24,
                        geometric_p=0.3, srams=1)
    free = Compiled(nl, level_split=10)
    assert max(max(c) for c in free.stage_counts()) > 1
    comp = Compiled(nl, level_split=10, max_boomerang_stages=1)
    for stages, eps in zip(comp.stage_counts(), comp.endpoint_counts()):
        assert all(st == 1 or ep == 1 for st, ep in zip(stages, eps))
    assert sum(comp.parts_per_stage()) > sum(free.parts_per_stage())
    s = comp.flatten(8)
    s.validate()
    cycles = 40
    st = stimulus.build_states(s, cycles, seed=8, p=0.3)
    oracle.run(s, st, cycles, sram=np.zeros(s.sram_storage_size, np.uint32))
    direct_check(s, st, cycles, 8, 0.3)


@pytest.mark.parametrize("xprop", [False, True])
def test_split_partitions_vs_direct(xprop):
    """Partition cut (Compiled.split): the largest partition of every stage
    mapped again as two halves of its endpoints; the design has more
    partitions, passes the validator, and its values (and X) equal the
    direct AIG simulation."""
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.random(seed=24, gates=20000, depth=30, dffs=1500, inputs=48, outputs=24,
                        geometric_p=0.3, srams=2)
    comp = Compiled(nl, level_split=5)
    pps = comp.parts_per_stage()
    eps = comp.endpoint_counts()
    parts = [(m, int(np.argmax(e))) for m, e in enumerate(eps) if max(e) > 1]
    assert parts
    sc, ok = comp.split(parts)
    assert ok.any()
    assert sum(sc.parts_per_stage()) == sum(pps) + int(ok.sum())
    assert [sum(c) for c in sc.endpoint_counts()] == [sum(c) for c in eps]
    s = sc.flatten(4, xprop=xprop)
    s.validate()
    cycles = 50
    m, pick = (None, None)
    if xprop:
        m, pick = stimulus.dff_initial_xmask(s, 0.1, seed=4)
    st = stimulus.build_states(s, cycles, seed=7, p=0.3, initial_xmask=m)
    sram = np.zeros(s.sram_storage_size, np.uint32)
    sram_x = np.full(s.sram_storage_size, 0xFFFFFFFF, np.uint32) if xprop else None
    oracle.run(s, st, cycles, sram=sram, sram_x=sram_x)
    direct_check(s, st, cycles, 7, 0.3, xprop=xprop,
                 dff_init_x=None if pick is None else pick.astype(np.uint8))
    with pytest.raises(RuntimeError):
        comp.split([(0, 0), (0, 0)])


def _wide_and_verilog(n):
    """A module whose one output is the AND of n inputs, as a balanced tree
    of AND2 cells (depth ceil(log2 n))."""
    lines = [f"module wide(x, y);", f"  input [{n - 1}:0] x;", "  output y;"]
    level = [f"x[{i}]" for i in range(n)]
    k = 0
    while len(level) > 1:
        nxt = []
        for i in range(0, len(level) - 1, 2):
            w = f"w{k}"
            k += 1
            lines.append(f"  wire {w};")
            lines.append(f"  AND2_00_0 g{k} (.A({level[i]}), .B({level[i + 1]}), .Y({w}));")
            nxt.append(w)
        if len(level) % 2:
            nxt.append(level[-1])
        level = nxt
    lines.append(f"  assign y = {level[0]};")
    lines.append("endmodule")
    return "\n".join(lines) + "\n"


def test_an_endpoint_wider_than_a_partition_is_an_error_and_the_compiler_recovers():
    """One output over 9,000 inputs in a single major stage needs more than
    the 8,192 input bits of a partition: mapping fails in a worker task and
    the compile raises (it must not end the process).  The same compiler
    then compiles a netlist that fits, and its script is exact."""
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    wide = Netlist.from_verilog(text=_wide_and_verilog(9000))
    with pytest.raises(RuntimeError, match="does not fit a partition"):
        Compiled(wide, level_split=0)
    fits = Netlist.from_verilog(text=_wide_and_verilog(4000))
    s = Compiled(fits, level_split=0).flatten(1)
    cycles = 6
    st = stimulus.build_states(s, cycles, seed=2, p=0.999)
    oracle.run(s, st, cycles)
    direct_check(s, st, cycles, 2, 0.999)


def test_blocks_data_outlives_its_script_object():
    """blocks_data is a view of the host library's buffer: it keeps that
    buffer alive after the script object that made it is gone (it used to
    dangle, so reading it then touched freed memory)."""
    import gc
    from jacquard_amd.netlist import Netlist
    from jacquard_amd.script import Compiled
    nl = Netlist.random(seed=5, gates=20000, depth=24, dffs=512, inputs=32, outputs=16,
                        geometric_p=0.3)
    comp = Compiled(nl, level_split=4)
    views = [comp.flatten(16).blocks_data for _ in range(3)]
    gc.collect()
    # same-size blocks, so the allocator hands out any freed buffer again
    churn = [np.full(views[0].size, 0xA5A5A5A5, np.uint32) for _ in range(16)]
    ref = comp.flatten(16)
    for v in views:
        assert np.array_equal(v, ref.blocks_data)
    del churn
