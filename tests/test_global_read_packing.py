"""The global-read packing of flattened scripts against a restatement of
its rule (the contract of map_global_read_to_rounds, src/flatten.rs:
343-400: at most 32 bits per thread).

For every partition the (word | flag, mask) entries are read back from the
script's global-read section and merged per word, which gives the packer's
input.  The restatement then places them as a plain scan over all 256
threads would: entries by decreasing bit count (stable), each to the thread
with the fewest entries, then the most bits that still leave room, then the
lowest index; an entry that fits no thread is split, its lower half of set
bits first.  The script's per-thread entry lists must be exactly those."""
import numpy as np
import pytest

from jacquard_amd.netlist import Netlist
from jacquard_amd.script import Compiled


def _restated(entries):
    ent = sorted(entries, key=lambda e: -bin(e[1]).count("1"))  # stable
    work = list(reversed(ent))
    th = [[] for _ in range(256)]
    size = np.zeros(256, np.int64)
    bits = np.zeros(256, np.int64)
    while work:
        idx, mask = work.pop()
        pc = bin(mask).count("1")
        ok = bits + pc <= 32
        if not ok.any():
            m, lo = mask, 0
            for _ in range(pc // 2):
                low = m & -m
                lo |= low
                m ^= low
            assert lo, "a one-bit entry always fits an empty thread"
            work.append((idx, m))
            work.append((idx, lo))
            continue
        # fewest entries, then most bits, then lowest index
        key = np.where(ok, size * 64 - bits, np.iinfo(np.int64).max)
        t = int(np.argmin(key))
        th[t].append((idx, mask))
        size[t] += 1
        bits[t] += pc
    return th


def _sections(script):
    bd = np.asarray(script.blocks_data, np.uint32)
    for rec in np.asarray(script.parts, np.uint64):
        off = int(rec[3])
        rounds = int(bd[off + 6])
        gr = bd[off + 256: off + 256 + 512 * rounds].reshape(rounds, 256, 2)
        th = [[(int(gr[r, t, 0]), int(gr[r, t, 1])) for r in range(rounds) if gr[r, t, 1]]
              for t in range(256)]
        yield th


@pytest.mark.parametrize("kw,split,cap", [
    (dict(seed=3, gates=90000, depth=24, dffs=1024, inputs=64, outputs=32,
          geometric_p=0.0, invert_prob=0.4), 4, 0),
    (dict(seed=8, gates=60000, depth=30, dffs=2048, inputs=64, outputs=32,
          geometric_p=0.3, invert_prob=0.5, srams=2), 6, 600),
])
def test_global_read_sections_follow_the_packing_rule(kw, split, cap):
    nl = Netlist.random(**kw)
    comp = Compiled(nl, level_split=split, max_level_nodes=cap)
    s = comp.flatten(max(comp.parts_per_stage()))
    checked = wide = 0
    for th in _sections(s):
        merged = {}
        for lst in th:
            for idx, mask in lst:
                merged[idx] = merged.get(idx, 0) | mask
        ent = sorted(merged.items())
        assert th == _restated(ent)
        checked += 1
        wide += max(len(lst) for lst in th) > 1
    assert checked > 20 and wide > 5, "partitions with several rounds were checked"
