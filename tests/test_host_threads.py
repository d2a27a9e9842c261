"""The host compile pipeline's parallel paths give the same script as one
thread.

Every major stage's grouping is a task and so is the mapping of every group
it closes, and partition merge and split map their candidates in parallel;
flattening places every partition of a major stage in parallel, records its
map writes and applies them in partition order, then builds and copies the
sections in parallel.  With OMP_NUM_THREADS=1 all of that runs in program
order, so comparing a run with one thread against runs with several checks
that no result depends on which thread finishes first: every byte of the
script, the three maps, the SRAM offsets, the DFF data positions, the
partition records, X capability and the injected delays, and the merged and
split designs' scripts."""
import hashlib
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

_CHILD = r"""
import hashlib, json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from jacquard_amd.netlist import Netlist
from jacquard_amd.script import Compiled

def h(*arrs):
    m = hashlib.sha256()
    for a in arrs:
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()

def dh(d):
    return h(np.array(sorted(d.items()), np.uint64).reshape(-1, 2))

out = {}
for name, kw, split, nb, xp in [
        ("x_sram_delays", dict(seed=3, gates=30000, depth=24, dffs=512, inputs=32,
                               outputs=16, geometric_p=0.3, srams=4, delay=(5, 40)),
         4, 24, True),
        ("wide", dict(seed=5, gates=200000, depth=40, dffs=4096, inputs=64, outputs=64,
                      geometric_p=0.3, invert_prob=0.5, srams=2), 5, 64, True)]:
    nl = Netlist.random(**kw)
    c = Compiled(nl, level_split=split)
    s = c.flatten(nb, xprop=xp)
    s.inject_timing()
    pps = c.parts_per_stage()
    mg, merged = c.merged([(m, 0, 1) for m in range(len(pps)) if pps[m] >= 2])
    sp, split = c.split([(m, 0) for m in range(len(pps))])
    out[name] = [h(s.blocks_data), h(s.blocks_start), h(s.input_layout),
                 dh(s.input_map), dh(s.output_map), dh(s.staged_io_map),
                 h(s.parts), h(s.partition_x_capable), h(s.sram_offsets),
                 h(s.dff_data_pos), s.num_parts, s.reg_io_state_size, pps,
                 h(mg.flatten(nb).blocks_data), h(np.asarray(merged, np.uint8)),
                 h(sp.flatten(nb).blocks_data), h(np.asarray(split, np.uint8))]
print(json.dumps(out))
"""


def _run(threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, "-c", _CHILD, ROOT], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_one_thread_and_many_threads_build_the_same_script():
    one = _run(1)
    assert one["wide"][10] > 100, "enough partitions to spread over threads"
    for threads in (3, 6):
        assert _run(threads) == one, f"OMP_NUM_THREADS={threads}"
