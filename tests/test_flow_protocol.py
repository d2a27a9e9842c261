"""The dataflow synchronisation of value-only scan launches, checked as a
protocol on the CPU.

Before such a launch, small kernels derive a producer table from the script
(jq_scan.hip, build_flow_tables): for block b and major stage m, the latest
progress it needs of every other block q, c * M + d, where d = s_q + 1 for
a word that q wrote in major stage s_q of the same cycle (an `out` read,
bit 31 of the global-read entry) and d = s_q + 1 - M for a word of the
previous cycle (an `in` read).  Block b publishes c * M + m + 1 once it has
finished major stage m of cycle c.

These tests restate that rule from the script's bytes and run random
schedules that respect only the rule (no barrier): at every step a block
whose producers have progressed far enough is picked at random and runs
its partitions of the next major stage.  Every read must see the write the
barrier schedule gives it (the same cycle's for an `out` read, the
previous cycle's for an `in` read), and some block must always be able to
run.  A rule that waits one stage too little, or not at all for the
previous cycle's words, must let a read run ahead of its write in some
schedule, so the check has teeth."""
import functools

import numpy as np
import pytest

from jacquard_amd.netlist import Netlist
from jacquard_amd.script import Compiled


@functools.lru_cache(maxsize=None)
def _script(gates, depth, split, cap, blocks):
    # narrow partitions (a small per-level cap) spread every major stage
    # over several blocks, so most reads cross blocks
    nl = Netlist.random(seed=31, gates=gates, depth=depth, dffs=512, inputs=32,
                        outputs=16, geometric_p=0.3, invert_prob=0.5)
    return Compiled(nl, level_split=split, max_level_nodes=cap).flatten(blocks)


def _chains(script):
    """{(m, b): [(reads, writes)]}: per partition, the global-read words
    as (word, from_out) and the writeout words."""
    bd = np.asarray(script.blocks_data, np.uint32)
    bs = np.asarray(script.blocks_start, np.uint64)
    nb, M = script.num_blocks, script.num_major_stages
    out = {}
    for m in range(M):
        for b in range(nb):
            off = int(bs[m * nb + b])
            parts = []
            while True:
                ns, last, nio, io, rounds = (int(bd[off + i]) for i in (0, 1, 2, 3, 6))
                if ns == 0:
                    break
                gr = bd[off + 256: off + 256 + 512 * rounds].reshape(rounds, 256, 2)
                reads = [(int(w) & 0x7FFFFFFF, bool(int(w) >> 31))
                         for w, mask in gr.reshape(-1, 2) if mask]
                parts.append((reads, list(range(io, io + nio))))
                if last:
                    break
                off += 256 + 512 * rounds + 5120 * (ns + 2)
            out[(m, b)] = parts
    return out


def _rule(chains, nb, M, out_lag=1, in_reads=True):
    """The producer table: {(b, m): {q: d}}.  out_lag = 1 and in_reads
    are the rule the kernels build; out_lag = 0 waits one stage too little
    for same-cycle words, in_reads = False never waits for the previous
    cycle's."""
    owner = {}
    for (m, b), parts in chains.items():
        for _reads, writes in parts:
            for w in writes:
                assert w not in owner, "a writeout word written twice"
                owner[w] = (b, m)
    deps = {}
    for (m, b), parts in chains.items():
        row = {}
        for reads, _writes in parts:
            for w, from_out in reads:
                if w not in owner:
                    continue  # a primary input
                q, sq = owner[w]
                if q == b:
                    continue  # the block's own earlier work
                if not from_out and not in_reads:
                    continue
                d = sq + out_lag if from_out else sq + 1 - M
                row[q] = max(row.get(q, d), d)
        deps[(b, m)] = row
    return owner, deps


def _run(chains, nb, M, cycles, deps, owner, rng):
    """One random schedule; returns the reads that ran ahead of their write
    (an empty list when the rule held)."""
    prog = [0] * nb
    cursor = [0] * nb  # steps done: cycle c = k // M, major stage m = k % M
    written = set()
    bad = []
    total = cycles * M
    while min(cursor) < total:
        ready = []
        for b in range(nb):
            k = cursor[b]
            if k == total:
                continue
            c, m = divmod(k, M)
            if all(prog[q] >= c * M + d for q, d in deps[(b, m)].items()):
                ready.append(b)
        assert ready, "no block can run: the rule deadlocks"
        b = ready[rng.integers(len(ready))]
        c, m = divmod(cursor[b], M)
        for reads, writes in chains[(m, b)]:
            for w, from_out in reads:
                if w not in owner or owner[w][0] == b:
                    continue
                need = (c, w) if from_out else (c - 1, w)
                if need[0] >= 0 and need not in written:
                    bad.append((b, c, m, w, from_out))
            for w in writes:
                written.add((c, w))
        cursor[b] += 1
        prog[b] = c * M + m + 1
    return bad


@pytest.mark.parametrize("shape", [(40000, 24, 3, 300, 12), (40000, 24, 2, 400, 24),
                                   (30000, 30, 4, 250, 8)])
def test_random_schedules_respect_every_read(shape):
    s = _script(*shape)
    nb, M = s.num_blocks, s.num_major_stages
    chains = _chains(s)
    owner, deps = _rule(chains, nb, M)
    assert sum(len(r) for r in deps.values()) > 100, "most reads cross blocks"
    rng = np.random.default_rng(shape[-1])
    for _ in range(40):
        assert _run(chains, nb, M, 3, deps, owner, rng) == []


@pytest.mark.parametrize("mistake", [dict(out_lag=0), dict(in_reads=False)])
def test_a_mistaken_rule_lets_a_read_run_ahead(mistake):
    s = _script(40000, 24, 3, 300, 12)
    nb, M = s.num_blocks, s.num_major_stages
    chains = _chains(s)
    owner, wrong = _rule(chains, nb, M, **mistake)
    rng = np.random.default_rng(7)
    assert any(_run(chains, nb, M, 3, wrong, owner, rng) for _ in range(40))
