"""Host-side derived data of the batched form, checked against the script
format itself (src/flatten.rs:198-258): the lane layout of every
boomerang-stage section decodes back to the section's permutation and
flags, and the gather table reproduces the global read's bit placement
(csrc/kernel_v1_impl.cuh:87-151: per thread, rounds in order, mask bits
ascending, the n-th of B bits at word bit B-1-n)."""
import numpy as np

from helpers import CONFIG1, design
from jacquard_amd import batched as B


def _perm(sec, p):
    w = int(sec[1024 * ((p & 31) >> 3) + 4 * (p >> 5) + ((p >> 1) & 3)])
    return (w >> (16 * (p & 1))) & 0xFFFF


def _flag(sec, p, c):
    return (int(sec[4096 + 4 * (p >> 5) + c]) >> (p & 31)) & 1


def test_lane_sections_decode_to_the_script():
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    lb = B.lane_blocks_data(s)
    bd = s.blocks_data
    gates = B._lane_gate_positions()
    rng = np.random.default_rng(0)
    checked = 0
    for off in s.parts[:, 3].tolist():
        S, R = int(bd[off]), int(bd[off + 6])
        assert np.array_equal(lb[off:off + 256 + 512 * R], bd[off:off + 256 + 512 * R])
        base = off + 256 + 512 * R
        for k in range(S):
            a = base + 5120 * k
            sec, ls = bd[a:a + 5120], lb[a:a + 5120]
            for t in rng.choice(256, 12, replace=False).tolist() + [0, 63]:
                for m in range(32):
                    w = m >> 1
                    e = (int(ls[1024 * (w >> 2) + 4 * t + (w & 3)]) >> (16 * (m & 1))) & 0xFFFF
                    p = _perm(sec, t + 256 * m)
                    assert e == p ^ ((p >> 5) & 31)
                for g in range(31):
                    for c in range(3):
                        assert (int(ls[4096 + 4 * t + c]) >> g) & 1 == _flag(sec, int(gates[g, t]), c)
                if t < 64:
                    for i, p in enumerate((128 + t, 192 + t, 64 + t, t)):
                        for c in range(3):
                            assert (int(ls[4096 + 4 * t + 3]) >> (3 * i + c)) & 1 == _flag(sec, p, c)
                checked += 1
        # the writeout sections after the stages are left as they are
        tail = base + 5120 * S
        assert np.array_equal(lb[tail:tail + 10240], bd[tail:tail + 10240])
    assert checked > 0


def test_gather_table_places_bits_like_the_global_read():
    nl, comp = design("c1", level_split=6, **CONFIG1)
    s = comp.flatten(8)
    table, start = B.gather_table(s)
    bd = s.blocks_data
    rng = np.random.default_rng(1)
    rio = s.reg_io_state_size
    prev = rng.integers(0, 2**32, rio, dtype=np.uint64).astype(np.uint32)
    cur = rng.integers(0, 2**32, rio, dtype=np.uint32)
    for k, off in enumerate(s.parts[:, 3].tolist()):
        R = int(bd[off + 6])
        gr = bd[off + 256:off + 256 + 512 * R].reshape(R, 256, 2)
        tab = table[8192 * k:8192 * (k + 1)]
        for t in range(256):
            word = 0
            nb = 0
            for r in range(R):
                idx, mask = int(gr[r, t, 0]), int(gr[r, t, 1])
                src = cur if idx >> 31 else prev
                v = int(src[idx & 0x7FFFFFFF])
                for b in range(32):
                    if mask >> b & 1:
                        word = (word << 1) | ((v >> b) & 1)
                        nb += 1
            for q in range(32):
                e = int(tab[32 * t + q])
                if q >= nb:
                    assert e == B.NO_SOURCE
                    continue
                src = cur if e >> 31 else prev
                bit = (int(src[(e & 0x7FFFFFFF) >> 5]) >> (e & 31)) & 1
                assert bit == (word >> q) & 1, (k, t, q)
    for (m, b, idx, o, l), k in zip(s.parts.tolist(), range(len(s.parts))):
        assert start[m * s.num_blocks + b] <= k
