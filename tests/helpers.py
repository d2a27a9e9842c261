"""Shared builders for the parity tests."""
import numpy as np

from jacquard_amd import stimulus
from jacquard_amd.netlist import Netlist
from jacquard_amd.script import Compiled

_cache = {}


def design(key, **kw):
    """Cached (netlist, Compiled) for a generator parameter set."""
    split = kw.pop("level_split", 0)
    k = (key, tuple(sorted(kw.items())), split)
    if k not in _cache:
        nl = Netlist.random(**kw)
        _cache[k] = (nl, Compiled(nl, level_split=split))
    return _cache[k]


CONFIG1 = dict(seed=1, gates=2000, depth=24, dffs=256, inputs=32, outputs=16,
               geometric_p=0.0, invert_prob=0.2)


def direct_check(script, states, cycles, seed, p, xprop=False, dff_init_x=None):
    """Compare every PO and DFF bit of a finished snapshot buffer against the
    direct AIG simulation (mapping-independent closed form)."""
    import oracle
    nl = script.netlist
    A = nl.arrays()
    data, fl = stimulus.input_roles(script)
    vals = stimulus.lane_inputs(len(data), cycles, seed, p)
    pos2pin = {v: k for k, v in script.input_map.items() if A["drv"][k] in (2, 3)}
    pins, bits = [], []
    for k, pos in enumerate(data.tolist()):
        pins.append(pos2pin[pos])
        bits.append(vals[:cycles, k])
    for pos in fl.tolist():
        pins.append(pos2pin[pos])
        bits.append(np.ones(cycles, np.uint8))
    in_bits = np.stack(bits, 1).astype(np.uint8)
    obs, regs = oracle.aig_sim(A, cycles, pins, in_bits, A["pos"], xprop=xprop,
                               dff_init_x=dff_init_x, want_regs=True)
    ss = script.effective_state_size()
    rio = script.reg_io_state_size
    S = states.reshape(-1, ss)
    for k, lit in enumerate(A["pos"].tolist()):
        if lit <= 1:
            continue
        pos = script.output_map[lit]
        got = (S[1:, pos >> 5] >> np.uint32(pos & 31)) & 1
        assert (got == (obs[:, k] & 1)).all(), f"PO {k} value"
        if xprop:
            gx = (S[1:, rio + (pos >> 5)] >> np.uint32(pos & 31)) & 1
            assert (gx == (obs[:, k] >> 1)).all(), f"PO {k} X"
    for d, pos in enumerate(script.dff_data_pos.tolist()):
        if pos == 0xFFFFFFFF:
            continue
        got = (S[1:, pos >> 5] >> np.uint32(pos & 31)) & 1
        assert (got == (regs[:, d] & 1)).all(), f"DFF {d} value"
        if xprop:
            gx = (S[1:, rio + (pos >> 5)] >> np.uint32(pos & 31)) & 1
            assert (gx == (regs[:, d] >> 1)).all(), f"DFF {d} X"


class OneBitRounds:
    """The same script with every partition's global read respread one bit
    per round: thread t's (word, bit) sequence, in the order the pext
    accumulates it, becomes rounds 0, 1, ... of single-bit masks.  The
    accumulated words are unchanged, and partitions that read up to 32 bits
    per thread now use up to 32 rounds, which drives the interpreter's path
    for reads wider than its prefetched rounds."""

    def __init__(self, script):
        self._s = script
        self.max_rounds = 0
        bs = np.asarray(script.blocks_start, np.uint64)
        bd = np.asarray(script.blocks_data, np.uint32)
        pieces, starts = [], [0]
        for k in range(len(bs) - 1):
            pos, end = int(bs[k]), int(bs[k + 1])
            size = 0
            while pos < end:
                h = bd[pos:pos + 256].copy()
                if h[0] == 0:  # empty block
                    pieces.append(bd[pos:end])
                    size += end - pos
                    break
                R = int(h[6])
                gr = bd[pos + 256:pos + 256 + 512 * R].reshape(R, 256, 2)
                seqs = []
                for t in range(256):
                    seq = []
                    for r in range(R):
                        idx, m = int(gr[r, t, 0]), int(gr[r, t, 1])
                        seq += [(idx, 1 << b) for b in range(32) if (m >> b) & 1]
                    seqs.append(seq)
                rn = max(1, max(len(q) for q in seqs)) if R else 0
                ng = np.zeros((rn, 256, 2), np.uint32)
                for t, seq in enumerate(seqs):
                    for j, (idx, m) in enumerate(seq):
                        ng[j, t] = (idx, m)
                h[6] = rn
                self.max_rounds = max(self.max_rounds, rn)
                rest = pos + 256 + 512 * R
                tail = 5120 * (int(h[0]) + 2)
                pieces += [h, ng.reshape(-1), bd[rest:rest + tail]]
                size += 256 + ng.size + tail
                pos = rest + tail
                if h[1]:  # is_last
                    break
            starts.append(starts[-1] + size)
        pieces.append(bd[int(bs[-1]):])  # trailing padding
        self.blocks_data = np.concatenate(pieces).astype(np.uint32)
        self.blocks_start = np.asarray(starts, np.uint64)

    def __getattr__(self, name):
        return getattr(self._s, name)

    def validate(self, state_size=None):
        from jacquard_amd import _native as N
        ss = state_size if state_size is not None else self.effective_state_size()
        bs = np.ascontiguousarray(self.blocks_start, np.uint64)
        bd = np.ascontiguousarray(self.blocks_data, np.uint32)
        rc = N.host().jqh_validate(bs.ctypes.data, bs.size, bd.ctypes.data, bd.size,
                                   self.num_blocks, self.num_major_stages, ss,
                                   max(self.sram_storage_size, 0), 0,
                                   int(self.arrival_state_offset))
        if rc != 0:
            raise ValueError(N.host_error())
