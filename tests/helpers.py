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
