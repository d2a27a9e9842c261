"""Synthetic stimulus: per-lane toggling primary inputs.

Each data input bit k of lane (seed, thresh) toggles at cycle c >= 1 when
``hash(seed, c, k) < thresh`` (thresh = round(p * 2**32)); every input starts
at 0 and every posedge clock flag is 1 in every cycle, so each simulated
cycle is one active clock edge.  The hash is Chris Wellons' "lowbias32"
integer mixer applied twice; the batched kernel evaluates the same function
on the device, so plain, batched and CPU runs see identical inputs.
"""
import numpy as np

GOLDEN = np.uint32(0x9E3779B9)


def mix32(x):
    x = np.asarray(x, np.uint32)
    with np.errstate(over="ignore"):
        x = x ^ (x >> np.uint32(16))
        x = (x * np.uint32(0x7FEB352D)).astype(np.uint32)
        x = x ^ (x >> np.uint32(15))
        x = (x * np.uint32(0x846CA68B)).astype(np.uint32)
        x = x ^ (x >> np.uint32(16))
    return x


def toggle_hash(seed, cycles, bits):
    """hash for every (cycle, bit): uint32[len(cycles), len(bits)]."""
    with np.errstate(over="ignore"):
        h1 = mix32((np.uint32(seed) + GOLDEN * np.asarray(bits, np.uint32)).astype(np.uint32))
    return mix32(h1[None, :] ^ np.asarray(cycles, np.uint32)[:, None])


def threshold(p):
    return np.uint32(min(int(round(float(p) * 4294967296.0)), 0xFFFFFFFF))


def input_roles(script):
    """(data_bit_positions, clock_flag_positions) of the state's primary
    input area.  Data bits are the netlist's non-clock input pins in pin
    order (independent of partitioning); clock flags are posedge flags."""
    ins, _ = script.netlist.ports()
    clock_pins = {pin for (_, pin, flag) in ins if flag}
    flags = [flag for (_, _, flag) in ins if flag]
    A = script.netlist.arrays()
    data, fl = [], []
    for pin in script.input_layout.tolist():
        pos = script.input_map[pin]
        if A["drv"][pin] == 3:  # clock flag
            if pin in flags:
                fl.append(pos)
        elif pin not in clock_pins:
            data.append(pos)
    return np.asarray(data, np.uint32), np.asarray(fl, np.uint32)


def lane_inputs(n_data, cycles, seed, p, chunk=8192):
    """bool[cycles+1, n_data]: input values of snapshots 0..cycles."""
    thr = threshold(p)
    out = np.zeros((cycles + 1, n_data), bool)
    cur = np.zeros(n_data, bool)
    bits = np.arange(n_data, dtype=np.uint32)
    for c0 in range(1, cycles + 1, chunk):
        c1 = min(cycles + 1, c0 + chunk)
        t = toggle_hash(seed, np.arange(c0, c1, dtype=np.uint32), bits) < thr
        v = np.logical_xor.accumulate(t, axis=0) ^ cur[None, :]
        out[c0:c1] = v
        cur = v[-1]
    return out


def fill_input_words(states, state_size, data_pos, flag_pos, values):
    """OR the packed input bits into every snapshot of `states`."""
    S = states.reshape(-1, state_size)
    n = values.shape[0]
    for k, pos in enumerate(data_pos.tolist()):
        w, b = pos >> 5, pos & 31
        S[:n, w] |= values[:, k].astype(np.uint32) << np.uint32(b)
    for pos in flag_pos.tolist():
        S[:n, pos >> 5] |= np.uint32(1) << np.uint32(pos & 31)


def input_word_count(script):
    """Words of the state that hold primary inputs (and the constant-zero
    padding bit), src/flatten.rs:1163-1170."""
    n = len(script.input_layout)
    w = (n + 31) // 32
    return w + 1 if n == w * 32 else w


def input_words(script, cycles, seed=1, p=0.5, first_cycle=0):
    """uint32[cycles+1, input_word_count]: the input area of snapshots
    first_cycle .. first_cycle+cycles."""
    nw = input_word_count(script)
    data, fl = input_roles(script)
    vals = lane_inputs(len(data), first_cycle + cycles, seed, p)[first_cycle:]
    words = np.zeros((cycles + 1) * nw, np.uint32)
    fill_input_words(words, nw, data, fl, vals)
    return words.reshape(cycles + 1, nw)


def build_states(script, cycles, seed=1, p=0.5, initial_xmask=None):
    """(cycles+1) x effective_state_size snapshot buffer for the plain
    drop-in: input bits per snapshot, X masks (reference template, or an
    explicit snapshot-0 mask), zeroed arrival section."""
    ss = script.effective_state_size()
    rio = script.reg_io_state_size
    states = np.zeros((cycles + 1) * ss, np.uint32)
    iw = input_words(script, cycles, seed, p)
    states.reshape(-1, ss)[:, :iw.shape[1]] = iw
    if script.xprop_enabled:
        S = states.reshape(-1, ss)
        S[:, rio:2 * rio] = script.xmask_template()[None, :]
        if initial_xmask is not None:
            S[0, rio:2 * rio] = initial_xmask
    return states


def dff_initial_xmask(script, fraction, seed=3):
    """Explicit snapshot-0 X mask with a seeded `fraction` of the DFFs X
    (and the template X outside the input area)."""
    rng = np.random.default_rng(seed)
    m = script.xmask_template()
    pos = script.dff_data_pos
    valid = pos != 0xFFFFFFFF
    pick = valid & (rng.random(pos.size) < fraction)
    for q in pos[pick].tolist():
        m[q >> 5] |= np.uint32(1) << np.uint32(q & 31)
    return m, pick
