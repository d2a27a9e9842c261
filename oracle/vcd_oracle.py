"""Pure-Python restatement of the input-VCD value-change loop, and of the
output VCD's value-change section.

TEST INFRASTRUCTURE ONLY: the checker for the native streaming reader and
body writers (jacquard_amd/csrc/host/vcd.cpp), imported by tests, never by
the product package.

Follows parse_input_vcd's body loop (src/sim/vcd_io.rs:347-455): a snapshot
is pushed at the first new timestamp after one at which a clock edge
happened (and once at the end), with the offset entry (words so far, time of
the previous active timestamp; u64 max first); posedge flags are cleared
after each pushed snapshot; data pins change at once, a clock pin's change
sets its posedge flag on 0 -> 1 and is applied at the next new timestamp.
Pinned end to end by the reference's golden output VCDs
(tests/timing_test/gem_output_fixed.vcd, tests/assert_test/
assert_test_output.vcd), which the product reproduces byte for byte.
"""
import numpy as np

U64_MAX = (1 << 64) - 1


def parse_body(body, code_map, flag_bits, rio, max_clock_edges=None):
    """Same contract as jacquard_amd.vcd.parse_body."""
    state = [0] * rio

    def get(b):
        return (state[b >> 5] >> (b & 31)) & 1

    def flip(b):
        state[b >> 5] ^= 1 << (b & 31)

    snaps, offs = [], []
    last_time = U64_MAX
    now = 0
    active = True
    pending = []
    toks = body.split()
    i = 0
    while i < len(toks):
        tok = toks[i]
        i += 1
        if tok[0] == "#":
            t = int(tok[1:])
            if t == now:
                continue
            if active:
                snaps.append(list(state))
                offs.append((len(snaps) * rio, last_time))
                for f in flag_bits:
                    state[f >> 5] &= ~(1 << (f & 31))
                if max_clock_edges and len(offs) >= max_clock_edges:
                    break
                last_time = now
            now = t
            active = False
            for b in dict.fromkeys(pending):
                flip(b)
            pending = []
            continue
        if tok[0] == "$":
            if tok == "$comment":
                while i < len(toks) and toks[i] != "$end":
                    i += 1
                i += 1
            continue
        if tok[0] in "bBrR":
            if i >= len(toks):
                break
            value, code = tok[1:], toks[i]
            i += 1
            if tok[0] in "rR":
                continue
        else:
            value, code = tok[0], tok[1:]
        targets = code_map.get(code)
        if not targets:
            continue
        for pos, ch in enumerate(value):
            tgt = targets.get(pos)
            if tgt is None:
                continue
            bit, flag = tgt
            new = 1 if ch == "1" else 0
            old = get(bit)
            if old == new:
                continue
            if flag is None:
                flip(bit)
                continue
            if old == 0:
                active = True
                state[flag >> 5] |= 1 << (flag & 31)
            pending.append(bit)
    snaps.append(list(state))
    return np.array(snaps, np.uint64).astype(np.uint32).reshape(-1), offs


# ---------------------------------------------------------------- writers
# The value-change section of an output VCD (src/sim/vcd_io.rs:758-820),
# restated as plain loops: the checker for the native body writers
# (jqh_vcd_write_values / jqh_vcd_write_changes in csrc/host/vcd.cpp).

def id_code(i):
    """rust-vcd IdCode display: base-94 over '!'..'~', least significant
    digit first, each further digit offset by one."""
    s = []
    while True:
        s.append(chr(33 + i % 94))
        if i < 94:
            return "".join(s)
        i = i // 94 - 1


_SYM = {0: "0", 1: "1", 3: "x"}


def body_from_values(rows, timestamps):
    """rows[r][i] in {0, 1, 3}: every row gets '#t'; column i is written
    when it differs from the value it last wrote (nothing before row 0)."""
    out = []
    last = None
    for row, ts in zip(rows, timestamps):
        if last is None:
            last = [2] * len(row)
        out.append(f"#{ts}")
        for i, v in enumerate(row):
            v = int(v)
            if v == last[i]:
                continue
            last[i] = v
            out.append(f"{_SYM[v]}{id_code(i)}")
    return "".join(line + "\n" for line in out)


def body_from_changes(counts, records, timestamps):
    """Row r writes its counts[r] records (output << 2 | value) in order."""
    out = []
    k = 0
    for r, ts in enumerate(timestamps):
        out.append(f"#{ts}")
        for rec in records[k:k + int(counts[r])]:
            rec = int(rec)
            out.append(f"{_SYM[rec & 3]}{id_code(rec >> 2)}")
        k += int(counts[r])
    return "".join(line + "\n" for line in out)


def _round_half_away(x):
    """f64::round: halves away from zero."""
    import math
    return math.floor(x + 0.5) if x >= 0 else -math.floor(-x + 0.5)


def body_timed(rows, arrivals, timestamps, ps_per_unit):
    """Timed form (src/sim/vcd_io.rs:831-946): a change of column i at row r
    happens at timestamps[r] + round(arrivals[r][i] / ps_per_unit) (0 when
    ps_per_unit <= 0); changes are stably sorted by time and '#t' opens
    each distinct time."""
    trans = []
    last = None
    for r, (row, ts) in enumerate(zip(rows, timestamps)):
        if last is None:
            last = [2] * len(row)
        for i, v in enumerate(row):
            v = int(v)
            if v == last[i]:
                continue
            last[i] = v
            a = int(arrivals[r][i])
            units = int(_round_half_away(a / ps_per_unit)) if ps_per_unit > 0 else 0
            trans.append((int(ts) + units, i, v))
    trans.sort(key=lambda x: x[0])
    out = []
    cur = None
    for t, i, v in trans:
        if t != cur:
            out.append(f"#{t}")
            cur = t
        out.append(f"{_SYM[v]}{id_code(i)}")
    return "".join(line + "\n" for line in out)
