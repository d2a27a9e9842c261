"""Hand-built partition scripts for the parity tests.

Two of them restate the reference's own script recipes word for word:
  - minimal_xprop_script: build_minimal_xprop_script
    (src/sim/cpu_reference.rs:994-1091), one partition, one boomerang stage,
    a dummy end partition;
  - bench_script: build_bench_script (benches/xprop.rs:11-80), identity
    shuffles, pass-through trees, 0xFF00FF00 hooks, any stage count.
The third, tree_probe_script, is this suite's probe of the arrival model:
one partition, one boomerang stage, every tree word hooked to the writeout
of the same index, caller-chosen tree flags and delay slots.

The section layout is the script ISA (src/flatten.rs:198-258): a 256-word
header, 512 words per global-read round, 5120 per boomerang stage (4096
permutation words, then xora / xorb / orb / delay per thread word), then
the SRAM/duplicate and the clock-enable sections of 5120 words each.
"""
import numpy as np

M = 0xFFFFFFFF
HDR, ROUND, STAGE = 256, 512, 5120


def minimal_xprop_script(is_x, num_ios):
    """build_minimal_xprop_script (src/sim/cpu_reference.rs:994-1091)."""
    total = HDR + ROUND + STAGE + STAGE + STAGE + HDR
    s = np.zeros(total, np.uint32)
    s[0], s[1], s[2], s[3] = 1, 0, num_ios, 0
    s[4], s[5], s[6], s[7] = 0, 0, 1, 0
    s[8], s[9] = 1 if is_x else 0, num_ios
    s[128:256] = 0xFF00FF00
    s[128] = s[128] & 0xFFFF0000
    pi = HDR
    s[pi], s[pi + 1] = 0, 1
    pi += ROUND + 4096
    s[pi + 128 * 4 + 2] = M
    for hi in range(1, 8):
        hw = 1 << (7 - hi)
        s[pi + hw * 4 + 2] = M
    s[pi + 0 * 4 + 2] = M
    pi += 1024 + STAGE + 4096
    s[pi] = M
    pi += 1024
    s[pi], s[pi + 1] = 0, 1
    return s


def bench_script(is_x, num_ios, num_stages):
    """build_bench_script (benches/xprop.rs:11-80)."""
    total = HDR + ROUND + num_stages * STAGE + STAGE + STAGE + HDR
    s = np.zeros(total, np.uint32)
    s[0], s[1], s[2], s[3] = num_stages, 0, num_ios, 0
    s[4], s[5], s[6], s[7] = 0, 0, 1, 0
    s[8], s[9] = 1 if is_x else 0, num_ios
    s[128:256] = 0xFF00FF00
    s[128] = s[128] & 0xFFFF0000
    pi = HDR
    s[pi], s[pi + 1] = 0, 1
    pi += ROUND
    for _ in range(num_stages):
        pi += 4096
        s[pi + 128 * 4 + 2] = M
        for hi in range(1, 8):
            hw = 1 << (7 - hi)
            s[pi + hw * 4 + 2] = M
        s[pi + 0 * 4 + 2] = M
        pi += 1024
    pi += STAGE
    pi += 4096
    s[pi] = M
    pi += 1024
    s[pi], s[pi + 1] = 0, 1
    return s


def tree_probe_script(xora, xorb, orb, delay, num_ios=256, rounds=1):
    """One partition, one boomerang stage: writeout t takes stage-0 tree word
    t (t < num_ios), the clock enables are all ones, and thread t's tree flags
    and delay slot are xora[t], xorb[t], orb[t], delay[t].  The global read
    (rounds > 0) gives thread t state word t, all 32 bits."""
    total = HDR + ROUND * rounds + STAGE + STAGE + STAGE + HDR
    s = np.zeros(total, np.uint32)
    s[0], s[1], s[2], s[3] = 1, 0, num_ios, 0
    s[6] = rounds
    hooks = np.full(256, 0xFF00, np.uint32)
    hooks[:num_ios] = np.arange(num_ios, dtype=np.uint32)  # stage 0, slot t
    s[128:256] = hooks[0::2] | (hooks[1::2] << 16)
    pi = HDR
    if rounds:
        s[pi:pi + 512:2] = np.arange(256, dtype=np.uint32)
        s[pi + 1:pi + 512:2] = M
        pi += ROUND * rounds
    # permutation: layer bit 32t + k of the stage input <- state bit 32t + k
    for ko in range(4):
        for t in range(256):
            for ki in range(4):
                k = ko * 4 + ki
                lo, hi = 32 * t + 2 * k, 32 * t + 2 * k + 1
                s[pi + ko * 1024 + 4 * t + ki] = lo | (hi << 16)
    pi += 4096
    fl = np.stack([np.asarray(xora, np.uint32), np.asarray(xorb, np.uint32),
                   np.asarray(orb, np.uint32),
                   np.asarray(delay, np.uint32) & 0xFFFF], 1)
    s[pi:pi + 1024] = fl.reshape(-1)
    pi += 1024 + STAGE + 4096
    s[pi:pi + 1024:4] = M  # clock enable: inv = all ones, set0 = 0
    pi += 1024
    s[pi], s[pi + 1] = 0, 1
    return s


class OneBlockScript:
    """A FlattenedScriptV1-shaped view of one hand-built partition script as
    the only block of one major stage, with the 2048 padding words that
    flatten appends (src/flatten.rs:1341)."""

    def __init__(self, part, reg_io_state_size, xprop=False, arrivals=False):
        self.num_blocks = 1
        self.num_major_stages = 1
        self.blocks_data = np.concatenate([part, np.zeros(2048, np.uint32)])
        self.blocks_start = np.array([0, part.size], np.uint64)
        self.reg_io_state_size = reg_io_state_size
        self.sram_storage_size = 0
        self.xprop = self.xprop_enabled = xprop
        rio = reg_io_state_size
        self.arrival_state_offset = rio * (2 if xprop else 1) if arrivals else 0

    def effective_state_size(self):
        rio = self.reg_io_state_size
        return rio * (1 + int(self.xprop) + int(self.arrival_state_offset != 0))

    def validate(self, state_size=None):
        from jacquard_amd import _native as N
        ss = state_size if state_size is not None else self.effective_state_size()
        bs = np.ascontiguousarray(self.blocks_start, np.uint64)
        bd = np.ascontiguousarray(self.blocks_data, np.uint32)
        rc = N.host().jqh_validate(bs.ctypes.data, bs.size, bd.ctypes.data, bd.size,
                                   1, 1, ss, 0, 0, int(self.arrival_state_offset))
        if rc != 0:
            raise ValueError(N.host_error())
