"""Batched lane-sliced form: many independent stimulus lanes of one netlist
in a single persistent launch (jq_simulate_batched).

Lanes are sorted by cycle count and packed 32 to a group, so the lanes of
one group (which run in lockstep) have similar lengths; every group runs
its own cycle count behind its own barrier.  Each lane is described by
(seed, toggle probability, cycles) and sees exactly the stimulus that
``stimulus.build_states(script, cycles, seed, p)`` gives the plain path.
"""
import ctypes

import numpy as np

from . import _native as N
from . import stimulus


class BatchedRun:
    def __init__(self, script, seeds, probs, cycles, obs_lits=None):
        self.script = script
        self.seeds = np.asarray(seeds, np.uint32)
        self.probs = np.asarray(probs, np.float64)
        self.cycles = np.asarray(cycles, np.int64)
        n = len(self.seeds)
        assert len(self.probs) == n and len(self.cycles) == n
        if obs_lits is None:
            obs_lits = [l for l in script.netlist.arrays()["pos"].tolist() if l > 1]
        self.obs_lits = list(obs_lits)
        self.obs_bits = np.asarray([script.output_map[l] for l in self.obs_lits], np.uint32)
        # lane packing: longest first, 32 per group
        self.order = np.argsort(-self.cycles, kind="stable")
        self.G = (n + 31) // 32
        slot_seed = np.zeros(self.G * 32, np.uint32)
        slot_thr = np.zeros(self.G * 32, np.uint32)
        self.gcycles = np.zeros(self.G, np.uint32)
        for k, lane in enumerate(self.order.tolist()):
            slot_seed[k] = self.seeds[lane]
            slot_thr[k] = stimulus.threshold(self.probs[lane])
            g = k // 32
            self.gcycles[g] = max(self.gcycles[g], self.cycles[lane])
        self.slot_seed, self.slot_thr = slot_seed, slot_thr
        self.max_cycles = int(self.gcycles.max()) if n else 0
        self.data_bits, self.flag_bits = stimulus.input_roles(script)

    def lane_cycles_total(self):
        return int(self.cycles.sum())

    def slot_cycles_total(self):
        return int(self.gcycles.sum()) * 32

    def prepare(self):
        """Allocate and initialise the device buffers."""
        import torch
        from .sim import to_device
        s = self.script
        rio = s.reg_io_state_size
        st = np.zeros((self.G, 2, rio * 32), np.uint32)
        st[:, :, self.flag_bits.astype(np.int64)] = 0xFFFFFFFF
        self.d = dict(
            bs=to_device(np.ascontiguousarray(s.blocks_start, np.uint64)),
            bd=to_device(np.ascontiguousarray(s.blocks_data, np.uint32)),
            states=to_device(st.reshape(-1)),
            sram=torch.zeros(max(self.G * 32 * s.sram_storage_size, 1),
                             dtype=torch.int32, device="cuda"),
            gcycles=to_device(self.gcycles),
            data_bits=to_device(self.data_bits),
            seed=to_device(self.slot_seed),
            thr=to_device(self.slot_thr),
            obs_bits=to_device(self.obs_bits),
            obs=torch.zeros(max(self.G * self.max_cycles * len(self.obs_bits), 1),
                            dtype=torch.int32, device="cuda"),
        )
        return self.d

    def launch(self, stream=None):
        import torch
        s = self.script
        d = self.d
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = N.hip().jq_simulate_batched(
            self.G, s.num_blocks, s.num_major_stages, d["bs"].data_ptr(),
            d["bd"].data_ptr(), s.reg_io_state_size, s.sram_storage_size,
            d["sram"].data_ptr(), d["gcycles"].data_ptr(), self.max_cycles,
            d["states"].data_ptr(), stimulus.input_word_count(s),
            d["data_bits"].data_ptr(), len(self.data_bits), d["seed"].data_ptr(),
            d["thr"].data_ptr(), d["obs_bits"].data_ptr(), len(self.obs_bits),
            d["obs"].data_ptr(), ctypes.c_void_p(st.cuda_stream))
        N.check(rc)

    def run(self, validate=True):
        import torch
        from .sim import device_error
        if validate:
            self.script.validate(self.script.reg_io_state_size)
        self.prepare()
        torch.cuda.synchronize()
        self.launch()
        torch.cuda.synchronize()
        if device_error():
            raise RuntimeError("device barrier timed out in the batched launch")
        return self.lane_outputs()

    def lane_outputs(self):
        """{lane: uint8[cycles, n_obs]} observed values after every cycle."""
        obs = self.d["obs"].cpu().numpy().view(np.uint32)
        obs = obs[:self.G * self.max_cycles * len(self.obs_bits)].reshape(
            self.G, self.max_cycles, len(self.obs_bits))
        out = {}
        for k, lane in enumerate(self.order.tolist()):
            g, l = divmod(k, 32)
            c = int(self.cycles[lane])
            out[lane] = ((obs[g, :c, :] >> np.uint32(l)) & 1).astype(np.uint8)
        return out


def plain_observations(script, states, cycles, obs_lits):
    """uint8[cycles, n_obs] of a finished plain snapshot buffer."""
    ss = script.effective_state_size()
    S = states.reshape(-1, ss)
    cols = []
    for lit in obs_lits:
        pos = script.output_map[lit]
        cols.append((S[1:cycles + 1, pos >> 5] >> np.uint32(pos & 31)) & 1)
    return np.stack(cols, 1).astype(np.uint8)
