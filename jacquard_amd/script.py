"""FlattenedScriptV1: the packed program of the Boolean-processor interpreter.

Mirrors the reference type (src/flatten.rs:189-341) field for field:
``blocks_start`` / ``blocks_data`` in the exact script ISA of
src/flatten.rs:198-258, the state layout (primary-input words first, then
partition writeouts in block-affinity order, X masks at
``xprop_state_offset``, arrivals at ``arrival_state_offset``), and the
input/output maps.  Construction goes through the native compiler
(level split -> partitions -> boomerang stages -> flatten).
"""
import ctypes

import numpy as np

from . import _native as N
from .netlist import Netlist

SRAM_WORDS = 1 << 13  # src/aigpdk.rs:22
THREADS = 256


class Compiled:
    """Partitioned + mapped design, flattenable for any block count and
    per-partition cost vector."""

    def __init__(self, netlist: Netlist, level_split=0, max_inputs=0,
                 max_outputs=0, max_level_nodes=0, max_part_nodes=0, multilevel=8,
                 max_boomerang_stages=0):
        """multilevel: highest boomerang-tree level at which AND gates are
        placed (0 = one logic level per boomerang stage);
        max_boomerang_stages: partitions mapped to more boomerang stages
        are split by endpoints (0 = no limit)."""
        self.netlist = netlist
        h = N.host().jqh_compile(netlist.handle, level_split, max_inputs,
                                 max_outputs, max_level_nodes, max_part_nodes,
                                 multilevel, max_boomerang_stages)
        if not h:
            raise RuntimeError("compile failed: " + N.host_error())
        self._h = h
        self.log = N.host().jqh_last_log().decode()

    @classmethod
    def _wrap(cls, netlist, handle, log=""):
        c = cls.__new__(cls)
        c.netlist = netlist
        c._h = handle
        c.log = log
        return c

    def merged(self, pairs):
        """A new Compiled with each (major, i, j) pair of partitions mapped
        as one partition over their joint endpoints (pairs of one major
        stage must be disjoint); a pair that does not fit one partition
        stays two.  Returns (compiled, merged bool[len(pairs)]).  Partition
        indices of the result follow its own order (decreasing boomerang
        stages), not this design's."""
        p = np.ascontiguousarray(np.asarray(pairs, np.uint32).reshape(-1, 3))
        ok = np.zeros(max(len(p), 1), np.uint8)
        h = N.host().jqh_compiled_merge(self.netlist.handle, self._h, p.ctypes.data,
                                        len(p), ok.ctypes.data)
        if not h:
            raise RuntimeError("merge failed: " + N.host_error())
        return Compiled._wrap(self.netlist, h, self.log), ok[:len(p)].astype(bool)

    def split(self, parts):
        """A new Compiled with each (major, i) partition mapped again as two
        partitions over the halves of its endpoints (a partition with one
        endpoint, or whose halves do not both fit, stays whole).  Returns
        (compiled, split bool[len(parts)]); partition indices of the result
        follow its own order."""
        p = np.ascontiguousarray(np.asarray(parts, np.uint32).reshape(-1, 2))
        ok = np.zeros(max(len(p), 1), np.uint8)
        h = N.host().jqh_compiled_split(self.netlist.handle, self._h, p.ctypes.data,
                                        len(p), ok.ctypes.data)
        if not h:
            raise RuntimeError("split failed: " + N.host_error())
        return Compiled._wrap(self.netlist, h, self.log), ok[:len(p)].astype(bool)

    def endpoint_counts(self):
        """Endpoints per partition, [major][partition]."""
        L = N.host()
        return [[L.jqh_compiled_part_endpoints(self._h, m, p) for p in range(n)]
                for m, n in enumerate(self.parts_per_stage())]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.host().jqh_compiled_free(h)
            except Exception:  # interpreter shutdown: the library is gone
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def num_major_stages(self):
        return N.host().jqh_compiled_num_major(self._h)

    def parts_per_stage(self):
        L = N.host()
        return [L.jqh_compiled_num_parts(self._h, m)
                for m in range(self.num_major_stages())]

    def stage_counts(self):
        """Boomerang stages per partition, [major][partition]."""
        L = N.host()
        return [[L.jqh_compiled_part_stages(self._h, m, p)
                 for p in range(n)]
                for m, n in enumerate(self.parts_per_stage())]

    def flatten(self, num_blocks, xprop=False, costs=None):
        return FlattenedScriptV1(self, num_blocks, xprop=xprop, costs=costs)

    def flatten_shared(self, num_blocks):
        """2-state flattening with the heuristic cost, built once per block
        count and shared; callers must not modify it (no timing injection)."""
        cache = self.__dict__.setdefault("_shared", {})
        if num_blocks not in cache:
            cache[num_blocks] = self.flatten(num_blocks)
        return cache[num_blocks]


class _ScriptHandle:
    """Owns a flattened script's native handle.  blocks_data is a view of the
    handle's buffer and holds this owner too, so the view stays valid after
    the script object itself is gone."""

    def __init__(self, h):
        self.h = h

    def __del__(self):
        h, self.h = self.h, None
        if h:
            try:
                N.host().jqh_script_free(h)
            except Exception:  # interpreter shutdown: the library is gone
                pass


class FlattenedScriptV1:
    def __init__(self, compiled: Compiled, num_blocks, xprop=False, costs=None):
        self.compiled = compiled
        self.netlist = compiled.netlist
        L = N.host()
        cptr = None
        if costs is not None:
            flat = np.ascontiguousarray(np.concatenate(
                [np.asarray(c, np.float64) for c in costs]), np.float64)
            self._costs = flat
            cptr = flat.ctypes.data
        h = L.jqh_flatten(self.netlist.handle, compiled.handle, num_blocks,
                          int(bool(xprop)), cptr)
        if not h:
            raise RuntimeError("flatten failed: " + N.host_error())
        self._owner = _ScriptHandle(h)
        self._h = h
        info = (ctypes.c_uint64 * 15)()
        L.jqh_script_info(h, info)
        self.num_blocks = int(info[0])
        self.num_major_stages = int(info[1])
        self.reg_io_state_size = int(info[2])
        self.sram_storage_size = int(info[3])
        self.xprop_enabled = bool(info[6])
        self.xprop_state_offset = int(info[7]) if self.xprop_enabled else 0
        self.timing_enabled = bool(info[13])
        self.timing_arrivals_enabled = False
        self.arrival_state_offset = 0
        self.clock_period_ps = 1000
        nbs, nbd = int(info[4]), int(info[5])
        self.blocks_start = np.ctypeslib.as_array(
            ctypes.cast(L.jqh_script_blocks_start(h), ctypes.POINTER(ctypes.c_uint64)),
            (nbs,)).copy()
        self._bd_ptr = L.jqh_script_blocks_data(h)
        buf = (ctypes.c_uint32 * nbd).from_address(self._bd_ptr)
        buf._owner = self._owner  # the view keeps the native buffer alive
        self.blocks_data = np.ctypeslib.as_array(buf)
        nl = int(info[8])
        self.input_layout = np.ctypeslib.as_array(
            ctypes.cast(L.jqh_script_input_layout(h), ctypes.POINTER(ctypes.c_uint32)),
            (nl,)).copy()
        ptrs = [ctypes.c_void_p() for _ in range(6)]
        L.jqh_script_maps(h, *[ctypes.byref(p) for p in ptrs])

        def arr(p, n):
            if n == 0:
                return np.zeros(0, np.uint32)
            return np.ctypeslib.as_array(
                ctypes.cast(p, ctypes.POINTER(ctypes.c_uint32)), (n,)).copy()

        nin, nout, nst = int(info[9]), int(info[10]), int(info[11])
        self.input_map = dict(zip(arr(ptrs[0], nin).tolist(), arr(ptrs[1], nin).tolist()))
        self.output_map = dict(zip(arr(ptrs[2], nout).tolist(), arr(ptrs[3], nout).tolist()))
        # the largest of the three maps (every word written for a later
        # major stage), read by none of this package's paths: kept as two
        # arrays and made a dict when first read
        self._staged = (arr(ptrs[4], nst), arr(ptrs[5], nst))
        self.num_parts = int(info[12])
        self.partition_x_capable = (
            arr(L.jqh_script_part_x(h), self.num_parts).astype(bool)
            if self.xprop_enabled else np.zeros(0, bool))
        if self.num_parts:
            self.parts = np.ctypeslib.as_array(
                ctypes.cast(L.jqh_script_parts(h), ctypes.POINTER(ctypes.c_uint64)),
                (self.num_parts, 5)).copy()
        else:
            self.parts = np.zeros((0, 5), np.uint64)
        nsr = int(info[14])
        self.sram_offsets = arr(L.jqh_script_sram_offsets(h), nsr)
        nd = self.netlist.counts()["num_dffs"]
        self.dff_data_pos = arr(L.jqh_script_dff_data_pos(h), nd)

    # ------------------------------------------------------------------
    @property
    def staged_io_map(self):
        """pin -> state bit of every value written for a later major stage
        of the same cycle (src/flatten.rs:694-858)."""
        if not hasattr(self, "_staged_map"):
            pins, pos = self._staged
            self._staged_map = dict(zip(pins.tolist(), pos.tolist()))
        return self._staged_map

    @property
    def num_srams(self):
        return self.sram_storage_size // SRAM_WORDS

    def effective_state_size(self):
        """src/flatten.rs:1581-1590."""
        size = self.reg_io_state_size
        if self.xprop_enabled:
            size += self.reg_io_state_size
        if self.timing_arrivals_enabled:
            size += self.reg_io_state_size
        return size

    def inject_timing(self):
        """Fill the per-word delay slots from the netlist's gate delays
        (src/flatten.rs:1812-1846)."""
        if N.host().jqh_inject_timing(self.netlist.handle, self._h) != 0:
            raise RuntimeError(N.host_error())
        self.timing_enabled = True

    def enable_timing_arrivals(self):
        """src/flatten.rs:1596-1603."""
        if not self.timing_enabled:
            raise RuntimeError("Cannot enable timing arrivals without SDF timing data")
        self.timing_arrivals_enabled = True
        self.arrival_state_offset = self.reg_io_state_size * (1 + int(self.xprop_enabled))

    def build_timing_constraint_buffer(self, clock_period_ps, setup_ps, hold_ps):
        """[clock_ps, per-word (min setup << 16 | min hold)...]
        (src/flatten.rs:1891-1924, src/sim/setup.rs:480-497)."""
        out = np.zeros(1 + self.reg_io_state_size, np.uint32)
        N.host().jqh_constraint_buffer(self._h, min(int(clock_period_ps), 0xFFFFFFFF),
                                       setup_ps, hold_ps, out.ctypes.data)
        self.clock_period_ps = int(clock_period_ps)
        return out

    def validate(self, state_size=None):
        """Check every access the interpreter will make (raises on error)."""
        ss = state_size if state_size is not None else self.effective_state_size()
        bs = np.ascontiguousarray(self.blocks_start, np.uint64)
        rc = N.host().jqh_validate(
            bs.ctypes.data, bs.size, self._bd_ptr, self.blocks_data.size,
            self.num_blocks, self.num_major_stages, ss,
            max(self.sram_storage_size, 0), 0, int(self.arrival_state_offset))
        if rc != 0:
            raise ValueError(N.host_error())

    def part_features(self):
        """24 script features per partition, in blocks_data order."""
        out = np.zeros((self.num_parts, 24), np.float32)
        parts = np.ascontiguousarray(self.parts, np.uint64)
        rc = N.host().jqh_part_features(self._bd_ptr, self.blocks_data.size,
                                        parts.ctypes.data, self.num_parts,
                                        out.ctypes.data)
        if rc != 0:
            raise RuntimeError(N.host_error())
        return out

    def part_index(self):
        """Index of the first partition of every (major stage, block) in
        blocks_data order (the profiled kernel's sample slot)."""
        idx = np.zeros(self.num_major_stages * self.num_blocks, np.uint32)
        parts = self.parts
        k = 0
        for m in range(self.num_major_stages):
            for b in range(self.num_blocks):
                idx[m * self.num_blocks + b] = k
                while k < len(parts) and parts[k, 0] == m and parts[k, 1] == b:
                    k += 1
        return idx

    def xmask_template(self):
        """X mask of the state layout: 1 everywhere except input_map bits
        (src/sim/vcd_io.rs:470-474)."""
        t = np.full(self.reg_io_state_size, 0xFFFFFFFF, np.uint32)
        pos = np.fromiter(self.input_map.values(), np.uint64, len(self.input_map))
        np.bitwise_and.at(t, (pos >> 5).astype(np.int64),
                          ~(np.uint32(1) << (pos & 31).astype(np.uint32)))
        return t

    def algorithmic_bytes_per_cycle(self):
        """Algorithmic HBM bytes of one lane-cycle, from the script alone:
        script words + gathered state words + writeout read/write + SRAM
        accesses (+X lane, +arrival write and constraint read per writeout
        word when timing arrivals are on)."""
        bd = self.blocks_data
        total_script = int(self.blocks_start[-1])
        g = 0
        io = 0
        sr = 0
        tm = 0
        for off, length in self.parts[:, 3:5]:
            h = bd[int(off):int(off) + 256]
            R = int(h[6])
            x = 2 if (self.xprop_enabled and h[8]) else 1
            gr = bd[int(off) + 256:int(off) + 256 + 512 * R].reshape(R, 256, 2)
            g += x * int(np.count_nonzero(gr[:, :, 1]))
            io += x * 2 * int(h[2])
            sr += x * 3 * int(h[4])
            if self.timing_arrivals_enabled:
                tm += 2 * int(h[2])
        return 4 * (total_script + g + io + sr + tm)

    def algorithmic_lds_bytes_per_cycle(self):
        """LDS bytes of one lane-cycle in the interpreter's access pattern
        (csrc/kernel_v1_impl.cuh): per boomerang stage the 8192 shuffle bit
        gathers, the stage write, the hier0-3 reads/writes and the tail write
        (880 words) and 2 words per writeout hook fired; per partition the
        SRAM/dup and clock-enable permutations (2 x 8192), the writeout
        read/writes (3 x 256) and the global-read store (256); doubled for
        X-capable partitions.  Timing adds u16 arrival traffic of about 700
        per stage plus one per writeout word."""
        bd = self.blocks_data
        total = 0
        for off, _length in self.parts[:, 3:5]:
            h = bd[int(off):int(off) + 256]
            S = int(h[0])
            x = 2 if (self.xprop_enabled and h[8]) else 1
            hooks = h[128:256].astype(np.uint32)
            nh = int(np.count_nonzero((hooks & 0xFFFF) != 0xFFFF)
                     + np.count_nonzero((hooks >> 16) != 0xFFFF))
            words = S * (8192 + 880) + 2 * nh + 2 * 8192 + 3 * 256 + 256
            total += 4 * x * words
            if self.timing_arrivals_enabled:
                total += 2 * (S * 700 + int(h[2]))
        return total


def compile_design(netlist, num_blocks, level_split=0, xprop=False, **kw):
    """Convenience: partition + flatten with the heuristic cost."""
    c = Compiled(netlist, level_split=level_split, **kw)
    return c.flatten(num_blocks, xprop=xprop)
