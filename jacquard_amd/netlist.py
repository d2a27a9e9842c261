"""Netlists (AIGs) for the simulation backend.

The host keeps the reference's netlist -> AIG contract (src/aig.rs:41-346):
pins 1..num_pins in topological order, pin 0 tied to 0, literals
``pin << 1 | invert``, endpoints = primary outputs, DFFs and 13-bit x 32-bit
SRAMs.  Netlists come from the seeded synthetic generators or from a flat
AIGPDK structural Verilog file.
"""
import ctypes

import numpy as np

from . import _native as N

DRV_TIE0, DRV_AND, DRV_INPUT, DRV_CLOCK_FLAG, DRV_DFF, DRV_SRAM = range(6)


class Netlist:
    def __init__(self, handle):
        if not handle:
            raise RuntimeError("netlist construction failed: " + N.host_error())
        self._h = handle
        self._arrays = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.host().jqh_aig_free(h)
            except Exception:  # interpreter shutdown: the library is gone
                pass
            self._h = None

    # ------------------------------------------------------------ builders
    @classmethod
    def random(cls, seed=1, gates=2000, depth=24, dffs=256, inputs=32,
               outputs=16, geometric_p=0.0, invert_prob=0.2, srams=0,
               sram_addr_bits=10, delay=(0, 0)):
        """Seeded random synchronous AIG (one clock, posedge DFFs).

        geometric_p == 0 draws the second fan-in's level uniformly from all
        earlier levels, otherwise the level distance is Geometric(p).  srams
        adds single-port 2**sram_addr_bits x 32 macros with random address,
        data and write-enable nets.  delay=(lo, hi) draws integer per-gate
        delays in ps."""
        h = N.host().jqh_aig_random(seed, gates, depth, dffs, inputs, outputs,
                                    geometric_p, invert_prob, srams,
                                    sram_addr_bits, delay[0], delay[1])
        return cls(h)

    @classmethod
    def multiplier(cls, seed=7, tiles=64, width=16, control_gates=2_000_000,
                   control_depth=128, dffs=1 << 20, inputs=512, outputs=256):
        """Structured netlist: array multipliers with pipeline registers plus
        a random control AIG."""
        h = N.host().jqh_aig_multiplier(seed, tiles, width, control_gates,
                                        control_depth, dffs, inputs, outputs)
        return cls(h)

    @classmethod
    def from_verilog(cls, path=None, text=None):
        if text is None:
            with open(path, "rb") as f:
                text = f.read()
        if isinstance(text, str):
            text = text.encode()
        return cls(N.host().jqh_aig_from_verilog(text, len(text)))

    # ------------------------------------------------------------ queries
    def counts(self):
        out = (ctypes.c_uint64 * 6)()
        N.host().jqh_aig_counts(self._h, out)
        return dict(num_pins=out[0], num_and=out[1], num_pos=out[2],
                    num_dffs=out[3], num_srams=out[4], has_delays=bool(out[5]))

    @property
    def handle(self):
        return self._h

    def arrays(self):
        """Flat numpy view of the netlist (drivers, AND operands, DFFs,
        SRAMs, primary outputs, delays)."""
        if self._arrays is not None:
            return self._arrays
        c = self.counts()
        n = c["num_pins"] + 1
        nd, ns = c["num_dffs"], c["num_srams"]
        A = dict(
            drv=np.zeros(n, np.uint8), and_a=np.zeros(n, np.uint32),
            and_b=np.zeros(n, np.uint32), ref=np.zeros(n, np.uint32),
            pos=np.zeros(c["num_pos"], np.uint32),
            dff_d=np.zeros(nd, np.uint32), dff_en=np.zeros(nd, np.uint32),
            dff_q=np.zeros(nd, np.uint32),
            sram_r_addr=np.zeros(ns * 13, np.uint32),
            sram_r_en=np.zeros(ns, np.uint32),
            sram_rd=np.zeros(ns * 32, np.uint32),
            sram_w_addr=np.zeros(ns * 13, np.uint32),
            sram_w_en=np.zeros(ns * 32, np.uint32),
            sram_w_data=np.zeros(ns * 32, np.uint32),
            delay=np.zeros(n, np.uint16),
        )
        keys = ["drv", "and_a", "and_b", "ref", "pos", "dff_d", "dff_en",
                "dff_q", "sram_r_addr", "sram_r_en", "sram_rd", "sram_w_addr",
                "sram_w_en", "sram_w_data", "delay"]
        N.host().jqh_aig_export(self._h, *[A[k].ctypes.data for k in keys])
        if not c["has_delays"]:
            A["delay"] = None
        self._arrays = A
        return A

    def set_delays(self, delay):
        d = np.ascontiguousarray(delay, dtype=np.uint16)
        assert d.size == self.counts()["num_pins"] + 1
        N.host().jqh_aig_set_delays(self._h, d.ctypes.data)
        self._arrays = None

    def ports(self):
        """Input and output port bits in declaration order:
        (inputs: [(name, pin, posedge_flag_pin_or_0)], outputs: [(name, lit)])."""
        L = N.host()
        cnt = (ctypes.c_uint64 * 2)()
        L.jqh_aig_port_counts(self._h, cnt)
        ins, outs = [], []
        for i in range(cnt[0]):
            pin, flag = ctypes.c_uint32(), ctypes.c_uint32()
            name = L.jqh_aig_in_port(self._h, i, ctypes.byref(pin), ctypes.byref(flag))
            ins.append((name.decode(), pin.value, flag.value))
        for i in range(cnt[1]):
            lit = ctypes.c_uint32()
            name = L.jqh_aig_out_port(self._h, i, ctypes.byref(lit))
            outs.append((name.decode(), lit.value))
        return ins, outs

    def dff_names(self):
        """Instance name of every DFF, in DFF order ("" where the netlist
        has none, as for generated netlists)."""
        L = N.host()
        return [L.jqh_aig_dff_name(self._h, j).decode()
                for j in range(self.counts()["num_dffs"])]

    def module_name(self):
        return N.host().jqh_aig_module(self._h).decode()

    def x_capable_pins(self):
        """DFF Q / SRAM read pins and their forward cone, with the DFF fixpoint
        (src/aig.rs:3276-3367)."""
        A = self.arrays()
        n = A["drv"].size
        x = np.zeros(n, bool)
        x[A["dff_q"]] = True
        rd = A["sram_rd"][A["sram_rd"] != 0]
        x[rd] = True
        ands = np.nonzero(A["drv"] == DRV_AND)[0]
        a = A["and_a"][ands] >> 1
        b = A["and_b"][ands] >> 1
        while True:
            for k, p in enumerate(ands):
                if not x[p] and (x[a[k]] or x[b[k]]):
                    x[p] = True
            d = A["dff_d"] >> 1
            newq = x[d] & ~x[A["dff_q"]]
            if not newq.any():
                break
            x[A["dff_q"][newq]] = True
        return x
