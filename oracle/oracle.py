"""ctypes wrapper of the CPU oracle (libjq_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never by the product package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

EVENT_DTYPE = np.dtype([("event_type", "<u4"), ("message_id", "<u4"),
                        ("cycle", "<u4"), ("reserved", "<u4"),
                        ("data", "<u4", (4,))])


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE])
    # the reference's SPI flash model, where the reference tree is present
    if os.path.exists("/root/reference/csrc/spiflash_model.cc"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    d = os.environ.get("JACQUARD_ORACLE_LIB_DIR")  # the sanitizer build
    path = os.path.join(d or HERE, "libjq_oracle.so")
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    L.jqo_simulate_block_v1.argtypes = [vp, sz, vp, vp, vp]
    L.jqo_simulate_block_v1_xprop.argtypes = [vp, sz, vp, vp, vp, vp, vp, vp, ctypes.c_int]
    L.jqo_block_kernel.argtypes = [vp, sz, vp, vp, vp, vp, vp, ctypes.c_uint32,
                                   ctypes.c_int32, ctypes.c_uint32, vp, sz, vp]
    L.jqo_run.argtypes = [sz, sz, vp, vp, vp, vp, sz, sz, vp, vp, ctypes.c_int32,
                          vp, sz, vp]
    L.jqo_run_v1.argtypes = [sz, sz, vp, vp, vp, sz, sz, vp]
    L.jqo_run_v1_parallel.argtypes = [sz, sz, vp, vp, vp, sz, sz, vp, ctypes.c_int]
    L.jqo_aig_sim.argtypes = [vp, sz, sz, vp, vp, sz, vp, vp, ctypes.c_int, vp, vp, vp]
    for f in ("jqo_simulate_block_v1", "jqo_simulate_block_v1_xprop",
              "jqo_block_kernel", "jqo_run", "jqo_run_v1", "jqo_run_v1_parallel",
              "jqo_aig_sim"):
        getattr(L, f).restype = ctypes.c_int
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def simulate_block_v1(script, input_state, output_state, sram):
    rc = lib().jqo_simulate_block_v1(_p(script), script.size, _p(input_state),
                                     _p(output_state), _p(sram))
    return rc


def simulate_block_v1_xprop(script, in_v, out_v, in_x, out_x, sram, sram_x,
                            cpu_quirk=False):
    return lib().jqo_simulate_block_v1_xprop(
        _p(script), script.size, _p(in_v), _p(out_v), _p(in_x), _p(out_x),
        _p(sram), _p(sram_x), int(cpu_quirk))


def run(script, states, num_cycles, *, sram=None, sram_x=None,
        constraints=None, arrival_state_offset=0, event_capacity=0,
        state_size=None):
    """Kernel-semantics run over a (num_cycles+1) x state_size buffer, in
    place.  Returns (events structured array, total event count)."""
    ss = state_size or script.effective_state_size()
    bs = np.ascontiguousarray(script.blocks_start, np.uint64)
    bd = np.ascontiguousarray(script.blocks_data, np.uint32)
    if sram is None:
        sram = np.zeros(max(script.sram_storage_size, 1), np.uint32)
    ev = np.zeros(max(event_capacity, 1), EVENT_DTYPE)
    cnt = ctypes.c_size_t(0)
    rc = lib().jqo_run(script.num_blocks, script.num_major_stages, _p(bs), _p(bd),
                       _p(sram), _p(sram_x), num_cycles, ss, _p(states),
                       _p(constraints), int(arrival_state_offset),
                       ev.ctypes.data if event_capacity else None, event_capacity,
                       ctypes.byref(cnt))
    if rc:
        raise RuntimeError("oracle: malformed script")
    return ev[:min(cnt.value, event_capacity)], cnt.value


def run_v1(script, states, num_cycles, sram=None, state_size=None, threads=0):
    """Literal 2-state restatement (sanity_check_cpu loop); threads>0 runs
    the blocks of a major stage in parallel."""
    ss = state_size or script.reg_io_state_size
    bs = np.ascontiguousarray(script.blocks_start, np.uint64)
    bd = np.ascontiguousarray(script.blocks_data, np.uint32)
    if sram is None:
        sram = np.zeros(max(script.sram_storage_size, 1), np.uint32)
    if threads:
        rc = lib().jqo_run_v1_parallel(script.num_blocks, script.num_major_stages,
                                       _p(bs), _p(bd), _p(sram), num_cycles, ss,
                                       _p(states), int(threads))
    else:
        rc = lib().jqo_run_v1(script.num_blocks, script.num_major_stages, _p(bs),
                              _p(bd), _p(sram), num_cycles, ss, _p(states))
    if rc:
        raise RuntimeError("oracle: malformed script")


class _Aig(ctypes.Structure):
    _fields_ = [("num_pins", ctypes.c_uint32), ("drv_type", ctypes.c_void_p),
                ("and_a", ctypes.c_void_p), ("and_b", ctypes.c_void_p),
                ("num_dffs", ctypes.c_uint32), ("dff_d_iv", ctypes.c_void_p),
                ("dff_en_iv", ctypes.c_void_p), ("dff_q", ctypes.c_void_p),
                ("num_srams", ctypes.c_uint32), ("sram_r_addr_iv", ctypes.c_void_p),
                ("sram_r_en_iv", ctypes.c_void_p), ("sram_rd_pin", ctypes.c_void_p),
                ("sram_w_addr_iv", ctypes.c_void_p), ("sram_w_en_iv", ctypes.c_void_p),
                ("sram_w_data_iv", ctypes.c_void_p)]


def aig_sim(arrays, num_cycles, in_pins, in_bits, obs_lits, xprop=False,
            dff_init_x=None, want_regs=False, dff_init=None):
    """Direct cycle simulation of the AIG (no script), from all-zero
    registers or from the register values `dff_init` (uint8 per DFF).
    Returns (obs uint8[cycles, n_obs] with bit0 value / bit1 X, regs or
    None)."""
    A = arrays
    g = _Aig(A["drv"].size - 1, A["drv"].ctypes.data, A["and_a"].ctypes.data,
             A["and_b"].ctypes.data, A["dff_d"].size, A["dff_d"].ctypes.data,
             A["dff_en"].ctypes.data, A["dff_q"].ctypes.data, A["sram_r_en"].size,
             A["sram_r_addr"].ctypes.data, A["sram_r_en"].ctypes.data,
             A["sram_rd"].ctypes.data, A["sram_w_addr"].ctypes.data,
             A["sram_w_en"].ctypes.data, A["sram_w_data"].ctypes.data)
    in_pins = np.ascontiguousarray(in_pins, np.uint32)
    in_bits = np.ascontiguousarray(in_bits, np.uint8)
    obs = np.ascontiguousarray(obs_lits, np.uint32)
    out = np.zeros((num_cycles, obs.size), np.uint8)
    regs = np.zeros((num_cycles, A["dff_d"].size), np.uint8) if want_regs else None
    dx = None if dff_init_x is None else np.ascontiguousarray(dff_init_x, np.uint8)
    dv = None if dff_init is None else np.ascontiguousarray(dff_init, np.uint8)
    assert dv is None or dv.size == A["dff_d"].size
    rc = lib().jqo_aig_sim(ctypes.byref(g), num_cycles, in_pins.size, _p(in_pins),
                           _p(in_bits), obs.size, _p(obs), _p(out), int(xprop),
                           _p(dx), _p(dv), _p(regs))
    if rc:
        raise RuntimeError("aig_sim failed")
    return out, regs
