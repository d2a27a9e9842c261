"""ctypes loaders for the in-tree native libraries.

The HIP library is the product path: if it is missing the caller gets a
NativeLibraryMissing error naming the build command; nothing falls back to
a CPU implementation.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_i32 = ctypes.c_int32
c_vp = ctypes.c_void_p
c_cp = ctypes.c_char_p
P = ctypes.POINTER


class NativeLibraryMissing(RuntimeError):
    pass


_host = None
_hip = None


def _load(name):
    # JACQUARD_HOST_LIB_DIR points the host library at another build of the
    # same sources (the sanitizer build of the CPU tests); the HIP library
    # always comes from the in-tree build
    d = os.environ.get("JACQUARD_HOST_LIB_DIR") if name == "libjacquard_host.so" else None
    path = os.path.join(d or LIBDIR, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -m jacquard_amd.build` "
            f"(or __graft_entry__.build())")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def host():
    """libjacquard_host.so: generators, script compiler, validator."""
    global _host
    if _host is not None:
        return _host
    L = _load("libjacquard_host.so")
    sig = {
        "jqh_last_error": (c_cp, []),
        "jqh_last_log": (c_cp, []),
        "jqh_aig_random": (c_vp, [c_u64, c_u32, c_u32, c_u32, c_u32, c_u32,
                                  ctypes.c_double, ctypes.c_double, c_u32,
                                  c_u32, c_u32, c_u32]),
        "jqh_aig_multiplier": (c_vp, [c_u64, c_u32, c_u32, c_u32, c_u32,
                                      c_u32, c_u32, c_u32]),
        "jqh_aig_from_verilog": (c_vp, [c_cp, ctypes.c_size_t]),
        "jqh_aig_free": (None, [c_vp]),
        "jqh_aig_counts": (None, [c_vp, P(c_u64)]),
        "jqh_aig_export": (None, [c_vp] + [c_vp] * 15),
        "jqh_aig_po_name": (c_cp, [c_vp, c_u32]),
        "jqh_aig_port_counts": (None, [c_vp, P(c_u64)]),
        "jqh_aig_in_port": (c_cp, [c_vp, c_u32, P(c_u32), P(c_u32)]),
        "jqh_aig_out_port": (c_cp, [c_vp, c_u32, P(c_u32)]),
        "jqh_aig_module": (c_cp, [c_vp]),
        "jqh_aig_dff_name": (c_cp, [c_vp, c_u32]),
        "jqh_aig_set_delays": (None, [c_vp, c_vp]),
        "jqh_compile": (c_vp, [c_vp, c_u32, c_u32, c_u32, c_u32, c_u32, c_u32, c_u32]),
        "jqh_compiled_free": (None, [c_vp]),
        "jqh_compiled_num_major": (c_u32, [c_vp]),
        "jqh_compiled_num_parts": (c_u32, [c_vp, c_u32]),
        "jqh_compiled_part_stages": (c_u32, [c_vp, c_u32, c_u32]),
        "jqh_compiled_part_endpoints": (c_u32, [c_vp, c_u32, c_u32]),
        "jqh_compiled_merge": (c_vp, [c_vp, c_vp, c_vp, c_u64, c_vp]),
        "jqh_compiled_split": (c_vp, [c_vp, c_vp, c_vp, c_u64, c_vp]),
        "jqh_flatten": (c_vp, [c_vp, c_vp, c_u32, ctypes.c_int, c_vp]),
        "jqh_script_free": (None, [c_vp]),
        "jqh_script_info": (None, [c_vp, P(c_u64)]),
        "jqh_script_blocks_start": (c_vp, [c_vp]),
        "jqh_script_blocks_data": (c_vp, [c_vp]),
        "jqh_script_input_layout": (c_vp, [c_vp]),
        "jqh_script_maps": (None, [c_vp] + [P(c_vp)] * 6),
        "jqh_script_part_x": (c_vp, [c_vp]),
        "jqh_script_parts": (c_vp, [c_vp]),
        "jqh_script_sram_offsets": (c_vp, [c_vp]),
        "jqh_script_dff_data_pos": (c_vp, [c_vp]),
        "jqh_inject_timing": (ctypes.c_int, [c_vp, c_vp]),
        "jqh_vcd_parse_body": (c_vp, [c_cp, c_u64, c_cp, c_vp, c_u32, c_vp, c_vp, c_vp,
                                      c_vp, c_vp, c_u32, c_u32, c_u64]),
        "jqh_vcd_num_words": (c_u64, [c_vp]),
        "jqh_vcd_states": (c_vp, [c_vp]),
        "jqh_vcd_num_offsets": (c_u64, [c_vp]),
        "jqh_vcd_offsets": (c_vp, [c_vp]),
        "jqh_vcd_timestamps": (c_vp, [c_vp]),
        "jqh_vcd_free": (None, [c_vp]),
        "jqh_vcd_last_error": (c_cp, []),
        "jqh_vcd_write_values": (c_vp, [c_vp, c_u64, c_u32, c_vp, c_cp, c_vp]),
        "jqh_vcd_write_changes": (c_vp, [c_vp, c_vp, c_u64, c_u64, c_vp, c_u32, c_cp,
                                         c_vp]),
        "jqh_vcd_write_timed": (c_vp, [c_vp, c_vp, c_u64, c_u32, c_vp, ctypes.c_double,
                                       c_cp, c_vp]),
        "jqh_text_data": (c_vp, [c_vp]),
        "jqh_text_size": (c_u64, [c_vp]),
        "jqh_text_free": (None, [c_vp]),
        "jqh_constraint_buffer": (None, [c_vp, c_u32, ctypes.c_uint16,
                                         ctypes.c_uint16, c_vp]),
        "jqh_inject_delays": (c_u64, [c_vp, c_u64, c_vp, c_vp, c_vp, c_u64,
                                      c_vp, c_vp, c_u64]),
        "jqh_constraint_words": (c_u32, [c_u32, c_u64, c_vp, c_vp, c_vp, c_vp,
                                         c_u64, c_vp]),
        "jqh_validate": (ctypes.c_int, [c_vp, c_u64, c_vp, c_u64, c_u64,
                                        c_u64, c_u64, c_u64, c_u64,
                                        c_i32]),
        "jqh_part_features": (ctypes.c_int, [c_vp, c_u64, c_vp, c_u64,
                                             c_vp]),
        "jqh_lane_blocks_data": (ctypes.c_int, [c_vp, c_u64, c_vp, c_u64,
                                                c_vp]),
        "jqh_gather_table": (ctypes.c_int, [c_vp, c_u64, c_vp, c_u64, c_u32,
                                            c_u32, c_vp, c_vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _host = L
    return L


HIP_SYMBOLS = [
    "simulate_v1_noninteractive_simple_scan_hip",
    "simulate_v1_noninteractive_timed_hip",
    "jq_simulate_scan",
    "jq_simulate_scan_profiled",
    "jq_simulate_batched",
    "jq_simulate_batched_profiled",
    "jq_simulate_batched_chunk",
    "jq_simulate_batched_stimulus",
    "jq_max_resident_blocks",
    "jq_output_changes",
    "jq_lds_read_probe",
    "jq_cosim_flash_state_bytes",
    "jq_cosim_edges",
    "jq_device_error_flag",
    "jq_scan_sync_mode",
    "jq_wait_limit",
    "jq_mlp_forward",
    "jq_mlp_train_step",
    "jq_mlp_num_params",
    "jq_mlp_workspace_floats",
    "jq_mlp_noalloc_max_samples",
    "jq_mlp_train_step_noalloc",
    "jq_last_error",
]


def hip():
    """libjacquard_hip.so: the gfx950 kernels behind the C ABI."""
    global _hip
    if _hip is not None:
        return _hip
    L = _load("libjacquard_hip.so")
    sig = {
        "simulate_v1_noninteractive_simple_scan_hip": (
            None, [c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64, c_vp,
                   c_i32]),
        "simulate_v1_noninteractive_timed_hip": (
            None, [c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_u64, c_u64, c_vp,
                   c_vp, c_vp, c_i32]),
        "jq_simulate_scan": (ctypes.c_int, [c_u64, c_u64, c_vp, c_vp, c_vp,
                                            c_vp, c_u64, c_u64, c_vp, c_vp,
                                            c_vp, c_u32, c_i32, c_vp]),
        "jq_simulate_scan_profiled": (ctypes.c_int, [c_u64, c_u64, c_vp, c_vp,
                                                     c_vp, c_vp, c_u64, c_u64,
                                                     c_vp, c_vp, c_u32, c_vp,
                                                     c_u32, c_vp]),
        "jq_simulate_batched": (ctypes.c_int, [
            c_u64, c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_u32, c_u32, c_vp,
            c_vp, c_vp, c_u32, c_vp, c_u32, c_u32, c_vp, c_vp, c_vp, c_vp,
            c_vp, c_vp, c_vp, c_vp, c_u32, c_u32, c_vp, c_vp]),
        "jq_simulate_batched_profiled": (ctypes.c_int, [
            c_u64, c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_u32, c_u32, c_vp,
            c_vp, c_vp, c_u32, c_vp, c_u32, c_u32, c_vp, c_vp, c_vp, c_vp,
            c_vp, c_vp, c_vp, c_vp, c_u32, c_u32, c_vp, c_u32, c_vp, c_u32,
            c_vp]),
        "jq_simulate_batched_chunk": (ctypes.c_int, [
            c_u64, c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_u32, c_u32, c_vp,
            c_vp, c_vp, c_u32, c_vp, c_u32, c_u32, c_vp, c_vp, c_vp, c_vp,
            c_vp, c_vp, c_vp, c_u32, c_vp, c_u32, c_u32, c_vp, c_vp]),
        "jq_simulate_batched_stimulus": (ctypes.c_int, [
            c_u64, c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_u32, c_u32, c_vp,
            c_vp, c_vp, c_u32, c_u32, c_vp, c_vp, c_u32, c_vp, c_vp, c_vp,
            c_vp, c_vp, c_u32, c_vp, c_u32, c_u32, c_vp, c_vp]),
        "jq_output_changes": (ctypes.c_int, [c_vp, c_vp, c_u32, c_vp, c_u32,
                                             c_u32, c_vp, c_vp, c_vp, c_vp]),
        "jq_lds_read_probe": (ctypes.c_int, [c_u32, c_u32, c_vp, c_vp]),
        "jq_cosim_flash_state_bytes": (ctypes.c_size_t, []),
        "jq_cosim_edges": (ctypes.c_int, [
            c_u64, c_u64, c_vp, c_vp, c_vp, c_u64, c_vp, c_vp, c_vp, c_u32,
            c_u32, c_u32, c_vp, c_vp, c_vp, c_u32, c_vp, c_vp, c_vp, c_vp, c_vp,
            c_vp, c_vp, c_u32, ctypes.c_int32, c_vp, c_vp]),
        "jq_max_resident_blocks": (ctypes.c_int, [ctypes.c_int,
                                                  P(ctypes.c_int)]),
        "jq_mlp_forward": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, c_vp, ctypes.c_int,
                                          c_vp, c_vp]),
        "jq_mlp_train_step": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int,
                                             c_vp, c_vp, ctypes.c_int,
                                             ctypes.c_float, c_vp, c_vp]),
        "jq_mlp_num_params": (ctypes.c_size_t, [ctypes.c_int] * 3),
        "jq_mlp_workspace_floats": (ctypes.c_size_t, [ctypes.c_int] * 4),
        "jq_mlp_noalloc_max_samples": (ctypes.c_size_t, []),
        "jq_mlp_train_step_noalloc": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_int,
                                                     c_vp, c_vp, ctypes.c_int,
                                                     ctypes.c_float, c_vp, c_vp]),
        "jq_last_error": (c_cp, []),
        "jq_device_error_flag": (ctypes.c_int, [P(c_u32), ctypes.c_int]),
        "jq_scan_sync_mode": (ctypes.c_int, [ctypes.c_int]),
        "jq_wait_limit": (c_u64, [c_u64]),
    }
    for name, (res, args) in sig.items():
        try:
            f = getattr(L, name)
        except AttributeError:
            continue  # reported by missing_hip_symbols(); calling it raises
        f.restype = res
        f.argtypes = args
    _hip = L
    return L


def missing_hip_symbols():
    L = hip()
    out = []
    for name in HIP_SYMBOLS:
        try:
            getattr(L, name)
        except AttributeError:
            out.append(name)
    return out


def check(rc, lib=None):
    if rc != 0:
        lib = lib or hip()
        msg = lib.jq_last_error().decode()
        raise RuntimeError(f"jacquard HIP backend error {rc}: {msg}")


def host_error():
    return host().jqh_last_error().decode()
