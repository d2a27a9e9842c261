//! Raw bindings to libjacquard_hip.so, the MI355X (gfx950) interpreter of
//! Jacquard's packed Boolean-processor scripts.
//!
//! The two `simulate_v1_noninteractive_*_hip` functions are the reference's
//! launch symbols with the reference's signatures (`csrc/kernel_v1.hip.cpp:34-100`),
//! so `sim_hip` (`src/bin/jacquard.rs:1258-1318`) calls them unchanged; the
//! `jq_*` functions are this backend's additions (status-returning launches,
//! the batched lane-sliced form, timing samples, the cost model's network,
//! output change extraction, co-simulation edges).  Every pointer is a
//! device pointer owned by the caller; launches are asynchronous on the
//! given stream (the reference symbols on the null stream); `jq_*`
//! functions return 0 or a negative code whose message `jq_last_error`
//! gives.  The declarations mirror include/jacquard_hip.h one for one.
#![no_std]

use core::ffi::{c_char, c_int, c_void};

// The C side's `usize` is a 64-bit integer, as the reference's Rust usize
// on its 64-bit hosts.
const _: () = assert!(core::mem::size_of::<usize>() == 8);

extern "C" {
    /// Replaces csrc/kernel_v1.hip.cpp:34-66 (CUDA twin csrc/kernel_v1.cu:17-47): the
    /// reference's launch symbol, same signature and semantics; exits the
    /// process on a launch error, as the reference does.
    pub fn simulate_v1_noninteractive_simple_scan_hip(
        num_blocks: usize,
        num_major_stages: usize,
        blocks_start: *const usize,
        blocks_data: *const u32,
        sram_data: *mut u32,
        sram_xmask: *mut u32,
        num_cycles: usize,
        state_size: usize,
        states_noninteractive: *mut u32,
        arrival_state_offset: i32,
    );
    /// Replaces csrc/kernel_v1.hip.cpp:69-100 (CUDA twin csrc/kernel_v1.cu:50-78):
    /// as above plus setup/hold checks into the 1024-record event buffer.
    pub fn simulate_v1_noninteractive_timed_hip(
        num_blocks: usize,
        num_major_stages: usize,
        blocks_start: *const usize,
        blocks_data: *const u32,
        sram_data: *mut u32,
        sram_xmask: *mut u32,
        num_cycles: usize,
        state_size: usize,
        states_noninteractive: *mut u32,
        timing_constraints: *const u32,
        event_buffer: *mut u8,
        arrival_state_offset: i32,
    );
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_simulate_scan(
        num_blocks: usize,
        num_major_stages: usize,
        blocks_start: *const usize,
        blocks_data: *const u32,
        sram_data: *mut u32,
        sram_xmask: *mut u32,
        num_cycles: usize,
        state_size: usize,
        states: *mut u32,
        timing_constraints: *const u32,
        event_buffer: *mut u8,
        event_capacity: u32,
        arrival_state_offset: i32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_simulate_scan_profiled(
        num_blocks: usize,
        num_major_stages: usize,
        blocks_start: *const usize,
        blocks_data: *const u32,
        sram_data: *mut u32,
        sram_xmask: *mut u32,
        num_cycles: usize,
        state_size: usize,
        states: *mut u32,
        part_index: *const u32,
        num_parts: u32,
        samples: *mut u32,
        sample_cycles: u32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_simulate_batched(
        num_groups: usize,
        blocks_per_group: usize,
        num_major_stages: usize,
        blocks_start: *const usize,
        blocks_data: *const u32,
        gather_table: *const u32,
        gather_start: *const u32,
        reg_io_state_size: u32,
        sram_storage_size: u32,
        sram_data: *mut u32,
        states: *mut u32,
        data_bits: *const u32,
        n_data: u32,
        flag_bits: *const u32,
        n_flag: u32,
        n_lanes: u32,
        lane_cycles: *const u32,
        lane_seed: *const u32,
        lane_thresh: *const u32,
        lane_info: *mut u32,
        slots: *mut u32,
        queue: *mut u32,
        group_cycles: *mut u32,
        obs_bits: *const u32,
        num_obs: u32,
        obs_capacity: u32,
        obs_out: *mut u32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_simulate_batched_profiled(
        num_groups: usize,
        blocks_per_group: usize,
        num_major_stages: usize,
        blocks_start: *const usize,
        blocks_data: *const u32,
        gather_table: *const u32,
        gather_start: *const u32,
        reg_io_state_size: u32,
        sram_storage_size: u32,
        sram_data: *mut u32,
        states: *mut u32,
        data_bits: *const u32,
        n_data: u32,
        flag_bits: *const u32,
        n_flag: u32,
        n_lanes: u32,
        lane_cycles: *const u32,
        lane_seed: *const u32,
        lane_thresh: *const u32,
        lane_info: *mut u32,
        slots: *mut u32,
        queue: *mut u32,
        group_cycles: *mut u32,
        obs_bits: *const u32,
        num_obs: u32,
        obs_capacity: u32,
        obs_out: *mut u32,
        num_parts: u32,
        samples: *mut u32,
        sample_cycles: u32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_simulate_batched_chunk(
        num_groups: usize,
        blocks_per_group: usize,
        num_major_stages: usize,
        blocks_start: *const usize,
        blocks_data: *const u32,
        gather_table: *const u32,
        gather_start: *const u32,
        reg_io_state_size: u32,
        sram_storage_size: u32,
        sram_data: *mut u32,
        states: *mut u32,
        data_bits: *const u32,
        n_data: u32,
        flag_bits: *const u32,
        n_flag: u32,
        n_lanes: u32,
        lane_cycles: *const u32,
        lane_seed: *const u32,
        lane_thresh: *const u32,
        lane_info: *mut u32,
        slots: *mut u32,
        queue: *mut u32,
        group_cycles: *mut u32,
        max_group_cycles: u32,
        obs_bits: *const u32,
        num_obs: u32,
        obs_capacity: u32,
        obs_out: *mut u32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_simulate_batched_stimulus(
        num_groups: usize,
        blocks_per_group: usize,
        num_major_stages: usize,
        blocks_start: *const usize,
        blocks_data: *const u32,
        gather_table: *const u32,
        gather_start: *const u32,
        reg_io_state_size: u32,
        sram_storage_size: u32,
        sram_data: *mut u32,
        states: *mut u32,
        in_bits: *const u32,
        n_in: u32,
        in_words: u32,
        stimulus: *const u32,
        stimulus_offset: *const usize,
        n_lanes: u32,
        lane_cycles: *const u32,
        lane_info: *mut u32,
        slots: *mut u32,
        queue: *mut u32,
        group_cycles: *mut u32,
        max_group_cycles: u32,
        obs_bits: *const u32,
        num_obs: u32,
        obs_capacity: u32,
        obs_out: *mut u32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_max_resident_blocks(
        kind: c_int,
        out_blocks: *mut c_int,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_device_error_flag(
        out: *mut u32,
        clear: c_int,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_scan_sync_mode(mode: c_int) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_wait_limit(polls: usize) -> usize;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_output_changes(
        states: *const u32,
        snapshot_offsets: *const usize,
        num_snapshots: u32,
        output_bits: *const u32,
        num_outputs: u32,
        xmask_offset: u32,
        counts: *mut u32,
        records: *mut u32,
        record_offsets: *const usize,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_lds_read_probe(
        num_blocks: u32,
        iters: u32,
        sink: *mut u32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_cosim_flash_state_bytes() -> usize;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_cosim_edges(
        num_blocks: usize,
        num_major_stages: usize,
        blocks_start: *const usize,
        blocks_data: *const u32,
        sram_data: *mut u32,
        state_size: usize,
        states: *mut u32,
        edge_ops: *const u32,
        edge_op_start: *const u32,
        num_sched: u32,
        sched_offset: u32,
        num_edges: u32,
        flash_state: *mut c_void,
        flash_pins: *const u32,
        flash_data: *const u8,
        flash_data_size: u32,
        uart_state: *mut c_void,
        uart_params: *const c_void,
        uart_channels: *mut c_void,
        bus_trace: *mut c_void,
        bus_trace_params: *const c_void,
        timing_constraints: *const u32,
        event_buffer: *mut u8,
        event_capacity: u32,
        arrival_state_offset: i32,
        trace: *mut u32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_mlp_num_params(
        in_: c_int,
        h1: c_int,
        h2: c_int,
    ) -> usize;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_mlp_workspace_floats(
        in_: c_int,
        h1: c_int,
        h2: c_int,
        n: c_int,
    ) -> usize;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_mlp_forward(
        params: *const f32,
        in_: c_int,
        h1: c_int,
        h2: c_int,
        x: *const f32,
        n: c_int,
        y: *mut f32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_mlp_train_step(
        params: *mut f32,
        grads: *mut f32,
        workspace: *mut f32,
        in_: c_int,
        h1: c_int,
        h2: c_int,
        x: *const f32,
        target: *const f32,
        n: c_int,
        lr: f32,
        loss_out: *mut f32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_mlp_noalloc_max_samples() -> usize;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_mlp_train_step_noalloc(
        params: *mut f32,
        grads: *mut f32,
        in_: c_int,
        h1: c_int,
        h2: c_int,
        x: *const f32,
        target: *const f32,
        n: c_int,
        lr: f32,
        loss_out: *mut f32,
        stream: *mut c_void,
    ) -> c_int;
    /// No counterpart in the reference; see include/jacquard_hip.h.
    pub fn jq_last_error() -> *const c_char;
}

/// Turns a `jq_*` status code into a `Result`, with the library's message
/// for a failure.
///
/// # Safety
/// Must be called on the thread that made the failing call (the message is
/// per thread), before another call of this library on that thread.
pub unsafe fn check(rc: c_int) -> Result<(), &'static core::ffi::CStr> {
    if rc == 0 {
        Ok(())
    } else {
        Err(core::ffi::CStr::from_ptr(jq_last_error()))
    }
}
