/*
 * jacquard_hip.h -- C ABI of the MI355X (gfx950) Boolean-processor backend.
 *
 * libjacquard_hip.so exports the reference's two launch symbols with the
 * same signatures and semantics, so the existing host binding links against
 * it unchanged, plus new entry points for status-returning launches, the
 * batched lane-sliced form, per-partition timing samples and the learned
 * cost model's MLP.
 *
 * Conventions shared by every entry point:
 *   - all pointers are device pointers owned by the caller; the library
 *     allocates no memory on the caller's behalf except where stated;
 *   - launches are asynchronous on the given stream (the reference symbols
 *     use the null stream); the caller synchronises;
 *   - "usize" is 64-bit, as the reference's Rust usize.
 */
#ifndef JACQUARD_HIP_H
#define JACQUARD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef uint64_t usize;
typedef uint32_t u32;
typedef int32_t i32;
typedef uint8_t u8;
typedef uint16_t u16;

/* ---- Reference drop-in symbols ------------------------------------------
 * Replaces csrc/kernel_v1.hip.cpp:34-66 (and the CUDA twin
 * csrc/kernel_v1.cu:17-47).  Runs num_cycles cycles of the packed script
 * over states_noninteractive[(num_cycles+1) * state_size] in place: snapshot
 * c is the input of cycle c, snapshot c+1 (primary inputs pre-filled by the
 * caller) receives its outputs.  No timing checks.  On a launch error prints
 * a message and exits with EXIT_FAILURE, as the reference does. */
void simulate_v1_noninteractive_simple_scan_hip(
    usize num_blocks, usize num_major_stages, const usize *blocks_start,
    const u32 *blocks_data, u32 *sram_data, u32 *sram_xmask, usize num_cycles,
    usize state_size, u32 *states_noninteractive, i32 arrival_state_offset);

/* Replaces csrc/kernel_v1.hip.cpp:69-100 (CUDA twin csrc/kernel_v1.cu:50-78).
 * As above, plus setup/hold checks: timing_constraints = [clock_ps,
 * per-state-word (setup<<16 | hold)...] or NULL, event_buffer = the 16-byte
 * header + 1024 x 32-byte Event layout of csrc/event_buffer.h:35-50 or NULL.
 * Events beyond 1024 set the overflow flag. */
void simulate_v1_noninteractive_timed_hip(
    usize num_blocks, usize num_major_stages, const usize *blocks_start,
    const u32 *blocks_data, u32 *sram_data, u32 *sram_xmask, usize num_cycles,
    usize state_size, u32 *states_noninteractive,
    const u32 *timing_constraints, u8 *event_buffer, i32 arrival_state_offset);

/* ---- Status-returning form (no counterpart in the reference) ------------
 * Same computation as the timed symbol with an explicit stream and an event
 * capacity of the caller's choosing (event_buffer must hold
 * 16 + 32*event_capacity bytes).  Returns 0 or a negative error code; the
 * message is available from jq_last_error().  For hosts that must not exit
 * on error (tests, services).  A call on a stream that is being captured
 * records the launch into the caller's graph as an ordinary kernel node
 * with grid barriers (it fails if the blocks cannot all be resident at
 * once); each replay runs the same cycles again over the same buffers, so
 * the caller restores what a replay must start from (snapshot 0, SRAM
 * contents, the event buffer's header). */
int jq_simulate_scan(usize num_blocks, usize num_major_stages,
                     const usize *blocks_start, const u32 *blocks_data,
                     u32 *sram_data, u32 *sram_xmask, usize num_cycles,
                     usize state_size, u32 *states,
                     const u32 *timing_constraints, u8 *event_buffer,
                     u32 event_capacity, i32 arrival_state_offset,
                     void *stream);

/* Per-partition timing samples (no counterpart in the reference; feeds the
 * learned cost model).  Runs like jq_simulate_scan and additionally writes,
 * for every partition execution of the first sample_cycles cycles, its
 * duration in GPU clock ticks into samples[cycle * num_parts + part], where
 * part is the partition's order of appearance in blocks_data (stage-major,
 * then block, then position in block).  samples is u32[sample_cycles *
 * num_parts]. */
int jq_simulate_scan_profiled(usize num_blocks, usize num_major_stages,
                              const usize *blocks_start, const u32 *blocks_data,
                              u32 *sram_data, u32 *sram_xmask,
                              usize num_cycles, usize state_size, u32 *states,
                              const u32 *part_index, u32 num_parts,
                              u32 *samples, u32 sample_cycles, void *stream);

/* ---- Batched lane-sliced form (no counterpart in the reference) ---------
 * Simulates n_lanes independent stimulus lanes of one script (value lane:
 * X masks and arrivals are not computed in this form) in a single
 * persistent launch.  Lanes run 32 to a group: state words are stored
 * bit-transposed, one u32 "lane-word" per state bit whose bit l belongs to
 * the lane in slot l of the group.  Each group runs behind its own barrier;
 * when a lane completes its cycles its slot takes the next lane of a device
 * queue (its state bit cleared, clock flags set, its SRAM zeroed), so lanes
 * of uneven length keep every slot busy.
 *   blocks_per_group : the script's num_blocks; the launch uses
 *                      num_groups * blocks_per_group co-resident workgroups
 *   blocks_data      : the script with every boomerang-stage section in the
 *                      lane layout (same size and offsets): words
 *                      1024*(w>>2) + 4t + (w&3), w < 16, hold thread t's
 *                      operand LDS indices m = 2w, 2w+1 (u16 each: the
 *                      row-swizzled index p ^ ((p>>5)&31) of the previous
 *                      layer position p = perm[t + 256m]); words 4096+4t+c,
 *                      c < 3, hold the xora/xorb/orb bit of thread t's 31
 *                      subtree gates (level-1 position 4096+t+256g for
 *                      g < 16, then levels 2..5); word 4096+4t+3 (t < 64)
 *                      the flags of positions 128+t, 192+t, 64+t, t at bits
 *                      3i..3i+2.  Headers, global-read rounds, SRAM/dup and
 *                      clock-enable sections are unchanged.
 *   gather_table     : u32[num_parts * 8192], the script's global reads
 *                      resolved per partition (blocks_data order): entry q
 *                      is the state bit landing at first-layer position q
 *                      (bit 31: read from the current snapshot) or
 *                      0xFFFFFFFF for a zero position
 *   gather_start     : u32[num_major_stages * blocks_per_group], index of
 *                      the first partition of each (stage, block) in it
 *   states           : num_groups * 2 * (reg_io_state_size*32) u32: two
 *                      ping-pong snapshots per group, both zero except the
 *                      clock-flag lane-words (~0)
 *   sram_data        : num_groups * 32 * sram_storage_size u32, zero
 *   data_bits[n_data]: state bit of each toggling primary input
 *   flag_bits[n_flag]: state bit of each clock flag
 *   lane_cycles/lane_seed/lane_thresh : per lane, in queue order; input bit
 *                      k of a lane toggles before its cycle c (c >= 1) when
 *                      mix32(mix32(seed + 0x9E3779B9*k) ^ c) < thresh,
 *                      mix32 = the "lowbias32" mixer
 *   slots            : num_groups * 32 u32, the lane of every slot at start
 *                      (0xFFFFFFFF = empty); updated in place
 *   queue            : one u32, the first lane not in a slot; updated
 *   lane_info        : 2 u32 per lane, written: group*32 + slot, and the
 *                      group cycle at which the lane started (pre-set by the
 *                      caller for the lanes placed in slots)
 *   group_cycles     : num_groups u32, read: the group cycle each group
 *                      starts from (0 for a fresh run, see
 *                      jq_simulate_batched_chunk); written: the group
 *                      cycle each group stopped at
 *   obs_bits[num_obs]: state bits captured after every group cycle into
 *                      obs_out[(g*obs_capacity + cycle)*num_obs + k]; lane
 *                      (g, l, start) sees its cycle c at group cycle start+c
 * Returns 0 or a negative error code (message from jq_last_error()). */
int jq_simulate_batched(usize num_groups, usize blocks_per_group,
                        usize num_major_stages, const usize *blocks_start,
                        const u32 *blocks_data, const u32 *gather_table,
                        const u32 *gather_start, u32 reg_io_state_size,
                        u32 sram_storage_size, u32 *sram_data, u32 *states,
                        const u32 *data_bits, u32 n_data, const u32 *flag_bits,
                        u32 n_flag, u32 n_lanes, const u32 *lane_cycles,
                        const u32 *lane_seed, const u32 *lane_thresh,
                        u32 *lane_info, u32 *slots, u32 *queue,
                        u32 *group_cycles, const u32 *obs_bits, u32 num_obs,
                        u32 obs_capacity, u32 *obs_out, void *stream);

/* Per-partition timing samples of the batched form (no counterpart in the
 * reference; the learned cost model trains on them).  Same launch as
 * jq_simulate_batched; additionally, for the first sample_cycles cycles of
 * group 0, writes the clock ticks of every partition execution into
 * samples[cycle * num_parts + part] (part = order of appearance in
 * blocks_data, as in gather_table). */
int jq_simulate_batched_profiled(
    usize num_groups, usize blocks_per_group, usize num_major_stages,
    const usize *blocks_start, const u32 *blocks_data, const u32 *gather_table,
    const u32 *gather_start, u32 reg_io_state_size, u32 sram_storage_size,
    u32 *sram_data, u32 *states, const u32 *data_bits, u32 n_data,
    const u32 *flag_bits, u32 n_flag, u32 n_lanes, const u32 *lane_cycles,
    const u32 *lane_seed, const u32 *lane_thresh, u32 *lane_info, u32 *slots,
    u32 *queue, u32 *group_cycles, const u32 *obs_bits, u32 num_obs,
    u32 obs_capacity, u32 *obs_out, u32 num_parts, u32 *samples,
    u32 sample_cycles, void *stream);

/* Resumable form of jq_simulate_batched (no counterpart in the reference):
 * the same launch, except that group_cycles is read as well as written --
 * each group resumes at group_cycles[g] (0 for a fresh run), its slots'
 * lanes at the start cycles recorded in lane_info -- and with
 * max_group_cycles > 0 each group stops after that many group cycles (0:
 * every lane runs to its end).  Successive launches over the same buffers
 * therefore run one batched job in steps, and give the same lanes the
 * same results as one launch. */
int jq_simulate_batched_chunk(
    usize num_groups, usize blocks_per_group, usize num_major_stages,
    const usize *blocks_start, const u32 *blocks_data, const u32 *gather_table,
    const u32 *gather_start, u32 reg_io_state_size, u32 sram_storage_size,
    u32 *sram_data, u32 *states, const u32 *data_bits, u32 n_data,
    const u32 *flag_bits, u32 n_flag, u32 n_lanes, const u32 *lane_cycles,
    const u32 *lane_seed, const u32 *lane_thresh, u32 *lane_info, u32 *slots,
    u32 *queue, u32 *group_cycles, u32 max_group_cycles, const u32 *obs_bits,
    u32 num_obs, u32 obs_capacity, u32 *obs_out, void *stream);

/* Batched lanes of recorded stimulus (no counterpart in the reference):
 * as jq_simulate_batched_chunk, with every lane's inputs taken from packed
 * input words instead of the hash.  Lane q (queue order) has
 * lane_cycles[q] + 1 snapshots of in_words words at
 * stimulus[stimulus_offset[q] + k * in_words], k = 0..lane_cycles[q]: the
 * first in_words words of the snapshots that parse_input_vcd
 * (src/sim/vcd_io.rs:347-455) writes for the lane's VCD, clock flags
 * included.  The state bits in_bits[0..n_in) (every primary input and
 * clock-flag bit) are set from snapshot k before cycle k - 1 runs; a lane
 * entering a slot gets its snapshot 0.  The caller sets the snapshot-0
 * input bits of the lanes placed in slots at start in `states`. */
int jq_simulate_batched_stimulus(
    usize num_groups, usize blocks_per_group, usize num_major_stages,
    const usize *blocks_start, const u32 *blocks_data, const u32 *gather_table,
    const u32 *gather_start, u32 reg_io_state_size, u32 sram_storage_size,
    u32 *sram_data, u32 *states, const u32 *in_bits, u32 n_in, u32 in_words,
    const u32 *stimulus, const usize *stimulus_offset, u32 n_lanes,
    const u32 *lane_cycles, u32 *lane_info, u32 *slots, u32 *queue,
    u32 *group_cycles, u32 max_group_cycles, const u32 *obs_bits, u32 num_obs,
    u32 obs_capacity, u32 *obs_out, void *stream);

/* Occupancy query used by the autotuner: the largest number of resident
 * workgroups (all co-resident, as the persistent launches require) for the
 * scan kernel (kind 0) or the batched kernel (kind 1) on the current device. */
int jq_max_resident_blocks(int kind, int *out_blocks);

/* Device wait time-out flag (no counterpart in the reference): *out is
 * nonzero when a bounded wait (a grid barrier or a dataflow wait) of a scan,
 * co-simulation or batched launch gave up (jq_wait_limit) since the flag
 * was last cleared.  Launches return before their kernels finish, so
 * callers read the flag after synchronising; the results of a launch that
 * timed out are not valid.  With clear != 0 the flag is reset, and when it
 * was set every barrier slot of the kind that timed out is zeroed as well
 * (abort words included), so later launches start from a consistent
 * barrier state; clear only while no launch runs on the device. */
int jq_device_error_flag(u32 *out, int clear);

/* Bound of the device waits (no counterpart in the reference, whose grid
 * barrier waits without a bound).  Every grid barrier and dataflow wait of
 * the scan, co-simulation and batched launches gives up after `polls`
 * polls (about 1-2 us each; default 2^23, several seconds): it raises the
 * flag of jq_device_error_flag and marks its launch aborted; the launch's
 * other waits notice within 256 polls and every workgroup leaves at its
 * next cycle, so a launch whose workgroups are not all resident ends soon
 * after one time-out.  polls > 0 sets the bound of the current device's
 * launches (capped at 2^32 - 1; each device keeps its own; set it only while
 * no launch runs on that device); 0 only queries.  Returns the previous
 * bound, or 0 if the device could not be reached.  For hosts that want a
 * shorter bound and for tests of the time-out path. */
usize jq_wait_limit(usize polls);

/* Synchronisation of value-only scan launches (no counterpart in the
 * reference, whose kernel always takes a grid barrier after each major
 * stage, csrc/kernel_v1_impl.cuh:565-626).  mode 1 (the default): every
 * launch without timing first derives from the script which blocks each
 * (block, major stage) reads, and a block then waits only for those
 * blocks' progress; launches of fewer than 8 cycles and scripts with more
 * than 256 blocks or 120 major stages still use grid barriers.  The
 * derived table is the one allocation these launches make: 4 bytes per
 * state word plus 256 bytes per (block, major stage), taken with
 * hipMallocAsync on the launch's stream and released with hipFreeAsync
 * after it (if the allocation fails, the launch uses grid barriers, and so
 * does a launch issued while its stream is being captured into a graph).
 * mode 0: grid barriers everywhere.  mode < 0
 * only queries.  Returns the previous mode.  Results are identical in both
 * modes; set it only while no scan launch is being issued. */
int jq_scan_sync_mode(int mode);

/* ---- Output change extraction (no counterpart in the reference) ---------
 * Replaces the host-side scan of every snapshot in write_output_vcd_impl
 * (src/sim/vcd_io.rs:758-820): for each of num_snapshots emitted snapshots
 * (word offsets snapshot_offsets[k] into states, in emission order), the
 * outputs whose value differs from the previous emitted snapshot's (every
 * output at k = 0).  output_bits[i] is output i's state bit, 0xFFFFFFFF for
 * a constant 0 or 0xFFFFFFFE for a constant 1; a value is 3 (X) when bit
 * xmask_offset*32 + state bit of the snapshot is set (xmask_offset = 0: no X
 * lane).  With records == NULL, writes counts[k]; with records, writes the
 * changes of snapshot k in output order at records[record_offsets[k] ...]
 * as (i << 2 | value).  Records are compacted on the device, so only they
 * need to cross PCIe. */
int jq_output_changes(const u32 *states, const usize *snapshot_offsets,
                      u32 num_snapshots, const u32 *output_bits, u32 num_outputs,
                      u32 xmask_offset, u32 *counts, u32 *records,
                      const usize *record_offsets, void *stream);

/* ---- LDS read-bandwidth probe (no counterpart in the reference) ---------
 * Measures the device's LDS read peak, the figure the interpreter kernels'
 * LDS roofline is quoted against: num_blocks workgroups of 256 threads each
 * read 64 KB of LDS iters times with conflict-free 16-byte reads
 * (num_blocks * 256 * 128 * iters bytes in all) and write one checksum word
 * per thread to sink (u32[num_blocks * 256]).  Asynchronous on `stream`;
 * the caller times it. */
int jq_lds_read_probe(u32 num_blocks, u32 iters, u32 *sink, void *stream);

/* ---- Reactive co-simulation edges ---------------------------------------
 * Replaces the Metal co-simulation's per-edge GPU pipeline
 * (src/sim/cosim_metal.rs:547-651 encoding csrc/kernel_v1.metal:678-1025):
 * for num_edges edges, schedule edge k = (sched_offset + e) % num_sched:
 *   1. the output snapshot states[ss .. 2ss) is copied to the input snapshot
 *      states[0 .. ss) and the edge's bit operations are applied in order
 *      (edge_ops: u32 pairs (state bit, value), edge k's pairs at
 *      [edge_op_start[k], edge_op_start[k+1]); the reference's BitOp
 *      layout, csrc/kernel_v1.metal:664-667);
 *   2. the SPI flash's data outputs d_i are written to its data-in bits;
 *   3. one cycle of the scan kernel over the two snapshots;
 *   4. the SPI flash model steps on the edge's outputs, then the UART
 *      decoders and the Wishbone bus trace (gpu_io_step,
 *      csrc/kernel_v1.metal:1103-1231);
 *   5. with trace != NULL, both snapshots are copied to trace[e*2ss ..].
 * The batch is one persistent cooperative launch (num_blocks workgroups, as
 * jq_simulate_scan).  state_size = ss (2-state, no timing).  flash_state:
 * the reference's 48-byte FlashState (jq_cosim_flash_state_bytes(),
 * csrc/kernel_v1.metal:731-749) or NULL for no flash; flash_pins: 11 u32 =
 * data-in bits d0..d3, SPI clock output bit, chip-select output bit,
 * data-out bits d0..d3 (0xFFFFFFFF: not connected), has_flash; flash_data:
 * the flash image of flash_data_size bytes.  uart_state / uart_params /
 * uart_channels: the reference's UartDecoderState[4], UartParams and
 * UartChannel[4] (csrc/kernel_v1.metal:1027-1053), or NULL;
 * bus_trace / bus_trace_params: its WbTraceChannel (entries follow the
 * 16-byte header) and WbTraceParams (:1061-1101), or NULL.
 * timing_constraints / event_buffer / event_capacity / arrival_state_offset:
 * as jq_simulate_scan, with state_size the effective size (values, then
 * arrivals); events carry cycle 0, as in the reference's co-simulation
 * dispatches (src/sim/cosim_metal.rs:591-600).  All pointers
 * are device pointers; the host keeps FlashState.in_reset and the reset bit
 * operations current between calls.  Bit positions past the snapshot are
 * ignored.  Asynchronous on `stream`. */
size_t jq_cosim_flash_state_bytes(void);
int jq_cosim_edges(usize num_blocks, usize num_major_stages,
                   const usize *blocks_start, const u32 *blocks_data,
                   u32 *sram_data, usize state_size, u32 *states,
                   const u32 *edge_ops, const u32 *edge_op_start, u32 num_sched,
                   u32 sched_offset, u32 num_edges, void *flash_state,
                   const u32 *flash_pins, const u8 *flash_data,
                   u32 flash_data_size, void *uart_state, const void *uart_params,
                   void *uart_channels, void *bus_trace,
                   const void *bus_trace_params, const u32 *timing_constraints,
                   u8 *event_buffer, u32 event_capacity, i32 arrival_state_offset,
                   u32 *trace, void *stream);

/* ---- Learned cost model MLP (no counterpart in the reference) -----------
 * Fully connected FP32 network in -> h1 -> h2 -> 1 with ReLU (in <= 32,
 * h1, h2 <= 64), row-major weights W1[h1][in], b1[h1], W2[h2][h1], b2[h2],
 * W3[h2], b3[1] packed in that order in `params`
 * (jq_mlp_num_params floats).  Forward: y[n] for x[n][in].  Train step: the
 * mean-squared-error gradient of the batch (deterministic: per-workgroup
 * partial sums in `workspace`, jq_mlp_workspace_floats floats, reduced in a
 * fixed order) written to `grads`, then params -= lr * grads, and the batch
 * loss to *loss_out (device pointer, may be NULL).  All on `stream`. */
size_t jq_mlp_num_params(int in, int h1, int h2);
size_t jq_mlp_workspace_floats(int in, int h1, int h2, int n);
int jq_mlp_forward(const float *params, int in, int h1, int h2,
                   const float *x, int n, float *y, void *stream);
int jq_mlp_train_step(float *params, float *grads, float *workspace, int in,
                      int h1, int h2, const float *x, const float *target,
                      int n, float lr, float *loss_out, void *stream);

/* Allocation-free train step (no counterpart in the reference; for callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * training loops, harnesses that forbid hipMalloc).  Same computation and
 * results as jq_mlp_train_step, with the workspace taken from a static
 * device array of the library (1.6 MB per GPU, held until the process
 * ends) instead of an argument.  n may not exceed
 * jq_mlp_noalloc_max_samples() (4096); larger batches return an error.
 * Calls, and the graph replays that contain them, share one workspace per
 * GPU: calls that overlap on one GPU (different streams, no ordering) race
 * and give wrong gradients.  Nothing in the workspace needs clearing: every
 * call writes all of it that it reads. */
size_t jq_mlp_noalloc_max_samples(void);
int jq_mlp_train_step_noalloc(float *params, float *grads, int in, int h1,
                              int h2, const float *x, const float *target,
                              int n, float lr, float *loss_out, void *stream);

/* Last error message of this thread ("" if none). */
const char *jq_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
