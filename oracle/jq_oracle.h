/*
 * jq_oracle.h -- CPU oracle for the packed Boolean-processor interpreter.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (jacquard_amd/) links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / the timed
 * CPU baseline.  The product path fails loudly when the HIP library is
 * missing instead of falling back to this.
 *
 * Every function restates a reference function (cited per function in
 * jq_oracle.c).  All arithmetic is integer; results are bit-exact.
 */
#ifndef JQ_ORACLE_H
#define JQ_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Size of one AIGPDK SRAM in u32 entries (src/aigpdk.rs:20-22). */
#define JQO_SRAM_SIZE 8192u

/* Event record, 32 bytes (csrc/event_buffer.h:35-41). */
typedef struct {
  uint32_t event_type;
  uint32_t message_id;
  uint32_t cycle;
  uint32_t reserved;
  uint32_t data[4];
} jqo_event;

#define JQO_EVENT_SETUP 4u
#define JQO_EVENT_HOLD 5u

/* Literal restatement of cpu_reference.rs simulate_block_v1 (2-state).
 * Returns 0 on success, -1 when the script walk does not end exactly at
 * script_len (the reference asserts this). */
int jqo_simulate_block_v1(const uint32_t *script, size_t script_len,
                          const uint32_t *input_state, uint32_t *output_state,
                          uint32_t *sram_data);

/* Restatement of cpu_reference.rs simulate_block_v1_xprop.  X masks live in
 * separate arrays indexed like the value arrays.  cpu_quirk != 0 reproduces
 * the reference CPU's early return (a non-X-capable partition runs the rest
 * of the block in 2-state mode); cpu_quirk == 0 follows the GPU kernel's
 * per-partition rule. */
int jqo_simulate_block_v1_xprop(const uint32_t *script, size_t script_len,
                                const uint32_t *input_state,
                                uint32_t *output_state,
                                const uint32_t *input_xmask,
                                uint32_t *output_xmask, uint32_t *sram_data,
                                uint32_t *sram_xmask, int cpu_quirk);

/* One block of one major stage with the GPU kernel's full semantics:
 * value lane, per-partition X lane (X words at header word 9 offset inside
 * the same snapshot), arrival tracking and setup/hold events.
 *   constraints : per-state-word constraint words (no clock word) or NULL
 *   clock_ps    : clock period in ps
 *   arrival_state_offset : 0 disables arrival write-back
 *   events/ev_cap/ev_count : event sink (NULL to disable); events beyond
 *                 ev_cap are counted in *ev_count but not stored.
 * Returns 0 on success, -1 on a malformed script. */
int jqo_block_kernel(const uint32_t *script, size_t script_len,
                     const uint32_t *input_state, uint32_t *output_state,
                     uint32_t *sram_data, uint32_t *sram_xmask,
                     const uint32_t *constraints, uint32_t clock_ps,
                     int32_t arrival_state_offset, uint32_t cycle,
                     jqo_event *events, size_t ev_cap, size_t *ev_count);

/* Whole-run driver with the contract of
 * simulate_v1_noninteractive_timed_hip (csrc/kernel_v1.hip.cpp:69-100):
 * states holds (num_cycles+1) snapshots of state_size words, snapshot c is
 * the input of cycle c and snapshot c+1 (pre-filled by the caller) receives
 * its outputs.  timing_constraints = [clock_ps, words...] or NULL.
 * Returns 0 or -1. */
int jqo_run(size_t num_blocks, size_t num_major_stages,
            const uint64_t *blocks_start, const uint32_t *blocks_data,
            uint32_t *sram_data, uint32_t *sram_xmask, size_t num_cycles,
            size_t state_size, uint32_t *states,
            const uint32_t *timing_constraints, int32_t arrival_state_offset,
            jqo_event *events, size_t ev_cap, size_t *ev_count);

/* Same, but with the literal 2-state restatement (sanity_check_cpu loop,
 * cpu_reference.rs:745-790).  Used as the timed CPU baseline. */
int jqo_run_v1(size_t num_blocks, size_t num_major_stages,
               const uint64_t *blocks_start, const uint32_t *blocks_data,
               uint32_t *sram_data, size_t num_cycles, size_t state_size,
               uint32_t *states);

/* Block-parallel variant of jqo_run_v1 (blocks of one major stage are
 * independent); nthreads <= 0 uses all cores.  Reported separately. */
int jqo_run_v1_parallel(size_t num_blocks, size_t num_major_stages,
                        const uint64_t *blocks_start,
                        const uint32_t *blocks_data, uint32_t *sram_data,
                        size_t num_cycles, size_t state_size,
                        uint32_t *states, int nthreads);

/* ---------------- independent direct AIG simulator ----------------------
 * A mapping-free cycle simulator of the AIG itself, used as a closed-form
 * cross-check of script interpretation.  Pins 1..num_pins; pin 0 is Tie0.
 * drv_type: 0 Tie0, 1 AND, 2 input port, 3 clock flag, 4 DFF Q, 5 SRAM rd.
 * and_a/and_b: AND operands as pin<<1|inv.  Values are 3-valued when
 * xprop != 0 (value, xmask) using the kernel's AND formula. */
typedef struct {
  uint32_t num_pins;
  const uint8_t *drv_type;
  const uint32_t *and_a;
  const uint32_t *and_b;
  uint32_t num_dffs;
  const uint32_t *dff_d_iv;
  const uint32_t *dff_en_iv;
  const uint32_t *dff_q;
  uint32_t num_srams;
  const uint32_t *sram_r_addr_iv; /* num_srams*13 */
  const uint32_t *sram_r_en_iv;   /* num_srams    */
  const uint32_t *sram_rd_pin;    /* num_srams*32 (0 = unused) */
  const uint32_t *sram_w_addr_iv; /* num_srams*13 */
  const uint32_t *sram_w_en_iv;   /* num_srams*32 */
  const uint32_t *sram_w_data_iv; /* num_srams*32 */
} jqo_aig;

/* Simulate num_cycles cycles.
 *   in_bits  : num_cycles * num_inputs input values (0/1), in_pins order
 *   obs_pins : pins whose combinational value is recorded every cycle
 *   obs_out  : num_cycles * num_obs values (bit0 value, bit1 X)
 *   dff_init_x : per-DFF initial X flag (NULL = all known)
 *   dff_init_v : per-DFF initial value, bit 0 (NULL = all 0); with it a run
 *                can start from any register state, e.g. a snapshot taken
 *                in the middle of a long simulation
 *   reg_out  : if non-NULL, num_cycles * num_dffs DFF values after commit
 * Returns 0. */
int jqo_aig_sim(const jqo_aig *g, size_t num_cycles, size_t num_inputs,
                const uint32_t *in_pins, const uint8_t *in_bits,
                size_t num_obs, const uint32_t *obs_pins, uint8_t *obs_out,
                int xprop, const uint8_t *dff_init_x, const uint8_t *dff_init_v,
                uint8_t *reg_out);

#ifdef __cplusplus
}
#endif
#endif
