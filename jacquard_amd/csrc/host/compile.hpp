// Script compiler: AIG -> major stages -> partitions -> boomerang stages ->
// FlattenedScriptV1 (the packed program the device interpreter runs).
//
// The output format is the reference's script ISA bit for bit
// (src/flatten.rs:198-258, 863-1130, 1133-1573); the partitioning and
// boomerang mapping algorithms are this project's own (any valid mapping
// yields the same simulated values).
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "aig.hpp"

namespace jq {

constexpr int kBoomStages = 13;                  // src/pe.rs:27
constexpr uint32_t kWidth = 1u << kBoomStages;   // 8192 bits
constexpr uint32_t kThreads = kWidth >> 5;       // 256 words

// One boomerang stage: hier[0] (8192 inputs) ... hier[13] (1 output).  A
// level left empty is all-unused.  write_outs lists after-layout words
// (0..255) whose values are written out after this stage.
struct BoomStage {
  std::vector<uint32_t> hier[kBoomStages + 1];
  std::vector<uint16_t> write_outs;
  uint32_t at(int h, uint32_t j) const {
    return hier[h].empty() ? kNone : hier[h][j];
  }
};

struct Partition {
  std::vector<uint32_t> endpoints;  // indices into MajorStage endpoint space
  std::vector<BoomStage> stages;
  size_t num_nodes = 0;             // AND gates mapped into it
};

// One level-split slice of the AIG (mirrors StagedAIG, src/staging.rs:12-21).
// Endpoint space: [staged_out pins | original endpoint groups].
struct MajorStage {
  std::vector<uint32_t> staged_out;  // pins produced here, read later
  std::vector<uint32_t> orig_endpoints;
  std::vector<Partition> parts;
  size_t num_endpoints() const { return staged_out.size() + orig_endpoints.size(); }
};

struct CompileOptions {
  uint32_t num_blocks = 256;
  // level-split period in logic levels (0 = a single major stage)
  uint32_t level_split = 0;
  bool xprop = false;
  bool timing = false;
  // partition sizing knobs (soft limits used by the greedy grouping)
  uint32_t max_inputs = 6144;
  uint32_t max_outputs = 2048;
  uint32_t max_level_nodes = 2600;
  uint32_t max_part_nodes = 24000;
  // highest tree level at which the mapper places AND gates (0: level-1
  // mapper only, one logic level per boomerang stage)
  uint32_t multilevel = 8;
  // a partition mapped to more boomerang stages than this is split in two
  // by endpoints, like one that does not fit (0: no limit)
  uint32_t max_boomerang_stages = 0;
};

struct Script {
  uint32_t num_blocks = 0, num_major_stages = 0;
  std::vector<uint64_t> blocks_start;
  std::vector<uint32_t> blocks_data;
  uint32_t reg_io_state_size = 0;
  uint32_t sram_storage_size = 0;  // u32 entries
  std::vector<uint32_t> input_layout;
  // the reference's three maps as dense arrays, kNone where a key has no
  // entry: indexed by pin (input_map, staged_io_map) or literal (output_map)
  std::vector<uint32_t> input_map;      // pin -> bit
  std::vector<uint32_t> output_map;     // literal -> bit
  std::vector<uint32_t> staged_io_map;  // pin -> bit
  std::vector<uint32_t> sram_offsets;   // per SRAM index
  bool xprop = false;
  uint32_t xprop_state_offset = 0;
  std::vector<uint8_t> part_x;  // flat over stages/blocks/parts
  // timing
  bool timing = false;
  std::vector<std::pair<uint64_t, std::vector<uint32_t>>> delay_patch;
  std::vector<uint32_t> dff_data_pos;  // per DFF: state bit of its D (or ~0)
  // per-partition records for cost models: stage, block, offset, length
  struct PartRec {
    uint32_t major, block, index_in_stage;
    uint64_t offset, length;
  };
  std::vector<PartRec> parts;
};

// Full pipeline state kept between the partitioning and flatten steps so
// block assignment can be redone with another cost model.
struct Compiled {
  std::vector<MajorStage> stages;
  std::vector<std::vector<uint8_t>> part_x;  // per stage, per partition
  uint32_t level_split = 1;                  // split period actually used
  CompileOptions opt;                        // options it was compiled with
};

// Level-split + partition + map.  Throws std::runtime_error on failure.
void compile_partitions(const Aig &g, const CompileOptions &opt, Compiled &out,
                        std::string *log = nullptr);

// Partition merge (the step of src/pe.rs:855-1063, with the acceptance
// decision left to the caller): `pairs` holds n_pairs (major, i, j) triples
// naming two partitions of one major stage of `in`; each pair is mapped as
// one partition over the union of their endpoints.  Pairs of a stage must
// be disjoint.  A pair whose union does not fit one partition keeps its two
// partitions.  `out` receives the design with every fitting pair replaced
// by its merged partition (partitions re-sorted by decreasing boomerang
// stages, as compile_partitions leaves them); merged[k] = 1 when pair k
// merged.  Throws std::runtime_error on a malformed pair list.
void merge_partitions(const Aig &g, const Compiled &in, const uint32_t *pairs,
                      size_t n_pairs, Compiled &out, std::vector<uint8_t> &merged);

// Partition cut, the converse: `parts` holds n_parts (major, i) pairs naming
// partitions of `in`; each is mapped again as two partitions, the first and
// the second half of its endpoints.  A partition with one endpoint, or whose
// halves do not both fit a partition, stays whole.  `out` receives the
// design with every split partition replaced by its halves (re-sorted by
// decreasing boomerang stages); split[k] = 1 when partition k was split.
// Throws std::runtime_error on a malformed list (a partition named twice).
void split_partitions(const Aig &g, const Compiled &in, const uint32_t *parts,
                      size_t n_parts, Compiled &out, std::vector<uint8_t> &split);

// Block assignment ("masonry", src/flatten.rs:1207-1226) with per-partition
// costs (nullptr = the reference heuristic stages+3) and flatten into the
// script format.  costs is indexed [major stage][partition].
void flatten(const Aig &g, const Compiled &c, const CompileOptions &opt,
             const std::vector<std::vector<double>> *costs, Script &s);

// Fill delay slots from per-pin delays (inject_timing_to_script,
// src/flatten.rs:1812-1846): slot k at blocks_data offset offsets[k] gets
// the largest max(rise, fall) over its pins pins[pin_start[k] ..
// pin_start[k+1]), clamped to 65535; pins past n_pins are skipped, offsets
// past n_bd are not written, and with no delays or no slots nothing is
// written.  Returns the number of nonzero slots written.
size_t inject_delays(uint32_t *bd, size_t n_bd, const uint64_t *offsets,
                     const uint32_t *pin_start, const uint32_t *pins, size_t n_patch,
                     const uint16_t *rise_ps, const uint16_t *fall_ps, size_t n_pins);
void inject_timing(const Aig &g, Script &s);

// One DFF's capture constraint (DFFConstraint, src/flatten.rs:77-95).
struct DffConstraint {
  uint16_t setup_ps, hold_ps;
  int16_t clock_arrival_ps;  // capture clock arrival, folded into setup/hold
  uint32_t data_state_pos;   // state bit of its D input, ~0 = none
};

// effective_setup_hold (src/flatten.rs:115-122): setup - arrival and
// hold + arrival, each clamped to [0, 65535].
void effective_setup_hold(const DffConstraint &c, uint16_t *setup, uint16_t *hold);

// build_timing_constraint_buffer (src/flatten.rs:1891-1924): out[w] =
// (min effective setup << 16) | min effective hold over the DFFs whose D
// lands in state word w (0 = no DFF); constraints without a position or
// past num_words are skipped.  Returns the clock period saturated to u32.
uint32_t constraint_words(uint32_t num_words, uint64_t clock_period_ps,
                          const DffConstraint *c, size_t n, uint32_t *out);

// [clock_ps, words...] for a compiled script with one setup/hold pair for
// every DFF.
std::vector<uint32_t> constraint_buffer(const Script &s, uint32_t clock_ps,
                                        uint16_t setup_ps, uint16_t hold_ps);

}  // namespace jq
