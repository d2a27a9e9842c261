// Seeded synthetic netlist generators for the benchmark workloads.
#pragma once

#include <cstdint>
#include <vector>

#include "aig.hpp"

namespace jq {

// SplitMix64: small, seedable, identical on every platform.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  uint32_t below(uint32_t n) { return (uint32_t)((next() >> 32) * n >> 32); }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct RandomAigParams {
  uint64_t seed = 1;
  uint32_t num_gates = 2000;   // AND gates
  uint32_t depth = 24;         // logic levels
  uint32_t num_dffs = 256;
  uint32_t num_inputs = 32;    // primary inputs (excluding the clock)
  uint32_t num_outputs = 16;
  // fan-in locality: 0 = uniform over earlier levels, else geometric(p) over
  // level distance
  double geometric_p = 0.0;
  double invert_prob = 0.2;    // probability an operand literal is inverted
  uint32_t num_srams = 0;      // 1024x32 single-port macros
  uint32_t sram_addr_bits = 10;
  // timing: per-gate integer delays drawn from [delay_lo, delay_hi] ps
  uint32_t delay_lo = 0, delay_hi = 0;
};

// Pins created: num_inputs input ports, one clock port (input) and its
// posedge flag, then DFF Qs, SRAM read-data pins, then AND gates level by
// level.  Every DFF is clocked by the posedge flag.
void gen_random_aig(const RandomAigParams &p, Aig &g);

struct MultiplierParams {
  uint64_t seed = 7;
  uint32_t num_tiles = 64;       // 16x16 AND-array multipliers
  uint32_t width = 16;
  uint32_t control_gates = 2000000;
  uint32_t control_depth = 128;
  uint32_t num_dffs = 1u << 20;  // total DFFs
  uint32_t num_inputs = 512;
  uint32_t num_outputs = 256;
};

// Structured netlist: num_tiles array multipliers (partial-product AND array
// + ripple adders built from AND/INV, pipeline DFF on every product bit and
// operand bit) plus a random control AIG.
void gen_multiplier_aig(const MultiplierParams &p, Aig &g);

// Port view of a generated netlist: inputs in[0..n-1] then the clock.
void fill_port_view(Aig &g, uint32_t num_inputs, uint32_t clk, uint32_t clk_flag);

}  // namespace jq
