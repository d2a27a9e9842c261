// Seeded synthetic AIG generators (benchmark workloads).
#include "netgen.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace jq {

namespace {

// Per-level node pools with an "unused first" picker so that nearly every
// generated gate has a consumer (dead gates would be stripped by the
// partitioner and shrink the workload below its stated size).
struct LevelPools {
  std::vector<std::vector<uint32_t>> nodes, unused;
  uint32_t pick(uint32_t lvl, Rng &r) {
    auto &u = unused[lvl];
    if (!u.empty()) {
      uint32_t k = r.below((uint32_t)u.size());
      uint32_t v = u[k];
      u[k] = u.back();
      u.pop_back();
      return v;
    }
    auto &n = nodes[lvl];
    return n[r.below((uint32_t)n.size())];
  }
};

uint32_t lit_of(uint32_t pin, double inv_p, Rng &r) {
  return pin << 1 | (r.uniform() < inv_p ? 1u : 0u);
}

}  // namespace

std::vector<uint8_t> Aig::x_capable_pins() const {
  std::vector<uint8_t> x(num_pins + 1, 0);
  for (auto &d : dffs) x[d.q] = 1;
  for (auto &s : srams)
    for (int k = 0; k < 32; ++k)
      if (s.rd_pin[k]) x[s.rd_pin[k]] = 1;
  for (;;) {
    for (uint32_t p = 1; p <= num_pins; ++p)
      if (!x[p] && drv[p] == kAnd && (x[a[p] >> 1] || x[b[p] >> 1])) x[p] = 1;
    bool changed = false;
    for (auto &d : dffs)
      if (x[d.d_iv >> 1] && !x[d.q]) {
        x[d.q] = 1;
        changed = true;
      }
    if (!changed) break;
  }
  return x;
}

void gen_random_aig(const RandomAigParams &p, Aig &g) {
  if (p.depth < 1 || p.num_gates < p.depth)
    throw std::invalid_argument("gen_random_aig: need num_gates >= depth >= 1");
  Rng r(p.seed);
  g = Aig();
  LevelPools pools;
  pools.nodes.resize(p.depth + 1);
  pools.unused.resize(p.depth + 1);
  for (uint32_t i = 0; i < p.num_inputs; ++i) {
    uint32_t pin = g.add_pin(kInput, i);
    pools.nodes[0].push_back(pin);
  }
  uint32_t clk = g.add_pin(kInput, p.num_inputs);  // clock port itself
  (void)clk;
  uint32_t clk_flag = g.add_pin(kClockFlag, p.num_inputs);  // posedge flag
  g.dffs.resize(p.num_dffs);
  for (uint32_t i = 0; i < p.num_dffs; ++i) {
    g.dffs[i].q = g.add_pin(kDffQ, i);
    pools.nodes[0].push_back(g.dffs[i].q);
  }
  g.srams.resize(p.num_srams);
  for (uint32_t s = 0; s < p.num_srams; ++s)
    for (int k = 0; k < 32; ++k) {
      g.srams[s].rd_pin[k] = g.add_pin(kSramQ, s);
      pools.nodes[0].push_back(g.srams[s].rd_pin[k]);
    }
  pools.unused[0] = pools.nodes[0];
  std::vector<uint32_t> fanout(1, 0);
  fanout.resize(g.num_pins + 1, 0);

  const double lq = p.geometric_p > 0 ? std::log(1.0 - p.geometric_p) : 0.0;
  for (uint32_t lvl = 1; lvl <= p.depth; ++lvl) {
    uint32_t width = p.num_gates / p.depth + (lvl <= p.num_gates % p.depth);
    for (uint32_t k = 0; k < width; ++k) {
      uint32_t out = 0;
      for (int tries = 0; tries < 64; ++tries) {
        uint32_t x = pools.pick(lvl - 1, r);
        uint32_t lb;
        if (p.geometric_p > 0) {
          double u = r.uniform();
          uint32_t d = 1 + (uint32_t)std::floor(std::log(1.0 - u) / lq);
          lb = d >= lvl ? 0 : lvl - d;
        } else {
          lb = r.below(lvl);
        }
        uint32_t y = pools.pick(lb, r);
        if (x == y) continue;
        size_t before = g.num_pins;
        uint32_t o = g.add_and(lit_of(x, p.invert_prob, r),
                               lit_of(y, p.invert_prob, r));
        if (g.num_pins == before) continue;  // structural duplicate
        out = o >> 1;
        fanout.push_back(0);
        ++fanout[x];
        ++fanout[y];
        break;
      }
      if (!out) throw std::runtime_error("gen_random_aig: could not place gate");
      pools.nodes[lvl].push_back(out);
      pools.unused[lvl].push_back(out);
    }
  }
  if (!p.delay_hi) {
    g.delay_ps.clear();
  } else {
    g.delay_ps.assign(g.num_pins + 1, 0);
    for (uint32_t q = 1; q <= g.num_pins; ++q)
      if (g.drv[q] == kAnd)
        g.delay_ps[q] = (uint16_t)(p.delay_lo + r.below(p.delay_hi - p.delay_lo + 1));
  }
  // sinks first (gates without consumers), deepest levels first, then
  // random gates from the upper half
  std::vector<uint32_t> sinks;
  for (uint32_t lvl = p.depth; lvl >= 1; --lvl)
    for (uint32_t q : pools.nodes[lvl])
      if (!fanout[q]) sinks.push_back(q);
  size_t sink_i = 0;
  auto endpoint_pin = [&]() -> uint32_t {
    if (sink_i < sinks.size()) return sinks[sink_i++];
    uint32_t lvl = p.depth / 2 + 1 + r.below(p.depth - p.depth / 2);
    return pools.nodes[lvl][r.below((uint32_t)pools.nodes[lvl].size())];
  };
  for (uint32_t i = 0; i < p.num_dffs; ++i) {
    g.dffs[i].d_iv = lit_of(endpoint_pin(), 0.5, r);
    g.dffs[i].en_iv = clk_flag << 1;
  }
  for (uint32_t s = 0; s < p.num_srams; ++s) {
    Sram &m = g.srams[s];
    auto anypin = [&]() {
      uint32_t lvl = 1 + r.below(p.depth);
      return pools.nodes[lvl][r.below((uint32_t)pools.nodes[lvl].size())];
    };
    for (int k = 0; k < kSramAddrBits; ++k) {
      uint32_t l = (uint32_t)k < p.sram_addr_bits ? lit_of(anypin(), 0.5, r) : 0;
      m.r_addr_iv[k] = l;
      m.w_addr_iv[k] = l;  // single port: one shared address
    }
    uint32_t we = lit_of(anypin(), 0.5, r);
    uint32_t wclk = g.add_and(we, clk_flag << 1);
    m.r_en_iv = clk_flag << 1;
    for (int k = 0; k < 32; ++k) {
      m.w_en_iv[k] = wclk;
      m.w_data_iv[k] = lit_of(anypin(), 0.5, r);
    }
  }
  if (!g.delay_ps.empty()) g.delay_ps.resize(g.num_pins + 1, 0);
  std::vector<uint8_t> used(g.num_pins + 1, 0);
  for (uint32_t i = 0; i < p.num_outputs; ++i) {
    uint32_t q = 0;
    for (int t = 0; t < 16; ++t) {
      q = endpoint_pin();
      if (!used[q]) break;
    }
    used[q] = 1;
    g.pos.push_back(q << 1);
    g.po_names.push_back("out" + std::to_string(i));
  }
  fill_port_view(g, p.num_inputs, clk, clk_flag);
}

void fill_port_view(Aig &g, uint32_t num_inputs, uint32_t clk, uint32_t clk_flag) {
  g.module_name = "top";
  for (uint32_t i = 0; i < num_inputs; ++i) {
    g.in_port_names.push_back("in[" + std::to_string(i) + "]");
    g.in_port_pins.push_back(i + 1);
    g.in_port_flag.push_back(0);
  }
  g.in_port_names.push_back("clk");
  g.in_port_pins.push_back(clk);
  g.in_port_flag.push_back(clk_flag);
  g.out_port_names = g.po_names;
  g.out_port_lits = g.pos;
}

namespace {

// Two-input helpers on literals.
struct Logic {
  Aig &g;
  uint32_t AND(uint32_t x, uint32_t y) { return g.add_and(x, y); }
  uint32_t OR(uint32_t x, uint32_t y) { return g.add_and(x ^ 1, y ^ 1) ^ 1; }
  uint32_t XOR(uint32_t x, uint32_t y) {
    return OR(g.add_and(x, y ^ 1), g.add_and(x ^ 1, y));
  }
  void full_add(uint32_t x, uint32_t y, uint32_t z, uint32_t &s, uint32_t &c) {
    uint32_t t = XOR(x, y);
    s = XOR(t, z);
    c = OR(AND(x, y), AND(t, z));
  }
};

}  // namespace

void gen_multiplier_aig(const MultiplierParams &p, Aig &g) {
  Rng r(p.seed);
  g = Aig();
  std::vector<uint32_t> pis;
  for (uint32_t i = 0; i < p.num_inputs; ++i) pis.push_back(g.add_pin(kInput, i));
  uint32_t clk = g.add_pin(kInput, p.num_inputs);
  uint32_t clk_flag = g.add_pin(kClockFlag, p.num_inputs);
  Logic L{g};
  // operand registers of all tiles, then product registers; the control
  // block's registers take the remaining DFF budget
  const uint32_t W = p.width;
  uint32_t tile_regs = 2 * W + 2 * W;  // operands + product
  std::vector<uint32_t> opq;
  uint32_t dff_budget = p.num_dffs;
  auto new_dff = [&]() {
    Dff d;
    d.q = g.add_pin(kDffQ, (uint32_t)g.dffs.size());
    d.en_iv = clk_flag << 1;
    g.dffs.push_back(d);
    return (uint32_t)g.dffs.size() - 1;
  };
  // estimate gates of one multiplier: W*W partial products + (W-1)*W full
  // adders of ~9 ANDs
  uint32_t per_mult = W * W + (W - 1) * W * 9;
  uint64_t tile_gate_budget =
      p.num_tiles ? ((uint64_t)p.control_gates * 7) / p.num_tiles : 0;  // 16M-2M
  uint32_t mults_per_tile =
      std::max<uint32_t>(1, (uint32_t)(tile_gate_budget / per_mult));
  uint32_t num_mults = p.num_tiles * mults_per_tile;
  if ((uint64_t)num_mults * tile_regs > dff_budget / 2)
    num_mults = std::max<uint32_t>(1, dff_budget / 2 / tile_regs);
  // control block first: random logic over PIs and its own registers
  uint32_t ctrl_dffs = dff_budget - num_mults * tile_regs;
  std::vector<uint32_t> ctrl_q;
  for (uint32_t i = 0; i < ctrl_dffs; ++i) ctrl_q.push_back(g.dffs[new_dff()].q);
  std::vector<uint32_t> mult_dff0(num_mults);
  for (uint32_t m = 0; m < num_mults; ++m) {
    mult_dff0[m] = (uint32_t)g.dffs.size();
    for (uint32_t i = 0; i < tile_regs; ++i) new_dff();
  }
  // random control logic, levelized
  std::vector<std::vector<uint32_t>> lv(p.control_depth + 1);
  for (auto x : pis) lv[0].push_back(x);
  for (auto x : ctrl_q) lv[0].push_back(x);
  for (uint32_t m = 0; m < num_mults; ++m)
    for (uint32_t i = 2 * W; i < tile_regs; ++i)
      lv[0].push_back(g.dffs[mult_dff0[m] + i].q);  // products feed control
  for (uint32_t l = 1; l <= p.control_depth; ++l) {
    uint32_t w = p.control_gates / p.control_depth;
    for (uint32_t k = 0; k < w; ++k) {
      for (int t = 0; t < 16; ++t) {
        uint32_t x = lv[l - 1][r.below((uint32_t)lv[l - 1].size())];
        uint32_t ll = l - 1 - std::min<uint32_t>(l - 1, r.below(4));
        uint32_t y = lv[ll][r.below((uint32_t)lv[ll].size())];
        if (x == y) continue;
        size_t before = g.num_pins;
        uint32_t o = g.add_and(x << 1 | r.below(2), y << 1 | r.below(2));
        if (g.num_pins == before) continue;
        lv[l].push_back(o >> 1);
        break;
      }
    }
  }
  auto ctrl_pin = [&]() {
    uint32_t l = p.control_depth / 2 + r.below(p.control_depth / 2) + 1;
    return lv[l][r.below((uint32_t)lv[l].size())] << 1;
  };
  for (uint32_t i = 0; i < ctrl_dffs; ++i) g.dffs[i].d_iv = ctrl_pin();
  // multipliers: operand regs load from control logic, product regs load
  // the array product of the operand regs
  for (uint32_t m = 0; m < num_mults; ++m) {
    uint32_t d0 = mult_dff0[m];
    std::vector<uint32_t> A(W), B(W);
    for (uint32_t i = 0; i < W; ++i) {
      A[i] = g.dffs[d0 + i].q << 1;
      B[i] = g.dffs[d0 + W + i].q << 1;
      g.dffs[d0 + i].d_iv = ctrl_pin();
      g.dffs[d0 + W + i].d_iv = ctrl_pin();
    }
    // carry-save array: row j adds A*B[j] shifted by j
    std::vector<uint32_t> acc(2 * W, 0);
    for (uint32_t j = 0; j < W; ++j) {
      uint32_t carry = 0;
      for (uint32_t i = 0; i < W; ++i) {
        uint32_t pp = L.AND(A[i], B[j]);
        uint32_t s, c;
        L.full_add(acc[i + j], pp, carry, s, c);
        acc[i + j] = s;
        carry = c;
      }
      acc[W + j] = carry;
    }
    for (uint32_t i = 0; i < 2 * W; ++i) g.dffs[d0 + 2 * W + i].d_iv = acc[i];
  }
  for (uint32_t i = 0; i < p.num_outputs; ++i) {
    g.pos.push_back(ctrl_pin());
    g.po_names.push_back("out" + std::to_string(i));
  }
  // DFFs whose D folded to a constant would be unobservable; keep them
  // driven by a control pin instead
  for (auto &d : g.dffs)
    if (d.d_iv <= 1) d.d_iv = ctrl_pin();
  fill_port_view(g, p.num_inputs, clk, clk_flag);
}

}  // namespace jq
