// C ABI of the host library (libjacquard_host.so), consumed by the Python
// package through ctypes.  Handles are opaque pointers; arrays are returned
// as (pointer, length) views valid until the handle is freed.
#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "aig.hpp"
#include "compile.hpp"
#include "netgen.hpp"
#include "verilog.hpp"

using namespace jq;

namespace {
thread_local std::string g_err;
thread_local std::string g_log;

template <class F>
int guard(F f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

struct ScriptH {
  Script s;
  std::vector<uint32_t> in_pos, in_pin, out_lit, out_pos, staged_pin, staged_pos;
  std::vector<uint32_t> part_x;
  std::vector<uint64_t> part_rec;  // 5 per partition
};
}  // namespace

extern "C" {

const char *jqh_last_error() { return g_err.c_str(); }
const char *jqh_last_log() { return g_log.c_str(); }

// ---------------------------------------------------------------- AIG
void *jqh_aig_random(uint64_t seed, uint32_t gates, uint32_t depth, uint32_t dffs,
                     uint32_t inputs, uint32_t outputs, double geometric_p,
                     double invert_prob, uint32_t srams, uint32_t sram_addr_bits,
                     uint32_t delay_lo, uint32_t delay_hi) {
  auto *g = new Aig();
  RandomAigParams p;
  p.seed = seed;
  p.num_gates = gates;
  p.depth = depth;
  p.num_dffs = dffs;
  p.num_inputs = inputs;
  p.num_outputs = outputs;
  p.geometric_p = geometric_p;
  p.invert_prob = invert_prob;
  p.num_srams = srams;
  p.sram_addr_bits = sram_addr_bits;
  p.delay_lo = delay_lo;
  p.delay_hi = delay_hi;
  if (guard([&] { gen_random_aig(p, *g); })) {
    delete g;
    return nullptr;
  }
  return g;
}

void *jqh_aig_multiplier(uint64_t seed, uint32_t tiles, uint32_t width,
                         uint32_t control_gates, uint32_t control_depth,
                         uint32_t dffs, uint32_t inputs, uint32_t outputs) {
  auto *g = new Aig();
  MultiplierParams p;
  p.seed = seed;
  p.num_tiles = tiles;
  p.width = width;
  p.control_gates = control_gates;
  p.control_depth = control_depth;
  p.num_dffs = dffs;
  p.num_inputs = inputs;
  p.num_outputs = outputs;
  if (guard([&] { gen_multiplier_aig(p, *g); })) {
    delete g;
    return nullptr;
  }
  return g;
}

void *jqh_aig_from_verilog(const char *text, size_t len) {
  auto *g = new Aig();
  if (guard([&] { read_aigpdk_verilog(std::string(text, len), *g); })) {
    delete g;
    return nullptr;
  }
  return g;
}

void jqh_aig_free(void *h) { delete static_cast<Aig *>(h); }

// counts: [num_pins, num_and, num_pos, num_dffs, num_srams, has_delays]
void jqh_aig_counts(void *h, uint64_t *out) {
  auto *g = static_cast<Aig *>(h);
  out[0] = g->num_pins;
  out[1] = g->num_and();
  out[2] = g->pos.size();
  out[3] = g->dffs.size();
  out[4] = g->srams.size();
  out[5] = g->delay_ps.empty() ? 0 : 1;
}

// Export the netlist as flat arrays (caller-allocated, sizes from counts).
void jqh_aig_export(void *h, uint8_t *drv, uint32_t *and_a, uint32_t *and_b,
                    uint32_t *ref, uint32_t *pos, uint32_t *dff_d, uint32_t *dff_en,
                    uint32_t *dff_q, uint32_t *sram_r_addr, uint32_t *sram_r_en,
                    uint32_t *sram_rd, uint32_t *sram_w_addr, uint32_t *sram_w_en,
                    uint32_t *sram_w_data, uint16_t *delay) {
  auto *g = static_cast<Aig *>(h);
  std::memcpy(drv, g->drv.data(), g->drv.size());
  std::memcpy(and_a, g->a.data(), g->a.size() * 4);
  std::memcpy(and_b, g->b.data(), g->b.size() * 4);
  std::memcpy(ref, g->ref.data(), g->ref.size() * 4);
  std::memcpy(pos, g->pos.data(), g->pos.size() * 4);
  for (size_t i = 0; i < g->dffs.size(); ++i) {
    dff_d[i] = g->dffs[i].d_iv;
    dff_en[i] = g->dffs[i].en_iv;
    dff_q[i] = g->dffs[i].q;
  }
  for (size_t s = 0; s < g->srams.size(); ++s) {
    const Sram &m = g->srams[s];
    for (int k = 0; k < kSramAddrBits; ++k) {
      sram_r_addr[s * 13 + k] = m.r_addr_iv[k];
      sram_w_addr[s * 13 + k] = m.w_addr_iv[k];
    }
    sram_r_en[s] = m.r_en_iv;
    for (int k = 0; k < 32; ++k) {
      sram_rd[s * 32 + k] = m.rd_pin[k];
      sram_w_en[s * 32 + k] = m.w_en_iv[k];
      sram_w_data[s * 32 + k] = m.w_data_iv[k];
    }
  }
  if (delay && !g->delay_ps.empty())
    std::memcpy(delay, g->delay_ps.data(), g->delay_ps.size() * 2);
}

const char *jqh_aig_po_name(void *h, uint32_t i) {
  auto *g = static_cast<Aig *>(h);
  return i < g->po_names.size() ? g->po_names[i].c_str() : "";
}

// port view sizes: [n_in_ports, n_out_ports]
void jqh_aig_port_counts(void *h, uint64_t *out) {
  auto *g = static_cast<Aig *>(h);
  out[0] = g->in_port_names.size();
  out[1] = g->out_port_names.size();
}
const char *jqh_aig_in_port(void *h, uint32_t i, uint32_t *pin, uint32_t *flag) {
  auto *g = static_cast<Aig *>(h);
  *pin = g->in_port_pins[i];
  *flag = g->in_port_flag[i];
  return g->in_port_names[i].c_str();
}
const char *jqh_aig_out_port(void *h, uint32_t i, uint32_t *lit) {
  auto *g = static_cast<Aig *>(h);
  *lit = g->out_port_lits[i];
  return g->out_port_names[i].c_str();
}
const char *jqh_aig_module(void *h) { return static_cast<Aig *>(h)->module_name.c_str(); }
// Instance name of DFF j ("" when the netlist carries none).
const char *jqh_aig_dff_name(void *h, uint32_t j) {
  auto *g = static_cast<Aig *>(h);
  return j < g->dff_names.size() ? g->dff_names[j].c_str() : "";
}

void jqh_aig_set_delays(void *h, const uint16_t *delay) {
  auto *g = static_cast<Aig *>(h);
  g->delay_ps.assign(delay, delay + g->num_pins + 1);
}

// ---------------------------------------------------------------- compile
void *jqh_compile(void *aig, uint32_t level_split, uint32_t max_inputs,
                  uint32_t max_outputs, uint32_t max_level_nodes,
                  uint32_t max_part_nodes, uint32_t multilevel,
                  uint32_t max_boomerang_stages) {
  auto *g = static_cast<Aig *>(aig);
  auto *c = new Compiled();
  CompileOptions o;
  o.level_split = level_split;
  o.multilevel = multilevel;
  o.max_boomerang_stages = max_boomerang_stages;
  if (max_inputs) o.max_inputs = max_inputs;
  if (max_outputs) o.max_outputs = max_outputs;
  if (max_level_nodes) o.max_level_nodes = max_level_nodes;
  if (max_part_nodes) o.max_part_nodes = max_part_nodes;
  if (guard([&] { compile_partitions(*g, o, *c, &g_log); })) {
    delete c;
    return nullptr;
  }
  return c;
}

void jqh_compiled_free(void *h) { delete static_cast<Compiled *>(h); }

// A new compiled design with the pairs (major, i, j) of `compiled` merged
// (merge_partitions); merged_out[k] = 1 for each pair that merged.
void *jqh_compiled_merge(void *aig, void *compiled, const uint32_t *pairs,
                         uint64_t n_pairs, uint8_t *merged_out) {
  auto *g = static_cast<Aig *>(aig);
  auto *c = static_cast<Compiled *>(compiled);
  auto *o = new Compiled();
  std::vector<uint8_t> merged;
  if (guard([&] { merge_partitions(*g, *c, pairs, n_pairs, *o, merged); })) {
    delete o;
    return nullptr;
  }
  for (size_t k = 0; k < merged.size(); ++k) merged_out[k] = merged[k];
  return o;
}

// A new compiled design with the partitions (major, i) of `compiled` each
// split in two (split_partitions); split_out[k] = 1 for each one split.
void *jqh_compiled_split(void *aig, void *compiled, const uint32_t *parts,
                         uint64_t n_parts, uint8_t *split_out) {
  auto *g = static_cast<Aig *>(aig);
  auto *c = static_cast<Compiled *>(compiled);
  auto *o = new Compiled();
  std::vector<uint8_t> split;
  if (guard([&] { split_partitions(*g, *c, parts, n_parts, *o, split); })) {
    delete o;
    return nullptr;
  }
  for (size_t k = 0; k < split.size(); ++k) split_out[k] = split[k];
  return o;
}

uint32_t jqh_compiled_part_endpoints(void *h, uint32_t m, uint32_t p) {
  return (uint32_t)static_cast<Compiled *>(h)->stages[m].parts[p].endpoints.size();
}

uint32_t jqh_compiled_num_major(void *h) {
  return (uint32_t)static_cast<Compiled *>(h)->stages.size();
}
uint32_t jqh_compiled_num_parts(void *h, uint32_t m) {
  return (uint32_t)static_cast<Compiled *>(h)->stages[m].parts.size();
}
uint32_t jqh_compiled_part_stages(void *h, uint32_t m, uint32_t p) {
  return (uint32_t)static_cast<Compiled *>(h)->stages[m].parts[p].stages.size();
}

// costs: concatenated per major stage, per partition (nullptr = heuristic)
void *jqh_flatten(void *aig, void *compiled, uint32_t num_blocks, int xprop,
                  const double *costs) {
  auto *g = static_cast<Aig *>(aig);
  auto *c = static_cast<Compiled *>(compiled);
  auto *h = new ScriptH();
  CompileOptions o;
  o.num_blocks = num_blocks;
  o.xprop = xprop != 0;
  std::vector<std::vector<double>> cv;
  if (costs) {
    size_t k = 0;
    for (auto &ms : c->stages) {
      cv.emplace_back(costs + k, costs + k + ms.parts.size());
      k += ms.parts.size();
    }
  }
  if (guard([&] { flatten(*g, *c, o, costs ? &cv : nullptr, h->s); })) {
    delete h;
    return nullptr;
  }
  // the maps as (key, value) arrays in key order; the dense arrays they
  // come from are not needed after flattening
  auto pairs = [](std::vector<uint32_t> &dense, std::vector<uint32_t> &key,
                  std::vector<uint32_t> &val) {
    for (size_t k = 0; k < dense.size(); ++k)
      if (dense[k] != kNone) {
        key.push_back((uint32_t)k);
        val.push_back(dense[k]);
      }
    std::vector<uint32_t>().swap(dense);
  };
  pairs(h->s.input_map, h->in_pin, h->in_pos);
  pairs(h->s.output_map, h->out_lit, h->out_pos);
  pairs(h->s.staged_io_map, h->staged_pin, h->staged_pos);
  for (auto x : h->s.part_x) h->part_x.push_back(x);
  for (auto &r : h->s.parts) {
    h->part_rec.push_back(r.major);
    h->part_rec.push_back(r.block);
    h->part_rec.push_back(r.index_in_stage);
    h->part_rec.push_back(r.offset);
    h->part_rec.push_back(r.length);
  }
  return h;
}

void jqh_script_free(void *h) { delete static_cast<ScriptH *>(h); }

// info: [num_blocks, num_major_stages, reg_io_state_size, sram_storage_size,
//        len(blocks_start), len(blocks_data), xprop, xprop_state_offset,
//        len(input_layout), len(input_map), len(output_map), len(staged),
//        num_parts, timing, num_srams]
void jqh_script_info(void *h, uint64_t *o) {
  auto *x = static_cast<ScriptH *>(h);
  auto *s = &x->s;
  o[0] = s->num_blocks;
  o[1] = s->num_major_stages;
  o[2] = s->reg_io_state_size;
  o[3] = s->sram_storage_size;
  o[4] = s->blocks_start.size();
  o[5] = s->blocks_data.size();
  o[6] = s->xprop;
  o[7] = s->xprop_state_offset;
  o[8] = s->input_layout.size();
  o[9] = x->in_pin.size();
  o[10] = x->out_lit.size();
  o[11] = x->staged_pin.size();
  o[12] = s->parts.size();
  o[13] = s->timing;
  o[14] = s->sram_offsets.size();
}

const uint64_t *jqh_script_blocks_start(void *h) {
  return static_cast<ScriptH *>(h)->s.blocks_start.data();
}
uint32_t *jqh_script_blocks_data(void *h) {
  return static_cast<ScriptH *>(h)->s.blocks_data.data();
}
const uint32_t *jqh_script_input_layout(void *h) {
  return static_cast<ScriptH *>(h)->s.input_layout.data();
}
void jqh_script_maps(void *h, const uint32_t **in_pin, const uint32_t **in_pos,
                     const uint32_t **out_lit, const uint32_t **out_pos,
                     const uint32_t **st_pin, const uint32_t **st_pos) {
  auto *x = static_cast<ScriptH *>(h);
  *in_pin = x->in_pin.data();
  *in_pos = x->in_pos.data();
  *out_lit = x->out_lit.data();
  *out_pos = x->out_pos.data();
  *st_pin = x->staged_pin.data();
  *st_pos = x->staged_pos.data();
}
const uint32_t *jqh_script_part_x(void *h) { return static_cast<ScriptH *>(h)->part_x.data(); }
const uint64_t *jqh_script_parts(void *h) { return static_cast<ScriptH *>(h)->part_rec.data(); }
const uint32_t *jqh_script_sram_offsets(void *h) {
  return static_cast<ScriptH *>(h)->s.sram_offsets.data();
}
const uint32_t *jqh_script_dff_data_pos(void *h) {
  return static_cast<ScriptH *>(h)->s.dff_data_pos.data();
}

int jqh_inject_timing(void *aig, void *h) {
  return guard([&] { inject_timing(*static_cast<Aig *>(aig), static_cast<ScriptH *>(h)->s); });
}

// out must hold 1 + reg_io_state_size words
void jqh_constraint_buffer(void *h, uint32_t clock_ps, uint16_t setup, uint16_t hold,
                           uint32_t *out) {
  auto v = constraint_buffer(static_cast<ScriptH *>(h)->s, clock_ps, setup, hold);
  std::memcpy(out, v.data(), v.size() * 4);
}

// inject_timing_to_script's slot rule over caller-given arrays (see
// inject_delays); returns the number of nonzero slots written.
uint64_t jqh_inject_delays(uint32_t *bd, uint64_t n_bd, const uint64_t *offsets,
                           const uint32_t *pin_start, const uint32_t *pins, uint64_t n_patch,
                           const uint16_t *rise_ps, const uint16_t *fall_ps, uint64_t n_pins) {
  return inject_delays(bd, n_bd, offsets, pin_start, pins, n_patch, rise_ps, fall_ps, n_pins);
}

// build_timing_constraint_buffer over per-DFF constraints; writes num_words
// words to out and returns the saturated clock period.
uint32_t jqh_constraint_words(uint32_t num_words, uint64_t clock_period_ps,
                              const uint16_t *setup_ps, const uint16_t *hold_ps,
                              const int16_t *clock_arrival_ps,
                              const uint32_t *data_state_pos, uint64_t n, uint32_t *out) {
  std::vector<DffConstraint> cs(n);
  for (uint64_t i = 0; i < n; ++i)
    cs[i] = {setup_ps[i], hold_ps[i], clock_arrival_ps[i], data_state_pos[i]};
  return constraint_words(num_words, clock_period_ps, cs.data(), n, out);
}

// Walks every partition of a script exactly as the device interpreter does
// and checks that every state, SRAM, LDS and script access stays in bounds
// (the interpreter itself trusts the script, as the reference kernel does).
// Returns 0, or -1 with the reason in jqh_last_error().
int jqh_validate(const uint64_t *bs, uint64_t n_bs, const uint32_t *bd, uint64_t n_bd,
                 uint64_t num_blocks, uint64_t num_major, uint64_t state_size,
                 uint64_t sram_size, uint64_t /*reserved*/, int32_t arrival_offset) {
  return guard([&] {
    auto bad = [](const std::string &m) { throw std::runtime_error("script: " + m); };
    if (n_bs != num_blocks * num_major + 1) bad("blocks_start length");
    for (uint64_t k = 0; k < num_blocks * num_major; ++k) {
      uint64_t st = bs[k], en = bs[k + 1];
      if (st > en || en > n_bd) bad("block bounds");
      uint64_t pi = st;
      for (;;) {
        if (pi + 256 > en) bad("header past block end");
        const uint32_t *h = bd + pi;
        uint32_t S = h[0];
        if (S == 0) {
          pi += 256;
          break;
        }
        uint32_t num_ios = h[2], io = h[3], ns = h[4], so = h[5], R = h[6], nd = h[7];
        bool x = h[8] != 0;
        uint64_t xoff = h[9];
        if (num_ios > 256 || ns * 4 + nd > 256 || ns + nd > num_ios) bad("writeout counts");
        uint64_t len = 256 + 512ull * R + 5120ull * S + 10240;
        if (pi + len > en) bad("partition past block end");
        if ((uint64_t)io + num_ios > state_size) bad("writeout range");
        if (x && xoff + io + num_ios > state_size) bad("X writeout range");
        if (arrival_offset && (uint64_t)arrival_offset + io + num_ios > state_size)
          bad("arrival range");
        if ((uint64_t)so + (uint64_t)ns * kSramWords > sram_size) bad("SRAM range");
        for (uint32_t t = 0; t < 256; ++t) {
          uint32_t hk = h[128 + t / 2];
          hk = (t & 1) ? hk >> 16 : hk & 0xffff;
          if ((hk >> 8) < S && t >= num_ios) bad("hook beyond num_ios");
        }
        const uint32_t *g = h + 256;
        for (uint32_t t = 0; t < 256; ++t) {
          uint32_t bits = 0;
          for (uint64_t r = 0; r < R; ++r) {
            uint32_t idx = g[512 * r + 2 * t], m = g[512 * r + 2 * t + 1];
            if (!m) continue;
            bits += (uint32_t)__builtin_popcount(m);
            uint64_t w = idx & 0x7fffffffu;
            if (w >= state_size || (x && xoff + w >= state_size)) bad("global read index");
          }
          if (bits > 32) bad("more than 32 global-read bits in one thread");
        }
        const uint32_t *sec = g + 512 * R;
        for (uint32_t s = 0; s < S; ++s, sec += 5120)
          for (uint32_t k = 0; k < 4096; ++k)
            if ((sec[k] & 0xffff) >= kWidth || (sec[k] >> 16) >= kWidth) bad("shuffle index");
        for (int which = 0; which < 2; ++which, sec += 5120) {
          uint32_t nt = which == 0 ? ns * 4 + nd : num_ios;
          for (uint32_t o = 0; o < 4; ++o)
            for (uint32_t t = 0; t < nt; ++t)
              for (uint32_t q = 0; q < 4; ++q) {
                uint32_t w = sec[1024 * o + 4 * t + q];
                if ((w & 0xffff) >= num_ios * 32 || (w >> 16) >= num_ios * 32)
                  bad("writeout permutation index");
              }
        }
        pi += len;
        if (h[1]) break;
      }
      if (pi != en) bad("block does not end at its last partition");
    }
  });
}

// 24 script features per partition for the learned cost model.  parts holds
// 5 u64 per partition (major, block, index, offset, length); out is
// float[n * 24].
int jqh_part_features(const uint32_t *bd, uint64_t n_bd, const uint64_t *parts,
                      uint64_t n, float *out) {
  return guard([&] {
    for (uint64_t p = 0; p < n; ++p) {
      uint64_t off = parts[5 * p + 3];
      if (off + 256 > n_bd) throw std::runtime_error("features: bad offset");
      const uint32_t *h = bd + off;
      uint32_t S = h[0], num_ios = h[2], ns = h[4], R = h[6], nd = h[7];
      float f[24] = {0};
      f[0] = (float)S;
      f[1] = (float)num_ios;
      f[2] = (float)ns;
      f[3] = (float)nd;
      f[4] = (float)R;
      f[5] = h[8] ? 1.f : 0.f;
      const uint32_t *g = h + 256;
      uint32_t entries = 0, bits = 0, staged = 0, maxbits = 0;
      for (uint32_t t = 0; t < 256; ++t) {
        uint32_t tb = 0;
        for (uint32_t r = 0; r < R; ++r) {
          uint32_t idx = g[512 * r + 2 * t], m = g[512 * r + 2 * t + 1];
          if (!m) continue;
          ++entries;
          tb += (uint32_t)__builtin_popcount(m);
          staged += idx >> 31;
        }
        bits += tb;
        maxbits = std::max(maxbits, tb);
      }
      f[6] = (float)entries;
      f[7] = (float)bits;
      f[8] = (float)maxbits;
      f[9] = entries ? (float)staged / entries : 0.f;
      uint32_t hooks = 0;
      for (uint32_t t = 0; t < 256; ++t) {
        uint32_t hk = h[128 + t / 2];
        hk = (t & 1) ? hk >> 16 : hk & 0xffff;
        hooks += (hk >> 8) < S;
      }
      f[10] = (float)hooks;
      const uint32_t *sec = g + 512 * R;
      uint64_t nzp = 0, pass = 0, xa = 0, xb = 0, dly = 0, conflicts = 0;
      for (uint32_t s = 0; s < S; ++s, sec += 5120) {
        for (uint32_t k = 0; k < 4096; ++k) {
          nzp += (sec[k] & 0xffff) != 0;
          nzp += (sec[k] >> 16) != 0;
        }
        // LDS bank pressure proxy: distinct words touched by one gather step
        for (uint32_t k = 0; k < 16; ++k) {
          uint32_t seen[8] = {0};
          for (uint32_t t = 0; t < 256; ++t) {
            uint32_t w = sec[1024 * (k / 4) + 4 * t + (k % 4)] & 0xffff;
            seen[(w >> 5) & 7] |= 1u << ((w >> 8) & 31);
          }
          for (auto v : seen) conflicts += (uint64_t)__builtin_popcount(v);
        }
        const uint32_t *fl = sec + 4096;
        for (uint32_t t = 0; t < 256; ++t) {
          xa += (uint64_t)__builtin_popcount(fl[4 * t]);
          xb += (uint64_t)__builtin_popcount(fl[4 * t + 1]);
          pass += (uint64_t)__builtin_popcount(fl[4 * t + 2]);
          dly += fl[4 * t + 3] != 0;
        }
      }
      f[11] = (float)nzp / 8192.f;
      f[12] = (float)pass / 4096.f;
      f[13] = (float)xa / 4096.f;
      f[14] = (float)xb / 4096.f;
      f[15] = (float)dly;
      f[16] = S ? (float)conflicts / (S * 16.f) : 0.f;
      uint64_t nz_sd = 0, nz_ck = 0, ck_placed = 0;
      for (uint32_t k = 0; k < 4096; ++k) {
        nz_sd += (sec[k] & 0xffff) != 0;
        nz_ck += (sec[5120 + k] & 0xffff) != 0;
      }
      for (uint32_t t = 0; t < 256; ++t)
        ck_placed += (uint64_t)(32 - __builtin_popcount(sec[5120 + 4096 + 4 * t + 1]));
      f[17] = (float)nz_sd;
      f[18] = (float)nz_ck;
      f[19] = (float)ck_placed;
      f[20] = (float)parts[5 * p + 4] / 10000.f;
      f[21] = (float)S * (float)R;
      f[22] = (float)bits / 256.f;
      f[23] = num_ios ? (float)hooks / num_ios : 0.f;
      std::memcpy(out + 24 * p, f, sizeof(f));
    }
  });
}


// Batched-form layout of every boomerang-stage section of a script: out is
// a copy of blocks_data in which each stage section (4096 permutation words
// + 256 x {xora, xorb, orb, delay}, src/flatten.rs:1027-1063) is replaced
// by the lane layout read by the batched kernel:
//   words 1024*(w>>2) + 4t + (w&3), w < 16: thread t's operand indices
//     m = 2w (low u16) and 2w+1 (high u16), the row-swizzled LDS index
//     q ^ ((q>>5)&31) of previous-layer position q = perm[t + 256m];
//   words 4096 + 4t + c, c < 3: xora/xorb/orb of thread t's 31 subtree
//     gates at bit g: level-1 position 4096+t+256g (g < 16), level 2
//     2048+t+256(g-16) (g < 24), level 3 1024+... (g < 28), level 4
//     512+... (g < 30), level 5 256+t (g = 30);
//   word 4096 + 4t + 3, t < 64: flags of positions 128+t, 192+t, 64+t, t
//     at bits 3i..3i+2.
int jqh_lane_blocks_data(const uint32_t *bd, uint64_t n_bd, const uint64_t *parts,
                         uint64_t n, uint32_t *out) {
  return guard([&] {
    std::memcpy(out, bd, n_bd * sizeof(uint32_t));
    std::vector<uint64_t> secs;
    for (uint64_t p = 0; p < n; ++p) {
      uint64_t off = parts[5 * p + 3];
      if (off + 256 > n_bd) throw std::runtime_error("lane layout: bad offset");
      uint32_t S = bd[off], R = bd[off + 6];
      uint64_t base = off + 256 + 512ull * R;
      if (base + 5120ull * (S + 2) > n_bd) throw std::runtime_error("lane layout: truncated");
      for (uint32_t k = 0; k < S; ++k) secs.push_back(base + 5120ull * k);
    }
    auto perm = [](const uint32_t *sec, uint32_t q) -> uint32_t {
      uint32_t w = sec[1024 * ((q & 31) >> 3) + 4 * (q >> 5) + ((q >> 1) & 3)];
      return (q & 1) ? (w >> 16) : (w & 0xffffu);
    };
    auto flag = [](const uint32_t *sec, uint32_t q, int c) -> uint32_t {
      return (sec[4096 + 4 * (q >> 5) + c] >> (q & 31)) & 1u;
    };
    auto sw = [](uint32_t q) { return q ^ ((q >> 5) & 31u); };
    static const uint32_t kBase[5] = {4096, 2048, 1024, 512, 256};
    static const uint32_t kCnt[5] = {16, 8, 4, 2, 1};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)secs.size(); ++i) {
      const uint32_t *sec = bd + secs[i];
      uint32_t *o = out + secs[i];
      for (uint32_t t = 0; t < 256; ++t) {
        for (uint32_t w = 0; w < 16; ++w) {
          uint32_t lo = sw(perm(sec, t + 256 * (2 * w)));
          uint32_t hi = sw(perm(sec, t + 256 * (2 * w + 1)));
          o[1024 * (w >> 2) + 4 * t + (w & 3)] = lo | (hi << 16);
        }
        uint32_t f[3] = {0, 0, 0};
        uint32_t g = 0;
        for (int lv = 0; lv < 5; ++lv)
          for (uint32_t k = 0; k < kCnt[lv]; ++k, ++g)
            for (int c = 0; c < 3; ++c) f[c] |= flag(sec, kBase[lv] + t + 256 * k, c) << g;
        uint32_t w3 = 0;
        if (t < 64) {
          const uint32_t pos[4] = {128 + t, 192 + t, 64 + t, t};
          for (int j = 0; j < 4; ++j)
            for (int c = 0; c < 3; ++c) w3 |= flag(sec, pos[j], c) << (3 * j + c);
        }
        o[4096 + 4 * t + 0] = f[0];
        o[4096 + 4 * t + 1] = f[1];
        o[4096 + 4 * t + 2] = f[2];
        o[4096 + 4 * t + 3] = w3;
      }
    }
  });
}


// Gather table of the batched form: for partition p (rows of `parts`, in
// blocks_data order) 8192 u32, entry q = the state bit that lands at
// first-layer position q after the global read (bit 31: from the current
// snapshot), 0xFFFFFFFF where the position stays zero.  Thread t's
// (idx, mask) entries of rounds 0..R-1 are taken in order, each mask's set
// bits ascending; the n-th of its B bits lands at position 32t + B-1-n
// (csrc/kernel_v1_impl.cuh:87-151).  start[m*num_blocks + b] = index of the
// first partition of block b in major stage m (0 for an empty block).
int jqh_gather_table(const uint32_t *bd, uint64_t n_bd, const uint64_t *parts, uint64_t n,
                     uint32_t num_blocks, uint32_t num_major, uint32_t *table,
                     uint32_t *start) {
  return guard([&] {
    for (uint64_t k = 0; k < (uint64_t)num_blocks * num_major; ++k) start[k] = 0;
    for (uint64_t p = n; p-- > 0;) {
      uint64_t key = parts[5 * p + 0] * num_blocks + parts[5 * p + 1];
      if (key >= (uint64_t)num_blocks * num_major)
        throw std::runtime_error("gather table: bad partition row");
      start[key] = (uint32_t)p;  // the lowest index wins
    }
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < (int64_t)n; ++p) {
      uint32_t *tab = table + (size_t)p * 8192;
      for (int q = 0; q < 8192; ++q) tab[q] = 0xFFFFFFFFu;
      uint64_t off = parts[5 * p + 3];
      uint32_t R = bd[off + 6];
      const uint32_t *g = bd + off + 256;
      for (uint32_t t = 0; t < 256; ++t) {
        uint32_t B = 0;
        for (uint32_t r = 0; r < R; ++r) B += (uint32_t)__builtin_popcount(g[512 * r + 2 * t + 1]);
        uint32_t nth = 0;
        for (uint32_t r = 0; r < R; ++r) {
          uint32_t idx = g[512 * r + 2 * t], m = g[512 * r + 2 * t + 1];
          while (m) {
            uint32_t k = (uint32_t)__builtin_ctz(m);
            m &= m - 1;
            tab[32 * t + (B - 1 - nth)] = (((idx & 0x7FFFFFFFu) * 32u + k) | (idx & 0x80000000u));
            ++nth;
          }
        }
      }
    }
  });
}

}  // extern "C"
