// Streaming reader of an input VCD's value-change section into packed
// state snapshots: the body loop of parse_input_vcd (src/sim/vcd_io.rs:
// 347-455), in one pass over the text with no per-token allocation.
//
// Semantics (the same as the reference reader):
//   - a snapshot is pushed at a timestamp that follows an active clock edge
//     (and once more at the end); its offset entry is (words pushed so far,
//     time of the previous active timestamp, u64 max for the first);
//   - posedge flags are cleared after each pushed snapshot;
//   - data pins change immediately; a clock pin's change sets its posedge
//     flag (on a 0 -> 1 change) and is applied at the next new timestamp;
//   - 'x' / 'z' read as 0; real values and $comment blocks are skipped.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint64_t kU64Max = ~0ull;

struct VcdBody {
  std::vector<uint32_t> states;  // snapshots, rio words each
  std::vector<uint64_t> off, ts;
};

struct Target {
  uint32_t bit, flag;  // state bit, posedge flag bit (kNone: data pin)
};

inline bool is_space(char c) {
  return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f';
}

thread_local std::string t_err;

}  // namespace

extern "C" {

// Parses the value-change section `body` (the text after $enddefinitions
// $end).  Codes: code i is codes[code_off[i] .. code_off[i+1]); its bit
// positions (index into a vector value, 0 = first character) map through
// entries [ent_start[i], ent_start[i+1]) of (ent_pos, ent_bit, ent_flag).
// flag_bits are the posedge flags cleared after each snapshot.  Stops after
// max_edges snapshots when max_edges > 0.  Returns a handle or NULL.
void *jqh_vcd_parse_body(const char *body, uint64_t len, const char *codes,
                         const uint32_t *code_off, uint32_t n_codes,
                         const uint32_t *ent_start, const uint32_t *ent_pos,
                         const uint32_t *ent_bit, const uint32_t *ent_flag,
                         const uint32_t *flag_bits, uint32_t n_flag, uint32_t rio,
                         uint64_t max_edges) {
  t_err.clear();
  if (rio == 0) {
    t_err = "vcd: empty state";
    return nullptr;
  }
  // code -> dense per-position targets
  std::unordered_map<std::string_view, uint32_t> code_idx;
  code_idx.reserve(n_codes * 2 + 1);
  std::vector<std::vector<Target>> targets(n_codes);
  for (uint32_t i = 0; i < n_codes; ++i) {
    code_idx.emplace(std::string_view(codes + code_off[i], code_off[i + 1] - code_off[i]), i);
    for (uint32_t e = ent_start[i]; e < ent_start[i + 1]; ++e) {
      if (ent_bit[e] >= 32ull * rio || (ent_flag[e] != kNone && ent_flag[e] >= 32ull * rio)) {
        t_err = "vcd: state bit out of range";
        return nullptr;
      }
      auto &v = targets[i];
      if (v.size() <= ent_pos[e]) v.resize(ent_pos[e] + 1, Target{kNone, kNone});
      v[ent_pos[e]] = Target{ent_bit[e], ent_flag[e]};
    }
  }
  for (uint32_t k = 0; k < n_flag; ++k)
    if (flag_bits[k] >= 32ull * rio) {
      t_err = "vcd: flag bit out of range";
      return nullptr;
    }
  auto *out = new VcdBody;
  std::vector<uint32_t> state(rio, 0u);
  std::vector<uint32_t> delayed;
  std::vector<uint8_t> in_delayed(32ull * rio, 0);
  uint64_t last_active_time = kU64Max, vcd_time = 0;
  bool last_active = true;
  auto flip = [&](uint32_t b) { state[b >> 5] ^= 1u << (b & 31); };
  auto get = [&](uint32_t b) { return (state[b >> 5] >> (b & 31)) & 1u; };

  const char *p = body, *end = body + len;
  auto next_token = [&](std::string_view &tok) {
    while (p < end && is_space(*p)) ++p;
    if (p >= end) return false;
    const char *s = p;
    while (p < end && !is_space(*p)) ++p;
    tok = std::string_view(s, p - s);
    return true;
  };
  std::string_view tok;
  bool stop = false;
  while (!stop && next_token(tok)) {
    const char c0 = tok[0];
    if (c0 == '#') {
      uint64_t t = 0;
      for (size_t k = 1; k < tok.size(); ++k) {
        const uint64_t d = (uint64_t)(tok[k] - '0');
        if (tok[k] < '0' || tok[k] > '9' || t > (kU64Max - d) / 10) {
          t_err = "vcd: bad timestamp " + std::string(tok);
          delete out;
          return nullptr;
        }
        t = t * 10 + d;
      }
      if (t == vcd_time) continue;
      if (last_active) {
        out->states.insert(out->states.end(), state.begin(), state.end());
        out->off.push_back(out->states.size());
        out->ts.push_back(last_active_time);
        for (uint32_t k = 0; k < n_flag; ++k) state[flag_bits[k] >> 5] &= ~(1u << (flag_bits[k] & 31));
        if (max_edges && out->off.size() >= max_edges) {
          stop = true;
          break;
        }
        last_active_time = vcd_time;
      }
      vcd_time = t;
      last_active = false;
      for (uint32_t b : delayed) {
        flip(b);
        in_delayed[b] = 0;
      }
      delayed.clear();
      continue;
    }
    if (c0 == '$') {
      if (tok == "$comment") {
        while (next_token(tok) && tok != "$end") {
        }
      }
      continue;
    }
    std::string_view bits, code;
    if (c0 == 'b' || c0 == 'B' || c0 == 'r' || c0 == 'R') {
      bits = tok.substr(1);
      if (!next_token(code)) break;
      if (c0 == 'r' || c0 == 'R') continue;
    } else {
      bits = tok.substr(0, 1);
      code = tok.substr(1);
    }
    auto it = code_idx.find(code);
    if (it == code_idx.end()) continue;
    const auto &tg = targets[it->second];
    const size_t n = bits.size() < tg.size() ? bits.size() : tg.size();
    for (size_t pos = 0; pos < n; ++pos) {
      const Target e = tg[pos];
      if (e.bit == kNone) continue;
      const uint32_t nv = bits[pos] == '1' ? 1u : 0u;
      const uint32_t old = get(e.bit);
      if (old == nv) continue;
      if (e.flag != kNone) {
        if (old == 0) {
          last_active = true;
          state[e.flag >> 5] |= 1u << (e.flag & 31);
        }
        if (!in_delayed[e.bit]) {
          in_delayed[e.bit] = 1;
          delayed.push_back(e.bit);
        }
      } else {
        flip(e.bit);
      }
    }
  }
  // the final snapshot (after an edge limit: the state with flags cleared)
  out->states.insert(out->states.end(), state.begin(), state.end());
  return out;
}

// ---- writers: the value-change section of an output VCD
// (src/sim/vcd_io.rs:758-820).  Every row gets a "#<time>" line; a column
// is written ("0", "1" or "x" then its identifier code) only when its value
// differs from the one it last wrote (none before the first row).  Code i
// is codes[code_off[i] .. code_off[i+1]).  The text goes into a handle
// (jqh_text_data / _size / _free) or, for NULL, the error message.

namespace {

struct VcdText {
  std::string s;
};

inline void put_u64(std::string &o, uint64_t v) {
  char b[24];
  int n = 0;
  do {
    b[n++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (n) o.push_back(b[--n]);
}

inline void put_change(std::string &o, uint32_t v, const char *codes,
                       const uint64_t *code_off, uint32_t i) {
  o.push_back(v == 3 ? 'x' : (char)('0' + v));
  o.append(codes + code_off[i], code_off[i + 1] - code_off[i]);
  o.push_back('\n');
}

}  // namespace

// vals[r * n_cols + i] in {0, 1, 3 (X)}: column i's value at row r, emitted
// at timestamps[r].
void *jqh_vcd_write_values(const uint8_t *vals, uint64_t n_rows, uint32_t n_cols,
                           const uint64_t *timestamps, const char *codes,
                           const uint64_t *code_off) {
  t_err.clear();
  auto *out = new VcdText();
  try {
    std::vector<uint8_t> last(n_cols, 2);
    out->s.reserve(n_rows * 12);  // the time lines; changes grow it
    for (uint64_t r = 0; r < n_rows; ++r) {
      out->s.push_back('#');
      put_u64(out->s, timestamps[r]);
      out->s.push_back('\n');
      const uint8_t *row = vals + r * n_cols;
      for (uint32_t i = 0; i < n_cols; ++i) {
        const uint8_t v = row[i];
        if (v != 0 && v != 1 && v != 3) throw std::runtime_error("vcd: value not 0, 1 or 3");
        if (v == last[i]) continue;
        last[i] = v;
        put_change(out->s, v, codes, code_off, i);
      }
    }
  } catch (const std::exception &e) {
    t_err = e.what();
    delete out;
    return nullptr;
  }
  return out;
}

// Change records (output << 2 | value): row r owns counts[r] of them, in
// order, emitted at timestamps[r] (the device change extraction's output).
void *jqh_vcd_write_changes(const uint32_t *counts, const uint32_t *records,
                            uint64_t n_records, uint64_t n_rows,
                            const uint64_t *timestamps, uint32_t n_cols,
                            const char *codes, const uint64_t *code_off) {
  t_err.clear();
  auto *out = new VcdText();
  try {
    out->s.reserve(n_rows * 8 + n_records * 4);
    uint64_t k = 0;
    for (uint64_t r = 0; r < n_rows; ++r) {
      out->s.push_back('#');
      put_u64(out->s, timestamps[r]);
      out->s.push_back('\n');
      if (counts[r] > n_records - k) throw std::runtime_error("vcd: counts exceed the records");
      for (uint64_t e = k + counts[r]; k < e; ++k) {
        const uint32_t rec = records[k], i = rec >> 2, v = rec & 3;
        if (i >= n_cols || v == 2) throw std::runtime_error("vcd: bad change record");
        put_change(out->s, v, codes, code_off, i);
      }
    }
  } catch (const std::exception &e) {
    t_err = e.what();
    delete out;
    return nullptr;
  }
  return out;
}

// Timed form (src/sim/vcd_io.rs:831-946): a change of column i at row r
// (against the value it last changed to; none before row 0) happens at
// timestamps[r] + round(arrival_ps[r * n_cols + i] / ps_per_unit), halves
// away from zero (0 when ps_per_unit <= 0); changes are stably sorted by
// that time and a "#<time>" line opens each distinct time.  Times are kept
// in 128 bits, so a shift past 2^64 - 1 prints as the sum.
void *jqh_vcd_write_timed(const uint8_t *vals, const uint16_t *arrival_ps, uint64_t n_rows,
                          uint32_t n_cols, const uint64_t *timestamps, double ps_per_unit,
                          const char *codes, const uint64_t *code_off) {
  t_err.clear();
  auto *out = new VcdText();
  try {
    struct Change {
      unsigned __int128 t;
      uint32_t col;
      uint8_t v;
    };
    std::vector<Change> ch;
    std::vector<uint8_t> last(n_cols, 2);
    for (uint64_t r = 0; r < n_rows; ++r)
      for (uint32_t i = 0; i < n_cols; ++i) {
        const uint8_t v = vals[r * n_cols + i];
        if (v != 0 && v != 1 && v != 3) throw std::runtime_error("vcd: value not 0, 1 or 3");
        if (v == last[i]) continue;
        last[i] = v;
        uint64_t units = 0;
        if (ps_per_unit > 0) {
          const double x = (double)arrival_ps[r * n_cols + i] / ps_per_unit;
          units = (uint64_t)std::floor(x + 0.5);  // x >= 0
        }
        ch.push_back({(unsigned __int128)timestamps[r] + units, i, v});
      }
    std::stable_sort(ch.begin(), ch.end(),
                     [](const Change &a, const Change &b) { return a.t < b.t; });
    out->s.reserve(ch.size() * 6);
    bool open = false;
    unsigned __int128 cur = 0;
    for (const Change &c : ch) {
      if (!open || c.t != cur) {
        out->s.push_back('#');
        char b[48];
        int n = 0;
        unsigned __int128 t = c.t;
        do {
          b[n++] = (char)('0' + (int)(t % 10));
          t /= 10;
        } while (t);
        while (n) out->s.push_back(b[--n]);
        out->s.push_back('\n');
        cur = c.t;
        open = true;
      }
      put_change(out->s, c.v, codes, code_off, c.col);
    }
  } catch (const std::exception &e) {
    t_err = e.what();
    delete out;
    return nullptr;
  }
  return out;
}

const char *jqh_text_data(void *h) { return static_cast<VcdText *>(h)->s.data(); }
uint64_t jqh_text_size(void *h) { return static_cast<VcdText *>(h)->s.size(); }
void jqh_text_free(void *h) { delete static_cast<VcdText *>(h); }

uint64_t jqh_vcd_num_words(void *h) { return static_cast<VcdBody *>(h)->states.size(); }
const uint32_t *jqh_vcd_states(void *h) { return static_cast<VcdBody *>(h)->states.data(); }
uint64_t jqh_vcd_num_offsets(void *h) { return static_cast<VcdBody *>(h)->off.size(); }
const uint64_t *jqh_vcd_offsets(void *h) { return static_cast<VcdBody *>(h)->off.data(); }
const uint64_t *jqh_vcd_timestamps(void *h) { return static_cast<VcdBody *>(h)->ts.data(); }
void jqh_vcd_free(void *h) { delete static_cast<VcdBody *>(h); }
const char *jqh_vcd_last_error(void) { return t_err.c_str(); }

}  // extern "C"
