// AIG model used by the script compiler.
//
// Mirrors the reference's AIG contract (src/aig.rs:41-346): pins are numbered
// 1..num_pins in topological order, pin 0 is tied to 0, literals are
// pin<<1|invert, AND gates are structurally hashed, and endpoints are primary
// outputs, DFFs (D, clock enable, Q) and 13-bit x 32-bit synchronous SRAMs.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace jq {

enum Driver : uint8_t {
  kTie0 = 0,
  kAnd = 1,
  kInput = 2,
  kClockFlag = 3,
  kDffQ = 4,
  kSramQ = 5,
};

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kSramAddrBits = 13;          // src/aigpdk.rs:20
constexpr uint32_t kSramWords = 1u << 13;  // src/aigpdk.rs:22

struct Dff {
  uint32_t d_iv = 0;
  uint32_t en_iv = 0;
  uint32_t q = 0;
};

struct Sram {
  uint32_t r_addr_iv[kSramAddrBits] = {};
  uint32_t r_en_iv = 0;
  uint32_t rd_pin[32] = {};  // 0 = unused bit
  uint32_t w_addr_iv[kSramAddrBits] = {};
  uint32_t w_en_iv[32] = {};
  uint32_t w_data_iv[32] = {};
};

// Structural-hash table of AND gates: key (lower operand literal << 32 |
// higher) -> pin.  Open addressing with linear probing over a power-of-two
// array, grown at half load: one probe sequence in two flat arrays instead
// of a heap node per gate.  Keys are never 0 (both operands are literals
// >= 2) and an AND's pin is never 0, so 0 marks empty slots and absence.
class AndTable {
 public:
  uint32_t find(uint64_t key) const {
    if (keys_.empty()) return 0;
    for (size_t i = slot(key);; i = (i + 1) & (keys_.size() - 1)) {
      if (keys_[i] == key) return vals_[i];
      if (keys_[i] == 0) return 0;
    }
  }
  void insert(uint64_t key, uint32_t pin) {
    if (2 * (used_ + 1) > keys_.size()) grow();
    size_t i = slot(key);
    while (keys_[i] != 0 && keys_[i] != key) i = (i + 1) & (keys_.size() - 1);
    used_ += keys_[i] == 0;
    keys_[i] = key;
    vals_[i] = pin;
  }

 private:
  std::vector<uint64_t> keys_;
  std::vector<uint32_t> vals_;
  size_t used_ = 0;
  // the splitmix64 output function (Steele, Lea and Flood, OOPSLA 2014)
  static uint64_t mix(uint64_t k) {
    k ^= k >> 30;
    k *= 0xbf58476d1ce4e5b9ull;
    k ^= k >> 27;
    k *= 0x94d049bb133111ebull;
    return k ^ (k >> 31);
  }
  size_t slot(uint64_t key) const { return (size_t)mix(key) & (keys_.size() - 1); }
  void grow() {
    std::vector<uint64_t> ok;
    std::vector<uint32_t> ov;
    ok.swap(keys_);
    ov.swap(vals_);
    keys_.assign(ok.empty() ? 1024 : 2 * ok.size(), 0);
    vals_.assign(keys_.size(), 0);
    for (size_t j = 0; j < ok.size(); ++j)
      if (ok[j]) {
        size_t i = slot(ok[j]);
        while (keys_[i] != 0) i = (i + 1) & (keys_.size() - 1);
        keys_[i] = ok[j];
        vals_[i] = ov[j];
      }
  }
};

struct Aig {
  uint32_t num_pins = 0;
  std::vector<uint8_t> drv{kTie0};
  std::vector<uint32_t> a{0}, b{0};  // AND operands (literals)
  std::vector<uint32_t> ref{0};      // port / dff / sram index of a source
  AndTable and_cache;

  std::vector<uint32_t> pos;  // primary outputs (literals), unique
  std::vector<std::string> po_names;
  // Port view for VCD I/O, in module declaration order (vector bits from
  // the declared MSB side): every output bit (name, literal; several bits may
  // share one PO literal) and every input bit (name, pin, posedge flag pin or
  // 0 when the port is not a clock).
  std::vector<std::string> out_port_names;
  std::vector<uint32_t> out_port_lits;
  std::vector<std::string> in_port_names;
  std::vector<uint32_t> in_port_pins, in_port_flag;
  std::string module_name;
  std::vector<Dff> dffs;
  // instance name of each DFF cell (netlists read from Verilog; empty for
  // generated netlists)
  std::vector<std::string> dff_names;
  std::vector<Sram> srams;
  // per-pin integer gate delay in ps (timing), empty when not loaded
  std::vector<uint16_t> delay_ps;

  uint32_t add_pin(Driver d, uint32_t r = 0) {
    ++num_pins;
    drv.push_back(d);
    a.push_back(0);
    b.push_back(0);
    ref.push_back(r);
    return num_pins;
  }

  // Structural hashing and constant folding as add_and_gate
  // (src/aig.rs:327-346).
  uint32_t add_and(uint32_t x, uint32_t y) {
    if (x == 0 || y == 0) return 0;
    if (x == 1) return y;
    if (y == 1) return x;
    if (x > y) std::swap(x, y);
    uint64_t key = (uint64_t)x << 32 | y;
    if (uint32_t hit = and_cache.find(key)) return hit << 1;
    uint32_t p = add_pin(kAnd);
    a[p] = x;
    b[p] = y;
    and_cache.insert(key, p);
    return p << 1;
  }

  size_t num_and() const {
    size_t n = 0;
    for (auto d : drv) n += (d == kAnd);
    return n;
  }

  // endpoint group index space: [POs | DFFs | SRAMs] (flatten.rs:1268)
  size_t num_endpoints() const { return pos.size() + dffs.size() + srams.size(); }

  // Calls f(pin) for every fan-in pin of endpoint group e
  // (EndpointGroup::for_each_input, src/aig.rs:134-168).
  template <class F>
  void for_each_endpoint_input(size_t e, F f) const {
    auto g = [&](uint32_t iv) {
      if ((iv >> 1) >= 1) f(iv >> 1);
    };
    if (e < pos.size()) {
      g(pos[e]);
      return;
    }
    e -= pos.size();
    if (e < dffs.size()) {
      g(dffs[e].en_iv);
      g(dffs[e].d_iv);
      return;
    }
    e -= dffs.size();
    const Sram &s = srams[e];
    g(s.r_en_iv);
    for (int i = 0; i < kSramAddrBits; ++i) {
      g(s.r_addr_iv[i]);
      g(s.w_addr_iv[i]);
    }
    for (int i = 0; i < 32; ++i) {
      g(s.w_en_iv[i]);
      g(s.w_data_iv[i]);
    }
  }

  // X-capable pins: DFF Q and SRAM read data are X sources, forward cone,
  // DFF fixpoint (src/aig.rs:3276-3367).
  std::vector<uint8_t> x_capable_pins() const;
};

}  // namespace jq
