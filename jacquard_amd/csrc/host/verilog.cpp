// Flat structural Verilog reader for AIGPDK netlists (see verilog.hpp).
#include "verilog.hpp"

#include <cctype>
#include <functional>
#include <map>
#include <stdexcept>
#include <unordered_map>

namespace jq {

namespace {

struct Lexer {
  const std::string &s;
  size_t i = 0;
  explicit Lexer(const std::string &t) : s(t) {}
  void skip() {
    for (;;) {
      while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
      if (s.compare(i, 2, "//") == 0) {
        while (i < s.size() && s[i] != '\n') ++i;
      } else if (s.compare(i, 2, "/*") == 0) {
        size_t e = s.find("*/", i + 2);
        i = e == std::string::npos ? s.size() : e + 2;
      } else if (s.compare(i, 2, "(*") == 0) {
        size_t e = s.find("*)", i + 2);
        i = e == std::string::npos ? s.size() : e + 2;
      } else {
        return;
      }
    }
  }
  std::string next() {
    skip();
    if (i >= s.size()) return "";
    char c = s[i];
    if (c == '\\') {  // escaped identifier
      size_t st = i++;
      while (i < s.size() && !std::isspace((unsigned char)s[i])) ++i;
      return s.substr(st, i - st);
    }
    if (std::isalnum((unsigned char)c) || c == '_' || c == '$' || c == '\'') {
      size_t st = i;
      while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' ||
                              s[i] == '$' || s[i] == '\''))
        ++i;
      return s.substr(st, i - st);
    }
    ++i;
    return std::string(1, c);
  }
  std::string peek() {
    size_t save = i;
    std::string t = next();
    i = save;
    return t;
  }
  void expect(const std::string &t) {
    std::string x = next();
    if (x != t) throw std::runtime_error("verilog: expected '" + t + "' got '" + x + "'");
  }
};

struct Cell {
  std::string type, name;
  std::map<std::string, std::string> pins;  // pin -> net bit name or const
};

bool is_const(const std::string &n, uint32_t &v) {
  size_t q = n.find('\'');
  if (q == std::string::npos) return false;
  char base = (char)std::tolower((unsigned char)n[q + 1]);
  std::string digits = n.substr(q + 2);
  unsigned long long x = 0;
  int b = base == 'h' ? 16 : base == 'b' ? 2 : base == 'o' ? 8 : 10;
  x = std::stoull(digits, nullptr, b);
  v = (uint32_t)(x & 1);
  return true;
}

}  // namespace

void read_aigpdk_verilog(const std::string &text, Aig &g) {
  g = Aig();
  Lexer lx(text);
  if (lx.next() != "module") throw std::runtime_error("verilog: no module");
  g.module_name = lx.next();
  std::vector<std::string> header;
  lx.expect("(");
  for (std::string t = lx.next(); t != ")"; t = lx.next())
    if (t != ",") header.push_back(t);
  lx.expect(";");
  struct Decl {
    bool vec = false;
    int msb = 0, lsb = 0;
    int dir = 0;  // 1 input, 2 output
  };
  std::map<std::string, Decl> decls;
  std::unordered_map<std::string, std::string> alias;
  std::vector<Cell> cells;
  auto bitname = [](const std::string &n, int i) { return n + "[" + std::to_string(i) + "]"; };
  auto read_net = [&]() -> std::string {
    std::string n = lx.next();
    if (lx.peek() == "[") {
      lx.next();
      std::string idx = lx.next();
      lx.expect("]");
      return n + "[" + idx + "]";
    }
    return n;
  };
  for (;;) {
    std::string t = lx.next();
    if (t.empty()) throw std::runtime_error("verilog: missing endmodule");
    if (t == "endmodule") break;
    if (t == "input" || t == "output" || t == "wire" || t == "inout") {
      Decl d;
      d.dir = t == "input" ? 1 : t == "output" ? 2 : t == "inout" ? 1 : 0;
      if (lx.peek() == "[") {
        lx.next();
        d.msb = std::stoi(lx.next());
        lx.expect(":");
        d.lsb = std::stoi(lx.next());
        lx.expect("]");
        d.vec = true;
      }
      for (;;) {
        std::string n = lx.next();
        auto &e = decls[n];
        if (d.dir) {
          e.dir = d.dir;
          e.vec = d.vec;
          e.msb = d.msb;
          e.lsb = d.lsb;
        } else if (!e.dir) {
          e = d;
        }
        std::string sep = lx.next();
        if (sep == ";") break;
        if (sep != ",") throw std::runtime_error("verilog: bad declaration");
      }
      continue;
    }
    if (t == "assign") {
      std::string lhs = read_net();
      lx.expect("=");
      std::string rhs = read_net();
      lx.expect(";");
      alias[lhs] = rhs;
      continue;
    }
    // cell instance: TYPE NAME ( .PIN(net), ... );
    Cell c;
    c.type = t;
    c.name = lx.next();  // instance name
    lx.expect("(");
    for (;;) {
      std::string x = lx.next();
      if (x == ")") break;
      if (x == ",") continue;
      if (x != ".") throw std::runtime_error("verilog: positional connections unsupported");
      std::string pin = lx.next();
      lx.expect("(");
      std::string net = read_net();
      lx.expect(")");
      c.pins[pin] = net;
    }
    lx.expect(";");
    cells.push_back(std::move(c));
  }
  auto bits_of = [&](const std::string &n) {
    std::vector<std::string> v;
    const Decl &d = decls[n];
    if (!d.vec) {
      v.push_back(n);
    } else {
      int step = d.msb >= d.lsb ? -1 : 1;
      for (int i = d.msb;; i += step) {
        v.push_back(bitname(n, i));
        if (i == d.lsb) break;
      }
    }
    return v;
  };
  // inputs
  std::unordered_map<std::string, uint32_t> net_lit;
  std::unordered_map<std::string, uint32_t> input_pin;
  uint32_t port_idx = 0;
  for (auto &p : header) {
    if (decls[p].dir != 1) continue;
    for (auto &b : bits_of(p)) {
      uint32_t pin = g.add_pin(kInput, port_idx++);
      input_pin[b] = pin;
      net_lit[b] = pin << 1;
      g.in_port_names.push_back(b);
      g.in_port_pins.push_back(pin);
      g.in_port_flag.push_back(0);
    }
  }
  // drivers
  std::unordered_map<std::string, size_t> driver;
  for (size_t k = 0; k < cells.size(); ++k) {
    const Cell &c = cells[k];
    for (const char *o : {"Y", "Q"}) {
      auto it = c.pins.find(o);
      if (it != c.pins.end()) driver[it->second] = k;
    }
  }
  auto resolve = [&](std::string n) {
    for (int hop = 0; hop < 1000 && alias.count(n); ++hop) n = alias[n];
    return n;
  };
  std::unordered_map<std::string, uint32_t> flag_of;  // clock bit -> flag pin
  auto clock_flag = [&](const std::string &net0) {
    std::string net = resolve(net0);
    auto it = input_pin.find(net);
    if (it == input_pin.end())
      throw std::runtime_error("verilog: DFF clock must be an input port: " + net);
    auto f = flag_of.find(net);
    if (f != flag_of.end()) return f->second;
    uint32_t pin = g.add_pin(kClockFlag, g.ref[it->second]);
    flag_of[net] = pin;
    for (size_t i = 0; i < g.in_port_names.size(); ++i)
      if (g.in_port_names[i] == net) g.in_port_flag[i] = pin;
    return pin;
  };
  // DFF Q pins first
  std::vector<size_t> dff_cells;
  for (size_t k = 0; k < cells.size(); ++k) {
    const Cell &c = cells[k];
    if (c.type == "DFF" || c.type == "DFFSR") {
      Dff d;
      d.q = g.add_pin(kDffQ, (uint32_t)g.dffs.size());
      auto q = c.pins.find("Q");
      if (q != c.pins.end()) net_lit[q->second] = d.q << 1;
      g.dffs.push_back(d);
      g.dff_names.push_back(c.name);
      dff_cells.push_back(k);
    }
  }
  std::function<uint32_t(const std::string &)> lit = [&](const std::string &n0) -> uint32_t {
    uint32_t cv;
    if (is_const(n0, cv)) return cv;
    std::string n = resolve(n0);
    if (is_const(n, cv)) return cv;
    auto it = net_lit.find(n);
    if (it != net_lit.end()) return it->second;
    auto d = driver.find(n);
    if (d == driver.end()) throw std::runtime_error("verilog: undriven net " + n);
    const Cell &c = cells[d->second];
    uint32_t r;
    if (c.type == "INV") {
      r = lit(c.pins.at("A")) ^ 1;
    } else if (c.type == "BUF") {
      r = lit(c.pins.at("A"));
    } else if (c.type.size() == 9 && c.type.compare(0, 4, "AND2") == 0) {
      uint32_t ia = c.type[5] - '0', ib = c.type[6] - '0', iy = c.type[8] - '0';
      r = g.add_and(lit(c.pins.at("A")) ^ ia, lit(c.pins.at("B")) ^ ib) ^ iy;
    } else {
      throw std::runtime_error("verilog: unsupported cell " + c.type);
    }
    net_lit[n] = r;
    return r;
  };
  for (size_t j = 0; j < dff_cells.size(); ++j) {
    const Cell &c = cells[dff_cells[j]];
    Dff &d = g.dffs[j];
    uint32_t din = lit(c.pins.at("D"));
    uint32_t clken = clock_flag(c.pins.at("CLK")) << 1;
    if (c.type == "DFFSR") {
      // AIGPDK DFFSR: R=0 clears, S=0 presets (src/aig.rs:348-381)
      uint32_t s = c.pins.count("S") ? lit(c.pins.at("S")) : 1;
      uint32_t rr = c.pins.count("R") ? lit(c.pins.at("R")) : 1;
      din = g.add_and(din ^ 1, s) ^ 1;
      clken = g.add_and(clken ^ 1, s) ^ 1;
      din = g.add_and(din, rr);
      clken = g.add_and(clken ^ 1, rr) ^ 1;
    }
    d.d_iv = din;
    d.en_iv = clken;
  }
  for (auto &p : header) {
    if (decls[p].dir != 2) continue;
    for (auto &b : bits_of(p)) {
      uint32_t l = lit(b);
      g.out_port_names.push_back(b);
      g.out_port_lits.push_back(l);
      bool have = false;
      for (auto x : g.pos) have |= x == l;
      if (!have) {
        g.pos.push_back(l);
        g.po_names.push_back(b);
      }
    }
  }
}

}  // namespace jq
