// Script compiler: level split, partitioning, boomerang mapping, flatten.
//
// Format fidelity (what must match the reference bit for bit) is cited per
// step; the partitioning/mapping policy is this project's own.
#include "compile.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace jq {

namespace {

constexpr uint32_t kL1 = kWidth / 2;  // 4096 level-1 slots

// ---------------------------------------------------------------------------
// Levelization and level split (semantics of StagedAIG::from_split,
// src/staging.rs:93-179: nodes at or below a split that are consumed above it
// become staged outputs, readable later in the same cycle).
struct Levels {
  std::vector<uint32_t> lvl;
  uint32_t max_lvl = 0;
};

Levels levelize(const Aig &g) {
  Levels L;
  L.lvl.assign(g.num_pins + 1, 0);
  for (uint32_t p = 1; p <= g.num_pins; ++p)
    if (g.drv[p] == kAnd) {
      L.lvl[p] = 1 + std::max(L.lvl[g.a[p] >> 1], L.lvl[g.b[p] >> 1]);
      L.max_lvl = std::max(L.max_lvl, L.lvl[p]);
    }
  return L;
}

// ---------------------------------------------------------------------------
// Level-1 list-scheduling boomerang mapper.  Every boomerang stage evaluates
// one wave of ready AND gates in hier[1] (4096 slots, operands duplicated
// freely in hier[0]) and carries every still-needed value as a pass-through.
// Outputs are packed densely at the front of the stage that produces them and
// written out from there.
struct MapInput {
  std::vector<uint32_t> nodes;    // AND pins, ascending (topological)
  std::vector<uint32_t> inputs;   // available pins
  std::vector<uint32_t> outputs;  // pins that must be written out
};

struct Scratch {
  std::vector<int32_t> loc;  // pin -> local index (or -1)
  std::vector<uint32_t> stamp;
  uint32_t cur = 1;
  explicit Scratch(size_t n) : loc(n, -1), stamp(n, 0) {}
};

bool map_partition(const Aig &g, const MapInput &in, Scratch &sc,
                   std::vector<BoomStage> &out, uint32_t max_write_words) {
  out.clear();
  // local numbering: inputs first, then nodes
  std::vector<uint32_t> pins;
  pins.reserve(in.inputs.size() + in.nodes.size());
  for (auto p : in.inputs) pins.push_back(p);
  for (auto p : in.nodes) pins.push_back(p);
  const uint32_t ni = (uint32_t)in.inputs.size();
  const uint32_t n = (uint32_t)pins.size();
  for (uint32_t i = 0; i < n; ++i) sc.loc[pins[i]] = (int32_t)i;
  struct Guard {
    Scratch &s;
    const std::vector<uint32_t> &p;
    ~Guard() {
      for (auto x : p) s.loc[x] = -1;
    }
  } guard{sc, pins};

  std::vector<uint32_t> rem(n, 0);        // unscheduled consumers
  std::vector<uint8_t> is_out(n, 0), written(n, 0), sched(n, 0);
  std::vector<uint8_t> pending(n, 0);     // operands not yet available
  std::vector<std::vector<uint32_t>> cons(n);
  std::vector<uint32_t> opa(n, 0), opb(n, 0);
  for (uint32_t i = ni; i < n; ++i) {
    uint32_t p = pins[i];
    int32_t la = sc.loc[g.a[p] >> 1], lb = sc.loc[g.b[p] >> 1];
    if (la < 0 || lb < 0) return false;  // cone not closed
    opa[i] = (uint32_t)la;
    opb[i] = (uint32_t)lb;
    cons[la].push_back(i);
    ++rem[la];
    if (lb != la) {
      cons[lb].push_back(i);
      ++rem[lb];
    }
  }
  for (auto p : in.outputs) {
    int32_t l = sc.loc[p];
    if (l < 0) return false;
    is_out[l] = 1;
  }
  // height (longest path to an output) as priority
  std::vector<uint32_t> height(n, 0);
  for (uint32_t i = n; i-- > ni;) {
    for (auto c : cons[i]) height[i] = std::max(height[i], height[c] + 1);
    (void)0;
  }
  std::vector<uint8_t> avail(n, 0);
  std::vector<uint32_t> avail_list;
  for (uint32_t i = 0; i < ni; ++i) {
    avail[i] = 1;
    avail_list.push_back(i);
  }
  for (uint32_t i = ni; i < n; ++i) pending[i] = (uint8_t)(opa[i] == opb[i] ? 1 : 2);
  std::vector<uint32_t> ready;
  for (uint32_t i = ni; i < n; ++i) {
    // operands that are inputs are available from the start
    uint8_t k = 0;
    if (opa[i] >= ni) ++k;
    if (opb[i] >= ni && opb[i] != opa[i]) ++k;
    pending[i] = k;
    if (k == 0) ready.push_back(i);
  }
  uint32_t remaining = n - ni;
  uint32_t total_words = 0;
  std::vector<uint32_t> taken(n, 0);  // consumers scheduled this stage
  while (true) {
    bool any_unwritten = false;
    for (auto v : avail_list)
      if (is_out[v] && !written[v]) {
        any_unwritten = true;
        break;
      }
    if (remaining == 0 && !any_unwritten) break;
    if (out.size() > 4096) return false;
    // select all ready nodes, then defer low-priority ones while over capacity
    std::sort(ready.begin(), ready.end(), [&](uint32_t x, uint32_t y) {
      if (height[x] != height[y]) return height[x] > height[y];
      return x < y;
    });
    std::vector<uint32_t> S = ready;
    for (auto x : S) {
      ++taken[opa[x]];
      if (opb[x] != opa[x]) ++taken[opb[x]];
    }
    auto carried = [&](uint32_t v) {
      return rem[v] > taken[v] || (is_out[v] && !written[v]);
    };
    uint32_t ncarry = 0;
    for (auto v : avail_list) ncarry += carried(v);
    // defer from the tail (lowest priority)
    size_t keep = S.size();
    for (size_t k = S.size(); k-- > 0 && keep + ncarry > kL1;) {
      uint32_t x = S[k];
      // deferring frees one slot; operands that stop being fully consumed
      // become carries
      int delta = -1;
      uint32_t ops[2] = {opa[x], opb[x]};
      int nops = opa[x] == opb[x] ? 1 : 2;
      for (int t = 0; t < nops; ++t)
        if (!carried(ops[t]) && taken[ops[t]] == rem[ops[t]]) ++delta;
      if (delta >= 0) continue;
      for (int t = 0; t < nops; ++t) --taken[ops[t]];
      S[k] = kNone;
      --keep;
      ncarry += (uint32_t)(delta + 1);
    }
    if (keep + ncarry > kL1) {
      for (auto x : S)
        if (x != kNone) {
          taken[opa[x]] = 0;
          taken[opb[x]] = 0;
        }
      return false;
    }
    S.erase(std::remove(S.begin(), S.end(), kNone), S.end());
    // layout: outputs produced/needed now first, then the rest
    std::vector<uint32_t> layout;
    layout.reserve(keep + ncarry);
    for (auto x : S)
      if (is_out[x]) layout.push_back(x);
    for (auto v : avail_list)
      if (is_out[v] && !written[v]) layout.push_back(v);
    uint32_t nout = (uint32_t)layout.size();
    for (auto x : S)
      if (!is_out[x]) layout.push_back(x);
    for (auto v : avail_list)
      if (carried(v) && !(is_out[v] && !written[v])) layout.push_back(v);
    for (auto x : S) {
      taken[opa[x]] = 0;
      taken[opb[x]] = 0;
    }
    BoomStage bs;
    bs.hier[0].assign(kWidth, kNone);
    bs.hier[1].assign(kL1, kNone);
    for (uint32_t j = 0; j < layout.size(); ++j) {
      uint32_t v = layout[j];
      if (v >= ni && !avail[v]) {  // computed here
        bs.hier[0][j] = g.a[pins[v]] >> 1;
        bs.hier[0][j + kL1] = g.b[pins[v]] >> 1;
      } else {
        bs.hier[0][j] = pins[v];
      }
      bs.hier[1][j] = pins[v];
    }
    uint32_t nw = (nout + 31) / 32;
    for (uint32_t w = 0; w < nw; ++w) bs.write_outs.push_back((uint16_t)(128 + w));
    total_words += nw;
    if (total_words > max_write_words) return false;
    out.push_back(std::move(bs));
    // advance state
    for (auto x : S) {
      sched[x] = 1;
      --remaining;
      --rem[opa[x]];
      if (opb[x] != opa[x]) --rem[opb[x]];
    }
    for (uint32_t j = 0; j < nout; ++j) written[layout[j]] = 1;
    for (auto v : avail_list) avail[v] = 0;
    avail_list.assign(layout.begin(), layout.end());
    for (auto v : avail_list) avail[v] = 1;
    std::vector<uint32_t> next_ready;
    for (auto r : ready)
      if (!sched[r]) next_ready.push_back(r);
    for (auto x : S)
      for (auto c : cons[x])
        if (--pending[c] == 0) next_ready.push_back(c);
    ready.swap(next_ready);
  }
  return true;
}

// ---------------------------------------------------------------------------
// Multi-level boomerang mapper.  A boomerang stage is a full binary tree:
// node (k, q) at level k >= 1 (q < 8192 >> k) reads its children (k-1, q)
// and (k-1, q + (8192 >> k)); level 0 is the shuffled layer, whose slots can
// hold any value present in the previous layer.  A node either computes the
// AND of its children (with inversions) or passes child (k-1, q) through.
// A gate whose cone, down to values already present, is d levels deep can
// therefore be computed at level d in one stage, with its operand gates
// recomputed at its children (every root owns its subtree; a value with
// fanout is simply computed again in each subtree that needs it) and
// present values routed up by pass-through chains.  Per stage:
//   1. present values that are outputs not yet written: level-1
//      pass-throughs at the front (dense write-out words);
//   2. roots in critical-path order at their cone depth (outputs at the
//      front of their level, others from the back);
//   3. present values still needed by gates not computed yet: level-1
//      pass-throughs; if they do not fit, roots are taken back, last first;
//   4. the 32-bit after-layout words holding written outputs are hooked.
// Every node of levels 1..13 is in the after-layout and readable by the
// next stage.
bool map_partition_ml(const Aig &g, const MapInput &in, Scratch &sc,
                      std::vector<BoomStage> &out, uint32_t max_write_words,
                      uint32_t max_level) {
  out.clear();
  std::vector<uint32_t> pins;
  pins.reserve(in.inputs.size() + in.nodes.size());
  for (auto p : in.inputs) pins.push_back(p);
  for (auto p : in.nodes) pins.push_back(p);
  const uint32_t ni = (uint32_t)in.inputs.size();
  const uint32_t n = (uint32_t)pins.size();
  for (uint32_t i = 0; i < n; ++i) sc.loc[pins[i]] = (int32_t)i;
  struct Guard {
    Scratch &s;
    const std::vector<uint32_t> &p;
    ~Guard() {
      for (auto x : p) s.loc[x] = -1;
    }
  } guard{sc, pins};
  std::vector<uint32_t> opa(n, 0), opb(n, 0);
  std::vector<std::vector<uint32_t>> cons(n);
  for (uint32_t i = ni; i < n; ++i) {
    uint32_t p = pins[i];
    int32_t la = sc.loc[g.a[p] >> 1], lb = sc.loc[g.b[p] >> 1];
    if (la < 0 || lb < 0) return false;
    opa[i] = (uint32_t)la;
    opb[i] = (uint32_t)lb;
    cons[la].push_back(i);
    if (lb != la) cons[lb].push_back(i);
  }
  std::vector<uint8_t> is_out(n, 0), written(n, 0), done(n, 0), present(n, 0);
  for (auto p : in.outputs) {
    int32_t l = sc.loc[p];
    if (l < 0) return false;
    is_out[l] = 1;
  }
  std::vector<uint32_t> height(n, 0);
  for (uint32_t i = n; i-- > ni;)
    for (auto c : cons[i]) height[i] = std::max(height[i], height[c] + 1);
  for (uint32_t i = 0; i < ni; ++i) done[i] = present[i] = 1;
  constexpr uint32_t kInf = 0xFFFFFFFFu;
  auto capl = [](int k) -> uint32_t { return k == 0 ? (uint32_t)kWidth : (uint32_t)kWidth >> k; };
  std::vector<std::vector<uint32_t>> occ(kBoomStages + 1);
  // per level, the length of the run of occupied slots at the front and at
  // the back: first-fit scans start past it.  Slots are freed only by a
  // rollback (slots that were free before the attempt, so never inside a
  // run), by a new stage and by a deferred root; the last two reset these.
  std::vector<uint32_t> front_run(kBoomStages + 1, 0), back_run(kBoomStages + 1, 0);
  std::vector<uint32_t> depth(n, kInf), now(n, 0);
  std::vector<std::pair<uint8_t, uint16_t>> log;
  // place value x at node (k, q) with its whole subtree (recursing through
  // a generic lambda rather than a std::function, so the calls inline)
  auto put_at = [&](auto &self, uint32_t x, int k, uint32_t q) -> bool {
    if (occ[k][q] != kNone) return false;
    if (k == 0) {
      if (!present[x]) return false;
      occ[0][q] = x;
      log.push_back({0, (uint16_t)q});
      return true;
    }
    if (!present[x] && (depth[x] == kInf || depth[x] > (uint32_t)k)) return false;
    occ[k][q] = x;
    log.push_back({(uint8_t)k, (uint16_t)q});
    if (present[x]) return self(self, x, k - 1, q);
    return self(self, opa[x], k - 1, q) && self(self, opb[x], k - 1, q + capl(k));
  };
  auto put = [&put_at](uint32_t x, int k, uint32_t q) { return put_at(put_at, x, k, q); };
  // undo a placement attempt that failed (its nodes were never counted)
  auto rollback = [&](size_t mark) {
    while (log.size() > mark) {
      auto e = log.back();
      log.pop_back();
      occ[e.first][e.second] = kNone;
    }
  };
  auto commit_now = [&](size_t mark) {
    for (size_t j = mark; j < log.size(); ++j) {
      uint32_t x = occ[log[j].first][log[j].second];
      if (log[j].first > 0 && !present[x]) ++now[x];
    }
  };
  uint32_t total_words = 0;
  while (true) {
    bool todo = false;
    for (uint32_t i = ni; i < n && !todo; ++i) todo = !done[i];
    for (uint32_t i = 0; i < n && !todo; ++i) todo = is_out[i] && !written[i];
    if (!todo) break;
    if (out.size() >= 4096) return false;
    for (int k = 0; k <= kBoomStages; ++k) occ[k].assign(capl(k), kNone);
    std::fill(front_run.begin(), front_run.end(), 0u);
    std::fill(back_run.begin(), back_run.end(), 0u);
    std::fill(now.begin(), now.end(), 0u);
    log.clear();
    // cone depth of every gate not computed yet, over present values
    for (uint32_t i = 0; i < ni; ++i) depth[i] = present[i] ? 0 : kInf;
    for (uint32_t i = ni; i < n; ++i) {
      if (done[i]) {
        depth[i] = present[i] ? 0 : kInf;
        continue;
      }
      uint32_t da = present[opa[i]] ? 0 : depth[opa[i]];
      uint32_t db = present[opb[i]] ? 0 : depth[opb[i]];
      depth[i] = (da == kInf || db == kInf) ? kInf : 1 + std::max(da, db);
    }
    std::vector<std::pair<uint8_t, uint16_t>> wpos;  // written outputs
    struct Root {
      uint32_t x;
      size_t mark, end;
      bool out;
      uint8_t k;
      uint16_t q;
    };
    std::vector<Root> roots;
    std::vector<uint8_t> wrote_now(n, 0);
    auto reset_stage = [&]() {
      for (int k = 0; k <= kBoomStages; ++k) std::fill(occ[k].begin(), occ[k].end(), kNone);
      std::fill(front_run.begin(), front_run.end(), 0u);
      std::fill(back_run.begin(), back_run.end(), 0u);
      std::fill(now.begin(), now.end(), 0u);
      log.clear();
      wpos.clear();
      roots.clear();
      std::fill(wrote_now.begin(), wrote_now.end(), 0);
    };
    // 1. unwritten present outputs: level-1 pass-throughs at the front
    auto place_present_outputs = [&]() -> bool {
      uint32_t front1 = 0;
      for (uint32_t i = 0; i < n; ++i) {
        if (!(is_out[i] && !written[i] && present[i])) continue;
        while (front1 < capl(1) && (occ[1][front1] != kNone || occ[0][front1] != kNone)) ++front1;
        if (front1 >= capl(1) || !put(i, 1, front1)) return false;
        wpos.push_back({1, (uint16_t)front1});
        wrote_now[i] = 1;
      }
      return true;
    };
    // first-fit root placement: outputs from the front of their level,
    // other roots from the back
    auto place_root = [&](uint32_t x, uint32_t max_tries) -> bool {
      int k = (int)depth[x];
      uint32_t cap = capl(k);
      size_t mark = log.size();
      uint32_t tries = 0;
      // occupied slots are skipped without counting as tries, so the scan
      // may start past the occupied run at its end
      uint32_t &run = is_out[x] ? front_run[k] : back_run[k];
      while (run < cap && occ[k][is_out[x] ? run : cap - 1 - run] != kNone) ++run;
      for (uint32_t j = run; j < cap && tries < max_tries; ++j) {
        uint32_t q = is_out[x] ? j : cap - 1 - j;
        if (occ[k][q] != kNone) continue;
        ++tries;
        if (put(x, k, q)) {
          roots.push_back({x, mark, log.size(), (bool)is_out[x], (uint8_t)k, (uint16_t)q});
          commit_now(mark);
          return true;
        }
        rollback(mark);
      }
      return false;
    };
    // 3. carries: present values still needed by gates not computed yet
    auto needed_later = [&](uint32_t v) {
      for (auto c : cons[v])
        if (!done[c] && !now[c]) return true;
      return false;
    };
    std::vector<uint8_t> at_l1(n, 0);
    auto carries_fit = [&](bool place) -> bool {
      std::fill(at_l1.begin(), at_l1.end(), 0);
      for (int k = 1; k <= kBoomStages; ++k)
        for (auto x : occ[k])
          if (x != kNone) at_l1[x] = 1;
      uint32_t need = 0;
      for (uint32_t v = 0; v < n; ++v) need += present[v] && !at_l1[v] && needed_later(v);
      uint32_t free_slots = 0;
      for (uint32_t q = 0; q < capl(1); ++q)
        free_slots += occ[1][q] == kNone && occ[0][q] == kNone;
      if (need > free_slots) return false;
      if (place) {
        uint32_t q = capl(1);
        for (uint32_t v = 0; v < n; ++v) {
          if (!(present[v] && !at_l1[v] && needed_later(v))) continue;
          while (q-- > 0)
            if (occ[1][q] == kNone && occ[0][q] == kNone) break;
          put(v, 1, q);
        }
      }
      return true;
    };
    // 2. the deepest D whose whole frontier (gates within depth D that are
    // outputs or feed deeper gates) fits, with its carries
    uint32_t dmax = 0;
    for (uint32_t i = ni; i < n; ++i)
      if (!done[i] && depth[i] != kInf) dmax = std::max(dmax, depth[i]);
    bool accepted = false;
    for (uint32_t D = std::min(dmax, max_level); D >= 2 && !accepted; --D) {
      std::vector<uint32_t> front;
      for (uint32_t i = ni; i < n; ++i) {
        if (done[i] || depth[i] == kInf || depth[i] > D) continue;
        bool f = is_out[i];
        for (auto c : cons[i])
          f |= !done[c] && (depth[c] == kInf || depth[c] > D);
        if (f) front.push_back(i);
      }
      std::stable_sort(front.begin(), front.end(), [&](uint32_t x, uint32_t y) {
        if (depth[x] != depth[y]) return depth[x] > depth[y];
        return height[x] > height[y];
      });
      reset_stage();
      if (!place_present_outputs()) return false;
      bool ok = true;
      for (uint32_t x : front)
        if (!place_root(x, 0xFFFFFFFFu)) {
          ok = false;
          break;
        }
      if (ok && carries_fit(true)) accepted = true;
    }
    if (!accepted) {
      // depth-1 roots by priority, as many as fit with their carries
      reset_stage();
      if (!place_present_outputs()) return false;
      std::vector<uint32_t> cand;
      for (uint32_t i = ni; i < n; ++i)
        if (!done[i] && depth[i] == 1) cand.push_back(i);
      std::stable_sort(cand.begin(), cand.end(),
                       [&](uint32_t x, uint32_t y) { return height[x] > height[y]; });
      for (uint32_t x : cand) place_root(x, 0xFFFFFFFFu);
      if (!carries_fit(false)) {
        // too many carries (wide inputs): defer roots, lowest priority
        // first, whose operands stay carried anyway; each frees one level-1
        // slot without adding a carry
        std::fill(at_l1.begin(), at_l1.end(), 0);
        for (int k = 1; k <= kBoomStages; ++k)
          for (auto x : occ[k])
            if (x != kNone) at_l1[x] = 1;
        uint32_t need = 0, free_slots = 0;
        for (uint32_t v = 0; v < n; ++v) need += present[v] && !at_l1[v] && needed_later(v);
        for (uint32_t q = 0; q < capl(1); ++q)
          free_slots += occ[1][q] == kNone && occ[0][q] == kNone;
        for (size_t r = roots.size(); r-- > 0 && need > free_slots;) {
          const uint32_t x = roots[r].x;
          bool free_defer = true;
          for (uint32_t v : {opa[x], opb[x]}) {
            if (!present[v] || at_l1[v]) continue;
            bool other = false;
            for (auto c : cons[v]) other |= c != x && !done[c] && !now[c];
            if (!other) free_defer = false;
          }
          if (!free_defer) continue;
          for (size_t j = roots[r].end; j-- > roots[r].mark;) {
            auto e = log[j];
            uint32_t y = occ[e.first][e.second];
            if (y == kNone) continue;
            if (e.first > 0 && !present[y]) --now[y];
            occ[e.first][e.second] = kNone;
          }
          roots.erase(roots.begin() + (long)r);
          std::fill(front_run.begin(), front_run.end(), 0u);
          std::fill(back_run.begin(), back_run.end(), 0u);
          ++free_slots;
        }
        if (need > free_slots) return false;
      }
      carries_fit(true);
    }
    // progress: something computed or written
    if (wpos.empty() && roots.empty()) return false;
    for (auto &r : roots)
      if (r.out) wpos.push_back({r.k, r.q});
    // 4. hooks
    BoomStage bs;
    for (int k = 0; k <= kBoomStages; ++k) {
      bool any = false;
      for (auto x : occ[k]) any |= x != kNone;
      if (!any) continue;
      bs.hier[k].assign(capl(k), kNone);
      for (uint32_t q = 0; q < capl(k); ++q)
        if (occ[k][q] != kNone) bs.hier[k][q] = pins[occ[k][q]];
    }
    std::vector<uint16_t> words;
    for (auto &w : wpos) words.push_back((uint16_t)((capl(w.first) + w.second) >> 5));
    std::sort(words.begin(), words.end());
    words.erase(std::unique(words.begin(), words.end()), words.end());
    bs.write_outs = words;
    total_words += (uint32_t)words.size();
    if (total_words > max_write_words) return false;
    // advance: computed gates, written outputs, present set of the next stage
    for (auto &r : roots)
      if (r.out) written[r.x] = 1;
    for (uint32_t i = 0; i < n; ++i)
      if (wrote_now[i]) written[i] = 1;
    for (uint32_t i = ni; i < n; ++i)
      if (now[i]) done[i] = 1;
    std::fill(present.begin(), present.end(), 0);
    for (int k = 1; k <= kBoomStages; ++k)
      for (auto x : occ[k])
        if (x != kNone) present[x] = 1;
    // an output computed inside a subtree but not as a root is written by a
    // later stage (step 1), so it must stay present: it is, at level >= 1
    out.push_back(std::move(bs));
  }
  return true;
}

// ---------------------------------------------------------------------------
// Endpoint bookkeeping for one major stage.
struct EndpointInfo {
  std::vector<uint32_t> outs;  // pins that must be written out
  uint32_t nsram = 0;
};

void endpoint_outputs(const Aig &g, const MajorStage &ms, size_t e,
                      EndpointInfo &info) {
  info.outs.clear();
  info.nsram = 0;
  if (e < ms.staged_out.size()) {
    info.outs.push_back(ms.staged_out[e]);
    return;
  }
  size_t oe = ms.orig_endpoints[e - ms.staged_out.size()];
  if (oe >= g.pos.size() + g.dffs.size()) info.nsram = 1;
  g.for_each_endpoint_input(oe, [&](uint32_t p) { info.outs.push_back(p); });
  if (oe >= g.pos.size() && oe < g.pos.size() + g.dffs.size()) {
    const Dff &d = g.dffs[oe - g.pos.size()];
    if (d.d_iv <= 1) info.outs.clear();  // constant-D DFF needs no slot
  }
}

// Per-thread mapping context (stamp arrays over all pins).
struct MapCtx {
  Scratch sc;
  std::vector<uint32_t> mark, seen, stack;
  uint32_t gstamp = 1, sstamp = 1;
  explicit MapCtx(size_t n) : sc(n), mark(n, 0), seen(n, 0) {}
};

// Map the endpoints `eps` of major stage `ms` (endpoint bookkeeping in
// `infos`; `avail_pin` says which pins earlier stages provide) as one
// partition.  Returns false when they do not fit one: more than 8192 input
// bits, too many SRAM / duplicate writeout words, or a cone the mappers
// cannot place.
bool map_group(const Aig &g, const CompileOptions &opt, const MajorStage &ms,
               const std::vector<EndpointInfo> &infos,
               const std::function<bool(uint32_t)> &avail_pin,
               const std::vector<uint32_t> &eps, MapCtx &cx, Partition &part) {
  Scratch &sc = cx.sc;
  std::vector<uint32_t> &mark = cx.mark, &seen = cx.seen, &stack = cx.stack;
  uint32_t &gstamp = cx.gstamp, &sstamp = cx.sstamp;
  MapInput mi;
  ++sstamp;
  uint32_t nsram = 0;
  for (auto e : eps) {
    nsram += infos[e].nsram;
    for (auto p : infos[e].outs) {
      if (p == 0) continue;
      if (seen[p] != sstamp) {
        seen[p] = sstamp;
        mi.outputs.push_back(p);
      }
    }
  }
  // cone closure
  ++gstamp;
  for (auto p : mi.outputs) stack.push_back(p);
  while (!stack.empty()) {
    uint32_t p = stack.back();
    stack.pop_back();
    if (p == 0 || mark[p] == gstamp) continue;
    mark[p] = gstamp;
    if (avail_pin(p)) {
      mi.inputs.push_back(p);
      continue;
    }
    mi.nodes.push_back(p);
    stack.push_back(g.a[p] >> 1);
    stack.push_back(g.b[p] >> 1);
  }
  std::sort(mi.nodes.begin(), mi.nodes.end());
  std::sort(mi.inputs.begin(), mi.inputs.end());
  // duplicate words: distinct activations of one pin beyond the first
  std::map<uint32_t, std::vector<uint32_t>> act;
  for (auto e : eps) {
    if (e < ms.staged_out.size()) {
      auto &v = act[ms.staged_out[e]];
      if (std::find(v.begin(), v.end(), 2u) == v.end()) v.push_back(2u);
      continue;
    }
    size_t oe = ms.orig_endpoints[e - ms.staged_out.size()];
    uint32_t pin = 0, key = 0;
    if (oe < g.pos.size()) {
      if (g.pos[oe] <= 1) continue;
      pin = g.pos[oe] >> 1;
      key = 2 | (g.pos[oe] & 1);
    } else if (oe < g.pos.size() + g.dffs.size()) {
      const Dff &d = g.dffs[oe - g.pos.size()];
      if (d.d_iv <= 1) continue;
      pin = d.d_iv >> 1;
      key = d.en_iv << 1 | (d.d_iv & 1);
    } else {
      continue;
    }
    auto &v = act[pin];
    if (std::find(v.begin(), v.end(), key) == v.end()) v.push_back(key);
  }
  uint32_t ndup = 0;
  for (auto &kv : act) ndup += (uint32_t)kv.second.size() - 1;
  uint32_t dup_words = (ndup + 31) / 32;
  if (mi.inputs.size() > kWidth || nsram * 4 + dup_words > 256 || nsram + dup_words >= 256)
    return false;
  std::vector<BoomStage> stages;
  // multi-level mapping first, the level-1 mapper as the fallback
  bool m = opt.multilevel &&
           map_partition_ml(g, mi, sc, stages, 256 - nsram - dup_words, opt.multilevel);
  if (!m) m = map_partition(g, mi, sc, stages, 256 - nsram - dup_words);
  if (!m) return false;
  part.endpoints = eps;
  part.stages = std::move(stages);
  part.num_nodes = mi.nodes.size();
  return true;
}

}  // namespace

// ---------------------------------------------------------------------------
void compile_partitions(const Aig &g, const CompileOptions &opt, Compiled &out,
                        std::string *log) {
  std::ostringstream lg;
  using clk = std::chrono::steady_clock;
  double t_group = 0, t_map = 0;
  auto secs = [](clk::time_point a) {
    return std::chrono::duration<double>(clk::now() - a).count();
  };
  Levels L = levelize(g);
  uint32_t split = opt.level_split ? opt.level_split : std::max<uint32_t>(1, L.max_lvl);
  uint32_t nmajor = std::max<uint32_t>(1, (L.max_lvl + split - 1) / split);
  auto stage_of = [&](uint32_t p) -> int {
    if (g.drv[p] != kAnd) return -1;
    return (int)std::min<uint32_t>((L.lvl[p] - 1) / split, nmajor - 1);
  };
  // endpoint stage = deepest fan-in stage (sources -> 0)
  size_t ne = g.num_endpoints();
  std::vector<int> ep_stage(ne, 0);
  for (size_t e = 0; e < ne; ++e) {
    int s = 0;
    g.for_each_endpoint_input(e, [&](uint32_t p) { s = std::max(s, stage_of(p)); });
    ep_stage[e] = s;
  }
  // last stage in which each node is consumed
  std::vector<int> last_use(g.num_pins + 1, -1);
  for (uint32_t p = 1; p <= g.num_pins; ++p)
    if (g.drv[p] == kAnd) {
      int s = stage_of(p);
      last_use[g.a[p] >> 1] = std::max(last_use[g.a[p] >> 1], s);
      last_use[g.b[p] >> 1] = std::max(last_use[g.b[p] >> 1], s);
    }
  for (size_t e = 0; e < ne; ++e)
    g.for_each_endpoint_input(e, [&](uint32_t p) {
      last_use[p] = std::max(last_use[p], ep_stage[e]);
    });
  out.stages.assign(nmajor, MajorStage());
  out.level_split = split;
  out.opt = opt;
  // the AND nodes of each major stage in ascending pin order (fan-ins
  // first), so that the per-stage passes below visit only their own nodes
  std::vector<uint32_t> stage_first(nmajor + 1, 0), stage_nodes;
  for (uint32_t p = 1; p <= g.num_pins; ++p) {
    int s = stage_of(p);
    if (s >= 0) ++stage_first[s + 1];
    if (s >= 0 && last_use[p] > s) out.stages[s].staged_out.push_back(p);
  }
  for (uint32_t m = 0; m < nmajor; ++m) stage_first[m + 1] += stage_first[m];
  stage_nodes.resize(stage_first[nmajor]);
  {
    std::vector<uint32_t> fill(stage_first.begin(), stage_first.end() - 1);
    for (uint32_t p = 1; p <= g.num_pins; ++p) {
      int s = stage_of(p);
      if (s >= 0) stage_nodes[fill[s]++] = p;
    }
  }
  for (size_t e = 0; e < ne; ++e) out.stages[ep_stage[e]].orig_endpoints.push_back((uint32_t)e);

  // stage-relative level of every AND node; each stage writes only its own
  std::vector<uint32_t> lvl_m(g.num_pins + 1, 0);
  size_t total_parts = 0, total_bstages = 0, total_nodes = 0;

  // per-thread mapping contexts (stamp arrays over all pins), made on first use
  int max_threads = 1;
#ifdef _OPENMP
  max_threads = omp_get_max_threads();
#endif
  std::vector<std::unique_ptr<MapCtx>> ctxs(max_threads);

  // Each major stage's greedy grouping is a task; once it has closed its
  // groups it adds one mapping task per group.  A stage's grouping reads
  // and writes only its own nodes' levels, and tracks membership with the
  // stamp arrays of the mapping context of the thread running it (its
  // stamps come from the same counters, so marks never collide).  Errors
  // are kept per stage and group and the first in stage order is rethrown.
  std::vector<std::vector<EndpointInfo>> infos_all(nmajor);
  std::vector<std::vector<std::vector<uint32_t>>> groups_all(nmajor);
  std::vector<std::vector<std::vector<Partition>>> built_all(nmajor);
  std::vector<std::vector<std::string>> err_all(nmajor);
  std::vector<std::string> group_err(nmajor);
  std::vector<double> group_s(nmajor, 0.0);
  auto ctx_here = [&]() -> MapCtx & {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    if (!ctxs[tid]) ctxs[tid] = std::make_unique<MapCtx>(g.num_pins + 1);
    return *ctxs[tid];
  };
  // map endpoints `eps` of stage m into res, in halves when they do not fit
  std::function<void(uint32_t, MapCtx &, const std::vector<uint32_t> &,
                     std::vector<Partition> &)>
      build = [&](uint32_t m, MapCtx &cx, const std::vector<uint32_t> &eps,
                  std::vector<Partition> &res) {
    const std::function<bool(uint32_t)> avail = [&stage_of, m](uint32_t p) {
      return stage_of(p) < (int)m;
    };
    Partition part;
    if (map_group(g, opt, out.stages[m], infos_all[m], avail, eps, cx, part) &&
        (!opt.max_boomerang_stages || eps.size() <= 1 ||
         part.stages.size() <= opt.max_boomerang_stages)) {
      res.push_back(std::move(part));
      return;
    }
    if (eps.size() <= 1)
      throw std::runtime_error("compile: a single endpoint does not fit a partition");
    std::vector<uint32_t> a(eps.begin(), eps.begin() + eps.size() / 2);
    std::vector<uint32_t> b(eps.begin() + eps.size() / 2, eps.end());
    build(m, cx, a, res);
    build(m, cx, b, res);
  };
  const auto t_all = clk::now();

#pragma omp parallel
#pragma omp single
  for (uint32_t m = 0; m < nmajor; ++m) {
#pragma omp task firstprivate(m)
  try {
    auto t0 = clk::now();
    MapCtx &cx = ctx_here();
    std::vector<uint32_t> &mark = cx.mark, &seen = cx.seen;  // group / collect stamps
    uint32_t &gstamp = cx.gstamp, &sstamp = cx.sstamp;
    ++gstamp;  // the first group's stamp: fresh
    const MajorStage &ms = out.stages[m];
    auto avail_pin = [&](uint32_t p) { return stage_of(p) < (int)m; };
    // stage-relative levels (cone depth above this stage's inputs)
    // endpoints: staged outputs then original endpoints
    size_t nep = ms.num_endpoints();
    std::vector<EndpointInfo> &infos = infos_all[m];
    infos.resize(nep);
    for (size_t e = 0; e < nep; ++e) endpoint_outputs(g, ms, e, infos[e]);

    // greedy grouping of endpoints; a group's cone is tracked with stamps
    std::vector<uint32_t> stack;
    std::vector<uint32_t> cur_nodes, cur_inputs;
    std::vector<uint32_t> lvl_count, lvl_extra, lvl_touched;
    uint32_t cur_outs = 0, cur_srams = 0;
    std::vector<uint32_t> add_nodes, add_inputs;
    auto collect = [&](size_t e) {
      add_nodes.clear();
      add_inputs.clear();
      ++sstamp;
      for (auto p : infos[e].outs) stack.push_back(p);
      while (!stack.empty()) {
        uint32_t p = stack.back();
        stack.pop_back();
        if (p == 0 || seen[p] == sstamp || mark[p] == gstamp) continue;
        seen[p] = sstamp;
        if (avail_pin(p)) {
          add_inputs.push_back(p);
          continue;
        }
        add_nodes.push_back(p);
        stack.push_back(g.a[p] >> 1);
        stack.push_back(g.b[p] >> 1);
      }
    };
    // stage-relative level of every node of this stage
    for (uint32_t k = stage_first[m]; k < stage_first[m + 1]; ++k) {
      const uint32_t p = stage_nodes[k];
      uint32_t la = avail_pin(g.a[p] >> 1) ? 0 : lvl_m[g.a[p] >> 1];
      uint32_t lb = avail_pin(g.b[p] >> 1) ? 0 : lvl_m[g.b[p] >> 1];
      lvl_m[p] = 1 + std::max(la, lb);
    }
    std::vector<std::vector<uint32_t>> &group_eps = groups_all[m];
    std::vector<uint32_t> ge;
    auto flush = [&]() {
      if (!ge.empty()) group_eps.push_back(ge);
      ge.clear();
      ++gstamp;
      cur_nodes.clear();
      cur_inputs.clear();
      lvl_count.clear();
      cur_outs = 0;
      cur_srams = 0;
    };
    for (size_t e = 0; e < nep; ++e) {
      collect(e);
      uint32_t nouts = (uint32_t)infos[e].outs.size();
      bool fits = true;
      if (!ge.empty()) {
        if (cur_inputs.size() + add_inputs.size() > opt.max_inputs) fits = false;
        if (cur_outs + nouts > opt.max_outputs) fits = false;
        if (cur_nodes.size() + add_nodes.size() > opt.max_part_nodes) fits = false;
        if (cur_srams + infos[e].nsram > 48) fits = false;
        if (fits) {
          // nodes the endpoint adds per stage-relative level, against the cap
          for (auto p : add_nodes) {
            const uint32_t l = lvl_m[p];
            if (l >= lvl_extra.size()) lvl_extra.resize(l + 1, 0);
            if (lvl_extra[l]++ == 0) lvl_touched.push_back(l);
          }
          for (uint32_t l : lvl_touched) {
            const uint32_t have = l < lvl_count.size() ? lvl_count[l] : 0;
            if (have + lvl_extra[l] > opt.max_level_nodes) fits = false;
            lvl_extra[l] = 0;
          }
          lvl_touched.clear();
        }
        if (!fits) {
          flush();
          collect(e);
        }
      }
      ge.push_back((uint32_t)e);
      for (auto p : add_nodes) {
        mark[p] = gstamp;
        cur_nodes.push_back(p);
        if (lvl_m[p] >= lvl_count.size()) lvl_count.resize(lvl_m[p] + 1, 0);
        ++lvl_count[lvl_m[p]];
      }
      for (auto p : add_inputs) {
        mark[p] = gstamp;
        cur_inputs.push_back(p);
      }
      cur_outs += nouts;
      cur_srams += infos[e].nsram;
    }
    flush();
    group_s[m] = secs(t0);
    // the grouping is done with the context's stamp arrays: from here on
    // the thread may run mapping tasks on the same context
    built_all[m].resize(group_eps.size());
    err_all[m].resize(group_eps.size());
    for (size_t gi = 0; gi < group_eps.size(); ++gi) {
#pragma omp task firstprivate(m, gi)
      {
        try {
          build(m, ctx_here(), groups_all[m][gi], built_all[m][gi]);
        } catch (const std::exception &e) {
          err_all[m][gi] = e.what();
        }
      }
    }
  } catch (const std::exception &e) {
    group_err[m] = e.what();
  }
  }
  // (the parallel region's closing barrier has waited for every task)
  for (uint32_t m = 0; m < nmajor; ++m) {
    if (!group_err[m].empty()) throw std::runtime_error(group_err[m]);
    for (const std::string &e : err_all[m])
      if (!e.empty()) throw std::runtime_error(e);
  }
  for (double t : group_s) t_group += t;
  t_map = secs(t_all);
  for (uint32_t m = 0; m < nmajor; ++m) {
    MajorStage &ms = out.stages[m];
    for (auto &v : built_all[m])
      for (auto &part : v) {
        total_bstages += part.stages.size();
        total_nodes += part.num_nodes;
        ms.parts.push_back(std::move(part));
      }
    // sort by decreasing number of boomerang stages (src/pe.rs:1068)
    std::stable_sort(ms.parts.begin(), ms.parts.end(),
                     [](const Partition &x, const Partition &y) {
                       return x.stages.size() > y.stages.size();
                     });
    total_parts += ms.parts.size();
    lg << "major stage " << m << ": " << ms.staged_out.size() << " staged outputs, "
       << ms.orig_endpoints.size() << " endpoints, " << ms.parts.size()
       << " partitions\n";
  }
  lg << "compile time: grouping " << t_group << " s summed over major stages, grouping and "
     << "mapping " << t_map << " s in all\n";
  lg << "total: " << total_parts << " partitions, " << total_bstages
     << " boomerang stages, " << total_nodes << " mapped AND evaluations ("
     << g.num_and() << " unique)\n";
  if (log) *log = lg.str();
}

void merge_partitions(const Aig &g, const Compiled &in, const uint32_t *pairs,
                      size_t n_pairs, Compiled &out, std::vector<uint8_t> &merged) {
  const uint32_t nmajor = (uint32_t)in.stages.size();
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> by_stage(nmajor);
  std::vector<std::vector<size_t>> pair_id(nmajor);
  for (size_t k = 0; k < n_pairs; ++k) {
    uint32_t m = pairs[3 * k], i = pairs[3 * k + 1], j = pairs[3 * k + 2];
    if (m >= nmajor || i >= in.stages[m].parts.size() || j >= in.stages[m].parts.size() ||
        i == j)
      throw std::runtime_error("merge: pair out of range");
    by_stage[m].push_back({i, j});
    pair_id[m].push_back(k);
  }
  Levels L = levelize(g);
  const uint32_t split = in.level_split ? in.level_split : 1;
  auto stage_of = [&](uint32_t p) -> int {
    if (g.drv[p] != kAnd) return -1;
    return (int)std::min<uint32_t>((L.lvl[p] - 1) / split, nmajor - 1);
  };
  out.stages.clear();
  out.stages.reserve(nmajor);
  out.part_x = in.part_x;
  out.level_split = in.level_split;
  out.opt = in.opt;
  merged.assign(n_pairs, 0);
  std::vector<std::unique_ptr<MapCtx>> ctxs;
  for (uint32_t m = 0; m < nmajor; ++m) {
    const MajorStage &ms = in.stages[m];
    MajorStage o;
    o.staged_out = ms.staged_out;
    o.orig_endpoints = ms.orig_endpoints;
    const auto &pl = by_stage[m];
    std::vector<uint8_t> used(ms.parts.size(), 0);
    for (auto &pr : pl) {
      if (used[pr.first] || used[pr.second])
        throw std::runtime_error("merge: a partition appears in two pairs of one stage");
      used[pr.first] = used[pr.second] = 1;
    }
    std::vector<EndpointInfo> infos;
    if (!pl.empty()) {
      infos.resize(ms.num_endpoints());
      for (size_t e = 0; e < infos.size(); ++e) endpoint_outputs(g, ms, e, infos[e]);
    }
    const std::function<bool(uint32_t)> avail = [&](uint32_t p) {
      return stage_of(p) < (int)m;
    };
    std::vector<Partition> res(pl.size());
    std::vector<uint8_t> ok(pl.size(), 0);
    std::vector<std::string> par_err(pl.size());
#pragma omp parallel
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
#pragma omp single
      {
        int nt = 1;
#ifdef _OPENMP
        nt = omp_get_num_threads();
#endif
        if ((int)ctxs.size() < nt) ctxs.resize(nt);
      }
#pragma omp for schedule(dynamic, 1)
      for (int64_t k = 0; k < (int64_t)pl.size(); ++k) {
        try {
          if (!ctxs[tid]) ctxs[tid] = std::make_unique<MapCtx>(g.num_pins + 1);
          std::vector<uint32_t> eps = ms.parts[pl[k].first].endpoints;
          const auto &b = ms.parts[pl[k].second].endpoints;
          eps.insert(eps.end(), b.begin(), b.end());
          std::sort(eps.begin(), eps.end());
          ok[k] = map_group(g, in.opt, ms, infos, avail, eps, *ctxs[tid], res[k]) ? 1 : 0;
        } catch (const std::exception &e) {
          par_err[k] = e.what();
        }
      }
    }
    for (const std::string &e : par_err)  // the first failure in pair order
      if (!e.empty()) throw std::runtime_error(e);
    std::vector<uint8_t> drop(ms.parts.size(), 0);
    for (size_t k = 0; k < pl.size(); ++k)
      if (ok[k]) {
        drop[pl[k].first] = drop[pl[k].second] = 1;
        merged[pair_id[m][k]] = 1;
      }
    for (size_t i = 0; i < ms.parts.size(); ++i)
      if (!drop[i]) o.parts.push_back(ms.parts[i]);
    for (size_t k = 0; k < pl.size(); ++k)
      if (ok[k]) o.parts.push_back(std::move(res[k]));
    std::stable_sort(o.parts.begin(), o.parts.end(),
                     [](const Partition &x, const Partition &y) {
                       return x.stages.size() > y.stages.size();
                     });
    out.stages.push_back(std::move(o));
  }
}

void split_partitions(const Aig &g, const Compiled &in, const uint32_t *parts,
                      size_t n_parts, Compiled &out, std::vector<uint8_t> &split) {
  const uint32_t nmajor = (uint32_t)in.stages.size();
  std::vector<std::vector<uint32_t>> by_stage(nmajor);
  std::vector<std::vector<size_t>> part_id(nmajor);
  for (size_t k = 0; k < n_parts; ++k) {
    uint32_t m = parts[2 * k], i = parts[2 * k + 1];
    if (m >= nmajor || i >= in.stages[m].parts.size())
      throw std::runtime_error("split: partition out of range");
    by_stage[m].push_back(i);
    part_id[m].push_back(k);
  }
  Levels L = levelize(g);
  const uint32_t lsplit = in.level_split ? in.level_split : 1;
  auto stage_of = [&](uint32_t p) -> int {
    if (g.drv[p] != kAnd) return -1;
    return (int)std::min<uint32_t>((L.lvl[p] - 1) / lsplit, nmajor - 1);
  };
  out.stages.clear();
  out.stages.reserve(nmajor);
  out.part_x = in.part_x;
  out.level_split = in.level_split;
  out.opt = in.opt;
  split.assign(n_parts, 0);
  std::vector<std::unique_ptr<MapCtx>> ctxs;
  for (uint32_t m = 0; m < nmajor; ++m) {
    const MajorStage &ms = in.stages[m];
    MajorStage o;
    o.staged_out = ms.staged_out;
    o.orig_endpoints = ms.orig_endpoints;
    const auto &pl = by_stage[m];
    std::vector<uint8_t> used(ms.parts.size(), 0);
    for (uint32_t i : pl) {
      if (used[i]) throw std::runtime_error("split: a partition is named twice");
      used[i] = 1;
    }
    std::vector<EndpointInfo> infos;
    if (!pl.empty()) {
      infos.resize(ms.num_endpoints());
      for (size_t e = 0; e < infos.size(); ++e) endpoint_outputs(g, ms, e, infos[e]);
    }
    const std::function<bool(uint32_t)> avail = [&](uint32_t p) {
      return stage_of(p) < (int)m;
    };
    std::vector<Partition> lo(pl.size()), hi(pl.size());
    std::vector<uint8_t> ok(pl.size(), 0);
    std::vector<std::string> par_err(pl.size());
#pragma omp parallel
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
#pragma omp single
      {
        int nt = 1;
#ifdef _OPENMP
        nt = omp_get_num_threads();
#endif
        if ((int)ctxs.size() < nt) ctxs.resize(nt);
      }
#pragma omp for schedule(dynamic, 1)
      for (int64_t k = 0; k < (int64_t)pl.size(); ++k) {
        try {
          if (!ctxs[tid]) ctxs[tid] = std::make_unique<MapCtx>(g.num_pins + 1);
          const auto &eps = ms.parts[pl[k]].endpoints;
          if (eps.size() < 2) continue;
          std::vector<uint32_t> a(eps.begin(), eps.begin() + eps.size() / 2);
          std::vector<uint32_t> b(eps.begin() + eps.size() / 2, eps.end());
          ok[k] = map_group(g, in.opt, ms, infos, avail, a, *ctxs[tid], lo[k]) &&
                          map_group(g, in.opt, ms, infos, avail, b, *ctxs[tid], hi[k])
                      ? 1
                      : 0;
        } catch (const std::exception &e) {
          par_err[k] = e.what();
        }
      }
    }
    for (const std::string &e : par_err)  // the first failure in partition order
      if (!e.empty()) throw std::runtime_error(e);
    std::vector<uint8_t> drop(ms.parts.size(), 0);
    for (size_t k = 0; k < pl.size(); ++k)
      if (ok[k]) {
        drop[pl[k]] = 1;
        split[part_id[m][k]] = 1;
      }
    for (size_t i = 0; i < ms.parts.size(); ++i)
      if (!drop[i]) o.parts.push_back(ms.parts[i]);
    for (size_t k = 0; k < pl.size(); ++k)
      if (ok[k]) {
        o.parts.push_back(std::move(lo[k]));
        o.parts.push_back(std::move(hi[k]));
      }
    std::stable_sort(o.parts.begin(), o.parts.end(),
                     [](const Partition &x, const Partition &y) {
                       return x.stages.size() > y.stages.size();
                     });
    out.stages.push_back(std::move(o));
  }
}

// ---------------------------------------------------------------------------
// Flatten (src/flatten.rs:470-1130 for one partition, 1133-1573 overall).
namespace {

struct PermFlags {
  std::vector<uint16_t> perm;  // 8192
  std::vector<uint32_t> inv, set0, extra;
  void init() {
    perm.assign(kWidth, 0);
    inv.assign(kThreads, 0);
    set0.assign(kThreads, 0xFFFFFFFFu);
    extra.assign(kThreads, 0);
  }
};

inline void set_bit(std::vector<uint32_t> &v, uint32_t pos, uint32_t bit) {
  if (bit)
    v[pos >> 5] |= 1u << (pos & 31);
  else
    v[pos >> 5] &= ~(1u << (pos & 31));
}

struct FlatPart {
  uint32_t num_srams = 0, num_normal = 0, num_dups = 0, num_writeouts = 0;
  uint32_t state_start = 0, sram_start = 0;
  std::vector<uint32_t> wo_pins;  // normal writeout bits (num_normal*32)
  std::unordered_map<uint32_t, uint32_t> pin2pos;
  // pin -> ordered (activation key, placed position or kNone)
  std::unordered_map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>> acts;
  uint32_t cnt_dup = 0;
  PermFlags sd, ck;  // sram/dup permutation, clock-enable permutation
};

// after-layout of one boomerang stage: index (8192>>h)+j for level h >= 1
inline uint32_t after_at(const BoomStage &s, uint32_t idx) {
  if (idx == 0) return kNone;
  int h = kBoomStages;
  while ((kWidth >> h) <= idx && h > 0) --h;  // level with base 8192>>h <= idx
  // base = 8192 >> h ... find h with base <= idx < 2*base
  h = 0;
  for (int lv = 1; lv <= kBoomStages; ++lv) {
    uint32_t base = kWidth >> lv;
    if (idx >= base && idx < 2 * base) {
      h = lv;
      break;
    }
  }
  return s.at(h, idx - (kWidth >> h));
}

uint32_t query_perm(const FlatPart &fp, uint32_t lit, uint32_t &perm,
                    uint32_t &inv, uint32_t &set0) {
  if (lit <= 1) {
    perm = 0;
    inv = lit;
    set0 = 1;
    return 0;
  }
  auto it = fp.pin2pos.find(lit >> 1);
  if (it == fp.pin2pos.end()) throw std::runtime_error("flatten: pin not written out");
  perm = it->second;
  inv = lit & 1;
  set0 = 0;
  return 0;
}

void place(PermFlags &pf, uint32_t pos, uint32_t perm, uint32_t inv, uint32_t set0) {
  pf.perm[pos] = (uint16_t)perm;
  set_bit(pf.inv, pos, inv);
  set_bit(pf.set0, pos, set0);
}

void place_ck(FlatPart &fp, uint32_t pos, uint32_t clken_lit, uint32_t data_inv) {
  uint32_t perm, inv, set0;
  query_perm(fp, clken_lit, perm, inv, set0);
  place(fp.ck, pos, perm, inv, set0);
  set_bit(fp.ck.extra, pos, data_inv);
}

// get_or_place_output_with_activation (src/flatten.rs:639-692)
uint32_t get_or_place(FlatPart &fp, uint32_t lit, uint32_t clken_lit) {
  auto &v = fp.acts.at(lit >> 1);
  uint32_t key = clken_lit << 1 | (lit & 1);
  size_t idx = 0;
  while (idx < v.size() && v[idx].first != key) ++idx;
  if (idx == v.size()) throw std::runtime_error("flatten: activation missing");
  if (v[idx].second != kNone) return v[idx].second;
  auto it = fp.pin2pos.find(lit >> 1);
  if (it == fp.pin2pos.end()) throw std::runtime_error("flatten: output not placed");
  uint32_t origpos = it->second, r;
  if (idx == 0) {
    place_ck(fp, origpos, clken_lit, lit & 1);
    r = origpos;
  } else {
    ++fp.cnt_dup;
    uint32_t dup_pos = (fp.num_writeouts - fp.num_srams) * 32 - fp.cnt_dup;
    uint32_t dup_perm_pos = (fp.num_srams * 4 + fp.num_dups) * 32 - fp.cnt_dup;
    if (dup_perm_pos >= kWidth) throw std::runtime_error("flatten: duplicate overflow");
    place(fp.sd, dup_perm_pos, origpos, 0, 0);
    place_ck(fp, dup_pos, clken_lit, lit & 1);
    r = dup_pos;
  }
  v[idx].second = r;
  return r;
}

// Open-addressing map for flatten's per-partition tables: linear probing
// over a power-of-two array kept at most half full, slots by Fibonacci
// hashing (the key times 2^64 / golden ratio, high bits).  The all-ones key
// is never stored: it marks empty slots.
template <class K, class V>
class FlatMap {
 public:
  void clear() {
    if (used_) std::fill(keys_.begin(), keys_.end(), kEmpty);
    used_ = 0;
  }
  const V *find(K k) const {
    if (keys_.empty()) return nullptr;
    for (size_t i = slot(k);; i = (i + 1) & (keys_.size() - 1)) {
      if (keys_[i] == k) return &vals_[i];
      if (keys_[i] == kEmpty) return nullptr;
    }
  }
  void set(K k, V v) {
    if (2 * (used_ + 1) > keys_.size()) grow();
    size_t i = slot(k);
    while (keys_[i] != kEmpty && keys_[i] != k) i = (i + 1) & (keys_.size() - 1);
    if (keys_[i] == kEmpty) {
      keys_[i] = k;
      ++used_;
    }
    vals_[i] = v;
  }

 private:
  static constexpr K kEmpty = (K)~(K)0;
  std::vector<K> keys_;
  std::vector<V> vals_;
  size_t used_ = 0;
  uint32_t shift_ = 64;
  size_t slot(K k) const {
    return (size_t)(((uint64_t)k * 0x9E3779B97F4A7C15ull) >> shift_);
  }
  void grow() {
    std::vector<K> ok;
    std::vector<V> ov;
    ok.swap(keys_);
    ov.swap(vals_);
    const size_t cap = ok.empty() ? 1024 : 2 * ok.size();
    shift_ = 64 - (uint32_t)__builtin_ctzll(cap);
    keys_.assign(cap, kEmpty);
    vals_.assign(cap, V());
    used_ = 0;
    for (size_t j = 0; j < ok.size(); ++j)
      if (ok[j] != kEmpty) set(ok[j], ov[j]);
  }
};

// Pack (word, mask) global reads into 256 threads x rounds, <= 32 bits per
// thread (contract of map_global_read_to_rounds, src/flatten.rs:343-400).
// Each entry goes to the thread with the fewest entries, then the most bits
// that still leave room for it, then the lowest index; threads are kept in
// 256-bit sets by (entries, bits), so the choice is a walk over those sets
// in that order instead of a scan of all threads.
void pack_global_reads(std::vector<std::pair<uint32_t, uint32_t>> entries,
                       std::vector<std::vector<std::pair<uint32_t, uint32_t>>> &th) {
  th.assign(kThreads, {});
  std::vector<uint32_t> bits(kThreads, 0);
  std::stable_sort(entries.begin(), entries.end(), [](auto &x, auto &y) {
    return __builtin_popcount(x.second) > __builtin_popcount(y.second);
  });
  std::vector<std::pair<uint32_t, uint32_t>> work(entries.rbegin(), entries.rend());
  // a thread holds at most 32 entries (each has a bit) and 32 bits
  constexpr uint32_t kLim = 33;
  static_assert(kThreads == 256, "four 64-bit words per thread set");
  std::vector<std::array<uint64_t, 4>> set(kLim * kLim, std::array<uint64_t, 4>{0, 0, 0, 0});
  std::array<uint32_t, kLim> with_size{};
  set[0] = {~0ull, ~0ull, ~0ull, ~0ull};
  with_size[0] = kThreads;
  while (!work.empty()) {
    auto e = work.back();
    work.pop_back();
    uint32_t pc = (uint32_t)__builtin_popcount(e.second);
    int best = -1;
    for (uint32_t sz = 0; sz < kLim && best < 0; ++sz) {
      if (!with_size[sz]) continue;
      for (int b = 32 - (int)pc; b >= 0 && best < 0; --b) {
        const auto &w = set[sz * kLim + (uint32_t)b];
        for (int q = 0; q < 4; ++q)
          if (w[q]) {
            best = q * 64 + __builtin_ctzll(w[q]);
            break;
          }
      }
    }
    if (best < 0) {
      // split the mask in two halves
      uint32_t m = e.second, lo = 0;
      for (uint32_t k = 0; k < pc / 2; ++k) {
        uint32_t low = m & (0u - m);
        lo |= low;
        m ^= low;
      }
      if (!lo) throw std::runtime_error("flatten: global read does not fit");
      work.push_back({e.first, m});
      work.push_back({e.first, lo});
      continue;
    }
    const uint32_t sz = (uint32_t)th[best].size(), bq = (uint32_t)best >> 6;
    const uint64_t bm = 1ull << (best & 63);
    set[sz * kLim + bits[best]][bq] &= ~bm;
    --with_size[sz];
    th[best].push_back(e);
    bits[best] += pc;
    set[(sz + 1) * kLim + bits[best]][bq] |= bm;
    ++with_size[sz + 1];
  }
}

}  // namespace

void flatten(const Aig &g, const Compiled &c, const CompileOptions &opt,
             const std::vector<std::vector<double>> *costs, Script &s) {
  s = Script();
  const uint32_t nb = opt.num_blocks;
  s.num_blocks = nb;
  s.num_major_stages = (uint32_t)c.stages.size();
  s.input_map.assign((size_t)g.num_pins + 1, kNone);
  s.staged_io_map.assign((size_t)g.num_pins + 1, kNone);
  s.output_map.assign(2 * ((size_t)g.num_pins + 1), kNone);
  for (uint32_t p = 1; p <= g.num_pins; ++p)
    if (g.drv[p] == kInput || g.drv[p] == kClockFlag) s.input_layout.push_back(p);
  for (uint32_t i = 0; i < s.input_layout.size(); ++i) s.input_map[s.input_layout[i]] = i;
  // The per-partition loops below run in parallel; each records its
  // errors, and the first partition's in stage order is rethrown, the one
  // a serial pass would have met first.
  auto rethrow_first = [](const std::vector<std::string> &err) {
    for (const std::string &e : err)
      if (!e.empty()) throw std::runtime_error(e);
  };
  // map writes of one partition's endpoints, applied in partition order
  // once the stage's partitions are placed, so a key written by two
  // partitions keeps the later one's value as in one serial pass
  enum : uint8_t { kWStaged, kWOut, kWIn, kWDff, kWSram };
  struct MapWrite {
    uint8_t kind;
    uint32_t key, val;
  };
  // the state bit a global read takes a pin from, from the previous
  // snapshot (an input or register output) or written earlier this cycle
  auto read_pos = [&s](uint32_t pin, bool &staged) -> uint32_t {
    if (pin < s.input_map.size() && s.input_map[pin] != kNone) {
      staged = false;
      return s.input_map[pin];
    }
    if (pin < s.staged_io_map.size() && s.staged_io_map[pin] != kNone) {
      staged = true;
      return s.staged_io_map[pin];
    }
    throw std::runtime_error("flatten: a read pin is neither an input nor written out");
  };
  const uint32_t const_zero_pos = (uint32_t)s.input_layout.size();
  uint32_t states_start = (uint32_t)((s.input_layout.size() + 31) / 32);
  if (const_zero_pos == states_start * 32) states_start += 1;
  uint32_t sum_state = states_start, sum_sram = 0;
  s.sram_offsets.assign(g.srams.size(), kNone);
  s.dff_data_pos.assign(g.dffs.size(), kNone);

  std::vector<std::vector<std::vector<uint32_t>>> blocks_parts(c.stages.size());
  std::vector<std::vector<FlatPart>> fparts(c.stages.size());

  for (size_t m = 0; m < c.stages.size(); ++m) {
    const MajorStage &ms = c.stages[m];
    auto &bp = blocks_parts[m];
    bp.assign(nb, {});
    // masonry (src/flatten.rs:1207-1226), cost per partition
    std::vector<double> tot(nb, 0.0);
    auto cost_of = [&](size_t i) {
      if (costs && m < costs->size() && i < (*costs)[m].size()) return (*costs)[m][i];
      return (double)ms.parts[i].stages.size() + 3.0;
    };
    std::vector<size_t> order(ms.parts.size());
    std::iota(order.begin(), order.end(), 0);
    if (costs)
      std::stable_sort(order.begin(), order.end(),
                       [&](size_t x, size_t y) { return cost_of(x) > cost_of(y); });
    for (size_t k = 0; k < order.size(); ++k) {
      size_t i = order[k];
      size_t put;
      if (k < nb) {
        put = k;
      } else {
        put = (size_t)(std::min_element(tot.begin(), tot.end()) - tot.begin());
      }
      bp[put].push_back((uint32_t)i);
      tot[put] += cost_of(i);
    }
    auto &fps = fparts[m];
    fps.assign(ms.parts.size(), FlatPart());
    std::vector<std::string> err(ms.parts.size());
    // init_afters_writeouts (src/flatten.rs:470-586); each partition only
    // writes its own FlatPart
#pragma omp parallel for schedule(dynamic, 8)
    for (size_t i = 0; i < ms.parts.size(); ++i) try {
      const Partition &pt = ms.parts[i];
      FlatPart &fp = fps[i];
      for (const BoomStage &bs : pt.stages)
        for (uint16_t wo : bs.write_outs)
          for (uint32_t k = 0; k < 32; ++k) fp.wo_pins.push_back(after_at(bs, wo * 32u + k));
      fp.num_normal = 0;
      for (const BoomStage &bs : pt.stages) fp.num_normal += (uint32_t)bs.write_outs.size();
      uint32_t ndupbits = 0;
      for (uint32_t e : pt.endpoints) {
        if (e < ms.staged_out.size()) {
          auto &v = fp.acts[ms.staged_out[e]];
          bool have = false;
          for (auto &kv : v) have |= kv.first == 2u;
          if (!have) v.push_back({2u, kNone});
          continue;
        }
        size_t oe = ms.orig_endpoints[e - ms.staged_out.size()];
        uint32_t pin, key;
        if (oe < g.pos.size()) {
          if (g.pos[oe] <= 1) continue;
          pin = g.pos[oe] >> 1;
          key = 2 | (g.pos[oe] & 1);
        } else if (oe < g.pos.size() + g.dffs.size()) {
          const Dff &d = g.dffs[oe - g.pos.size()];
          if (d.d_iv <= 1) continue;
          pin = d.d_iv >> 1;
          key = d.en_iv << 1 | (d.d_iv & 1);
        } else {
          ++fp.num_srams;
          continue;
        }
        auto &v = fp.acts[pin];
        bool have = false;
        for (auto &kv : v) have |= kv.first == key;
        if (!have) v.push_back({key, kNone});
      }
      for (auto &kv : fp.acts) ndupbits += (uint32_t)kv.second.size() - 1;
      fp.num_dups = (ndupbits + 31) / 32;
      fp.num_writeouts = fp.num_normal + fp.num_srams + fp.num_dups;
      if (fp.num_writeouts > 256 || fp.num_srams * 4 + fp.num_dups > 256)
        throw std::runtime_error("flatten: partition exceeds 256 writeout words");
      for (uint32_t k = 0; k < fp.wo_pins.size(); ++k)
        if (fp.wo_pins[k] != kNone) fp.pin2pos[fp.wo_pins[k]] = k;
    } catch (const std::exception &e) {
      err[i] = e.what();
    }
    rethrow_first(err);
    // state and SRAM allocation in block-affinity order
    // (src/flatten.rs:1247-1278)
    for (auto &blk : bp)
      for (uint32_t pid : blk) {
        FlatPart &fp = fps[pid];
        fp.state_start = sum_state;
        sum_state += fp.num_writeouts;
        fp.sram_start = sum_sram;
        sum_sram += fp.num_srams * kSramWords;
      }
    // make_inputs_outputs (src/flatten.rs:694-858): placements in each
    // partition's own FlatPart in parallel, map writes recorded
    std::vector<std::vector<MapWrite>> wr(ms.parts.size());
#pragma omp parallel for schedule(dynamic, 8)
    for (size_t i = 0; i < ms.parts.size(); ++i) try {
      const Partition &pt = ms.parts[i];
      FlatPart &fp = fps[i];
      std::vector<MapWrite> &w = wr[i];
      fp.sd.init();
      fp.ck.init();
      fp.cnt_dup = 0;
      uint32_t cur_sram = 0;
      for (uint32_t e : pt.endpoints) {
        if (e < ms.staged_out.size()) {
          uint32_t p = ms.staged_out[e];
          w.push_back({kWStaged, p, fp.state_start * 32 + get_or_place(fp, p << 1, 1)});
          continue;
        }
        size_t oe = ms.orig_endpoints[e - ms.staged_out.size()];
        if (oe < g.pos.size()) {
          uint32_t lit = g.pos[oe];
          if (lit <= 1) {
            w.push_back({kWOut, lit, 0});
            continue;
          }
          w.push_back({kWOut, lit, fp.state_start * 32 + get_or_place(fp, lit, 1)});
        } else if (oe < g.pos.size() + g.dffs.size()) {
          size_t di = oe - g.pos.size();
          const Dff &d = g.dffs[di];
          if (d.d_iv <= 1) {
            w.push_back({kWIn, d.q, const_zero_pos});
            continue;
          }
          uint32_t pos = fp.state_start * 32 + get_or_place(fp, d.d_iv, d.en_iv);
          w.push_back({kWOut, d.d_iv, pos});
          w.push_back({kWIn, d.q, pos});
          w.push_back({kWDff, (uint32_t)di, pos});
        } else {
          size_t si = oe - g.pos.size() - g.dffs.size();
          const Sram &r = g.srams[si];
          uint32_t rd_local = fp.num_writeouts - fp.num_srams + cur_sram;
          uint32_t rd_global = fp.state_start + rd_local;
          w.push_back({kWSram, (uint32_t)si, fp.sram_start + cur_sram * kSramWords});
          for (uint32_t k = 0; k < 32; ++k) {
            uint32_t d = r.rd_pin[k];
            if (!d) continue;
            w.push_back({kWIn, d, rd_global * 32 + k});
            w.push_back({kWOut, d << 1, rd_global * 32 + k});
            place_ck(fp, rd_local * 32 + k, r.r_en_iv, 0);
          }
          uint32_t st = cur_sram * 128, perm, inv, set0;
          for (int k = 0; k < kSramAddrBits; ++k) {
            query_perm(fp, r.r_addr_iv[k], perm, inv, set0);
            place(fp.sd, st + k, perm, inv, set0);
            query_perm(fp, r.w_addr_iv[k], perm, inv, set0);
            place(fp.sd, st + 16 + k, perm, inv, set0);
          }
          for (int k = 0; k < 32; ++k) {
            query_perm(fp, r.w_en_iv[k], perm, inv, set0);
            place(fp.sd, st + 32 + k, perm, inv, set0);
            query_perm(fp, r.w_data_iv[k], perm, inv, set0);
            place(fp.sd, st + 64 + k, perm, inv, set0);
          }
          ++cur_sram;
        }
      }
      if ((fp.cnt_dup + 31) / 32 != fp.num_dups)
        throw std::runtime_error("flatten: duplicate count mismatch");
    } catch (const std::exception &e) {
      err[i] = e.what();
    }
    rethrow_first(err);
    for (const std::vector<MapWrite> &w : wr)
      for (const MapWrite &x : w) {
        std::vector<uint32_t> &to = x.kind == kWStaged ? s.staged_io_map
                                    : x.kind == kWOut  ? s.output_map
                                    : x.kind == kWIn   ? s.input_map
                                    : x.kind == kWDff  ? s.dff_data_pos
                                                       : s.sram_offsets;
        if (x.key >= to.size()) throw std::runtime_error("flatten: map key out of range");
        to[x.key] = x.val;
      }
  }
  s.reg_io_state_size = sum_state;
  s.sram_storage_size = sum_sram;

  // X-capability per partition (reference rule src/flatten.rs:1424-1498 plus
  // two closures that keep the template X of non-X partitions out of every
  // X-capable read and out of primary outputs).
  std::vector<std::vector<uint8_t>> px(c.stages.size());
  if (opt.xprop) {
    std::vector<uint8_t> xp = g.x_capable_pins();
    // word ranges and reads per partition
    struct PI_ {
      uint32_t m, i, lo, hi;
      std::vector<uint32_t> reads;  // state words read
      bool po;
    };
    std::vector<PI_> all;
    std::vector<int32_t> word_owner(sum_state, -1);
    for (size_t m = 0; m < c.stages.size(); ++m) {
      px[m].assign(c.stages[m].parts.size(), 0);
      for (size_t i = 0; i < c.stages[m].parts.size(); ++i) {
        const Partition &pt = c.stages[m].parts[i];
        const FlatPart &fp = fparts[m][i];
        PI_ r{(uint32_t)m, (uint32_t)i, fp.state_start,
              fp.state_start + fp.num_writeouts, {}, false};
        bool x = false;
        for (const BoomStage &bs : pt.stages)
          for (int h = 0; h <= kBoomStages && !x; ++h)
            for (uint32_t pin : bs.hier[h])
              if (pin != kNone && xp[pin]) {
                x = true;
                break;
              }
        for (uint32_t pin : pt.stages[0].hier[0]) {
          if (pin == kNone) continue;
          bool staged;
          r.reads.push_back(read_pos(pin, staged) >> 5);
        }
        for (uint32_t e : pt.endpoints)
          if (e >= c.stages[m].staged_out.size() &&
              c.stages[m].orig_endpoints[e - c.stages[m].staged_out.size()] < g.pos.size())
            r.po = true;
        px[m][i] = x || r.po;
        for (uint32_t w = r.lo; w < r.hi; ++w) word_owner[w] = (int32_t)all.size();
        all.push_back(std::move(r));
      }
    }
    for (bool changed = true; changed;) {
      changed = false;
      for (auto &r : all) {
        for (uint32_t w : r.reads) {
          int32_t o = w < word_owner.size() ? word_owner[w] : -1;
          if (o < 0) continue;
          auto &ro = all[o];
          uint8_t &me = px[r.m][r.i], &them = px[ro.m][ro.i];
          if (them && !me) {
            me = 1;
            changed = true;
          }
          if (me && !them) {
            them = 1;
            changed = true;
          }
        }
      }
    }
    s.xprop = true;
    s.xprop_state_offset = sum_state;
  }

  // build_script per partition (src/flatten.rs:863-1130): the sections of
  // all stages' partitions are independent (the maps are only read), so
  // they are built in one parallel loop, then assembled in block order
  std::vector<std::pair<uint32_t, uint32_t>> work;  // (major stage, partition)
  std::vector<size_t> first(c.stages.size(), 0);    // work index of (m, 0)
  for (size_t m = 0; m < c.stages.size(); ++m) {
    first[m] = work.size();
    for (size_t i = 0; i < c.stages[m].parts.size(); ++i)
      work.push_back({(uint32_t)m, (uint32_t)i});
  }
  std::vector<std::vector<uint32_t>> scripts(work.size());
  std::vector<std::vector<std::pair<uint64_t, std::vector<uint32_t>>>> patches(work.size());
  std::vector<std::string> berr(work.size());
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t wi = 0; wi < work.size(); ++wi) try {
      const size_t m = work[wi].first, i = work[wi].second;
      const Partition &pt = c.stages[m].parts[i];
      const FlatPart &fp = fparts[m][i];
      std::vector<uint32_t> &sc = scripts[wi];
      sc.reserve(256 + 5120 * (pt.stages.size() + 2) + 1024);
      sc.push_back((uint32_t)pt.stages.size());
      sc.push_back(0);
      sc.push_back(fp.num_writeouts);
      sc.push_back(fp.state_start);
      sc.push_back(fp.num_srams);
      sc.push_back(fp.sram_start);
      sc.push_back(0);
      sc.push_back(fp.num_dups);
      sc.push_back(opt.xprop ? (uint32_t)px[m][i] : 0u);
      sc.push_back(opt.xprop ? s.xprop_state_offset : 0u);
      while (sc.size() < 128) sc.push_back(0);
      uint32_t last = kNone;
      for (size_t j = 0; j < pt.stages.size(); ++j)
        for (uint16_t wo : pt.stages[j].write_outs) {
          uint32_t cw = (uint32_t)j << 8 | wo;
          if (last == kNone) {
            last = cw;
          } else {
            sc.push_back(last | cw << 16);
            last = kNone;
          }
        }
      if (last != kNone) sc.push_back(last | 0xFFFFu << 16);
      while (sc.size() < 256) sc.push_back(0xFFFFFFFFu);
      if (sc.size() != 256) throw std::runtime_error("flatten: too many writeout hooks");
      // global reads
      // (word | flag, mask) per word read, in word order
      std::vector<std::pair<uint32_t, uint32_t>> ent;
      ent.reserve(kWidth);
      for (uint32_t pin : pt.stages[0].hier[0]) {
        if (pin == kNone) continue;
        bool staged;
        const uint32_t pos = read_pos(pin, staged);
        ent.push_back({(pos >> 5) | (staged ? 0x80000000u : 0u), 1u << (pos & 31)});
      }
      std::sort(ent.begin(), ent.end());
      {
        size_t o = 0;
        for (size_t k = 0; k < ent.size(); ++k)
          if (o && ent[o - 1].first == ent[k].first)
            ent[o - 1].second |= ent[k].second;
          else
            ent[o++] = ent[k];
        ent.resize(o);
      }
      std::vector<std::vector<std::pair<uint32_t, uint32_t>>> th;
      pack_global_reads(ent, th);
      uint32_t rounds = 0;
      for (auto &v : th) rounds = std::max(rounds, (uint32_t)v.size());
      sc[6] = rounds;
      size_t gstart = sc.size();
      sc.resize(gstart + 2 * (size_t)rounds * kThreads, 0);
      FlatMap<uint64_t, uint16_t> local;  // (flag,pos) -> local bit
      for (uint32_t t = 0; t < kThreads; ++t) {
        for (size_t r = 0; r < th[t].size(); ++r) {
          sc[gstart + kThreads * 2 * r + 2 * t] = th[t][r].first;
          sc[gstart + kThreads * 2 * r + 2 * t + 1] = th[t][r].second;
        }
        uint32_t bit_id = 0;
        for (size_t r = th[t].size(); r-- > 0;) {
          uint32_t idx = th[t][r].first, mask = th[t][r].second;
          uint64_t flag = idx >> 31;
          for (int k = 31; k >= 0; --k)
            if (mask >> k & 1) {
              uint64_t key = flag << 40 | (uint64_t)((idx & 0x7FFFFFFFu) * 32ull + k);
              local.set(key, (uint16_t)(t * 32 + bit_id));
              ++bit_id;
            }
        }
      }
      FlatMap<uint32_t, uint16_t> last_pos;  // pin -> position in the layer
      auto pos_of = [&last_pos](uint32_t pin) -> uint16_t {
        const uint16_t *v = last_pos.find(pin);
        if (!v) throw std::runtime_error("flatten: a tree operand has no position");
        return *v;
      };
      for (uint32_t pin : pt.stages[0].hier[0]) {
        if (pin == kNone || last_pos.find(pin)) continue;
        bool staged;
        const uint32_t pos = read_pos(pin, staged);
        const uint16_t *lb = local.find((uint64_t)staged << 40 | pos);
        if (!lb) throw std::runtime_error("flatten: a read bit has no local position");
        last_pos.set(pin, *lb);
      }
      for (size_t j = 0; j < pt.stages.size(); ++j) {
        const BoomStage &bs = pt.stages[j];
        std::vector<uint16_t> perm(kWidth, 0);
        for (uint32_t i2 = 0; i2 < kWidth; ++i2) {
          uint32_t pin = bs.hier[0][i2];
          if (pin != kNone) perm[i2] = pos_of(pin);
        }
        std::vector<uint32_t> xa(kThreads, 0), xb(kThreads, 0), ob(kThreads, 0);
        std::vector<std::vector<uint32_t>> tpins(kThreads);
        for (int h = 1; h <= kBoomStages; ++h) {
          if (bs.hier[h].empty()) continue;
          uint32_t len = kWidth >> h;
          for (uint32_t q = 0; q < len; ++q) {
            uint32_t o = bs.hier[h][q];
            if (o == kNone) continue;
            uint32_t pa = bs.at(h - 1, q), pb = bs.at(h - 1, q + len);
            uint32_t pos = len + q;
            tpins[pos >> 5].push_back(o);
            if (o == pa) {
              ob[pos >> 5] |= 1u << (pos & 31);
              continue;
            }
            if (g.drv[o] != kAnd) throw std::runtime_error("flatten: bad tree node");
            uint32_t la = g.a[o], lb = g.b[o];
            if ((la >> 1) == pa && (lb >> 1) == pb) {
            } else if ((lb >> 1) == pa && (la >> 1) == pb) {
              std::swap(la, lb);
            } else {
              throw std::runtime_error("flatten: tree operands misplaced");
            }
            if (la & 1) xa[pos >> 5] |= 1u << (pos & 31);
            if (lb & 1) xb[pos >> 5] |= 1u << (pos & 31);
          }
        }
        for (uint32_t k = 0; k < 4; ++k)
          for (uint32_t t = 0; t < kThreads; ++t) {
            uint32_t i2 = t * 32 + k * 8;
            for (uint32_t q = 0; q < 8; q += 2)
              sc.push_back((uint32_t)perm[i2 + q] | (uint32_t)perm[i2 + q + 1] << 16);
          }
        for (uint32_t t = 0; t < kThreads; ++t) {
          sc.push_back(xa[t]);
          sc.push_back(xb[t]);
          sc.push_back(ob[t]);
          if (!tpins[t].empty()) patches[wi].push_back({sc.size(), tpins[t]});
          sc.push_back(0);
        }
        last_pos.clear();
        for (int h = kBoomStages; h >= 1; --h) {
          if (bs.hier[h].empty()) continue;
          uint32_t len = kWidth >> h;
          for (uint32_t q = 0; q < len; ++q) {
            uint32_t o = bs.hier[h][q];
            if (o != kNone) last_pos.set(o, (uint16_t)(len + q));
          }
        }
      }
      auto emit = [&](const PermFlags &pf, bool third) {
        for (uint32_t k = 0; k < 4; ++k)
          for (uint32_t t = 0; t < kThreads; ++t) {
            uint32_t i2 = t * 32 + k * 8;
            for (uint32_t q = 0; q < 8; q += 2)
              sc.push_back((uint32_t)pf.perm[i2 + q] | (uint32_t)pf.perm[i2 + q + 1] << 16);
          }
        for (uint32_t t = 0; t < kThreads; ++t) {
          sc.push_back(pf.inv[t]);
          sc.push_back(pf.set0[t]);
          sc.push_back(third ? pf.extra[t] : 0);
          sc.push_back(0);
        }
      };
      emit(fp.sd, false);
      emit(fp.ck, true);
      // the partition's tables are not read again: freed on this thread
      { FlatPart done = std::move(fparts[m][i]); }
    } catch (const std::exception &e) {
      berr[wi] = e.what();
    }
  rethrow_first(berr);
  // every block's partitions in their order (a one-header dummy for an
  // empty block), the block's last header flagged: offsets in order, then
  // the copies in parallel
  uint64_t total = 0;
  for (size_t m = 0; m < c.stages.size(); ++m)
    for (uint32_t blk = 0; blk < nb; ++blk) {
      s.blocks_start.push_back(total);
      const auto &lst = blocks_parts[m][blk];
      if (lst.empty()) {
        total += kThreads;
        continue;
      }
      for (uint32_t pid : lst) {
        const size_t wi = first[m] + pid;
        for (auto &pp : patches[wi])
          s.delay_patch.push_back({total + pp.first, std::move(pp.second)});
        s.parts.push_back({(uint32_t)m, blk, pid, total, (uint64_t)scripts[wi].size()});
        if (opt.xprop) s.part_x.push_back(px[m][pid]);
        total += scripts[wi].size();
      }
    }
  s.blocks_start.push_back(total);
  s.blocks_data.assign(total + kThreads * 8, 0);
  const size_t nblk = c.stages.size() * (size_t)nb;
#pragma omp parallel for schedule(dynamic, 4)
  for (size_t q = 0; q < nblk; ++q) {
    const auto &lst = blocks_parts[q / nb][q % nb];
    uint32_t *dst = s.blocks_data.data() + s.blocks_start[q], *last = dst;
    for (uint32_t pid : lst) {
      std::vector<uint32_t> &sc = scripts[first[q / nb] + pid];
      last = dst;
      std::memcpy(dst, sc.data(), sc.size() * sizeof(uint32_t));
      dst += sc.size();
      std::vector<uint32_t>().swap(sc);
    }
    last[1] = 1;  // also the dummy header's "last" word
  }
  (void)c;
}

size_t inject_delays(uint32_t *bd, size_t n_bd, const uint64_t *offsets,
                     const uint32_t *pin_start, const uint32_t *pins, size_t n_patch,
                     const uint16_t *rise_ps, const uint16_t *fall_ps, size_t n_pins) {
  if (n_pins == 0 || n_patch == 0) return 0;  // no delays or no slots: untouched
  size_t patched = 0;
  for (size_t k = 0; k < n_patch; ++k) {
    uint32_t mx = 0;
    for (uint32_t j = pin_start[k]; j < pin_start[k + 1]; ++j) {
      const uint32_t pin = pins[j];
      if (pin < n_pins) mx = std::max<uint32_t>(mx, std::max(rise_ps[pin], fall_ps[pin]));
    }
    const uint32_t d = std::min<uint32_t>(mx, 65535);
    if (offsets[k] < n_bd) {
      bd[offsets[k]] = d;
      patched += d > 0;
    }
  }
  return patched;
}

void inject_timing(const Aig &g, Script &s) {
  if (g.delay_ps.empty()) return;
  std::vector<uint64_t> off;
  std::vector<uint32_t> start{0}, pins;
  for (auto &pp : s.delay_patch) {
    off.push_back(pp.first);
    pins.insert(pins.end(), pp.second.begin(), pp.second.end());
    start.push_back((uint32_t)pins.size());
  }
  inject_delays(s.blocks_data.data(), s.blocks_data.size(), off.data(), start.data(),
                pins.data(), off.size(), g.delay_ps.data(), g.delay_ps.data(),
                g.delay_ps.size());
  s.timing = true;
}

void effective_setup_hold(const DffConstraint &c, uint16_t *setup, uint16_t *hold) {
  const int32_t su = (int32_t)c.setup_ps - (int32_t)c.clock_arrival_ps;
  const int32_t ho = (int32_t)c.hold_ps + (int32_t)c.clock_arrival_ps;
  *setup = (uint16_t)std::min<int32_t>(std::max<int32_t>(su, 0), 65535);
  *hold = (uint16_t)std::min<int32_t>(std::max<int32_t>(ho, 0), 65535);
}

uint32_t constraint_words(uint32_t num_words, uint64_t clock_period_ps,
                          const DffConstraint *c, size_t n, uint32_t *out) {
  std::fill(out, out + num_words, 0u);
  for (size_t i = 0; i < n; ++i) {
    if (c[i].data_state_pos == kNone) continue;
    const uint32_t w = c[i].data_state_pos / 32;
    if (w >= num_words) continue;
    uint16_t su, ho;
    effective_setup_hold(c[i], &su, &ho);
    // per word the tightest setup and hold; 0 means "none yet"
    const uint16_t os = (uint16_t)(out[w] >> 16), oh = (uint16_t)(out[w] & 0xFFFF);
    const uint16_t ns = os == 0 ? su : std::min(os, su);
    const uint16_t nh = oh == 0 ? ho : std::min(oh, ho);
    out[w] = (uint32_t)ns << 16 | nh;
  }
  return (uint32_t)std::min<uint64_t>(clock_period_ps, 0xFFFFFFFFull);
}

std::vector<uint32_t> constraint_buffer(const Script &s, uint32_t clock_ps,
                                        uint16_t setup_ps, uint16_t hold_ps) {
  std::vector<DffConstraint> cs;
  cs.reserve(s.dff_data_pos.size());
  for (uint32_t pos : s.dff_data_pos) cs.push_back({setup_ps, hold_ps, 0, pos});
  std::vector<uint32_t> buf(1 + s.reg_io_state_size, 0);
  buf[0] = constraint_words(s.reg_io_state_size, clock_ps, cs.data(), cs.size(), buf.data() + 1);
  return buf;
}

}  // namespace jq
