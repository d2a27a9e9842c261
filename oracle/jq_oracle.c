/*
 * jq_oracle.c -- CPU oracle (test infrastructure only, see jq_oracle.h).
 *
 * Restates, in plain C:
 *   - src/sim/cpu_reference.rs:17-289   simulate_block_v1        (2-state)
 *   - src/sim/cpu_reference.rs:299-674  simulate_block_v1_xprop  (X lane)
 *   - src/sim/cpu_reference.rs:679-790  sanity_check_cpu[_xprop] (run loop)
 *   - csrc/kernel_v1_impl.cuh:32-626    kernel-only semantics: per-partition
 *     X lane inside one snapshot, u16 arrival tracking, setup/hold events.
 * plus an independent direct AIG simulator (no script involved) used as a
 * closed-form cross-check.
 *
 * One rule is pinned here on purpose where the reference backends disagree:
 * the per-word writeout arrival is reset to 0 at the start of every
 * partition (the CUDA/HIP kernel zeroes it once per launch,
 * kernel_v1_impl.cuh:589, the Metal kernel once per dispatch,
 * kernel_v1.metal:606).  This makes violation lists independent of how
 * partitions are packed into blocks.
 */
#include "jq_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint32_t u32;
typedef uint16_t u16;

/* Gather 32 bits per word from a 256-word (8192-bit) source through a
 * packed u16 permutation section of 4x1024 words
 * (flatten.rs:1027-1034, cpu_reference.rs:86-101). */
static inline u32 perm_word(const u32 *sec, int i, const u32 *src) {
  u32 r = 0;
  for (int ko = 0; ko < 4; ++ko) {
    for (int ki = 0; ki < 4; ++ki) {
      int k = ko * 4 + ki;
      u32 t = sec[ko * 1024 + i * 4 + ki];
      u32 i1 = t & 0xffffu, i2 = t >> 16;
      r |= ((src[i1 >> 5] >> (i1 & 31)) & 1u) << (k * 2);
      r |= ((src[i2 >> 5] >> (i2 & 31)) & 1u) << (k * 2 + 1);
    }
  }
  return r;
}

/* Same but the source has only `n` valid words; returns -1 on an index past
 * them (the reference indexes a Vec of length n and would panic). */
static inline int perm_word_checked(const u32 *sec, int i, const u32 *src,
                                    u32 n, u32 *out) {
  u32 r = 0;
  for (int ko = 0; ko < 4; ++ko) {
    for (int ki = 0; ki < 4; ++ki) {
      int k = ko * 4 + ki;
      u32 t = sec[ko * 1024 + i * 4 + ki];
      u32 i1 = t & 0xffffu, i2 = t >> 16;
      if ((i1 >> 5) >= n || (i2 >> 5) >= n) return -1;
      r |= ((src[i1 >> 5] >> (i1 & 31)) & 1u) << (k * 2);
      r |= ((src[i2 >> 5] >> (i2 & 31)) & 1u) << (k * 2 + 1);
    }
  }
  *out = r;
  return 0;
}

/* The boomerang AND-with-invert tree (cpu_reference.rs:118-151), value lane
 * only.  h[] holds the shuffled 256 words on entry and the after-layout on
 * exit. */
static void and_tree_v(u32 *h, const u32 *xa, const u32 *xb, const u32 *ob) {
  for (int i = 0; i < 128; ++i) {
    u32 a = h[i], b = h[128 + i];
    h[128 + i] = (a ^ xa[128 + i]) & ((b ^ xb[128 + i]) | ob[128 + i]);
  }
  for (int hi = 1; hi <= 7; ++hi) {
    int hw = 1 << (7 - hi);
    for (int i = 0; i < hw; ++i) {
      u32 a = h[hw * 2 + i], b = h[hw * 3 + i];
      h[hw + i] = (a ^ xa[hw + i]) & ((b ^ xb[hw + i]) | ob[hw + i]);
    }
  }
  u32 v1 = h[1], xora = xa[0], xorb = xb[0], orb = ob[0];
  u32 r8 = ((v1 << 16) ^ xora) & ((v1 ^ xorb) | orb) & 0xffff0000u;
  u32 r9 = ((r8 >> 8) ^ xora) & (((r8 >> 16) ^ xorb) | orb) & 0xff00u;
  u32 r10 = ((r9 >> 4) ^ xora) & (((r9 >> 8) ^ xorb) | orb) & 0xf0u;
  u32 r11 = ((r10 >> 2) ^ xora) & (((r10 >> 4) ^ xorb) | orb) & 0xcu;
  u32 r12 = ((r11 >> 1) ^ xora) & (((r11 >> 2) ^ xorb) | orb) & 0x2u;
  h[0] = r8 | r9 | r10 | r11 | r12;
}

/* 3-valued AND (docs/selective-x-propagation.md:45-58,
 * cpu_reference.rs:434-439). */
static inline void xand(u32 av, u32 ax, u32 bv, u32 bx, u32 xora, u32 xorb,
                        u32 orb, u32 mask, u32 *rv, u32 *rx) {
  u32 ae = av ^ xora;
  u32 be = (bv ^ xorb) | orb;
  u32 bex = bx & ~orb;
  *rv = (ae & be) & mask;
  *rx = ((ax | bex) & (ae | ax) & (be | bex)) & mask;
}

static void and_tree_vx(u32 *hv, u32 *hx, const u32 *xa, const u32 *xb,
                        const u32 *ob) {
  for (int i = 0; i < 128; ++i)
    xand(hv[i], hx[i], hv[128 + i], hx[128 + i], xa[128 + i], xb[128 + i],
         ob[128 + i], ~0u, &hv[128 + i], &hx[128 + i]);
  for (int hi = 1; hi <= 7; ++hi) {
    int hw = 1 << (7 - hi);
    for (int i = 0; i < hw; ++i)
      xand(hv[hw * 2 + i], hx[hw * 2 + i], hv[hw * 3 + i], hx[hw * 3 + i],
           xa[hw + i], xb[hw + i], ob[hw + i], ~0u, &hv[hw + i], &hx[hw + i]);
  }
  u32 v1 = hv[1], x1 = hx[1], xora = xa[0], xorb = xb[0], orb = ob[0];
  u32 r8v, r8x, r9v, r9x, r10v, r10x, r11v, r11x, r12v, r12x;
  xand(v1 << 16, x1 << 16, v1, x1, xora, xorb, orb, 0xffff0000u, &r8v, &r8x);
  xand(r8v >> 8, r8x >> 8, r8v >> 16, r8x >> 16, xora, xorb, orb, 0xff00u,
       &r9v, &r9x);
  xand(r9v >> 4, r9x >> 4, r9v >> 8, r9x >> 8, xora, xorb, orb, 0xf0u, &r10v,
       &r10x);
  xand(r10v >> 2, r10x >> 2, r10v >> 4, r10x >> 4, xora, xorb, orb, 0xcu,
       &r11v, &r11x);
  xand(r11v >> 1, r11x >> 1, r11v >> 2, r11x >> 2, xora, xorb, orb, 0x2u,
       &r12v, &r12x);
  hv[0] = r8v | r9v | r10v | r11v | r12v;
  hx[0] = r8x | r9x | r10x | r11x | r12x;
}

/* Arrival recurrence of one boomerang stage (kernel_v1_impl.cuh:202-350):
 * all 256 arrivals restart at 0, hier0 adds the word's delay, hier1..7 take
 * max(a,b) (or a for a pass-through word) plus the delay, u16 wrap-around;
 * word 0 (bit-level tail) keeps arrival 0. */
static void arrival_tree(u16 *arr, const u32 *ob, const u16 *dly) {
  for (int t = 0; t < 256; ++t) arr[t] = 0;
  for (int t = 128; t < 256; ++t) {
    u16 a = arr[t - 128], b = arr[t];
    u16 base = (ob[t] == 0xFFFFFFFFu) ? a : (a > b ? a : b);
    arr[t] = (u16)(base + dly[t]);
  }
  for (int hi = 1; hi <= 7; ++hi) {
    int hw = 1 << (7 - hi);
    for (int t = hw; t < 2 * hw; ++t) {
      u16 a = arr[t + hw], b = arr[t + 2 * hw];
      u16 base = (ob[t] == 0xFFFFFFFFu) ? a : (a > b ? a : b);
      arr[t] = (u16)(base + dly[t]);
    }
  }
  arr[0] = 0;
}

/* ------------------------------------------------------------------------ */
/* cpu_reference.rs:17-289                                                   */
int jqo_simulate_block_v1(const u32 *script, size_t script_len,
                          const u32 *input_state, u32 *output_state,
                          u32 *sram_data) {
  size_t pi = 0;
  u32 state[256], hier[256], xa[256], xb[256], ob[256];
  u32 wo[256], wo_ck[256], ck[256], perm[256];
  u16 hooks[256];
  for (;;) {
    if (pi + 256 > script_len) return -1;
    u32 num_stages = script[pi], is_last = script[pi + 1];
    u32 num_ios = script[pi + 2], io_offset = script[pi + 3];
    u32 num_srams = script[pi + 4], sram_offset = script[pi + 5];
    u32 rounds = script[pi + 6], num_dups = script[pi + 7];
    for (int i = 0; i < 128; ++i) {
      u32 t = script[pi + 128 + i];
      hooks[2 * i] = (u16)(t & 0xffffu);
      hooks[2 * i + 1] = (u16)(t >> 16);
    }
    if (num_stages == 0) {
      pi += 256;
      break;
    }
    pi += 256;
    if (num_ios > 256 || num_srams * 4 + num_dups > 256 ||
        num_srams + num_dups > num_ios)
      return -1;
    size_t need = pi + (size_t)rounds * 512 + (size_t)num_stages * 5120 +
                  2 * 5120;
    if (need > script_len) return -1;
    memset(wo, 0, sizeof(wo));
    memset(state, 0, sizeof(state));
    for (u32 r = 0; r < rounds; ++r) {
      for (int i = 0; i < 256; ++i) {
        u32 cur = state[i];
        u32 idx = script[pi + 2 * i], mask = script[pi + 2 * i + 1];
        if (mask == 0) continue;
        u32 value = (idx >> 31) ? output_state[idx ^ (1u << 31)]
                                : input_state[idx];
        while (mask) {
          cur <<= 1;
          u32 low = mask & (0u - mask);
          if (value & low) cur |= 1;
          mask ^= low;
        }
        state[i] = cur;
      }
      pi += 512;
    }
    for (u32 bs = 0; bs < num_stages; ++bs) {
      for (int i = 0; i < 256; ++i) hier[i] = perm_word(script + pi, i, state);
      pi += 4096;
      for (int i = 0; i < 256; ++i) {
        xa[i] = script[pi + 4 * i];
        xb[i] = script[pi + 4 * i + 1];
        ob[i] = script[pi + 4 * i + 2];
      }
      pi += 1024;
      and_tree_v(hier, xa, xb, ob);
      memcpy(state, hier, sizeof(state));
      for (int i = 0; i < 256; ++i) {
        u16 h = hooks[i];
        if ((u32)(h >> 8) == bs) {
          if ((u32)i >= num_ios) return -1;
          wo[i] = state[h & 255];
        }
      }
    }
    u32 nperm = num_srams * 4 + num_dups;
    for (u32 i = 0; i < nperm; ++i)
      if (perm_word_checked(script + pi, (int)i, wo, num_ios, &perm[i]))
        return -1;
    pi += 4096;
    for (u32 i = 0; i < nperm; ++i) {
      perm[i] &= ~script[pi + i * 4 + 1];
      perm[i] ^= script[pi + i * 4];
    }
    pi += 1024;
    for (u32 s = 0; s < num_srams; ++s) {
      u32 addrs = perm[s * 4];
      u32 ra = addrs & 0xffffu, wa = addrs >> 16;
      u32 en = perm[s * 4 + 1], data = perm[s * 4 + 2];
      if (ra >= JQO_SRAM_SIZE || wa >= JQO_SRAM_SIZE) return -1;
      u32 *ram = sram_data + sram_offset + (size_t)s * JQO_SRAM_SIZE;
      u32 r = ram[ra], w0 = ram[wa];
      wo[num_ios - num_srams + s] = r;
      ram[wa] = (w0 & ~en) | (data & en);
    }
    for (u32 i = 0; i < num_dups; ++i)
      wo[num_ios - num_srams - num_dups + i] = perm[num_srams * 4 + i];
    memcpy(wo_ck, wo, sizeof(wo));
    for (u32 i = 0; i < num_ios; ++i)
      if (perm_word_checked(script + pi, (int)i, wo_ck, num_ios, &ck[i]))
        return -1;
    pi += 4096;
    for (u32 i = 0; i < num_ios; ++i) {
      ck[i] &= ~script[pi + i * 4 + 1];
      ck[i] ^= script[pi + i * 4];
      wo[i] ^= script[pi + i * 4 + 2];
    }
    pi += 1024;
    for (u32 i = 0; i < num_ios; ++i) {
      u32 old = input_state[io_offset + i];
      output_state[io_offset + i] = (old & ~ck[i]) | (wo[i] & ck[i]);
    }
    if (is_last) break;
  }
  return pi == script_len ? 0 : -1;
}

/* ------------------------------------------------------------------------ */
/* cpu_reference.rs:299-674 (+ kernel per-partition rule when !cpu_quirk)    */
int jqo_simulate_block_v1_xprop(const u32 *script, size_t script_len,
                                const u32 *input_state, u32 *output_state,
                                const u32 *input_xmask, u32 *output_xmask,
                                u32 *sram_data, u32 *sram_xmask,
                                int cpu_quirk) {
  size_t pi = 0;
  u32 sv[256], sx[256], hv[256], hx[256], xa[256], xb[256], ob[256];
  u32 wv[256], wx[256], ckv[256], ckx[256], pv[256], px[256];
  u32 wv_ck[256], wx_ck[256];
  u16 hooks[256];
  for (;;) {
    if (pi + 256 > script_len) return -1;
    u32 num_stages = script[pi];
    u32 is_last = script[pi + 1];
    u32 num_ios = script[pi + 2], io_offset = script[pi + 3];
    u32 num_srams = script[pi + 4], sram_offset = script[pi + 5];
    u32 rounds = script[pi + 6], num_dups = script[pi + 7];
    int is_x = script[pi + 8] != 0;
    for (int i = 0; i < 128; ++i) {
      u32 t = script[pi + 128 + i];
      hooks[2 * i] = (u16)(t & 0xffffu);
      hooks[2 * i + 1] = (u16)(t >> 16);
    }
    if (num_stages == 0) {
      pi += 256;
      break;
    }
    if (!is_x) {
      if (cpu_quirk) {
        /* cpu_reference.rs:335-348: delegate the rest of the block. */
        return jqo_simulate_block_v1(script + pi, script_len - pi,
                                     input_state, output_state, sram_data);
      }
      /* Kernel rule: this partition alone runs 2-state; find its end by
       * walking it with a private copy of its single-partition script. */
      size_t part_len = 256 + (size_t)script[pi + 6] * 512 +
                        (size_t)num_stages * 5120 + 2 * 5120;
      if (pi + part_len > script_len) return -1;
      u32 *tmp = (u32 *)malloc(part_len * sizeof(u32));
      if (!tmp) return -1;
      memcpy(tmp, script + pi, part_len * sizeof(u32));
      tmp[1] = 1; /* stop after this partition */
      int rc = jqo_simulate_block_v1(tmp, part_len, input_state, output_state,
                                     sram_data);
      free(tmp);
      if (rc) return -1;
      pi += part_len;
      if (is_last) break;
      continue;
    }
    pi += 256;
    if (num_ios > 256 || num_srams * 4 + num_dups > 256 ||
        num_srams + num_dups > num_ios)
      return -1;
    if (pi + (size_t)rounds * 512 + (size_t)num_stages * 5120 + 2 * 5120 >
        script_len)
      return -1;
    memset(wv, 0, sizeof(wv));
    memset(wx, 0, sizeof(wx));
    memset(sv, 0, sizeof(sv));
    memset(sx, 0, sizeof(sx));
    for (u32 r = 0; r < rounds; ++r) {
      for (int i = 0; i < 256; ++i) {
        u32 cv = sv[i], cx = sx[i];
        u32 idx = script[pi + 2 * i], mask = script[pi + 2 * i + 1];
        if (mask == 0) continue;
        u32 value, xval;
        if (idx >> 31) {
          value = output_state[idx ^ (1u << 31)];
          xval = output_xmask[idx ^ (1u << 31)];
        } else {
          value = input_state[idx];
          xval = input_xmask[idx];
        }
        while (mask) {
          cv <<= 1;
          cx <<= 1;
          u32 low = mask & (0u - mask);
          if (value & low) cv |= 1;
          if (xval & low) cx |= 1;
          mask ^= low;
        }
        sv[i] = cv;
        sx[i] = cx;
      }
      pi += 512;
    }
    for (u32 bs = 0; bs < num_stages; ++bs) {
      for (int i = 0; i < 256; ++i) {
        hv[i] = perm_word(script + pi, i, sv);
        hx[i] = perm_word(script + pi, i, sx);
      }
      pi += 4096;
      for (int i = 0; i < 256; ++i) {
        xa[i] = script[pi + 4 * i];
        xb[i] = script[pi + 4 * i + 1];
        ob[i] = script[pi + 4 * i + 2];
      }
      pi += 1024;
      and_tree_vx(hv, hx, xa, xb, ob);
      memcpy(sv, hv, sizeof(sv));
      memcpy(sx, hx, sizeof(sx));
      for (int i = 0; i < 256; ++i) {
        u16 h = hooks[i];
        if ((u32)(h >> 8) == bs) {
          if ((u32)i >= num_ios) return -1;
          wv[i] = sv[h & 255];
          wx[i] = sx[h & 255];
        }
      }
    }
    u32 nperm = num_srams * 4 + num_dups;
    for (u32 i = 0; i < nperm; ++i) {
      if (perm_word_checked(script + pi, (int)i, wv, num_ios, &pv[i]) ||
          perm_word_checked(script + pi, (int)i, wx, num_ios, &px[i]))
        return -1;
    }
    pi += 4096;
    for (u32 i = 0; i < nperm; ++i) {
      u32 set0 = script[pi + i * 4 + 1], inv = script[pi + i * 4];
      pv[i] = (pv[i] & ~set0) ^ inv;
      px[i] &= ~set0;
    }
    pi += 1024;
    for (u32 s = 0; s < num_srams; ++s) {
      u32 addrs = pv[s * 4];
      u32 ra = addrs & 0xffffu, wa = addrs >> 16;
      u32 en = pv[s * 4 + 1], data = pv[s * 4 + 2], data_x = px[s * 4 + 2];
      if (ra >= JQO_SRAM_SIZE || wa >= JQO_SRAM_SIZE) return -1;
      size_t st = sram_offset + (size_t)s * JQO_SRAM_SIZE;
      wv[num_ios - num_srams + s] = sram_data[st + ra];
      wx[num_ios - num_srams + s] = sram_xmask[st + ra];
      u32 w0v = sram_data[st + wa], w0x = sram_xmask[st + wa];
      sram_data[st + wa] = (w0v & ~en) | (data & en);
      sram_xmask[st + wa] = (w0x & ~en) | (data_x & en);
    }
    for (u32 i = 0; i < num_dups; ++i) {
      wv[num_ios - num_srams - num_dups + i] = pv[num_srams * 4 + i];
      wx[num_ios - num_srams - num_dups + i] = px[num_srams * 4 + i];
    }
    memcpy(wv_ck, wv, sizeof(wv));
    memcpy(wx_ck, wx, sizeof(wx));
    for (u32 i = 0; i < num_ios; ++i) {
      if (perm_word_checked(script + pi, (int)i, wv_ck, num_ios, &ckv[i]) ||
          perm_word_checked(script + pi, (int)i, wx_ck, num_ios, &ckx[i]))
        return -1;
    }
    pi += 4096;
    for (u32 i = 0; i < num_ios; ++i) {
      u32 set0 = script[pi + i * 4 + 1], inv = script[pi + i * 4];
      u32 dinv = script[pi + i * 4 + 2];
      ckv[i] = (ckv[i] & ~set0) ^ inv;
      ckx[i] &= ~set0;
      wv[i] ^= dinv;
    }
    pi += 1024;
    for (u32 i = 0; i < num_ios; ++i) {
      u32 ov = input_state[io_offset + i], ox = input_xmask[io_offset + i];
      u32 c = ckv[i], cx = ckx[i];
      output_state[io_offset + i] = (ov & ~c) | (wv[i] & c);
      output_xmask[io_offset + i] =
          (ox & ~c & ~cx) | (wx[i] & c & ~cx) | cx;
    }
    if (is_last) break;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Kernel semantics: kernel_v1_impl.cuh:32-563                              */
static void push_event(jqo_event *events, size_t cap, size_t *count, u32 type,
                       u32 word, u32 cycle, int slack, u32 arrival,
                       u32 limit) {
  if (!count) return;
  size_t k = (*count)++;
  if (events && k < cap) {
    jqo_event *e = &events[k];
    e->event_type = type;
    e->message_id = word;
    e->cycle = cycle;
    e->reserved = 0;
    e->data[0] = word;
    e->data[1] = (u32)slack;
    e->data[2] = arrival;
    e->data[3] = limit;
  }
}

int jqo_block_kernel(const u32 *script, size_t script_len,
                     const u32 *input_state, u32 *output_state,
                     u32 *sram_data, u32 *sram_xmask, const u32 *constraints,
                     u32 clock_ps, int32_t arrival_state_offset, u32 cycle,
                     jqo_event *events, size_t ev_cap, size_t *ev_count) {
  size_t pi = 0;
  u32 sv[256], sx[256], xa[256], xb[256], ob[256];
  u32 wv[256], wx[256], ckv[256], ckx[256], pv[256], px[256];
  u16 dly[256], arr[256], wo_arr[256];
  u16 hooks[256];
  for (;;) {
    if (pi + 256 > script_len) return -1;
    const u32 *md = script + pi;
    u32 num_stages = md[0], is_last = md[1], num_ios = md[2];
    u32 io_offset = md[3], num_srams = md[4], sram_offset = md[5];
    u32 rounds = md[6], num_dups = md[7];
    int is_x = md[8] != 0;
    u32 xoff = md[9];
    if (num_stages == 0) {
      pi += 256;
      break;
    }
    for (int i = 0; i < 128; ++i) {
      hooks[2 * i] = (u16)(md[128 + i] & 0xffffu);
      hooks[2 * i + 1] = (u16)(md[128 + i] >> 16);
    }
    pi += 256;
    if (num_ios > 256 || num_srams * 4 + num_dups > 256 ||
        num_srams + num_dups > num_ios)
      return -1;
    if (pi + (size_t)rounds * 512 + (size_t)num_stages * 5120 + 2 * 5120 >
        script_len)
      return -1;
    /* pinned rule: writeout arrivals start at 0 for every partition */
    memset(wo_arr, 0, sizeof(wo_arr));
    memset(wv, 0, sizeof(wv));
    memset(wx, 0, sizeof(wx));
    memset(sv, 0, sizeof(sv));
    memset(sx, 0, sizeof(sx));
    for (u32 r = 0; r < rounds; ++r) {
      for (int i = 0; i < 256; ++i) {
        u32 idx = script[pi + 2 * i], mask = script[pi + 2 * i + 1];
        if (!mask) continue;
        const u32 *base = (idx >> 31) ? output_state : input_state;
        u32 w = idx & 0x7fffffffu;
        u32 value = base[w];
        u32 xval = is_x ? base[xoff + w] : 0;
        u32 cv = sv[i], cx = sx[i];
        while (mask) {
          cv <<= 1;
          cx <<= 1;
          u32 low = mask & (0u - mask);
          if (value & low) cv |= 1;
          if (xval & low) cx |= 1;
          mask ^= low;
        }
        sv[i] = cv;
        sx[i] = cx;
      }
      pi += 512;
    }
    for (u32 bs = 0; bs < num_stages; ++bs) {
      u32 hv[256], hx[256];
      for (int i = 0; i < 256; ++i) {
        hv[i] = perm_word(script + pi, i, sv);
        hx[i] = is_x ? perm_word(script + pi, i, sx) : 0;
      }
      pi += 4096;
      for (int i = 0; i < 256; ++i) {
        xa[i] = script[pi + 4 * i];
        xb[i] = script[pi + 4 * i + 1];
        ob[i] = script[pi + 4 * i + 2];
        dly[i] = (u16)(script[pi + 4 * i + 3] & 0xffffu);
      }
      pi += 1024;
      if (is_x)
        and_tree_vx(hv, hx, xa, xb, ob);
      else
        and_tree_v(hv, xa, xb, ob);
      arrival_tree(arr, ob, dly);
      memcpy(sv, hv, sizeof(sv));
      memcpy(sx, hx, sizeof(sx));
      for (int i = 0; i < 256; ++i) {
        u16 h = hooks[i];
        if ((u32)(h >> 8) == bs) {
          if ((u32)i >= num_ios) return -1;
          wv[i] = sv[h & 255];
          wx[i] = sx[h & 255];
          wo_arr[i] = arr[h & 255];
        }
      }
    }
    u32 nperm = num_srams * 4 + num_dups;
    for (u32 i = 0; i < nperm; ++i) {
      if (perm_word_checked(script + pi, (int)i, wv, num_ios, &pv[i])) return -1;
      px[i] = 0;
      if (is_x && perm_word_checked(script + pi, (int)i, wx, num_ios, &px[i]))
        return -1;
    }
    pi += 4096;
    for (u32 i = 0; i < nperm; ++i) {
      u32 set0 = script[pi + i * 4 + 1], inv = script[pi + i * 4];
      pv[i] = (pv[i] & ~set0) ^ inv;
      px[i] &= ~set0;
    }
    pi += 1024;
    for (u32 s = 0; s < num_srams; ++s) {
      u32 addrs = pv[s * 4];
      u32 ra = addrs & 0xffffu, wa = addrs >> 16;
      u32 en = pv[s * 4 + 1], data = pv[s * 4 + 2], data_x = px[s * 4 + 2];
      if (ra >= JQO_SRAM_SIZE || wa >= JQO_SRAM_SIZE) return -1;
      size_t st = sram_offset + (size_t)s * JQO_SRAM_SIZE;
      u32 r = sram_data[st + ra], w0 = sram_data[st + wa];
      u32 rx = 0, w0x = 0;
      if (is_x && sram_xmask) {
        rx = sram_xmask[st + ra];
        w0x = sram_xmask[st + wa];
      }
      wv[num_ios - num_srams + s] = r;
      sram_data[st + wa] = (w0 & ~en) | (data & en);
      if (is_x) {
        wx[num_ios - num_srams + s] = rx;
        if (sram_xmask) sram_xmask[st + wa] = (w0x & ~en) | (data_x & en);
      }
    }
    for (u32 i = 0; i < num_dups; ++i) {
      wv[num_ios - num_srams - num_dups + i] = pv[num_srams * 4 + i];
      if (is_x) wx[num_ios - num_srams - num_dups + i] = px[num_srams * 4 + i];
    }
    for (u32 i = 0; i < num_ios; ++i) {
      if (perm_word_checked(script + pi, (int)i, wv, num_ios, &ckv[i]))
        return -1;
      ckx[i] = 0;
      if (is_x && perm_word_checked(script + pi, (int)i, wx, num_ios, &ckx[i]))
        return -1;
    }
    pi += 4096;
    for (u32 i = 0; i < num_ios; ++i) {
      u32 set0 = script[pi + i * 4 + 1], inv = script[pi + i * 4];
      u32 dinv = script[pi + i * 4 + 2];
      ckv[i] = (ckv[i] & ~set0) ^ inv;
      ckx[i] &= ~set0;
      wv[i] ^= dinv;
    }
    pi += 1024;
    for (u32 i = 0; i < num_ios; ++i) {
      u32 w = io_offset + i;
      u32 c = ckv[i], cx = ckx[i];
      output_state[w] = (input_state[w] & ~c) | (wv[i] & c);
      if (is_x) {
        u32 ox = input_state[xoff + w];
        output_state[xoff + w] = (ox & ~c & ~cx) | (wx[i] & c & ~cx) | cx;
      }
      if (arrival_state_offset != 0)
        output_state[(u32)arrival_state_offset + w] = (u32)wo_arr[i];
      if (c != 0 && constraints) {
        u32 cons = constraints[w];
        if (cons != 0) {
          u16 setup = (u16)(cons >> 16), hold = (u16)(cons & 0xffffu);
          u16 a = wo_arr[i];
          if (a > 0 && (u32)a + (u32)setup > clock_ps)
            push_event(events, ev_cap, ev_count, JQO_EVENT_SETUP, w, cycle,
                       (int)clock_ps - (int)a - (int)setup, a, setup);
          if (a < hold)
            push_event(events, ev_cap, ev_count, JQO_EVENT_HOLD, w, cycle,
                       (int)a - (int)hold, a, hold);
        }
      }
    }
    if (is_last) break;
  }
  return pi <= script_len ? 0 : -1;
}

int jqo_run(size_t num_blocks, size_t num_major_stages,
            const uint64_t *blocks_start, const u32 *blocks_data,
            u32 *sram_data, u32 *sram_xmask, size_t num_cycles,
            size_t state_size, u32 *states, const u32 *timing_constraints,
            int32_t arrival_state_offset, jqo_event *events, size_t ev_cap,
            size_t *ev_count) {
  u32 clock_ps = timing_constraints ? timing_constraints[0] : 0;
  const u32 *cons = timing_constraints ? timing_constraints + 1 : NULL;
  for (size_t c = 0; c < num_cycles; ++c) {
    const u32 *in = states + c * state_size;
    u32 *out = states + (c + 1) * state_size;
    for (size_t s = 0; s < num_major_stages; ++s) {
      for (size_t b = 0; b < num_blocks; ++b) {
        uint64_t st = blocks_start[s * num_blocks + b];
        uint64_t en = blocks_start[s * num_blocks + b + 1];
        if (jqo_block_kernel(blocks_data + st, (size_t)(en - st), in, out,
                             sram_data, sram_xmask, cons, clock_ps,
                             arrival_state_offset, (u32)c, events, ev_cap,
                             ev_count))
          return -1;
      }
    }
  }
  return 0;
}

/* sanity_check_cpu loop, cpu_reference.rs:745-790. */
int jqo_run_v1(size_t num_blocks, size_t num_major_stages,
               const uint64_t *blocks_start, const u32 *blocks_data,
               u32 *sram_data, size_t num_cycles, size_t state_size,
               u32 *states) {
  for (size_t c = 0; c < num_cycles; ++c) {
    const u32 *in = states + c * state_size;
    u32 *out = states + (c + 1) * state_size;
    for (size_t s = 0; s < num_major_stages; ++s)
      for (size_t b = 0; b < num_blocks; ++b) {
        uint64_t st = blocks_start[s * num_blocks + b];
        uint64_t en = blocks_start[s * num_blocks + b + 1];
        if (jqo_simulate_block_v1(blocks_data + st, (size_t)(en - st), in, out,
                                  sram_data))
          return -1;
      }
  }
  return 0;
}

int jqo_run_v1_parallel(size_t num_blocks, size_t num_major_stages,
                        const uint64_t *blocks_start, const u32 *blocks_data,
                        u32 *sram_data, size_t num_cycles, size_t state_size,
                        u32 *states, int nthreads) {
  int bad = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  for (size_t c = 0; c < num_cycles && !bad; ++c) {
    const u32 *in = states + c * state_size;
    u32 *out = states + (c + 1) * state_size;
    for (size_t s = 0; s < num_major_stages; ++s) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
#endif
      for (long b = 0; b < (long)num_blocks; ++b) {
        uint64_t st = blocks_start[s * num_blocks + b];
        uint64_t en = blocks_start[s * num_blocks + b + 1];
        if (jqo_simulate_block_v1(blocks_data + st, (size_t)(en - st), in, out,
                                  sram_data))
          bad |= 1;
      }
    }
  }
  return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Direct AIG simulator.  Semantics follow the AIG definitions
 * (src/aig.rs:41-191, 2028-2055, 2217-2252): AND of two (pin<<1|inv)
 * operands; DFF Q <- D when the enable is 1; SRAM read-before-write with a
 * 13-bit address and per-bit write enable; read data register updated when
 * the read enable is 1.  X handling uses the kernel's formulas. */
typedef struct {
  uint8_t v, x;
} vx_t;

static inline vx_t lit(const vx_t *val, u32 iv) {
  vx_t r = val[iv >> 1];
  r.v ^= (uint8_t)(iv & 1);
  return r;
}

int jqo_aig_sim(const jqo_aig *g, size_t num_cycles, size_t num_inputs,
                const u32 *in_pins, const uint8_t *in_bits, size_t num_obs,
                const u32 *obs_pins, uint8_t *obs_out, int xprop,
                const uint8_t *dff_init_x, const uint8_t *dff_init_v,
                uint8_t *reg_out) {
  size_t n = (size_t)g->num_pins + 1;
  vx_t *val = (vx_t *)calloc(n, sizeof(vx_t));
  vx_t *reg = (vx_t *)calloc(g->num_dffs ? g->num_dffs : 1, sizeof(vx_t));
  vx_t *rd = (vx_t *)calloc(g->num_srams ? g->num_srams * 32 : 1, sizeof(vx_t));
  u32 *mem = (u32 *)calloc(g->num_srams ? (size_t)g->num_srams * JQO_SRAM_SIZE
                                        : 1,
                           sizeof(u32));
  u32 *memx = (u32 *)malloc((g->num_srams ? (size_t)g->num_srams * JQO_SRAM_SIZE
                                          : 1) *
                            sizeof(u32));
  if (!val || !reg || !rd || !mem || !memx) return -1;
  for (size_t i = 0; i < (g->num_srams ? (size_t)g->num_srams * JQO_SRAM_SIZE
                                       : 1);
       ++i)
    memx[i] = xprop ? 0xFFFFFFFFu : 0;
  for (u32 d = 0; d < g->num_dffs; ++d) {
    reg[d].x = (uint8_t)(xprop && dff_init_x && dff_init_x[d]);
    reg[d].v = (uint8_t)(dff_init_v ? dff_init_v[d] & 1 : 0);
  }
  /* SRAM read-data registers start known, as the reference's X template
   * clears every input_map bit (src/sim/vcd_io.rs:470-474); the memory
   * array itself starts all X (src/bin/jacquard.rs:1293-1295). */
  for (u32 s = 0; s < g->num_srams * 32; ++s) rd[s].x = 0;

  for (size_t c = 0; c < num_cycles; ++c) {
    /* sources */
    for (size_t k = 0; k < num_inputs; ++k) {
      val[in_pins[k]].v = in_bits[c * num_inputs + k] & 1;
      val[in_pins[k]].x = 0;
    }
    for (u32 d = 0; d < g->num_dffs; ++d) val[g->dff_q[d]] = reg[d];
    for (u32 s = 0; s < g->num_srams; ++s)
      for (int k = 0; k < 32; ++k) {
        u32 p = g->sram_rd_pin[s * 32 + k];
        if (p) val[p] = rd[s * 32 + k];
      }
    val[0].v = 0;
    val[0].x = 0;
    /* combinational evaluation in pin (topological) order */
    for (size_t p = 1; p < n; ++p) {
      if (g->drv_type[p] != 1) continue;
      vx_t a = lit(val, g->and_a[p]), b = lit(val, g->and_b[p]);
      val[p].v = a.v & b.v;
      if (xprop)
        val[p].x = (uint8_t)((a.x | b.x) & (a.v | a.x) & (b.v | b.x));
      else
        val[p].x = 0;
    }
    for (size_t k = 0; k < num_obs; ++k) {
      vx_t o = lit(val, obs_pins[k]);
      obs_out[c * num_obs + k] = (uint8_t)(o.v | (o.x << 1));
    }
    /* SRAM: read before write, both from this cycle's values */
    for (u32 s = 0; s < g->num_srams; ++s) {
      u32 ra = 0, wa = 0, en = 0, data = 0, datax = 0;
      for (int k = 0; k < 13; ++k) {
        ra |= (u32)lit(val, g->sram_r_addr_iv[s * 13 + k]).v << k;
        wa |= (u32)lit(val, g->sram_w_addr_iv[s * 13 + k]).v << k;
      }
      for (int k = 0; k < 32; ++k) {
        vx_t e = lit(val, g->sram_w_en_iv[s * 32 + k]);
        vx_t d = lit(val, g->sram_w_data_iv[s * 32 + k]);
        en |= (u32)e.v << k;
        data |= (u32)d.v << k;
        datax |= (u32)d.x << k;
      }
      u32 *m = mem + (size_t)s * JQO_SRAM_SIZE;
      u32 *mx = memx + (size_t)s * JQO_SRAM_SIZE;
      u32 r = m[ra], rx = mx[ra];
      u32 w0 = m[wa], w0x = mx[wa];
      m[wa] = (w0 & ~en) | (data & en);
      mx[wa] = (w0x & ~en) | (datax & en);
      vx_t ren = lit(val, g->sram_r_en_iv[s]);
      for (int k = 0; k < 32; ++k) {
        vx_t *q = &rd[s * 32 + k];
        uint8_t nv = (uint8_t)((r >> k) & 1), nx = (uint8_t)((rx >> k) & 1);
        uint8_t ov = q->v, ox = q->x;
        q->v = (uint8_t)((ov & !ren.v) | (nv & ren.v));
        q->x = xprop ? (uint8_t)((ox & !ren.v & !ren.x) |
                                 (nx & ren.v & !ren.x) | ren.x)
                     : 0;
      }
    }
    for (u32 d = 0; d < g->num_dffs; ++d) {
      vx_t en = lit(val, g->dff_en_iv[d]), dv = lit(val, g->dff_d_iv[d]);
      vx_t o = reg[d];
      reg[d].v = (uint8_t)((o.v & !en.v) | (dv.v & en.v));
      reg[d].x = xprop ? (uint8_t)((o.x & !en.v & !en.x) |
                                   (dv.x & en.v & !en.x) | en.x)
                       : 0;
    }
    if (reg_out)
      for (u32 d = 0; d < g->num_dffs; ++d)
        reg_out[c * g->num_dffs + d] = (uint8_t)(reg[d].v | (reg[d].x << 1));
  }
  free(val);
  free(reg);
  free(rd);
  free(mem);
  free(memx);
  return 0;
}
