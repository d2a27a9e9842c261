/*
 * abi_smoke.c -- the reference's launch symbols called from plain C, as the
 * reference's Rust host calls them (src/bin/jacquard.rs:1270-1318): device
 * buffers from the HIP runtime, one call, a synchronise, the snapshots read
 * back.  Only include/jacquard_hip.h and the HIP runtime API are used; no
 * C++ and no torch.
 *
 *   abi_smoke <job.bin> <result.bin>
 *
 * job.bin, little-endian: 11 u64 (magic 0x4a51534d4f4b4531, num_blocks,
 * num_major_stages, num_cycles, state_size, len(blocks_start),
 * len(blocks_data), len(sram), timed flag, len(constraints),
 * arrival_state_offset as i64), then blocks_start (u64), blocks_data,
 * states ((num_cycles+1)*state_size), sram, sram_xmask (len(sram) each) and
 * constraints (u32).  result.bin: states, sram, sram_xmask, then for a
 * timed job the 16 + 1024*32 byte event buffer.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "jacquard_hip.h"

#define MAGIC 0x4a51534d4f4b4531ull

static void die(const char *what) {
  fprintf(stderr, "abi_smoke: %s\n", what);
  exit(1);
}

static void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    fprintf(stderr, "abi_smoke: %s: %s\n", what, hipGetErrorString(e));
    exit(1);
  }
}

static void *read_n(FILE *f, size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p) die("out of memory");
  if (bytes && fread(p, 1, bytes, f) != bytes) die("short job file");
  return p;
}

static void *to_dev(const void *h, size_t bytes) {
  void *d = NULL;
  hip_ok(hipMalloc(&d, bytes ? bytes : 4), "hipMalloc");
  if (bytes) hip_ok(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice), "upload");
  return d;
}

static void write_back(FILE *f, const void *d, size_t bytes) {
  void *h = malloc(bytes ? bytes : 1);
  if (!h) die("out of memory");
  hip_ok(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost), "download");
  if (fwrite(h, 1, bytes, f) != bytes) die("short write");
  free(h);
}

int main(int argc, char **argv) {
  if (argc != 3) die("usage: abi_smoke <job.bin> <result.bin>");
  FILE *f = fopen(argv[1], "rb");
  if (!f) die("cannot open the job file");
  uint64_t h[11];
  if (fread(h, sizeof h, 1, f) != 1 || h[0] != MAGIC) die("bad job header");
  const usize num_blocks = h[1], num_major = h[2], num_cycles = h[3], ss = h[4];
  const size_t n_bs = h[5], n_bd = h[6], n_sram = h[7], n_cons = h[9];
  const int timed = h[8] != 0;
  const i32 arrival_off = (i32)(int64_t)h[10];
  const size_t n_states = (size_t)(num_cycles + 1) * ss;
  usize *bs = read_n(f, n_bs * 8);
  u32 *bd = read_n(f, n_bd * 4);
  u32 *st = read_n(f, n_states * 4);
  u32 *sram = read_n(f, n_sram * 4);
  u32 *sram_x = read_n(f, n_sram * 4);
  u32 *cons = read_n(f, n_cons * 4);
  fclose(f);

  usize *d_bs = to_dev(bs, n_bs * 8);
  u32 *d_bd = to_dev(bd, n_bd * 4);
  u32 *d_st = to_dev(st, n_states * 4);
  u32 *d_sram = to_dev(sram, n_sram * 4);
  u32 *d_sram_x = to_dev(sram_x, n_sram * 4);
  const size_t ev_bytes = 16 + 1024 * 32;
  u8 *d_ev = NULL;
  u32 *d_cons = NULL;
  if (timed) {
    d_cons = to_dev(cons, n_cons * 4);
    hip_ok(hipMalloc((void **)&d_ev, ev_bytes), "hipMalloc events");
    hip_ok(hipMemset(d_ev, 0, ev_bytes), "clear events");
    simulate_v1_noninteractive_timed_hip(num_blocks, num_major, d_bs, d_bd, d_sram,
                                         d_sram_x, num_cycles, ss, d_st, d_cons, d_ev,
                                         arrival_off);
  } else {
    simulate_v1_noninteractive_simple_scan_hip(num_blocks, num_major, d_bs, d_bd, d_sram,
                                               d_sram_x, num_cycles, ss, d_st,
                                               arrival_off);
  }
  hip_ok(hipDeviceSynchronize(), "synchronise");

  FILE *o = fopen(argv[2], "wb");
  if (!o) die("cannot open the result file");
  write_back(o, d_st, n_states * 4);
  write_back(o, d_sram, n_sram * 4);
  write_back(o, d_sram_x, n_sram * 4);
  if (timed) write_back(o, d_ev, ev_bytes);
  fclose(o);

  void *bufs[] = {d_bs, d_bd, d_st, d_sram, d_sram_x, d_cons, d_ev};
  for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; ++i)
    if (bufs[i]) hip_ok(hipFree(bufs[i]), "hipFree");
  free(bs);
  free(bd);
  free(st);
  free(sram);
  free(sram_x);
  free(cons);
  printf("abi_smoke: %llu cycles through the %s symbol\n", (unsigned long long)num_cycles,
         timed ? "timed" : "simple scan");
  return 0;
}
