// ccp's cell list of a device position array with no allocation: the
// counterpart of ccp_cell_list for callers that cannot allocate device
// memory (code under hipGraph capture, serving loops, harnesses that forbid
// hipMalloc). The caller passes the positions and the four output arrays;
// the only other device memory the kernels use is `scratch` below, a
// file-scope array of fixed size that each GPU holds for the life of the
// process.
//
// Result: sci/soi are the cell keys and bead indices sorted stably by cell,
// beg/end the slot range of every cell with empty cells 0/0 -- the arrays of
// ccp's sugrid::generate_arrays (sugrid.cu:125-133: calc_indexes, a stable
// radix sort of (cell, index), set_cells_empty, find_cells_limits), bit for
// bit the arrays ccp_cell_list returns for the same positions and geometry.
//
// Method (a counting sort that keeps its state in the output arrays):
//   1. end[c] = 0 (a fill kernel, no memset), flag = 0
//   2. key of bead i (ccp's floor(r / csl) + cps/2, clamped into the grid;
//      a bead outside raises the flag) into sci[i], end[key] += 1
//   3. exclusive scan of the counts: per-256-cell block sums in `scratch`,
//      one block scans them, then each block's cells get beg[c] = start and
//      end[c] = start (the scatter cursor)
//   4. soi[atomicAdd(end[key], 1)] = i (arbitrary order within a cell)
//   5. one lane per cell sorts its slots by bead index (the stable order),
//      writes sci[s] = c over them and sets empty cells to 0/0.
// Calls on one GPU share its scratch array: calls that overlap in time on
// one GPU race and must not be issued.

#include "ccpmi.h"
#include "host.hpp"

#include <mutex>

namespace ccpmi {
namespace {

constexpr int NA_BLOCK = 256;
constexpr uint32_t NA_MAX_CELLS = 1u << 24; // cps <= 255
constexpr uint32_t NA_MAX_BLOCKS = NA_MAX_CELLS / NA_BLOCK;
constexpr uint32_t NA_FLAG = NA_MAX_BLOCKS; // word of the outside flag
constexpr uint32_t NA_WORDS = NA_MAX_BLOCKS + 64;
constexpr int NA_MAX_DEVICES = 64;

// 65536 block sums and the flag: 256.25 KiB per GPU
static __device__ __align__(256) uint32_t scratch[NA_WORDS];

__global__ __launch_bounds__(NA_BLOCK) void k_na_clear(uint32_t n_c, uint32_t *end, uint32_t *sc) {
  const uint32_t c = blockIdx.x * NA_BLOCK + threadIdx.x;
  if (c < n_c) end[c] = 0u;
  if (c == 0) sc[NA_FLAG] = 0u;
}

// ccp's cell coordinate floor(x / csl) + cps/2 (calc_indexes,
// sugrid.cu:26-34), clamped into the grid
__device__ __forceinline__ int na_coord(float x, float csl, int cps, bool &outside) {
  int c = (int)floorf(x / csl) + cps / 2;
  if (c < 0) { c = 0; outside = true; }
  if (c >= cps) { c = cps - 1; outside = true; }
  return c;
}

__global__ __launch_bounds__(NA_BLOCK) void k_na_key(const float *xyz, uint32_t n, float csl, int cps, uint32_t *key,
                                                     uint32_t *count, uint32_t *sc) {
  const uint32_t i = blockIdx.x * NA_BLOCK + threadIdx.x;
  if (i >= n) return;
  bool outside = false;
  const int gx = na_coord(xyz[3 * (size_t)i], csl, cps, outside);
  const int gy = na_coord(xyz[3 * (size_t)i + 1], csl, cps, outside);
  const int gz = na_coord(xyz[3 * (size_t)i + 2], csl, cps, outside);
  const uint32_t k = (uint32_t)((gz * cps + gy) * cps + gx);
  key[i] = k;
  atomicAdd(count + k, 1u);
  if (outside) atomicAdd(sc + NA_FLAG, 1u);
}

// exclusive scan of v over the block's 256 lanes, and the block total
__device__ __forceinline__ uint32_t na_block_scan(uint32_t v, uint32_t &total) {
  __shared__ uint32_t s[NA_BLOCK];
  const int t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (int d = 1; d < NA_BLOCK; d <<= 1) {
    const uint32_t a = t >= d ? s[t - d] : 0u;
    __syncthreads();
    s[t] += a;
    __syncthreads();
  }
  total = s[NA_BLOCK - 1];
  const uint32_t incl = s[t];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(NA_BLOCK) void k_na_block_sums(const uint32_t *count, uint32_t n_c, uint32_t *sc) {
  const uint32_t c = blockIdx.x * NA_BLOCK + threadIdx.x;
  uint32_t total;
  na_block_scan(c < n_c ? count[c] : 0u, total);
  if (threadIdx.x == 0) sc[blockIdx.x] = total;
}

// one block: exclusive scan of the nblk block sums in place, each lane
// taking a contiguous run of them
__global__ __launch_bounds__(NA_BLOCK) void k_na_scan_sums(uint32_t nblk, uint32_t *sc) {
  const uint32_t per = (nblk + NA_BLOCK - 1) / NA_BLOCK;
  const uint32_t b0 = threadIdx.x * per, b1 = min(b0 + per, nblk);
  uint32_t part = 0;
  for (uint32_t b = b0; b < b1; ++b) part += sc[b];
  uint32_t total;
  uint32_t run = na_block_scan(part, total);
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t v = sc[b];
    sc[b] = run;
    run += v;
  }
}

__global__ __launch_bounds__(NA_BLOCK) void k_na_starts(uint32_t *count_cursor, uint32_t n_c, const uint32_t *sc,
                                                        uint32_t *beg) {
  const uint32_t c = blockIdx.x * NA_BLOCK + threadIdx.x;
  uint32_t total;
  const uint32_t ex = na_block_scan(c < n_c ? count_cursor[c] : 0u, total);
  if (c < n_c) {
    const uint32_t start = sc[blockIdx.x] + ex;
    beg[c] = start;
    count_cursor[c] = start;
  }
}

__global__ __launch_bounds__(NA_BLOCK) void k_na_scatter(const uint32_t *key, uint32_t n, uint32_t *cursor,
                                                         uint32_t *soi) {
  const uint32_t i = blockIdx.x * NA_BLOCK + threadIdx.x;
  if (i >= n) return;
  soi[atomicAdd(cursor + key[i], 1u)] = i;
}

__global__ __launch_bounds__(NA_BLOCK) void k_na_finish(uint32_t n_c, uint32_t *beg, uint32_t *end, uint32_t *soi,
                                                        uint32_t *sci) {
  const uint32_t c = blockIdx.x * NA_BLOCK + threadIdx.x;
  if (c >= n_c) return;
  const uint32_t s0 = beg[c], s1 = end[c];
  if (s0 == s1) {
    beg[c] = 0u;
    end[c] = 0u;
    return;
  }
  // insertion sort of the cell's bead indices (cells hold a few beads)
  for (uint32_t s = s0 + 1; s < s1; ++s) {
    const uint32_t v = soi[s];
    uint32_t t = s;
    while (t > s0 && soi[t - 1] > v) {
      soi[t] = soi[t - 1];
      --t;
    }
    soi[t] = v;
  }
  for (uint32_t s = s0; s < s1; ++s) sci[s] = c;
}

// the address of this GPU's scratch array, looked up once per device
std::mutex g_na_mutex;
uint32_t *g_na_ptr[NA_MAX_DEVICES] = {};

uint32_t *scratch_ptr() {
  int dev = 0;
  host::hip_check(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEVICES) throw host::Error(CCP_ERR_ARG, "device index beyond the supported range");
  std::lock_guard<std::mutex> lock(g_na_mutex);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    host::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = static_cast<uint32_t *>(p);
  }
  return g_na_ptr[dev];
}

unsigned na_blocks(uint32_t n) { return (n + NA_BLOCK - 1) / NA_BLOCK; }

} // namespace
} // namespace ccpmi

extern "C" int ccp_cell_list_noalloc(const float *xyz, uint32_t n, float csl, uint32_t cps, uint32_t *sci,
                                     uint32_t *soi, uint32_t *beg, uint32_t *end, void *stream) {
  using namespace ccpmi;
  try {
    if (!xyz || !sci || !soi || !beg || !end) throw host::Error(CCP_ERR_ARG, "null array");
    if (n == 0 || n >= (1u << 31)) throw host::Error(CCP_ERR_ARG, "number of objects out of range");
    if (!(csl > 0.0f)) throw host::Error(CCP_ERR_ARG, "cell side out of range");
    if (cps < 1 || (uint64_t)cps * cps * cps >= NA_MAX_CELLS)
      throw host::Error(CCP_ERR_ARG, "cells per side out of range: at most 255 (2^24 cells) without allocation");
    const uint32_t n_c = cps * cps * cps, nblk = na_blocks(n_c);
    hipStream_t s = static_cast<hipStream_t>(stream);
    uint32_t *sc = scratch_ptr();
    k_na_clear<<<nblk, NA_BLOCK, 0, s>>>(n_c, end, sc);
    k_na_key<<<na_blocks(n), NA_BLOCK, 0, s>>>(xyz, n, csl, (int)cps, sci, end, sc);
    k_na_block_sums<<<nblk, NA_BLOCK, 0, s>>>(end, n_c, sc);
    k_na_scan_sums<<<1, NA_BLOCK, 0, s>>>(nblk, sc);
    k_na_starts<<<nblk, NA_BLOCK, 0, s>>>(end, n_c, sc, beg);
    k_na_scatter<<<na_blocks(n), NA_BLOCK, 0, s>>>(sci, n, end, soi);
    k_na_finish<<<nblk, NA_BLOCK, 0, s>>>(n_c, beg, end, soi, sci);
    host::hip_check(hipGetLastError());
    uint32_t outside = 0;
    host::hip_check(hipMemcpyAsync(&outside, sc + NA_FLAG, sizeof(outside), hipMemcpyDeviceToHost, s));
    host::hip_check(hipStreamSynchronize(s));
    if (outside) throw host::Error(CCP_ERR_STATE, "bead left the cell grid");
    return CCP_OK;
  } catch (const host::Error &e) {
    ccpmi::host::set_null_error(e.what());
    return e.code;
  } catch (const std::exception &e) {
    ccpmi::host::set_null_error(e.what());
    return CCP_ERR_ARG;
  }
}
