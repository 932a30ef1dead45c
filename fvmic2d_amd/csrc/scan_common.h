// One-pass exclusive scan of the bucket counts (decoupled look-back): the
// per-tile body, shared by the single-instance sort (sort.hip) and the batched
// ensemble's persistent step (ensemble.hip).
#pragma once
#include "fvm_internal.h"

// exclusive scan of count[0..nk) into start[0..nk], start[nk] = total
#define SCAN_BS 256
#define SCAN_IPT 8
#define SCAN_TILE (SCAN_BS * SCAN_IPT)

// ---- one-pass scan (decoupled look-back) --------------------------------------------
// Each block takes the next tile number from a ticket, so every tile it looks
// back at was started earlier and will publish.  A tile publishes its aggregate
// (status 1) and then its inclusive prefix (status 2) as one 64-bit word
// (status << 32 | value); a successor's first wavefront reads its predecessors'
// words 64 at a time and sums them back to the nearest inclusive prefix.  The
// wait is bounded: after SCAN_SPIN_MAX polls the tile stops waiting, sums the
// counts before it directly (exact, only slower) and sets the sticky
// diagnostic flag *fail.  state[0] = ticket, state[2 + t] = word of tile t.
#define SCAN_SPIN_MAX (1 << 20)

// One tile of the scan, by a block of SCAN_BS threads.  `tile` must be handed
// out in increasing order (a ticket), so that every predecessor it looks back
// at was started earlier and will publish.
__device__ __forceinline__ void scan_tile(int32_t tile, const int32_t *__restrict__ in, int32_t *__restrict__ out,
                                          unsigned long long *__restrict__ state, int64_t n,
                                          int32_t *__restrict__ fail) {
  __shared__ int32_t s_wave[SCAN_BS / 64];
  __shared__ int32_t s_prefix;
  const int64_t base = (int64_t)tile * SCAN_TILE + (int64_t)threadIdx.x * SCAN_IPT;
  int32_t v[SCAN_IPT];
  int32_t t = 0;
#pragma unroll
  for (int i = 0; i < SCAN_IPT; i++) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    t += v[i];
  }
  // inclusive scan of the thread sums: inside each wavefront by shuffles, then
  // across the block's wavefronts through LDS
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int32_t incl = t;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == 63) s_wave[wv] = incl;
  __syncthreads();
  int32_t before = 0, agg = 0;
#pragma unroll
  for (int k = 0; k < SCAN_BS / 64; k++) {
    const int32_t wt = s_wave[k];
    before += k < wv ? wt : 0;
    agg += wt;
  }
  unsigned long long *word = state + 2;
  // status and value travel in one 64-bit word, so relaxed agent-scope atomics
  // suffice: a reader sees a whole word and reads nothing else the writer made
  // (an acquire would add a cache invalidation to every poll)
  if (threadIdx.x == 0)
    __hip_atomic_store(&word[tile], ((tile == 0 ? 2ull : 1ull) << 32) | (uint32_t)agg, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  if (tile > 0 && threadIdx.x < 64) {
    // wave 0 looks back over a window of 64 predecessors at a time: lane i
    // reads tile (j - i); the window ends at its nearest inclusive prefix
    const int lane = threadIdx.x;
    int32_t prefix = 0;
    int64_t j = tile - 1;
    bool failed = false;
    for (;;) {
      const int64_t idx = j - lane;
      unsigned long long w = idx >= 0 ? 0ull : (2ull << 32);  // before tile 0: inclusive 0
      int spins = 0;
      for (;;) {
        if (idx >= 0 && (w >> 32) == 0) w = __hip_atomic_load(&word[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__ballot((w >> 32) == 0) == 0) break;
        if (++spins > SCAN_SPIN_MAX) {
          failed = true;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (failed) break;
      const unsigned long long incl = __ballot((w >> 32) == 2);
      const int first = incl ? __ffsll((long long)incl) - 1 : 63;  // nearest inclusive lane
      int32_t v = lane <= first ? (int32_t)(uint32_t)w : 0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      prefix += v;
      if (incl) break;
      j -= 64;
    }
    if (failed) {
      // the bounded wait ran out: form the exclusive prefix from the counts
      // themselves (complete when a tile starts: written before the launch,
      // or by the work a persistent step's tile waited for), so the scan stays
      // exact; the flag only records that the slow path was taken
      int32_t s = 0;
      const int64_t end = (int64_t)tile * SCAN_TILE;
      for (int64_t k = lane; k < end; k += 64) s += in[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      prefix = s;
    }
    if (lane == 0) {
      if (failed) atomicOr(fail, 1);
      __hip_atomic_store(&word[tile], (2ull << 32) | (uint32_t)(prefix + agg), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      s_prefix = prefix;
    }
  } else if (threadIdx.x == 0) {
    s_prefix = 0;
  }
  __syncthreads();
  int32_t run = s_prefix + before + incl - t;  // exclusive prefix of this thread
#pragma unroll
  for (int i = 0; i < SCAN_IPT; i++) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  // the thread holding element n-1 ends with the inclusive prefix of all n: the total
  if (base <= n - 1 && n - 1 < base + SCAN_IPT) out[n] = run;
}
