// Wave64 device primitives shared by the HIP kernels of the engine: lane and
// ballot helpers, relaxed atomics, binary searches, and the membership test of
// the DFS (window_member), which the intersection entry point runs as well.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dsm {
namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t bcast(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t bcast64(uint64_t x) {
  uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T ld_relaxed(T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_release(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas_u32(uint32_t *p, uint32_t expect, uint32_t desire) {
  return __hip_atomic_compare_exchange_strong(p, &expect, desire, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}

// # elements < x in sorted a[0, n)
__device__ __forceinline__ uint32_t lower_bound(const uint32_t *a, uint32_t n, uint32_t x) {
  uint32_t lo = 0;
  while (n > 0) {
    uint32_t half = n >> 1;
    if (a[lo + half] < x) {
      lo += half + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  return lo;
}

__device__ __forceinline__ bool contains(const uint32_t *a, uint32_t n, uint32_t x) {
  if (n == 0) return false;
  uint32_t lo = 0;
  while (n > 1) {
    uint32_t half = n >> 1;
    lo = (a[lo + half] <= x) ? lo + half : lo;
    n -= half;
  }
  return a[lo] == x;
}

// The search window staged in LDS per membership test (entries).
constexpr uint32_t kWin = 256;

// Lower bounds of key0 (lanes 0-31) and key1 (lanes 32-63) in sorted L[0, d),
// by a 32-ary search per half-wave: every round probes 32 evenly spaced
// entries of the remaining range (one load per lane) and keeps the segment
// the key falls in, so a list of d entries takes about log32(d) + 1 dependent
// loads instead of log2(d).
__device__ __forceinline__ void bounds2(const uint32_t *L, uint32_t d, uint32_t key0, uint32_t key1,
                                        uint32_t &r0, uint32_t &r1, uint32_t &loads) {
  const uint32_t lane = lane_id();
  const uint32_t k = lane & 31;
  const uint32_t key = lane < 32 ? key0 : key1;
  const uint64_t halfmask = lane < 32 ? 0x00000000FFFFFFFFull : 0xFFFFFFFF00000000ull;
  uint32_t a = 0, b = d;  // the bound lies in [a, b]
  while (ballot(b > a)) {
    const uint32_t m = b - a;
    const uint32_t step = m > 32 ? (m + 31) / 32 : 1u;
    const uint32_t p = a + k * step;
    bool less = false;
    if (p < b) {
      less = L[p] < key;
      loads++;
    }
    const uint32_t c = (uint32_t)__popcll(ballot(less) & halfmask);
    if (b > a) {
      if (c == 0) {
        b = a;
      } else {
        const uint32_t na = a + (c - 1) * step + 1;
        const uint32_t nb = a + c * step < b ? a + c * step : b;
        a = na;
        b = nb;
      }
    }
  }
  r0 = __builtin_amdgcn_readlane(a, 0);
  r1 = __builtin_amdgcn_readlane(a, 32);
}

__device__ __forceinline__ bool lds_contains(const uint32_t *a, uint32_t n, uint32_t x) {
  if (n == 0) return false;
  uint32_t lo = 0;
  while (n > 1) {
    const uint32_t half = n >> 1;
    lo = (a[lo + half] <= x) ? lo + half : lo;
    n -= half;
  }
  return a[lo] == x;
}

// Membership of every lane's w (lanes with ok) in the sorted list L[0, d),
// the intersection step of the reference's intersect (gpu_match.cu:21-94:
// each lane binary-searches one element of the first set in the second). The
// lanes' values are sorted, so the live ones span [wmin, wmax]: the wave finds
// that window of L with one cooperative 32-ary search (bounds2) and, when it
// holds at most kWin entries, stages it in `win` (LDS, kWin entries) for the
// lanes' searches; a wider window is searched in place. Wave-uniform call;
// returns ok && (w in L).
__device__ __forceinline__ bool window_member(const uint32_t *L, uint32_t d, uint32_t w, bool ok, uint32_t *win,
                                              uint32_t &loads) {
  const uint32_t lane = lane_id();
  const uint64_t alive = ballot(ok);
  if (alive == 0) return false;
  const uint32_t wmin = __shfl(w, (int)__builtin_ctzll(alive));
  const uint32_t wmax = __shfl(w, (int)(63 - __builtin_clzll(alive)));
  uint32_t lo, hi;
  bounds2(L, d, wmin, wmax + 1, lo, hi, loads);
  const uint32_t W = hi - lo;
  if (W <= kWin) {
    for (uint32_t k = lane; k < W; k += 64) win[k] = L[lo + k];
    wave_sync();
    ok = ok && lds_contains(win, W, w);
    wave_sync();
  } else if (ok) {
    loads += 32u - (uint32_t)__builtin_clz(W);
    ok = contains(L + lo, W, w);
  }
  return ok;
}

}  // namespace
}  // namespace dsm
