// Sorted-set intersection on the device with the DFS's own membership test
// (window_member: a cooperative 32-ary bounds search for the window the 64
// lanes' values span, then LDS-staged or in-place binary searches). The
// reference's counterpart is intersect (gpu_match.cu:21-94), which the DFS
// calls with a lane per element of the first set. dsm_intersect_sorted is the
// entry point; it lets the primitive be checked against a set intersection on
// its own, outside the DFS.
#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/dsm.h"
#include "dsm_device.h"
#include "dsm_internal.h"

namespace dsm {
namespace {

constexpr int kIntersectWaves = 4;

// A wave per 64-entry chunk of A: the chunk's members of B, in order, at
// hits[chunk * 64 ..] and their number at cnt[chunk].
__global__ __launch_bounds__(kIntersectWaves * 64) void k_intersect(const uint32_t *A, uint32_t na, const uint32_t *B,
                                                                    uint32_t nb, uint32_t *hits, uint32_t *cnt) {
  __shared__ uint32_t win[kIntersectWaves][kWin];
  const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
  const uint32_t nch = (na + 63) / 64;
  for (uint32_t c = blockIdx.x * kIntersectWaves + wv; c < nch; c += gridDim.x * kIntersectWaves) {
    const uint32_t i = c * 64 + lane;
    const bool in = i < na;
    const uint32_t w = in ? A[i] : 0xFFFFFFFFu;
    uint32_t loads = 0;
    const bool ok = window_member(B, nb, w, in, win[wv], loads);
    const uint64_t m = ballot(ok);
    if (ok) hits[(size_t)c * 64 + mbcnt(m)] = w;
    if (lane == 0) cnt[c] = (uint32_t)__popcll(m);
  }
}

}  // namespace

uint32_t intersect_chunks(uint32_t na) { return (na + 63) / 64; }

hipError_t launch_intersect(const uint32_t *A, uint32_t na, const uint32_t *B, uint32_t nb, uint32_t *hits,
                            uint32_t *cnt, hipStream_t s) {
  const uint32_t nch = intersect_chunks(na);
  if (nch == 0) return hipSuccess;
  const uint32_t blocks = std::min<uint32_t>((nch + kIntersectWaves - 1) / kIntersectWaves, 2048);
  hipLaunchKernelGGL(k_intersect, dim3(blocks), dim3(kIntersectWaves * 64), 0, s, A, na, B, nb, hits, cnt);
  return hipGetLastError();
}

// Inputs of the intersection entry points: strictly ascending, no 0xFFFFFFFF.
bool sorted_set(const uint32_t *a, uint32_t n) {
  for (uint32_t i = 0; i < n; i++) {
    if (a[i] == 0xFFFFFFFFu) return false;
    if (i && a[i] <= a[i - 1]) return false;
  }
  return true;
}

// hits/cnt (device results of k_intersect, copied to the host) -> out.
uint32_t gather_hits(const std::vector<uint32_t> &hits, const std::vector<uint32_t> &cnt, uint32_t *out) {
  uint32_t k = 0;
  for (size_t c = 0; c < cnt.size(); c++)
    for (uint32_t j = 0; j < cnt[c]; j++) out[k++] = hits[c * 64 + j];
  return k;
}

}  // namespace dsm

using namespace dsm;

#define ICHK(expr)                                                                                  \
  do {                                                                                              \
    hipError_t _e = (expr);                                                                         \
    if (_e != hipSuccess) {                                                                         \
      set_global_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
      rc = DSM_EDEVICE;                                                                             \
      goto done;                                                                                    \
    }                                                                                               \
  } while (0)

extern "C" int dsm_intersect_sorted(int device, const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb,
                                    uint32_t *out, uint32_t *num_out) {
  if (!num_out || (na && !a) || (nb && !b) || (na && nb && !out)) {
    set_global_error("dsm_intersect_sorted: null argument");
    return DSM_EINVAL;
  }
  if (!sorted_set(a, na) || !sorted_set(b, nb)) {
    set_global_error("dsm_intersect_sorted: inputs must be strictly ascending and below 0xFFFFFFFF");
    return DSM_EINVAL;
  }
  *num_out = 0;
  if (na == 0 || nb == 0) return DSM_OK;
  int rc = DSM_OK;
  const uint32_t nch = intersect_chunks(na);
  std::vector<uint32_t> hits((size_t)nch * 64), cnt(nch);
  uint32_t *dA = nullptr, *dB = nullptr, *dH = nullptr, *dC = nullptr;
  ICHK(hipSetDevice(device));
  ICHK(hipMalloc(&dA, (size_t)na * 4));
  ICHK(hipMalloc(&dB, (size_t)nb * 4));
  ICHK(hipMalloc(&dH, (size_t)nch * 64 * 4));
  ICHK(hipMalloc(&dC, (size_t)nch * 4));
  ICHK(hipMemcpy(dA, a, (size_t)na * 4, hipMemcpyHostToDevice));
  ICHK(hipMemcpy(dB, b, (size_t)nb * 4, hipMemcpyHostToDevice));
  ICHK(launch_intersect(dA, na, dB, nb, dH, dC, nullptr));
  ICHK(hipMemcpy(cnt.data(), dC, (size_t)nch * 4, hipMemcpyDeviceToHost));
  ICHK(hipMemcpy(hits.data(), dH, (size_t)nch * 64 * 4, hipMemcpyDeviceToHost));
  *num_out = gather_hits(hits, cnt, out);
done:
  for (uint32_t *p : {dA, dB, dH, dC})
    if (p) (void)hipFree(p);
  return rc;
}

// ------------------------------------------------------------------------
// dsm_intersect_sorted_noalloc: the same intersection without any device
// allocation, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops). Its device memory is one file-scope array per GPU,
// `scratch`, of kScratchWords words; inputs, per-chunk results, their prefix
// offsets and the compacted output all live in it. The host reaches it through
// the address hipGetSymbolAddress returns for the current device (the array's
// name is only a host-side stand-in). Every call on one GPU shares that GPU's
// array, so calls that overlap on one GPU race; nothing in it needs clearing
// (every word a call reads was written by that call).

namespace dsm {
namespace {

constexpr size_t kScratchWords = size_t(1) << 22;  // 16 MiB per GPU
static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
uint32_t *g_scratch[kMaxDevices] = {};
std::once_flag g_scratch_once[kMaxDevices];
hipError_t g_scratch_rc[kMaxDevices] = {};

// Exclusive prefix of cnt[0, nch) into off, the total at *total (one block).
__global__ __launch_bounds__(1024) void k_chunk_offsets(const uint32_t *cnt, uint32_t nch, uint32_t *off,
                                                        uint32_t *total) {
  __shared__ uint32_t part[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < nch; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t x = i < nch ? cnt[i] : 0u;
    part[threadIdx.x] = x;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
      const uint32_t y = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
      __syncthreads();
      part[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < nch) off[i] = carry + part[threadIdx.x] - x;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

// Chunk c's members (hits[c * 64 ..], cnt[c] of them) to res[off[c] ..].
__global__ __launch_bounds__(256) void k_chunk_gather(const uint32_t *hits, const uint32_t *cnt,
                                                      const uint32_t *off, uint32_t nch, uint32_t *res) {
  const uint32_t lane = lane_id();
  for (uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < nch; c += (gridDim.x * blockDim.x) >> 6)
    if (lane < cnt[c]) res[off[c] + lane] = hits[(size_t)c * 64 + lane];
}

// The array's address on `device` (current), looked up once per device.
hipError_t scratch_of(int device, uint32_t **out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::call_once(g_scratch_once[device], [device]() {
    void *p = nullptr;
    g_scratch_rc[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (g_scratch_rc[device] == hipSuccess) g_scratch[device] = static_cast<uint32_t *>(p);
  });
  *out = g_scratch[device];
  return g_scratch_rc[device];
}

}  // namespace
}  // namespace dsm

extern "C" int dsm_intersect_sorted_noalloc(int device, const uint32_t *a, uint32_t na, const uint32_t *b,
                                            uint32_t nb, uint32_t *out, uint32_t *num_out) {
  if (!num_out || (na && !a) || (nb && !b) || (na && nb && !out)) {
    set_global_error("dsm_intersect_sorted_noalloc: null argument");
    return DSM_EINVAL;
  }
  if (!sorted_set(a, na) || !sorted_set(b, nb)) {
    set_global_error("dsm_intersect_sorted_noalloc: inputs must be strictly ascending and below 0xFFFFFFFF");
    return DSM_EINVAL;
  }
  *num_out = 0;
  if (na == 0 || nb == 0) return DSM_OK;
  const size_t nch = intersect_chunks(na);
  // layout: a | b | hits (64 per chunk) | cnt | off | total | result
  const size_t need = (size_t)na + nb + nch * 64 + 2 * nch + 1 + std::min(na, nb);
  if (need > kScratchWords) {
    set_global_error("dsm_intersect_sorted_noalloc: %zu scratch words needed, %zu available (na + nb + 66 per "
                     "64 entries of a + min(na, nb) + 1)", need, kScratchWords);
    return DSM_ELIMIT;
  }
  int rc = DSM_OK;
  uint32_t *base = nullptr, total = 0;
  ICHK(hipSetDevice(device));
  ICHK(scratch_of(device, &base));
  {
    uint32_t *dA = base, *dB = dA + na, *dH = dB + nb, *dC = dH + nch * 64, *dO = dC + nch, *dT = dO + nch,
             *dR = dT + 1;
    ICHK(hipMemcpy(dA, a, (size_t)na * 4, hipMemcpyHostToDevice));
    ICHK(hipMemcpy(dB, b, (size_t)nb * 4, hipMemcpyHostToDevice));
    ICHK(launch_intersect(dA, na, dB, nb, dH, dC, nullptr));
    hipLaunchKernelGGL(k_chunk_offsets, dim3(1), dim3(1024), 0, nullptr, dC, (uint32_t)nch, dO, dT);
    ICHK(hipGetLastError());
    hipLaunchKernelGGL(k_chunk_gather, dim3((uint32_t)std::min<size_t>((nch + 3) / 4, 2048)), dim3(256), 0,
                       nullptr, dH, dC, dO, (uint32_t)nch, dR);
    ICHK(hipGetLastError());
    ICHK(hipMemcpy(&total, dT, 4, hipMemcpyDeviceToHost));
    if (total) ICHK(hipMemcpy(out, dR, (size_t)total * 4, hipMemcpyDeviceToHost));
    *num_out = total;
  }
done:
  return rc;
}
