// HIP kernels of the cluster-encoding candidate refinement (gfx950, wave64).
//
// The reference runs one kernel per cluster with a warp per candidate of the
// cluster's core and a cudaDeviceSynchronize after each (encodeKernel,
// mergeKernel, combineMultipleClustersKernel, filter.cu:615-934, driven by
// encode(), filter.cu:936-1177), over vertex-major 32-bit encoding words
// (filter.cu:1546). Here every position is a bitmap row of its own beside the
// LDF/NLF rows, so a query vertex's refined candidates are one row the DFS
// reads directly (no compact pass), and one launch runs the same pass step of
// every standing query:
//
//   k_cluster_init   zero the position rows, load each core's candidate list
//                    (the LDF/NLF row) into its alive row
//   k_cluster_pass   a wave per (query pass, 64 bitmap words): each lane owns
//                    one 64-vertex word of the core's alive row, tests the
//                    pass's core condition for all 64 vertices with one load
//                    per condition row, writes the core and alive words, then
//                    the wave walks every candidate's neighbours (64 at a
//                    time; of a list longer than 64 entries only the label
//                    buckets of the pass's targets, from the label-sorted
//                    copy) and sets or clears the targets' bits; buckets of
//                    more than kClusterChunk entries are queued in chunks
//   k_cluster_heavy  waves claim those chunks from a counter
//
// A pass only reads rows of earlier clusters and only writes its own
// cluster's rows and its core's alive row, and the passes of one step belong
// to different queries, so the work items of a launch are independent. Every
// bit update is an atomic read-modify-write on its 64-bit word (the
// reference's plain read-modify-write of a core word, filter.cu:765, could
// drop a neighbouring bit set concurrently); a load first skips updates that
// would change nothing.
#include <algorithm>

#include "dsm_device.h"
#include "dsm_internal.h"

namespace dsm {
namespace {

constexpr int kClusterWaves = 4;

__device__ __forceinline__ uint64_t readlane64(uint64_t x, uint32_t l) {
  uint32_t lo = __builtin_amdgcn_readlane((uint32_t)x, l);
  uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(x >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ bool has(const uint64_t *bm, uint32_t words, uint32_t row, uint32_t v) {
  return (ld_relaxed(&bm[(size_t)row * words + (v >> 6)]) >> (v & 63)) & 1ull;
}

// Target t of a pass at neighbour w of a candidate (lane-parallel: every lane
// its own w). Encode, merge and new combines set the target's bit when its
// condition rows all hold w; combines into an existing cluster clear it when
// they do not.
__device__ __forceinline__ void update_target(const ClusterArgs &a, const ClusterOp &op, uint32_t kind, uint32_t t,
                                              uint32_t w) {
  bool cond = true;
  const uint32_t o = op.toff[t], c = op.tcnt[t];
  for (uint32_t k = 0; k < c && cond; k++) cond = has(a.bm, a.words, a.crow[o + k], w);
  unsigned long long *word = (unsigned long long *)&a.bm[(size_t)(op.core + 1 + t) * a.words + (w >> 6)];
  const unsigned long long bit = 1ull << (w & 63);
  if (kind == kOpCombineKeep) {
    if (!cond && (ld_relaxed(word) & bit)) atomicAnd(word, ~bit);
  } else if (cond && !(ld_relaxed(word) & bit)) {
    atomicOr(word, bit);
  }
}

// A neighbour of unknown label: the targets of its label (every condition
// row implies the target's label, and a combine into an existing cluster
// holds no bit of another label to clear).
__device__ __forceinline__ void visit_neighbour(const ClusterArgs &a, const ClusterOp &op, uint32_t kind,
                                                uint32_t w) {
  const uint8_t lw = a.g.lab[w];
  for (uint32_t t = 0; t < op.ntarget; t++)
    if (lw == op.tlab[t]) update_target(a, op, kind, t, w);
}

// A neighbour from label bucket li of a label-sorted list.
__device__ __forceinline__ void visit_bucket_entry(const ClusterArgs &a, const ClusterOp &op, uint32_t kind,
                                                   uint32_t li, uint32_t w) {
  for (uint32_t t = 0; t < op.ntarget; t++)
    if (op.tlix[t] == li) update_target(a, op, kind, t, w);
}

// The neighbours of candidate v that the targets of pass j can use. A list of
// more than kLabelSorted entries is read only in the label buckets of the
// targets (its label-sorted copy); a bucket of more than kClusterChunk entries
// is queued in chunks for k_cluster_heavy.
__device__ __forceinline__ void walk_candidate(const ClusterArgs &a, const ClusterOp &op, uint32_t kind, uint32_t j,
                                               uint32_t v) {
  const uint32_t lane = lane_id();
  const uint32_t st = a.g.start[v], dg = a.g.deg[v];
  if (dg <= kLabelSorted) {
    for (uint32_t i = lane; i < dg; i += 64) visit_neighbour(a, op, kind, a.g.pool[st + i]);
    return;
  }
  const uint32_t *row = a.g.loff + (size_t)v * (a.g.nbuckets + 1);
  for (uint32_t k = 0; k < op.nlix; k++) {
    const uint32_t li = op.lix[k];
    const uint32_t b0 = row[li], b1 = row[li + 1];
    if (b1 - b0 > kClusterChunk) {
      const uint32_t nch = (b1 - b0 + kClusterChunk - 1) / kClusterChunk;
      uint32_t first = 0;
      if (lane == 0) first = atomicAdd(&a.ctl[0], nch);
      first = bcast(first);
      for (uint32_t c = lane; c < nch; c += 64)
        if (first + c < a.heavy_cap) a.heavy[first + c] = ClusterHeavy{j, v, b0 + c * kClusterChunk, li};
      continue;
    }
    const uint32_t *L = a.g.lpool + st;
    for (uint32_t i = b0 + lane; i < b1; i += 64) visit_bucket_entry(a, op, kind, li, L[i]);
  }
}

__global__ __launch_bounds__(256) void k_cluster_init(uint64_t *bm, uint32_t words, uint32_t first_row,
                                                      uint32_t nrows, uint64_t *alive, const uint32_t *alive_src,
                                                      uint32_t nalive_src, uint32_t *ctl, uint32_t nctl) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (size_t)gridDim.x * blockDim.x;
  const size_t zero = (size_t)nrows * words, copy = (size_t)nalive_src * words;
  uint64_t *z = bm + (size_t)first_row * words;
  for (size_t i = tid; i < zero; i += stride) z[i] = 0;
  for (size_t i = tid; i < copy; i += stride) {
    const size_t r = i / words, wd = i - r * words;
    alive[(size_t)alive_src[2 * r] * words + wd] = bm[(size_t)alive_src[2 * r + 1] * words + wd];
  }
  for (size_t i = tid; i < nctl; i += stride) ctl[i] = 0;
}

__global__ __launch_bounds__(kClusterWaves * 64) void k_cluster_pass(ClusterArgs a) {
  const uint32_t lane = lane_id();
  const uint32_t wave = bcast((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t groups = (a.words + 63) >> 6;
  const uint64_t units = (uint64_t)a.nops * groups;
  for (uint64_t u = wave; u < units; u += nwaves) {
    const uint32_t j = bcast((uint32_t)(u / groups)), grp = bcast((uint32_t)(u - (uint64_t)j * groups));
    const ClusterOp &op = a.ops[j];
    const uint32_t kind = bcast(op.kind);
    const uint32_t wd = grp * 64 + lane;
    uint64_t aw = 0;
    if (wd < a.words) aw = a.alive[(size_t)op.alive * a.words + wd];
    // core condition, one word per lane: the vertices of this word that hold
    // every check row keep their place in the core's list
    if (aw) {
      uint64_t *core = &a.bm[(size_t)op.core * a.words + wd];
      if (kind == kOpEncode) {
        *core |= aw;
      } else {
        uint64_t pass = aw;
        for (uint32_t k = 0; k < op.ncheck && pass; k++) pass &= a.bm[(size_t)a.crow[op.check_off + k] * a.words + wd];
        if (kind == kOpCombineKeep) *core &= ~(aw & ~pass);
        else *core |= pass;
        a.alive[(size_t)op.alive * a.words + wd] = pass;
      }
    }
    // neighbours of every candidate the pass started with
    uint64_t lanes = ballot(aw != 0);
    while (lanes) {
      const uint32_t l = (uint32_t)__builtin_ctzll(lanes);
      lanes &= lanes - 1;
      uint64_t bits = readlane64(aw, l);
      const uint32_t base = (grp * 64 + l) * 64;
      while (bits) {
        const uint32_t v = base + (uint32_t)__builtin_ctzll(bits);
        bits &= bits - 1;
        walk_candidate(a, op, kind, j, v);
      }
    }
  }
}

__global__ __launch_bounds__(kClusterWaves * 64) void k_cluster_heavy(ClusterArgs a) {
  const uint32_t lane = lane_id();
  for (;;) {
    uint32_t e = 0;
    if (lane == 0) e = atomicAdd(&a.ctl[1], 1u);
    e = bcast(e);
    const uint32_t total = bcast(ld_relaxed(&a.ctl[0]));
    if (e >= total || e >= a.heavy_cap) break;
    const ClusterHeavy h = a.heavy[e];
    const uint32_t j = bcast(h.op), v = bcast(h.v), lo = bcast(h.lo), li = bcast(h.li);
    const ClusterOp &op = a.ops[j];
    const uint32_t kind = bcast(op.kind);
    const uint32_t st = a.g.start[v];
    const uint32_t b1 = a.g.loff[(size_t)v * (a.g.nbuckets + 1) + li + 1];
    const uint32_t hi = min(b1, lo + kClusterChunk);
    const uint32_t *L = a.g.lpool + st;
    for (uint32_t i = lo + lane; i < hi; i += 64) visit_bucket_entry(a, op, kind, li, L[i]);
  }
}

}  // namespace

hipError_t launch_cluster_init(uint64_t *bm, uint32_t words, uint32_t first_row, uint32_t nrows, uint64_t *alive,
                               const uint32_t *alive_src, uint32_t nalive_src, uint32_t *ctl, uint32_t nctl,
                               hipStream_t s) {
  const size_t work = std::max<size_t>((size_t)nrows * words, (size_t)nalive_src * words);
  const uint32_t blocks = (uint32_t)std::min<size_t>(std::max<size_t>((work + 255) / 256, 1), 4096);
  hipLaunchKernelGGL(k_cluster_init, dim3(blocks), dim3(256), 0, s, bm, words, first_row, nrows, alive, alive_src,
                     nalive_src, ctl, nctl);
  return hipGetLastError();
}

hipError_t launch_cluster_pass(const ClusterArgs &a, uint32_t max_blocks, hipStream_t s) {
  const uint64_t units = (uint64_t)a.nops * ((a.words + 63) / 64);
  const uint32_t blocks =
      (uint32_t)std::min<uint64_t>(std::max<uint64_t>((units + kClusterWaves - 1) / kClusterWaves, 1), max_blocks);
  hipLaunchKernelGGL(k_cluster_pass, dim3(blocks), dim3(kClusterWaves * 64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_cluster_heavy(const ClusterArgs &a, uint32_t blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_cluster_heavy, dim3(std::max<uint32_t>(blocks, 1)), dim3(kClusterWaves * 64), 0, s, a);
  return hipGetLastError();
}

}  // namespace dsm
