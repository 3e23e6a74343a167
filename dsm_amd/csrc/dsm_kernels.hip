// HIP kernels of the MI355X continuous subgraph matching engine (gfx950,
// wave64). Integer work only; no MFMA.
//
//   k_apply_merge / k_apply_commit (+ _big for hub lists)
//                                   in-place batch update of the slack CSR
//                                   (the reference rebuilds the whole graph on
//                                   the host, cpuGraph.cpp:521-607, and
//                                   re-uploads it, io.cu:279-302)
//   k_relabel (+ _big)              label changes of the update rule
//                                   (cpuGraph.cpp:564) + dirty marking
//   k_lbuild / k_lbuild_big / k_lbuild_huge_*
//                                   label-sorted copies of lists longer than
//                                   64 entries, rebuilt for the dirty set
//   k_filter_full / k_filter_light / k_filter_heavy
//                                   LDF/NLF candidate bitmaps
//                                   (filter.cu:394-566, compact filter.cu:1692)
//   k_seed                          device edge seeding (main.cu:268-329;
//                                   replaces the host ResTable and join.cu)
//   k_dfs                           persistent wave64 DFS with per-wave LDS
//                                   stacks, work donation and shared leaf sets
//                                   (gpu_match.cu:96-876 restated for CDNA)
#include "dsm_device.h"
#include "dsm_internal.h"

namespace dsm {
namespace {

constexpr int kApplyWaves = 4;
constexpr uint32_t kApplyWave = 8192;  // longer lists are copied by the whole grid
constexpr int kLight = 64;          // dirty vertices up to this degree: one lane each

__device__ __forceinline__ bool bit_of(const uint64_t *bm, uint32_t words, uint32_t sig, uint32_t v) {
  return (bm[(size_t)sig * words + (v >> 6)] >> (v & 63)) & 1ull;
}

// all 8 byte fields of x >= those of y (fields hold 0..15)
__device__ __forceinline__ bool ge_bytes(uint64_t x, uint64_t y) {
  constexpr uint64_t H = 0x8080808080808080ull;
  return (((x | H) - y) & H) == H;
}

// Marks x dirty in the update being applied: its bit only. The dirty list is
// compacted from the bits once all marks are in (k_dirty_compact), so that
// the marks of a hub's neighbours do not contend on a shared list counter. A
// set bit stays set for the update: read first, so that marking an already
// dirty vertex (the common case around hubs) costs a load, not an atomic.
__device__ __forceinline__ void mark_dirty(const DirtyArgs &d, uint32_t x) {
  const uint32_t bit = 1u << (x & 31);
  if (ld_relaxed(&d.bits[x >> 5]) & bit) return;
  atomicOr(&d.bits[x >> 5], bit);
}

// Only vertices with a label some query uses need their candidacy and their
// label-sorted copy kept current: no seed, DFS level or refilter reads the
// copy or the bits of any other vertex. A vertex that takes such a label later
// is marked by its own relabel, so a stale copy never gets read. (A label read
// here can race with the vertex's own relabel in the same kernel; either
// order marks what the relabel needs.)
__device__ __forceinline__ void mark_dirty_used(const DevGraph &g, const DirtyArgs &d, uint32_t x) {
  if (g.lab[x] != kNoLabel) mark_dirty(d, x);
}

// The dirty list from the dirty bits: a thread per bit word, a block prefix
// sum of the words' counts and one counter add per block; the list comes out
// in id order within each block's words.
__global__ __launch_bounds__(256) void k_dirty_compact(DirtyArgs d, uint32_t nwords) {
  __shared__ uint32_t s_sum[256];
  __shared__ uint32_t s_base;
  const uint32_t tid = threadIdx.x;
  for (uint32_t w0 = blockIdx.x * 256u; w0 < nwords; w0 += gridDim.x * 256u) {
    const uint32_t wi = w0 + tid;
    const uint32_t bits = wi < nwords ? d.bits[wi] : 0u;
    const uint32_t c = (uint32_t)__popc(bits);
    s_sum[tid] = c;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
      const uint32_t y = tid >= off ? s_sum[tid - off] : 0u;
      __syncthreads();
      s_sum[tid] += y;
      __syncthreads();
    }
    if (tid == 255) s_base = s_sum[255] ? atomicAdd(&d.actl->ndirty, s_sum[255]) : 0u;
    __syncthreads();
    uint32_t p = s_base + s_sum[tid] - c;
    uint32_t b = bits;
    while (b) {
      const uint32_t k = (uint32_t)__builtin_ctz(b);
      b &= b - 1;
      if (p < d.list_cap) d.list[p] = wi * 32u + k;
      p++;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ apply

constexpr uint32_t kBlockList = 65536;

// The lists of a *_big kernel: those of at most kBlockList entries sit at the
// front of its array (nblock of them), a block each; the longer ones at the
// back (nhuge, from index cap - 1 down), each split over every thread of the
// grid. f(entry, first, stride) gets the thread's first index into the list
// and the stride; both are multiples of 64 apart from the lane, so the lanes of
// a wave step through the list together.
template <typename F>
__device__ __forceinline__ void for_big_lists(const uint32_t *arr, uint32_t cap, uint32_t nblock, uint32_t nhuge,
                                              F f) {
  for (uint32_t r = blockIdx.x; r < nblock; r += gridDim.x) f(arr[r], threadIdx.x, blockDim.x);
  for (uint32_t r = 0; r < nhuge; r++)
    f(arr[cap - 1 - r], blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Queues entry v of a list of len entries on a *_big kernel's array: longer
// than kBlockList at the back, else at the front.
__device__ __forceinline__ void queue_big(uint32_t *arr, uint32_t cap, uint32_t *nblock, uint32_t *nhuge,
                                          uint32_t len, uint32_t v) {
  if (len > kBlockList) arr[cap - 1 - atomicAdd(nhuge, 1u)] = v;
  else arr[atomicAdd(nblock, 1u)] = v;
}

// Copies from[0, len) to to[0, len) (disjoint); the caller's threads are
// first, first + stride, ... (stride >= 4). When both ranges share their
// alignment modulo 16 bytes (the same offset of a pool in either copy of the
// graph), the bulk moves in 16-byte pieces, four in flight per thread, with
// fewer than four entries copied singly at each end; otherwise entry by
// entry, four in flight per thread.
__device__ __forceinline__ void copy_range(const uint32_t *__restrict__ from, uint32_t *__restrict__ to,
                                           uint32_t first, uint32_t stride, uint32_t len) {
  if ((((uintptr_t)from ^ (uintptr_t)to) & 15u) != 0) {
    uint32_t j = first;
    for (; j + 3 * stride < len; j += 4 * stride) {
      const uint32_t a = from[j], b = from[j + stride], c = from[j + 2 * stride], e = from[j + 3 * stride];
      to[j] = a;
      to[j + stride] = b;
      to[j + 2 * stride] = c;
      to[j + 3 * stride] = e;
    }
    for (; j < len; j += stride) to[j] = from[j];
    return;
  }
  const uint32_t lead = (uint32_t)((16u - ((uintptr_t)from & 15u)) & 15u) >> 2;
  const uint32_t head = lead < len ? lead : len;
  if (first < head) to[first] = from[first];
  const uint32_t nv = (len - head) >> 2;
  const uint4 *__restrict__ fv = reinterpret_cast<const uint4 *>(from + head);
  uint4 *__restrict__ tv = reinterpret_cast<uint4 *>(to + head);
  uint32_t j = first;
  for (; j + 3 * stride < nv; j += 4 * stride) {
    const uint4 a = fv[j], b = fv[j + stride], c = fv[j + 2 * stride], e = fv[j + 3 * stride];
    tv[j] = a;
    tv[j + stride] = b;
    tv[j + 2 * stride] = c;
    tv[j + 3 * stride] = e;
  }
  for (; j < nv; j += stride) tv[j] = fv[j];
  const uint32_t done = head + 4 * nv;
  if (first < len - done) to[done + first] = from[done + first];
}

// One wave per touched vertex: new list = (A \ D) u I written to scratch.
// I and D are sorted and hold at most kMaxVertexChanges entries each; every
// key has a single final operation, so I and D are disjoint.
__global__ __launch_bounds__(kApplyWaves * 64) void k_apply_merge(DevGraph g, BatchDev b,
                                                                  unsigned long long scratch_cap,
                                                                  uint32_t *scratch, ApplyCtl *actl) {
  __shared__ uint16_t s_icum[kApplyWaves][kMaxVertexChanges + 1];
  __shared__ uint16_t s_dcum[kApplyWaves][kMaxVertexChanges + 1];
  // up to kStageChanges inserts and deletes of the wave's vertex, staged in
  // LDS for the per-entry searches (more are searched in global memory)
  constexpr uint32_t kStageChanges = 128;
  __shared__ uint32_t s_I[kApplyWaves][kStageChanges], s_D[kApplyWaves][kStageChanges];
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  uint16_t *icum = s_icum[wv];
  uint16_t *dcum = s_dcum[wv];
  for (uint32_t t = blockIdx.x * kApplyWaves + wv; t < b.ntouched; t += gridDim.x * kApplyWaves) {
    const uint32_t x = b.tv[t];
    const uint32_t *A = g.pool + g.start[x];
    const uint32_t na = g.deg[x];
    const uint32_t i0 = b.ioff[t], ni = b.ioff[t + 1] - i0;
    const uint32_t d0 = b.doff[t], nd = b.doff[t + 1] - d0;
    const uint32_t *I = b.ins + i0;
    const uint32_t *D = b.del + d0;
    wave_sync();  // the previous vertex's stage is no longer read
    if (ni <= kStageChanges && nd <= kStageChanges) {
      for (uint32_t j = lane; j < ni; j += 64) s_I[wv][j] = I[j];
      for (uint32_t j = lane; j < nd; j += 64) s_D[wv][j] = D[j];
      wave_sync();
      I = s_I[wv];
      D = s_D[wv];
    }
    // exclusive prefix of "insert is new" and "delete hits"
    uint32_t run = 0;
    for (uint32_t c = 0; c < ni; c += 64) {
      const uint32_t j = c + lane;
      const bool fl = j < ni && !contains(A, na, I[j]);
      const uint64_t m = ballot(fl);
      if (j < ni) icum[j] = (uint16_t)(run + mbcnt(m));
      run += __popcll(m);
    }
    const uint32_t nnew = run;
    run = 0;
    for (uint32_t c = 0; c < nd; c += 64) {
      const uint32_t j = c + lane;
      const bool fl = j < nd && contains(A, na, D[j]);
      const uint64_t m = ballot(fl);
      if (j < nd) dcum[j] = (uint16_t)(run + mbcnt(m));
      run += __popcll(m);
    }
    const uint32_t nhit = run;
    if (lane == 0) {
      icum[ni] = (uint16_t)nnew;
      dcum[nd] = (uint16_t)nhit;
    }
    wave_sync();
    const uint32_t newdeg = na - nhit + nnew;
    if (na > kApplyWave) {
      // the prefix counts go to global memory; k_apply_merge_big copies
      for (uint32_t j = lane; j <= ni; j += 64) b.gicum[i0 + t + j] = icum[j];
      for (uint32_t j = lane; j <= nd; j += 64) b.gdcum[d0 + t + j] = dcum[j];
    }
    unsigned long long so = 0;
    if (lane == 0) so = atomicAdd(&actl->scratch_used, (unsigned long long)(na + nnew));
    so = bcast64(so);
    if (so + na + nnew > scratch_cap) {  // sized by the host so that this cannot happen
      if (lane == 0) {
        b.ndeg[t] = na;
        b.soff[t] = ~0ull;
        atomicAdd(&actl->merge_overflow, 1u);
      }
      continue;
    }
    uint32_t *out = scratch + so;
    if (na > kApplyWave) {
      if (lane == 0) {
        b.ndeg[t] = newdeg;
        b.soff[t] = so;
        queue_big(b.big_merge, b.ntouched, &actl->nmerge_big, &actl->nmerge_huge, na, t);
      }
      wave_sync();
      continue;
    }
    for (uint32_t j = lane; j < na; j += 64) {
      const uint32_t a = A[j];
      const uint32_t p = lower_bound(D, nd, a);
      if (p < nd && D[p] == a) continue;
      const uint32_t q = lower_bound(I, ni, a);
      out[j - dcum[p] + icum[q]] = a;
    }
    for (uint32_t j = lane; j < ni; j += 64) {
      if (icum[j + 1] == icum[j]) continue;  // already present
      const uint32_t i = I[j];
      const uint32_t r = lower_bound(A, na, i);
      const uint32_t p = lower_bound(D, nd, i);
      out[r - dcum[p] + icum[j]] = i;
    }
    if (lane == 0) {
      b.ndeg[t] = newdeg;
      b.soff[t] = so;
    }
    wave_sync();
  }
}

// The merge copies of the lists longer than kApplyWave (for_big_lists splits
// them), with the prefix counts k_apply_merge stored.
__global__ __launch_bounds__(256) void k_apply_merge_big(DevGraph g, BatchDev b, uint32_t *scratch,
                                                         const ApplyCtl *actl) {
  // the list's inserts and deletes with their prefix counts, staged in LDS:
  // every chunk of the list searches them
  __shared__ uint32_t s_I[kMaxVertexChanges], s_D[kMaxVertexChanges];
  __shared__ uint32_t s_icum[kMaxVertexChanges + 1], s_dcum[kMaxVertexChanges + 1];
  const uint32_t lane = lane_id();
  for_big_lists(b.big_merge, b.ntouched, actl->nmerge_big, actl->nmerge_huge, [&](uint32_t t, uint32_t tid,
                                                                                   uint32_t nth) {
    const uint32_t x = b.tv[t];
    const uint32_t na = g.deg[x];
    // waves of the share: the thread's wave index and the number of waves
    const uint32_t gw = tid >> 6, nw = nth >> 6;
    const uint32_t *A = g.pool + g.start[x];
    const uint32_t i0 = b.ioff[t], ni = b.ioff[t + 1] - i0;
    const uint32_t d0 = b.doff[t], nd = b.doff[t + 1] - d0;
    __syncthreads();  // the previous list's stage is no longer read
    for (uint32_t j = threadIdx.x; j <= ni; j += blockDim.x) {
      if (j < ni) s_I[j] = b.ins[i0 + j];
      s_icum[j] = b.gicum[i0 + t + j];
    }
    for (uint32_t j = threadIdx.x; j <= nd; j += blockDim.x) {
      if (j < nd) s_D[j] = b.del[d0 + j];
      s_dcum[j] = b.gdcum[d0 + t + j];
    }
    __syncthreads();
    const uint32_t *I = s_I;
    const uint32_t *D = s_D;
    const uint32_t *icum = s_icum;
    const uint32_t *dcum = s_dcum;
    uint32_t *out = scratch + b.soff[t];
    // waves take 64-entry chunks of A in turn; the deletes and inserts that
    // fall inside a chunk's key range are found once per chunk, so each entry
    // searches only that (usually empty) range
    for (uint32_t c0 = gw * 64; c0 < na; c0 += nw * 64) {
      const uint32_t c1 = c0 + 64 < na ? c0 + 64 : na;
      const uint32_t lo_key = A[c0], hi_key = A[c1 - 1];
      const uint32_t p0 = lower_bound(D, nd, lo_key), p1 = lower_bound(D, nd, hi_key + 1);
      const uint32_t q0 = lower_bound(I, ni, lo_key), q1 = lower_bound(I, ni, hi_key + 1);
      const uint32_t j = c0 + lane;
      if (j < c1) {
        const uint32_t a = A[j];
        const uint32_t p = p0 + lower_bound(D + p0, p1 - p0, a);
        if (!(p < nd && D[p] == a)) {
          const uint32_t q = q0 + lower_bound(I + q0, q1 - q0, a);
          out[j - dcum[p] + icum[q]] = a;
        }
      }
    }
    for (uint32_t j = tid; j < ni; j += nth) {
      if (icum[j + 1] == icum[j]) continue;
      const uint32_t i = I[j];
      const uint32_t rr = lower_bound(A, na, i);
      const uint32_t p = lower_bound(D, nd, i);
      out[rr - dcum[p] + icum[j]] = i;
    }
  });
}

// One wave per touched vertex: copy the merged list back into its slot, or
// into a new, larger slot taken from the pool's bump pointer. A vertex that
// finds no room keeps its old list and is flagged for a retry after the host
// grows the pool. With only_failed, only flagged vertices are processed.
__global__ __launch_bounds__(kApplyWaves * 64) void k_apply_commit(DevGraph g, BatchDev b,
                                                                   const uint32_t *scratch,
                                                                   unsigned long long pool_cap,
                                                                   ApplyCtl *actl, DirtyArgs d,
                                                                   int only_failed) {
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  for (uint32_t t = blockIdx.x * kApplyWaves + wv; t < b.ntouched; t += gridDim.x * kApplyWaves) {
    if (only_failed && b.fail[t] == 0) continue;
    const uint32_t x = b.tv[t];
    const uint32_t nd = b.ndeg[t];
    const unsigned long long so = b.soff[t];
    if (so == ~0ull) continue;
    uint32_t dst = g.start[x];
    if (nd > g.cap[x]) {
      const uint32_t newcap = nd + nd / 4 + 8;
      unsigned long long base = 0;
      if (lane == 0) base = atomicAdd(&actl->pool_used, (unsigned long long)newcap);
      base = bcast64(base);
      if (base + newcap > pool_cap) {
        if (lane == 0) {
          b.fail[t] = 1;
          atomicAdd(&actl->overflow, 1u);
        }
        continue;
      }
      dst = (uint32_t)base;
      if (lane == 0) {
        g.start[x] = dst;
        g.cap[x] = newcap;
        atomicAdd(&actl->relocated, 1u);
      }
    }
    if (nd > kApplyWave) {
      // k_apply_commit_big copies; the new degree is only read by later kernels
      if (lane == 0) queue_big(b.big_commit, b.ntouched, &actl->ncommit_big, &actl->ncommit_huge, nd, t);
    } else {
      copy_range(scratch + so, g.pool + dst, lane, 64, nd);
    }
    if (lane == 0) {
      g.deg[x] = nd;
      b.fail[t] = 0;
      mark_dirty_used(g, d, x);
    }
  }
}

// The commit copies of the lists longer than kApplyWave (for_big_lists splits
// them).
__global__ __launch_bounds__(256) void k_apply_commit_big(DevGraph g, BatchDev b, const uint32_t *scratch,
                                                          const ApplyCtl *actl) {
  for_big_lists(b.big_commit, b.ntouched, actl->ncommit_big, actl->ncommit_huge, [&](uint32_t t, uint32_t first,
                                                                                     uint32_t stride) {
    const uint32_t nd = b.ndeg[t];
    const uint32_t dst = g.start[b.tv[t]];
    const unsigned long long so = b.soff[t];
    copy_range(scratch + so, g.pool + dst, first, stride, nd);
  });
}

__global__ void k_big_reset(ApplyCtl *actl, int merge, int commit) {
  if (threadIdx.x == 0) {
    if (merge) actl->nmerge_big = actl->nmerge_huge = 0;
    if (commit) actl->ncommit_big = actl->ncommit_huge = 0;
  }
}

// One wave per relabel entry: set the label; when it changed, clear the
// vertex's bits in every bitmap of its old label and mark it and all its
// (post-update) neighbours dirty, since their neighbour-label counts changed.
constexpr uint32_t kRelabelWave = 2048;  // longer lists are marked by the whole grid (k_relabel_big)

__global__ __launch_bounds__(256) void k_relabel(DevGraph g, BatchDev b, FilterTables ft, DirtyArgs d) {
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  for (uint32_t r = blockIdx.x * 4 + wv; r < b.nrelabel; r += gridDim.x * 4) {
    const uint32_t x = b.rl_v[r];
    const uint32_t L = b.rl_l[r];
    const uint32_t old = g.lab[x];
    if (old == L) continue;
    if (lane == 0) {
      g.lab[x] = (uint8_t)L;
      atomicAdd(&d.actl->relabeled, 1u);
      if (L != kNoLabel) mark_dirty(d, x);
    }
    if (old != kNoLabel) {
      const uint64_t clr = ~(1ull << (x & 63));
      for (uint32_t s = ft.sig_lo[old] + lane; s < ft.sig_hi[old]; s += 64)
        atomicAnd((unsigned long long *)&ft.bm[(size_t)s * ft.words + (x >> 6)], (unsigned long long)clr);
    }
    const uint32_t *A = g.pool + g.start[x];
    const uint32_t na = g.deg[x];
    if (na > kRelabelWave) {
      if (lane == 0) queue_big(b.rl_big, b.nrelabel, &d.actl->nrelabel_big, &d.actl->nrelabel_huge, na, x);
      continue;
    }
    for (uint32_t j = lane; j < na; j += 64) mark_dirty_used(g, d, A[j]);
  }
}

// The neighbours of every relabelled vertex of degree > kRelabelWave, marked
// dirty (for_big_lists splits the lists).
__global__ __launch_bounds__(256) void k_relabel_big(DevGraph g, BatchDev b, DirtyArgs d) {
  for_big_lists(b.rl_big, b.nrelabel, d.actl->nrelabel_big, d.actl->nrelabel_huge, [&](uint32_t x, uint32_t first,
                                                                                         uint32_t stride) {
    const uint32_t na = g.deg[x];
    const uint32_t *A = g.pool + g.start[x];
    // four entries per thread in flight before their marks
    uint32_t j = first;
    for (; j + 3 * stride < na; j += 4 * stride) {
      const uint32_t x0 = A[j], x1 = A[j + stride], x2 = A[j + 2 * stride], x3 = A[j + 3 * stride];
      mark_dirty_used(g, d, x0);
      mark_dirty_used(g, d, x1);
      mark_dirty_used(g, d, x2);
      mark_dirty_used(g, d, x3);
    }
    for (; j < na; j += stride) mark_dirty_used(g, d, A[j]);
  });
}

// ----------------------------------------------------------------- filter

__device__ __forceinline__ bool sig_pass(const SigReq &r, uint32_t deg, const uint64_t c[4]) {
  return deg >= r.deg && ge_bytes(c[0], r.nlc[0]) && ge_bytes(c[1], r.nlc[1]) &&
         ge_bytes(c[2], r.nlc[2]) && ge_bytes(c[3], r.nlc[3]);
}

// Whole-graph pass, one wave per 64-vertex group: the group's adjacency lists
// are scanned as one logical array (lanes balanced across vertices), label
// counts accumulate in LDS, then every bitmap word of the group is written
// with one ballot.
__global__ __launch_bounds__(256) void k_filter_full(DevGraph g, FilterTables ft) {
  __shared__ uint32_t s_cnt[4][64][kMaxRel + 1];
  __shared__ uint32_t s_pre[4][65];
  __shared__ uint32_t s_st[4][64];
  __shared__ uint8_t s_lmap[256];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_lmap[i] = ft.lmap[i];
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  const uint32_t ngroups = (g.n + 63) / 64;
  for (uint32_t grp = blockIdx.x * 4 + wv; grp < ngroups; grp += gridDim.x * 4) {
    const uint32_t v = grp * 64 + lane;
    const bool valid = v < g.n;
    const uint32_t L = valid ? g.lab[v] : kNoLabel;
    const bool has = L != kNoLabel && ft.sig_lo[L] < ft.sig_hi[L];
    const uint32_t dv = valid ? g.deg[v] : 0;
    const uint32_t d = has ? dv : 0;
    s_st[wv][lane] = valid ? g.start[v] : 0;
    for (int l = 0; l <= kMaxRel; l++) s_cnt[wv][lane][l] = 0;
    // inclusive scan of d
    uint32_t incl = d;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off);
      if (lane >= (uint32_t)off) incl += y;
    }
    s_pre[wv][lane + 1] = incl;
    if (lane == 0) s_pre[wv][0] = 0;
    const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
    wave_sync();
    for (uint32_t e = lane; e < total; e += 64) {
      // owner: last k with s_pre[k] <= e
      uint32_t lo = 0, n = 64;
      while (n > 1) {
        const uint32_t half = n >> 1;
        lo = (s_pre[wv][lo + half] <= e) ? lo + half : lo;
        n -= half;
      }
      const uint32_t w = g.pool[s_st[wv][lo] + (e - s_pre[wv][lo])];
      const uint32_t li = s_lmap[g.lab[w]];
      if (li != 0xFF) atomicAdd(&s_cnt[wv][lo][li], 1u);
    }
    wave_sync();
    uint64_t c[4] = {0, 0, 0, 0};
    for (int l = 0; l < kMaxRel; l++) {
      const uint64_t x = s_cnt[wv][lane][l] < 15 ? s_cnt[wv][lane][l] : 15;
      c[l >> 3] |= x << ((l & 7) * 8);
    }
    for (uint32_t s = 0; s < ft.nsig; s++) {
      const SigReq r = ft.sigs[s];
      const bool pass = has && L == r.label && sig_pass(r, dv, c);
      const uint64_t word = ballot(pass);
      if (lane == 0) ft.bm[(size_t)s * ft.words + grp] = word;
    }
    wave_sync();
  }
}

__device__ __forceinline__ void set_bit(uint64_t *w, uint32_t v, bool on) {
  const uint64_t bit = 1ull << (v & 63);
  const bool cur = (*w & bit) != 0;  // only this thread changes this bit in this kernel
  if (cur == on) return;
  if (on) atomicOr((unsigned long long *)w, (unsigned long long)bit);
  else atomicAnd((unsigned long long *)w, (unsigned long long)~bit);
}

// Dirty vertices, one lane each; lists longer than kLight go to the heavy list.
__global__ __launch_bounds__(256) void k_filter_light(DevGraph g, FilterTables ft, DirtyArgs d) {
  __shared__ uint8_t s_cnt[256][kMaxRel];
  __shared__ uint8_t s_lmap[256];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_lmap[i] = ft.lmap[i];
  __syncthreads();
  const uint32_t nd = d.actl->ndirty < d.list_cap ? d.actl->ndirty : d.list_cap;
  uint8_t *cnt = s_cnt[threadIdx.x];
  unsigned long long entries = 0;  // list entries of the vertices this thread re-filters
  uint32_t bits = 0;               // candidate bits they receive (one per signature of the label)
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nd; i += gridDim.x * blockDim.x) {
    const uint32_t v = d.list[i];
    atomicAnd(&d.bits[v >> 5], ~(1u << (v & 31)));
    const uint32_t L = g.lab[v];
    const bool has = L != kNoLabel && ft.sig_lo[L] != ft.sig_hi[L];
    const uint32_t dv = has ? g.deg[v] : 0u;
    if (has) {
      entries += dv;
      bits += ft.sig_hi[L] - ft.sig_lo[L];
    }
    // long lists go to the heavy list, one counter add per wave
    const bool heavy = has && dv > (uint32_t)kLight;
    const uint64_t hm = ballot(heavy);
    if (hm) {
      const uint32_t leader = (uint32_t)__builtin_ctzll(hm);
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(&d.actl->nheavy, (uint32_t)__popcll(hm));
      base = __shfl(base, (int)leader);
      const uint32_t p = base + mbcnt(hm);
      if (heavy && p < d.list_cap) d.heavy[p] = v;
    }
    if (!has || heavy) continue;
    for (int l = 0; l < kMaxRel; l++) cnt[l] = 0;
    const uint32_t *A = g.pool + g.start[v];
    // four entries, then their four label reads, in flight before the counts
    uint32_t j = 0;
    for (; j + 4 <= dv; j += 4) {
      const uint32_t x0 = A[j], x1 = A[j + 1], x2 = A[j + 2], x3 = A[j + 3];
      const uint32_t l0 = g.lab[x0], l1 = g.lab[x1], l2 = g.lab[x2], l3 = g.lab[x3];
      const uint32_t li[4] = {s_lmap[l0], s_lmap[l1], s_lmap[l2], s_lmap[l3]};
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (li[u] != 0xFF && cnt[li[u]] < 15) cnt[li[u]]++;
    }
    for (; j < dv; j++) {
      const uint32_t li = s_lmap[g.lab[A[j]]];
      if (li != 0xFF && cnt[li] < 15) cnt[li]++;
    }
    uint64_t c[4];
    for (int k = 0; k < 4; k++) {
      uint64_t x = 0;
      for (int b = 0; b < 8; b++) x |= (uint64_t)cnt[k * 8 + b] << (b * 8);
      c[k] = x;
    }
    for (uint32_t s = ft.sig_lo[L]; s < ft.sig_hi[L]; s++)
      set_bit(&ft.bm[(size_t)s * ft.words + (v >> 6)], v, sig_pass(ft.sigs[s], dv, c));
  }
  for (int o = 32; o >= 1; o >>= 1) {
    entries += __shfl_xor(entries, o);
    bits += __shfl_xor(bits, o);
  }
  if (lane_id() == 0 && entries) atomicAdd(&d.actl->filter_entries, entries);
  if (lane_id() == 0 && bits) atomicAdd(&d.actl->filter_bits, bits);
}

// Dirty vertices with long lists (every one has a label-sorted copy, rebuilt
// for this batch before this kernel runs): the neighbour count of query label
// index li is the size of label group li, so no list is scanned. One wave per
// vertex; lanes test the bitmaps in parallel.
__global__ __launch_bounds__(256) void k_filter_heavy(DevGraph g, FilterTables ft, DirtyArgs d) {
  __shared__ uint64_t s_pack[4][4];
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  const uint32_t nh = d.actl->nheavy < d.list_cap ? d.actl->nheavy : d.list_cap;
  for (uint32_t h = blockIdx.x * 4 + wv; h < nh; h += gridDim.x * 4) {
    const uint32_t v = d.heavy[h];
    const uint32_t L = g.lab[v];
    const uint32_t dv = g.deg[v];
    const uint32_t *row = g.loff + (size_t)v * (g.nbuckets + 1);
    if (lane < 4) {
      uint64_t x = 0;
      for (int b = 0; b < 8; b++) {
        // groups 0 .. nbuckets-2 are the query labels; the last one, the rest
        const uint32_t li = lane * 8 + b;
        const uint32_t y = li + 1 < g.nbuckets ? row[li + 1] - row[li] : 0u;
        x |= (uint64_t)(y < 15 ? y : 15) << (b * 8);
      }
      s_pack[wv][lane] = x;
    }
    wave_sync();
    const uint64_t c[4] = {s_pack[wv][0], s_pack[wv][1], s_pack[wv][2], s_pack[wv][3]};
    for (uint32_t s = ft.sig_lo[L] + lane; s < ft.sig_hi[L]; s += 64)
      set_bit(&ft.bm[(size_t)s * ft.words + (v >> 6)], v, sig_pass(ft.sigs[s], dv, c));
    wave_sync();
  }
}

// ---------------------------------------------------- label-sorted copies

// Lanes of the wave whose bucket equals this lane's (among valid lanes): one
// ballot per bit of the bucket index, so a chunk with many distinct buckets
// costs the same as one with a single bucket.
__device__ __forceinline__ uint64_t bucket_peers(uint32_t b, bool valid, uint32_t bits) {
  uint64_t m = ballot(valid);
  for (uint32_t i = 0; i < bits; i++) {
    const bool set = (b >> i) & 1;
    const uint64_t bi = ballot(valid && set);
    m &= set ? bi : ~bi;
  }
  return m;
}

__device__ __forceinline__ uint32_t bucket_of(const DevGraph &g, const uint8_t *s_lmap, uint32_t x) {
  const uint32_t b = s_lmap[g.lab[x]];
  return b == 0xFF ? g.nbuckets - 1 : b;
}

// Adds the bucket histogram of A[lo, hi) to cnt (per-wave LDS counters; the
// lowest lane of each peer group adds the group's size).
__device__ __forceinline__ void bucket_count(const DevGraph &g, const uint8_t *s_lmap, const uint32_t *A,
                                             uint32_t lo, uint32_t hi, uint32_t bits, uint32_t *cnt,
                                             const uint8_t *bk_in, uint8_t *bk_out) {
  const uint32_t lane = lane_id();
  // four chunks at a time: their entries, then their label reads, are all in
  // flight before the first chunk's counts are taken
  constexpr uint32_t kU = 4;
  for (uint32_t c0 = lo; c0 < hi; c0 += 64 * kU) {
    uint32_t x[kU], b[kU];
#pragma unroll
    for (uint32_t u = 0; u < kU; u++) {
      const uint32_t j = c0 + u * 64 + lane;
      x[u] = j < hi && !bk_in ? A[j] : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < kU; u++) {
      const uint32_t j = c0 + u * 64 + lane;
      b[u] = j < hi ? (bk_in ? (uint32_t)bk_in[j] : bucket_of(g, s_lmap, x[u])) : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < kU; u++) {
      const uint32_t j = c0 + u * 64 + lane;
      const bool valid = j < hi;
      if (valid && bk_out) bk_out[j] = (uint8_t)b[u];
      const uint64_t pm = bucket_peers(b[u], valid, bits);
      if (valid && mbcnt(pm) == 0) cnt[b[u]] += (uint32_t)__popcll(pm);
      wave_sync();
    }
  }
}

// Scatters A[lo, hi) into out by bucket, cnt[b] holding the next position of
// bucket b; id order is kept inside each bucket. Four chunks' entries and
// buckets are loaded before the first chunk is placed.
__device__ __forceinline__ void bucket_scatter(const DevGraph &g, const uint8_t *s_lmap, const uint32_t *A,
                                               uint32_t lo, uint32_t hi, uint32_t bits, uint32_t *cnt,
                                               uint32_t *out, const uint8_t *bk_in) {
  const uint32_t lane = lane_id();
  constexpr uint32_t kU = 4;
  for (uint32_t c0 = lo; c0 < hi; c0 += 64 * kU) {
    uint32_t x[kU], b[kU];
#pragma unroll
    for (uint32_t u = 0; u < kU; u++) {
      const uint32_t j = c0 + u * 64 + lane;
      x[u] = j < hi ? A[j] : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < kU; u++) {
      const uint32_t j = c0 + u * 64 + lane;
      b[u] = j < hi ? (bk_in ? (uint32_t)bk_in[j] : bucket_of(g, s_lmap, x[u])) : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < kU; u++) {
      const bool valid = c0 + u * 64 + lane < hi;
      const uint64_t pm = bucket_peers(b[u], valid, bits);
      const uint32_t rank = mbcnt(pm);
      const uint32_t base = valid ? cnt[b[u]] : 0u;
      wave_sync();
      if (valid) {
        out[base + rank] = x[u];
        if (rank == 0) cnt[b[u]] = base + (uint32_t)__popcll(pm);
      }
      wave_sync();
    }
  }
}

// Exclusive prefix sum over the lanes (a wave-wide scan of shuffles); the
// inclusive total comes back in *total.
__device__ __forceinline__ uint32_t wave_excl_sum(uint32_t c, uint32_t *total) {
  const uint32_t lane = lane_id();
  uint32_t incl = c;
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off);
    if (lane >= off) incl += y;
  }
  *total = __builtin_amdgcn_readlane(incl, 63);
  return incl - c;
}

constexpr uint32_t kLbuildBlock = 4096;  // lists longer than this are sorted by a whole block
constexpr uint32_t kLbuildWaves = 16;    // waves of that block

// The label-sorted copy of v (degree d) by one wave: bucket histogram,
// offsets, then a stable scatter of the id-sorted list into lpool bucket by
// bucket.
__device__ __forceinline__ void build_one(const DevGraph &g, const uint8_t *s_lmap, uint32_t v, uint32_t d,
                                          uint32_t nb, uint32_t bits, uint32_t *cnt, uint8_t *bucket_bytes) {
  const uint32_t lane = lane_id();
  const uint32_t *A = g.pool + g.start[v];
  uint32_t *row = g.loff + (size_t)v * (nb + 1);
  if (lane < nb) cnt[lane] = 0;
  wave_sync();
  uint8_t *bk = bucket_bytes ? bucket_bytes + g.start[v] : nullptr;
  bucket_count(g, s_lmap, A, 0, d, bits, cnt, nullptr, bk);
  // bucket starts: an exclusive scan with a lane per bucket (nb <= kMaxRel + 1 < 64)
  uint32_t total;
  const uint32_t c = lane < nb ? cnt[lane] : 0u;
  const uint32_t first = wave_excl_sum(c, &total);
  if (lane < nb) {
    cnt[lane] = first;
    row[lane] = first;
  }
  if (lane == 0) row[nb] = total;
  wave_sync();
  bucket_scatter(g, s_lmap, A, 0, d, bits, cnt, g.lpool + g.start[v], bk);
}

// Appends the lanes' vertices with `take` to q (count at *cnt, room for cap),
// one counter add per wave.
__device__ __forceinline__ void queue_wave(uint32_t *q, uint32_t cap, uint32_t *cnt, uint32_t v, bool take) {
  const uint64_t m = ballot(take);
  if (m == 0) return;
  const uint32_t leader = (uint32_t)__builtin_ctzll(m);
  uint32_t at = 0;
  if (lane_id() == leader) at = atomicAdd(cnt, (uint32_t)__popcll(m));
  at = __shfl(at, (int)leader) + mbcnt(m);
  if (take && at < cap) q[at] = v;
}

// The listed vertices (or all), a lane each, sorted into the two build
// kernels' queues: degree in (kLabelSorted, kLbuildBlock] on hs.mid for
// k_lbuild (a wave each), longer on hs.big for k_lbuild_big (a block each).
// Vertices whose label no query uses are skipped (never read, see
// mark_dirty_used); a full build leaves their groups empty.
__global__ __launch_bounds__(256) void k_lbuild_pick(DevGraph g, const uint32_t *list, const uint32_t *count,
                                                     HugeSort hs) {
  const uint32_t nb = g.nbuckets;
  const uint32_t total = list ? (*count < g.n ? *count : g.n) : g.n;
  for (uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; base < total;
       base += gridDim.x * blockDim.x) {
    const uint32_t t = base + lane_id();
    uint32_t v = 0, d = 0;
    bool used = false;
    if (t < total) {
      v = list ? list[t] : t;
      d = g.deg[v];
      used = g.lab[v] != kNoLabel;
      if (!used && !list && d > kLabelSorted)
        for (uint32_t k = 0; k <= nb; k++) g.loff[(size_t)v * (nb + 1) + k] = 0;
    }
    queue_wave(hs.mid, hs.mid_cap, &hs.ctl[3], v, used && d > kLabelSorted && d <= kLbuildBlock);
    queue_wave(hs.big, hs.big_cap, &hs.ctl[2], v, used && d > kLbuildBlock);
  }
}

// A wave per vertex queued on hs.mid: its label-sorted copy (build_one).
__global__ __launch_bounds__(256) void k_lbuild(DevGraph g, const uint8_t *lmap, HugeSort hs) {
  __shared__ uint8_t s_lmap[256];
  __shared__ uint32_t s_cnt[4][kMaxRel + 2];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_lmap[i] = lmap[i];
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6;
  const uint32_t nb = g.nbuckets;
  const uint32_t bits = 32u - (uint32_t)__builtin_clz(nb > 1 ? nb - 1 : 1u);
  const uint32_t total = hs.ctl[3] < hs.mid_cap ? hs.ctl[3] : hs.mid_cap;
  for (uint32_t t = blockIdx.x * 4 + wv; t < total; t += gridDim.x * 4) {
    const uint32_t v = hs.mid[t];
    build_one(g, s_lmap, v, g.deg[v], nb, bits, s_cnt[wv], hs.bucket_bytes);
  }
}

// One block of kLbuildWaves waves per vertex of degree > kLbuildBlock: each
// wave takes a contiguous segment of the list; per-wave histograms give every
// wave its starting position inside every bucket, so the block-wide scatter
// stays stable (id order inside each bucket). Control flow around the block
// barriers depends only on the block's vertex, so it is uniform.
__global__ __launch_bounds__(kLbuildWaves * 64) void k_lbuild_big(DevGraph g, const uint8_t *lmap,
                                                                  const uint32_t *list, const uint32_t *count,
                                                                  HugeSort hs) {
  __shared__ uint8_t s_lmap[256];
  __shared__ uint32_t s_cnt[kLbuildWaves][kMaxRel + 2];
  __shared__ uint32_t s_queued;
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_lmap[i] = lmap[i];
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  const uint32_t nb = g.nbuckets;
  const uint32_t bits = 32u - (uint32_t)__builtin_clz(nb > 1 ? nb - 1 : 1u);
  // the vertices k_lbuild queued; past the queue's capacity, every listed
  // vertex is looked at again
  const uint32_t queued = hs.ctl[2];
  const bool from_queue = queued <= hs.big_cap;
  const uint32_t total = from_queue ? queued : list ? (*count < g.n ? *count : g.n) : g.n;
  for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
    const uint32_t v = from_queue ? hs.big[t] : list ? list[t] : t;
    const uint32_t d = g.deg[v];
    if (d <= kLbuildBlock || g.lab[v] == kNoLabel) continue;
    if (d > kHugeList) {
      // queue it for the segmented build when its slots fit, else sort here
      if (threadIdx.x == 0) {
        const uint32_t nseg = (d + kHugeSeg - 1) / kHugeSeg;
        uint32_t ok = 0;
        const uint32_t i = atomicAdd(&hs.ctl[0], 1u);
        if (i < hs.cap) {
          const uint32_t sb = atomicAdd(&hs.ctl[1], nseg);
          if (sb + nseg <= hs.segs) {
            hs.list[2 * i] = v;
            hs.list[2 * i + 1] = sb;
            ok = 1;
          } else {
            hs.list[2 * i] = 0xFFFFFFFFu;  // slot taken but unused
          }
        }
        s_queued = ok;
      }
      __syncthreads();
      const uint32_t queued = s_queued;
      __syncthreads();
      if (queued) continue;
    }
    const uint32_t *A = g.pool + g.start[v];
    uint32_t *row = g.loff + (size_t)v * (nb + 1);
    const uint32_t seg = (d + kLbuildWaves - 1) / kLbuildWaves;
    const uint32_t lo = wv * seg < d ? wv * seg : d;
    const uint32_t hi = lo + seg < d ? lo + seg : d;
    uint32_t *cnt = s_cnt[wv];
    uint8_t *bk = hs.bucket_bytes ? hs.bucket_bytes + g.start[v] : nullptr;
    if (lane < nb) cnt[lane] = 0;
    wave_sync();
    bucket_count(g, s_lmap, A, lo, hi, bits, cnt, nullptr, bk);
    __syncthreads();
    if (wv == 0) {
      // wave 0, a lane per bucket (nb <= kMaxRel + 1 < 64): the bucket's total
      // over the waves, the bucket starts by a scan over the lanes, then each
      // wave's start inside the bucket
      uint32_t tot = 0;
      if (lane < nb)
        for (uint32_t w = 0; w < kLbuildWaves; w++) tot += s_cnt[w][lane];
      uint32_t total;
      uint32_t run = wave_excl_sum(tot, &total);
      if (lane < nb) {
        row[lane] = run;
        for (uint32_t w = 0; w < kLbuildWaves; w++) {
          const uint32_t c = s_cnt[w][lane];
          s_cnt[w][lane] = run;
          run += c;
        }
      }
      if (lane == 0) row[nb] = total;
    }
    __syncthreads();
    bucket_scatter(g, s_lmap, A, lo, hi, bits, cnt, g.lpool + g.start[v], bk);
    __syncthreads();
  }
}

// Segmented build of the queued lists (longer than kHugeList): segment k of a
// list is handled by block (slot + k) mod gridDim. Pass 1 stores each
// segment's bucket counts; pass 2 derives every segment's start inside every
// bucket from them and scatters the segment (per-wave histograms order the
// waves inside the segment), so the copy stays stable.
__device__ __forceinline__ void huge_entry(const HugeSort &hs, uint32_t i, uint32_t &v, uint32_t &sb) {
  v = hs.list[2 * i];
  sb = hs.list[2 * i + 1];
}

__global__ __launch_bounds__(kLbuildWaves * 64) void k_lbuild_huge_count(DevGraph g, const uint8_t *lmap,
                                                                         HugeSort hs) {
  __shared__ uint8_t s_lmap[256];
  __shared__ uint32_t s_cnt[kLbuildWaves][kMaxRel + 2];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_lmap[i] = lmap[i];
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  const uint32_t nb = g.nbuckets;
  const uint32_t bits = 32u - (uint32_t)__builtin_clz(nb > 1 ? nb - 1 : 1u);
  const uint32_t nq = hs.ctl[0] < hs.cap ? hs.ctl[0] : hs.cap;
  for (uint32_t i = 0; i < nq; i++) {
    uint32_t v, sb;
    huge_entry(hs, i, v, sb);
    if (v == 0xFFFFFFFFu) continue;
    const uint32_t d = g.deg[v];
    const uint32_t nseg = (d + kHugeSeg - 1) / kHugeSeg;
    const uint32_t *A = g.pool + g.start[v];
    for (uint32_t k = 0; k < nseg; k++) {
      if ((sb + k) % gridDim.x != blockIdx.x) continue;
      const uint32_t s0 = k * kHugeSeg, s1 = s0 + kHugeSeg < d ? s0 + kHugeSeg : d;
      const uint32_t sub = (s1 - s0 + kLbuildWaves - 1) / kLbuildWaves;
      const uint32_t lo = s0 + wv * sub < s1 ? s0 + wv * sub : s1;
      const uint32_t hi = lo + sub < s1 ? lo + sub : s1;
      uint8_t *bk = hs.bucket_bytes ? hs.bucket_bytes + g.start[v] : nullptr;
      if (lane < nb) s_cnt[wv][lane] = 0;
      wave_sync();
      bucket_count(g, s_lmap, A, lo, hi, bits, s_cnt[wv], nullptr, bk);
      __syncthreads();
      if (threadIdx.x < nb) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < kLbuildWaves; w++) tot += s_cnt[w][threadIdx.x];
        hs.hist[(size_t)(sb + k) * (kMaxRel + 2) + threadIdx.x] = tot;
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(kLbuildWaves * 64) void k_lbuild_huge_scatter(DevGraph g, const uint8_t *lmap,
                                                                           HugeSort hs) {
  __shared__ uint8_t s_lmap[256];
  __shared__ uint32_t s_cnt[kLbuildWaves][kMaxRel + 2];
  __shared__ uint32_t s_tot[kMaxRel + 2], s_before[kMaxRel + 2];
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) s_lmap[i] = lmap[i];
  __syncthreads();
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  const uint32_t nb = g.nbuckets;
  const uint32_t bits = 32u - (uint32_t)__builtin_clz(nb > 1 ? nb - 1 : 1u);
  const uint32_t nq = hs.ctl[0] < hs.cap ? hs.ctl[0] : hs.cap;
  for (uint32_t i = 0; i < nq; i++) {
    uint32_t v, sb;
    huge_entry(hs, i, v, sb);
    if (v == 0xFFFFFFFFu) continue;
    const uint32_t d = g.deg[v];
    const uint32_t nseg = (d + kHugeSeg - 1) / kHugeSeg;
    const uint32_t *A = g.pool + g.start[v];
    uint32_t *row = g.loff + (size_t)v * (nb + 1);
    for (uint32_t k = 0; k < nseg; k++) {
      if ((sb + k) % gridDim.x != blockIdx.x) continue;
      // bucket totals over all segments, and this segment's share before it
      if (threadIdx.x < nb) {
        uint32_t tot = 0, before = 0;
        for (uint32_t k2 = 0; k2 < nseg; k2++) {
          const uint32_t c = hs.hist[(size_t)(sb + k2) * (kMaxRel + 2) + threadIdx.x];
          tot += c;
          if (k2 < k) before += c;
        }
        s_tot[threadIdx.x] = tot;
        s_before[threadIdx.x] = before;
      }
      const uint32_t s0 = k * kHugeSeg, s1 = s0 + kHugeSeg < d ? s0 + kHugeSeg : d;
      const uint32_t sub = (s1 - s0 + kLbuildWaves - 1) / kLbuildWaves;
      const uint32_t lo = s0 + wv * sub < s1 ? s0 + wv * sub : s1;
      const uint32_t hi = lo + sub < s1 ? lo + sub : s1;
      const uint8_t *bk = hs.bucket_bytes ? hs.bucket_bytes + g.start[v] : nullptr;
      if (lane < nb) s_cnt[wv][lane] = 0;
      wave_sync();
      bucket_count(g, s_lmap, A, lo, hi, bits, s_cnt[wv], bk, nullptr);
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t b = 0; b < nb; b++) {
          const uint32_t c = s_tot[b];
          s_tot[b] = run;  // bucket start
          if (k == 0) row[b] = run;
          run += c;
        }
        if (k == 0) row[nb] = run;
      }
      __syncthreads();
      if (threadIdx.x < nb) {
        uint32_t run = s_tot[threadIdx.x] + s_before[threadIdx.x];
        for (uint32_t w = 0; w < kLbuildWaves; w++) {
          const uint32_t c = s_cnt[w][threadIdx.x];
          s_cnt[w][threadIdx.x] = run;
          run += c;
        }
      }
      __syncthreads();
      bucket_scatter(g, s_lmap, A, lo, hi, bits, s_cnt[wv], g.lpool + g.start[v], bk);
      __syncthreads();
    }
  }
}

__global__ void k_huge_reset(HugeSort hs) {
  if (threadIdx.x == 0) {
    hs.ctl[0] = 0;
    hs.ctl[1] = 0;
    hs.ctl[2] = 0;
    hs.ctl[3] = 0;
  }
}

// ------------------------------------------------------------------- seed

// One thread per record (s,d): every plan whose seeded query edge (u0,u1) has
// s in C(u0) and d in C(u1) becomes a DFS root, provided (s,d) is an edge of
// the graph being matched.
__global__ __launch_bounds__(256) void k_seed(SeedArgs a) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < a.nrec; r += gridDim.x * blockDim.x) {
    const uint32_t s = a.rs[r], d = a.rd[r];
    if (s == d || s >= a.g.n || d >= a.g.n) continue;
    const uint32_t ls = a.g.lab[s], ld = a.g.lab[d];
    if (ls == kNoLabel || ld == kNoLabel) continue;
    const uint32_t key = (ls << 8) | ld;
    const uint32_t p0 = lower_bound(a.plan_key, a.nplans, key);
    const uint32_t p1 = lower_bound(a.plan_key, a.nplans, key + 1);
    if (p0 == p1) continue;
    uint32_t cnt = 0;
    for (uint32_t p = p0; p < p1; p++) {
      const Plan &pl = a.plans[p];
      cnt += bit_of(a.bm, a.words, pl.sig[0], s) && bit_of(a.bm, a.words, pl.sig[1], d);
    }
    if (cnt == 0) continue;
    // d in N(s): the direction the DFS also tests (w in the list of an
    // earlier image); the two coincide whenever the graph is symmetric
    if (!contains(a.g.pool + a.g.start[s], a.g.deg[s], d)) continue;
    const uint32_t base = atomicAdd(&a.ctl->nseeds.v, cnt);
    if (base + cnt > a.seed_cap) {
      atomicAdd(&a.ctl->seed_overflow.v, cnt);
      continue;
    }
    atomicAdd(&a.ctl->work_left.v, (int)cnt);
    uint32_t o = base;
    for (uint32_t p = p0; p < p1; p++) {
      const Plan &pl = a.plans[p];
      if (bit_of(a.bm, a.words, pl.sig[0], s) && bit_of(a.bm, a.words, pl.sig[1], d)) {
        a.seed_plan[o] = p;
        a.seed_s[o] = s;
        a.seed_d[o] = d;
        a.seed_tag[o] = a.tag;
        o++;
      }
    }
  }
}

__global__ void k_ctl_reset(DfsCtl *ctl, ApplyCtl *actl, unsigned long long *tuple_count,
                            unsigned long long *counts, uint32_t ncounts) {
  for (uint32_t i = threadIdx.x; counts && i < ncounts; i += blockDim.x) counts[i] = 0;
  if (threadIdx.x == 0) {
    if (tuple_count) *tuple_count = 0;
    if (ctl) {
      ctl->nseeds.v = 0;
      ctl->seed_head.v = 0;
      ctl->q_head.v = 0;
      ctl->q_tail.v = 0;
      ctl->work_left.v = 0;
      ctl->seed_overflow.v = 0;
      ctl->abort.v = 0;
      ctl->scanned.v = 0;
      ctl->loads.v = 0;
      ctl->mbytes.v = 0;
    }
    if (actl) {
      actl->scratch_used = 0;
      actl->overflow = 0;
      actl->merge_overflow = 0;
      actl->filter_entries = 0;
      actl->filter_bits = 0;
      actl->relocated = 0;
      actl->ndirty = 0;
      actl->nheavy = 0;
      actl->relabeled = 0;
      actl->nrelabel_big = 0;
      actl->nrelabel_huge = 0;
    }
  }
}

__global__ void k_scratch_reset(ApplyCtl *actl) {
  if (threadIdx.x == 0) actl->scratch_used = 0;
}

// -------------------------------------------------------------------- DFS

// Per-wave DFS state, in dynamic LDS: a fixed header, a window buffer for the
// cooperative intersections and, for each chunk level 2 .. n-2, 64 candidates
// with their prefetched list start and degree. Sized to the engine's largest
// query, so small queries keep high occupancy.
struct WaveHdr {
  uint32_t f[kMaxQ];         // data vertex of each position
  uint32_t fs[kMaxQ];        // its list start
  uint32_t fd[kMaxQ];        // its degree
  uint32_t srcj[kMaxQ];      // backward position whose list drives the scan
  uint32_t pos[kMaxQ];       // next source index
  uint32_t end[kMaxQ];       // one past the last source index
  uint32_t cn[kMaxQ];        // chunk size
  uint32_t ci[kMaxQ];        // next chunk entry
  uint32_t cbase[kMaxQ];     // source index of the chunk's first lane
  uint64_t cmask[kMaxQ];     // lanes of the chunk that passed (k-th set bit = cand[k])
  const uint32_t *src[kMaxQ]; // the list (or label bucket) the level scans
  uint16_t sig[kMaxQ];
  uint16_t back[kMaxQ];
  uint16_t inj[kMaxQ];
  uint8_t qv[kMaxQ];
  uint8_t lix[kMaxQ];
  uint32_t n, query, plan, tag;
  uint32_t fstate, fn;       // leaf set F of the current positions 0..n-3 (see leaf_count)
};
constexpr uint32_t kFCap = 2048;  // largest leaf set F kept in LDS
// F is indexed by id range: bucket b = id >> DfsArgs::fshift (fshift makes the
// graph's ids fit kFIdx buckets), fidx[b] = first entry of F at or past the
// bucket. A membership test reads two offsets and binary-searches only the
// bucket's entries (a few, ids being spread over the range) instead of
// log2 |F| dependent LDS reads over all of F.
constexpr uint32_t kFIdx = 512;
// One level above the leaf, the candidates of a 64-entry window count their
// leaf images inline until their summed scan or search work passes
// DfsArgs::leaf_budget entries; the candidate that passes it descends to the
// leaf level instead, where the scan can be split with idle waves (an inline
// count cannot), and the window resumes after it.
// states of the leaf set F
constexpr uint32_t kFStale = 0;  // not computed for the current positions
constexpr uint32_t kFReady = 1;  // fbuf[0, fn) holds F
constexpr uint32_t kFLarge = 2;  // the shortest fixed list exceeds kFCap: per-candidate descent
constexpr uint32_t kFNone = 3;   // the leaf's only backward neighbour is position n-2

struct WaveMem {
  WaveHdr *h;
  uint32_t *win;    // [kWin]
  uint32_t *cand;   // [levels][64]
  uint32_t *cst;    // [levels][64] list start of each candidate
  uint32_t *cdg;    // [levels][64] degree of each candidate
  uint32_t *fbuf;   // [kFCap] leaf set F, ascending
  uint16_t *fidx;   // [kFIdx + 1] bucket offsets into fbuf
};

__host__ __device__ constexpr uint32_t wave_lds_bytes(uint32_t chunk_levels) {
  return (uint32_t)((sizeof(WaveHdr) + 15) / 16 * 16) + kWin * 4 + kFCap * 4 + chunk_levels * 64 * 12 +
         ((kFIdx + 1) * 2 + 15) / 16 * 16;
}

// 4 B per entry of the full lists of the positions in `mask`: what a merge
// intersection over those lists reads (the work measure of the expansion).
__device__ __forceinline__ uint64_t list_bytes(const WaveHdr &S, uint32_t mask) {
  uint64_t b = 0;
  while (mask) {
    const uint32_t j = __builtin_ctz(mask);
    mask &= mask - 1;
    b += 4ull * S.fd[j];
  }
  return b;
}

// Among the lists of the positions in `mask`, the one with the fewest entries
// of label index li: for a vertex of degree > kLabelSorted, its label bucket in
// the label-sorted copy; otherwise the whole list. Lane j examines position j;
// a wave minimum picks the shortest, lowest position on ties. Wave-uniform
// results; an empty mask gives length 0.
__device__ __forceinline__ void pick_source(const WaveHdr &S, const DevGraph &g, uint32_t mask, uint32_t li,
                                            const uint32_t *&src, uint32_t &len, uint32_t &j) {
  const uint32_t lane = lane_id();
  uint64_t key = ~0ull;
  uint32_t off = 0, sorted = 0;
  if (lane < (uint32_t)kMaxQ && (mask >> lane & 1)) {
    const uint32_t d = S.fd[lane];
    uint32_t n = d;
    off = S.fs[lane];
    if (d > kLabelSorted) {
      const uint32_t *row = g.loff + (size_t)S.f[lane] * (g.nbuckets + 1);
      const uint32_t b0 = row[li], b1 = row[li + 1];
      n = b1 - b0;
      off += b0;
      sorted = 1;
    }
    key = ((uint64_t)n << 32) | lane;
  }
  for (int o = 32; o >= 1; o >>= 1) {
    const uint64_t y = __shfl_xor(key, o);
    key = y < key ? y : key;
  }
  j = key == ~0ull ? 0u : (uint32_t)(key & 0xFFFFFFFFu);
  len = key == ~0ull ? 0u : (uint32_t)(key >> 32);
  const uint32_t boff = __shfl(off, (int)j);
  const uint32_t bsorted = __shfl(sorted, (int)j);
  src = (bsorted ? g.lpool : g.pool) + boff;
}

// Sets up level l to scan source indices [lo, min(hi, length)) of the shortest
// backward list (pick_source); no backward neighbour leaves it empty.
__device__ __forceinline__ void init_level(WaveHdr &S, const DevGraph &g, uint32_t l, uint32_t lo, uint32_t hi) {
  const uint32_t *src;
  uint32_t best, bj;
  pick_source(S, g, S.back[l], S.lix[l], src, best, bj);
  wave_sync();
  if (lane_id() == 0) {
    S.srcj[l] = bj;
    S.src[l] = src;
    S.pos[l] = lo;
    S.end[l] = hi < best ? hi : best;
    S.cn[l] = 0;
    S.ci[l] = 0;
  }
  wave_sync();
}

// Tests source entries [p, min(p + 64, e)) of `src`, one per lane: the
// candidate bit of signature `sig`, injectivity against the positions in
// `inj` and membership in the list of every position in `oth`. The bit is
// probed first (it narrows the membership windows) or, with bit_probe_last,
// only for the entries that passed the membership tests (the bitmap words are
// scattered reads; the windows are contiguous). The source chunk is sorted, so the entries still
// alive span [wmin, wmax]; the wave finds that window of the other list
// cooperatively (bounds2) and, when it holds at most kWin entries, stages it
// in LDS for the lanes' searches.
__device__ __forceinline__ bool test_entries(const WaveMem &M, const DfsArgs &a, const uint32_t *src, uint32_t oth,
                                             uint32_t inj, uint32_t sig, uint32_t p, uint32_t e, uint32_t &w,
                                             uint32_t &loads) {
  WaveHdr &S = *M.h;
  const uint32_t lane = lane_id();
  const uint32_t i = p + lane;
  bool ok = i < e;
  w = ok ? src[i] : 0xFFFFFFFFu;
  if (!a.bit_probe_last) ok = ok && bit_of(a.bm, a.words, sig, w);
  while (inj) {
    const uint32_t j = __builtin_ctz(inj);
    inj &= inj - 1;
    ok = ok && S.f[j] != w;
  }
  while (oth) {
    const uint32_t j = __builtin_ctz(oth);
    oth &= oth - 1;
    if (ballot(ok) == 0) break;
    ok = window_member(a.g.pool + S.fs[j], S.fd[j], w, ok, M.win, loads);
  }
  if (a.bit_probe_last) ok = ok && bit_of(a.bm, a.words, sig, w);
  return ok;
}

// The level's own chunk test: source = its chosen list, membership in the
// other backward lists, injectivity and signature of level l.
__device__ __forceinline__ bool test_chunk(const WaveMem &M, const DfsArgs &a, uint32_t l, uint32_t p,
                                           uint32_t e, uint32_t &w, uint32_t &loads) {
  const WaveHdr &S = *M.h;
  return test_entries(M, a, S.src[l], S.back[l] & ~(1u << S.srcj[l]), S.inj[l], S.sig[l], p, e, w, loads);
}

// Leaf set F for positions 0..n-3 (leaf = n-1): the images the leaf may take
// judging only by its backward neighbours other than position n-2, i.e. the
// entries of the intersection of their lists (label bucket of the shortest)
// that carry the leaf's signature bit and differ from the earlier images of
// its label. F is sorted. Sets fstate: kFReady (F in fbuf), kFLarge (shortest
// such list longer than kFCap) or kFNone (no such neighbour).
__device__ void compute_leaf_set(const WaveMem &M, const DfsArgs &a, uint32_t leaf, uint32_t &scanned,
                                 uint32_t &loads) {
  WaveHdr &S = *M.h;
  const uint32_t lane = lane_id();
  const uint32_t pv = leaf - 1;
  const uint32_t bp = S.back[leaf] & ~(1u << pv);
  uint32_t state = kFReady, cnt = 0;
  if (bp == 0) {
    state = kFNone;
  } else {
    const uint32_t *src;
    uint32_t len, j;
    pick_source(S, a.g, bp, S.lix[leaf], src, len, j);
    if (len > kFCap) {
      state = kFLarge;
    } else {
      const uint32_t oth = bp & ~(1u << j);
      const uint32_t inj = S.inj[leaf] & ~(1u << pv);
      const uint32_t sig = S.sig[leaf];
      for (uint32_t p = 0; p < len; p += 64) {
        uint32_t w;
        const bool ok = test_entries(M, a, src, oth, inj, sig, p, len, w, loads);
        scanned += p + lane < len ? 1u : 0u;
        const uint64_t m = ballot(ok);
        if (ok) M.fbuf[cnt + mbcnt(m)] = w;
        cnt += (uint32_t)__popcll(m);
      }
    }
  }
  wave_sync();
  if (state == kFReady) {
    // bucket offsets: entry i opens every bucket after its predecessor's up
    // to its own; buckets past the last entry start at cnt
    const uint32_t sh = a.fshift;
    for (uint32_t i = lane; i < cnt; i += 64) {
      const uint32_t b = M.fbuf[i] >> sh;
      const uint32_t pb = i ? (M.fbuf[i - 1] >> sh) + 1 : 0u;
      for (uint32_t k = pb; k <= b; k++) M.fidx[k] = (uint16_t)i;
    }
    const uint32_t last = cnt ? (M.fbuf[cnt - 1] >> sh) + 1 : 0u;
    for (uint32_t k = last + lane; k <= kFIdx; k += 64) M.fidx[k] = (uint16_t)cnt;
  }
  wave_sync();
  if (lane == 0) {
    S.fstate = state;
    S.fn = cnt;
  }
  wave_sync();
}

// y in F (state kFReady), through the bucket index.
__device__ __forceinline__ bool in_leaf_set(const WaveMem &M, const DfsArgs &a, uint32_t y) {
  const uint32_t b = y >> a.fshift;
  const uint32_t lo = M.fidx[b], hi = M.fidx[b + 1];
  return lds_contains(M.fbuf + lo, hi - lo, y);
}

// Leaf images of the partial match whose position n-2 holds w (this lane's
// candidate), given F. With the leaf not adjacent to position n-2 the count is
// |F|. Otherwise it scans w's list (its label bucket when sorted copies exist):
// with F, the entries that lie in F; with no other backward neighbour
// (kFNone), the entries that carry the leaf's signature and differ from the
// earlier images of its label. Injectivity against w itself applies when the
// leaf has w's label. A lane whose scan would take more than a.lane_work
// steps leaves it to the wave (leaf_count_wave) and reports its list instead.

struct LeafList {
  const uint32_t *L;
  uint32_t len;
  bool by_f;  // search F's elements in L instead of scanning L
};

__device__ __forceinline__ bool leaf_accept(const WaveHdr &S, const DfsArgs &a, const WaveMem &M, uint32_t leaf,
                                            uint32_t st, uint32_t fn, bool injw, uint32_t w, uint32_t y) {
  if (injw && y == w) return false;
  if (st == kFReady) return in_leaf_set(M, a, y);
  if (!bit_of(a.bm, a.words, S.sig[leaf], y)) return false;
  uint32_t m = S.inj[leaf] & ~(1u << (leaf - 1));
  while (m) {
    const uint32_t j = __builtin_ctz(m);
    m &= m - 1;
    if (S.f[j] == y) return false;
  }
  return true;
}

__device__ __forceinline__ uint32_t leaf_count_lane(const WaveMem &M, const DfsArgs &a, uint32_t leaf, uint32_t w,
                                                    bool &defer, bool &scan, LeafList &out, uint32_t &scanned,
                                                    uint32_t &loads, uint32_t &wdeg) {
  const WaveHdr &S = *M.h;
  const uint32_t pv = leaf - 1;
  const uint32_t back = S.back[leaf];
  const bool injw = (S.inj[leaf] >> pv) & 1;
  const uint32_t st = S.fstate, fn = S.fn;
  defer = false;
  if (!((back >> pv) & 1)) {
    uint32_t c = fn;
    if (injw && in_leaf_set(M, a, w)) c--;
    return c;
  }
  const uint32_t ws = a.g.start[w], wd = a.g.deg[w];
  wdeg = wd;
  const uint32_t *L = a.g.pool + ws;
  uint32_t len = wd;
  if (wd > kLabelSorted) {
    const uint32_t *row = a.g.loff + (size_t)w * (a.g.nbuckets + 1);
    const uint32_t li = S.lix[leaf];
    const uint32_t b0 = row[li], b1 = row[li + 1];
    L = a.g.lpool + ws + b0;
    len = b1 - b0;
    loads += 2;
  }
  const uint32_t lg = 33u - (uint32_t)__builtin_clz(len | 1u);
  const bool by_f = st == kFReady && (uint64_t)fn * lg < len;
  if (!by_f && a.leaf_segments) {
    // scanned with every other lane's list as one segmented array (leaf_scan_segments)
    scan = true;
    out.L = L;
    out.len = len;
    out.by_f = false;
    return 0;
  }
  const uint32_t work = by_f ? fn * lg : len;
  if (work > a.lane_work) {
    defer = true;
    out.L = L;
    out.len = len;
    out.by_f = by_f;
    return 0;
  }
  uint32_t c = 0;
  if (by_f) {
    for (uint32_t k = 0; k < fn; k++) {
      const uint32_t x = M.fbuf[k];
      c += (!(injw && x == w) && contains(L, len, x)) ? 1u : 0u;
    }
    loads += fn * lg;
  } else {
    for (uint32_t k = 0; k < len; k++) c += leaf_accept(S, a, M, leaf, st, fn, injw, w, L[k]) ? 1u : 0u;
    scanned += len;
  }
  return c;
}

// The lists of the lanes with `has`, scanned as one segmented array: entry k
// of the concatenation is read by lane k mod 64, which finds the list it falls
// in by a binary search over the lanes' exclusive prefix sums (shuffles). Short
// lists thus share each 64-entry step instead of keeping their lanes in a
// serial loop, and consecutive entries of a list are read by consecutive
// lanes. Returns this lane's share of the count (the task sums the wave).
__device__ __forceinline__ uint32_t leaf_scan_segments(const WaveMem &M, const DfsArgs &a, uint32_t leaf, uint32_t w,
                                                       bool has, const LeafList &ll, uint32_t &scanned) {
  const WaveHdr &S = *M.h;
  const uint32_t lane = lane_id();
  const bool injw = (S.inj[leaf] >> (leaf - 1)) & 1;
  const uint32_t st = S.fstate, fn = S.fn;
  const uint32_t n = has ? ll.len : 0u;
  uint32_t incl = n;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if (lane >= (uint32_t)o) incl += y;
  }
  const uint32_t total = __shfl(incl, 63);
  const uint32_t excl = incl - n;
  const unsigned long long lp = has ? (unsigned long long)ll.L : 0ull;
  uint32_t c = 0;
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t k = base + lane;
    // the owner is the last lane whose list starts at or before k
    uint32_t o = 0;
    for (uint32_t step = 32; step; step >>= 1) {
      const uint32_t cand = o + step;
      if (__shfl(excl, (int)cand) <= k) o = cand;
    }
    const uint32_t start = __shfl(excl, (int)o);
    const uint32_t *L = (const uint32_t *)__shfl(lp, (int)o);
    const uint32_t wo = __shfl(w, (int)o);
    if (k < total) {
      c += leaf_accept(S, a, M, leaf, st, fn, injw, wo, L[k - start]) ? 1u : 0u;
      scanned++;
    }
  }
  return c;
}

// The deferred lanes' lists, one after another, 64 entries (or F elements) at
// a time across the wave; returns this lane's share of the count.
__device__ __forceinline__ uint32_t leaf_count_wave(const WaveMem &M, const DfsArgs &a, uint32_t leaf, uint32_t w,
                                                    bool defer, const LeafList &ll, uint32_t &scanned,
                                                    uint32_t &loads) {
  const WaveHdr &S = *M.h;
  const uint32_t lane = lane_id();
  const bool injw = (S.inj[leaf] >> (leaf - 1)) & 1;
  const uint32_t st = S.fstate, fn = S.fn;
  uint32_t c = 0;
  uint64_t dm = ballot(defer);
  while (dm) {
    const int src = (int)__builtin_ctzll(dm);
    dm &= dm - 1;
    const uint32_t *L = (const uint32_t *)__shfl((unsigned long long)ll.L, src);
    const uint32_t len = __shfl(ll.len, src);
    const uint32_t ww = __shfl(w, src);
    const bool by_f = __shfl((uint32_t)ll.by_f, src) != 0;
    if (by_f) {
      const uint32_t lg = 33u - (uint32_t)__builtin_clz(len | 1u);
      for (uint32_t k = lane; k < fn; k += 64) {
        const uint32_t x = M.fbuf[k];
        c += (!(injw && x == ww) && contains(L, len, x)) ? 1u : 0u;
        loads += lg;
      }
    } else {
      for (uint32_t k = lane; k < len; k += 64) {
        c += leaf_accept(S, a, M, leaf, st, fn, injw, ww, L[k]) ? 1u : 0u;
        scanned++;
      }
    }
  }
  return c;
}

__device__ void emit_tuples(const WaveHdr &S, const DfsArgs &a, uint32_t l, uint32_t w, bool ok) {
  const uint64_t m = ballot(ok);
  if (m == 0) return;
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(a.tuple_count, (unsigned long long)__popcll(m));
  base = bcast64(base);
  if (!ok) return;
  const unsigned long long idx = base + mbcnt(m);
  if (idx >= a.tuple_cap) return;
  uint32_t *t = a.tuples + idx * (kMaxQ + 2);
  t[0] = S.query;
  t[1] = S.tag;
  for (uint32_t k = S.n; k < (uint32_t)kMaxQ; k++) t[2 + k] = 0xFFFFFFFFu;
  for (uint32_t j = 0; j < l; j++) t[2 + S.qv[j]] = S.f[j];
  t[2 + S.qv[l]] = w;
}

// Hands half of some remaining work to the queue when waves are waiting.
//
// Levels shallower than the current one give away work when enough entries
// remain there (fetched-but-unconsumed chunk entries count: their source
// indices are recovered from the chunk mask): two, when the level is two or
// more above the leaf; sixteen one level above it. The level being scanned
// gives away work only when at least donate_min source entries remain. The
// receiver re-fetches and re-filters the donated source range [mid, end).
// Hands ranges of the wave's unscanned work to waiting waves: up to
// min(waiting, kDonateBurst) of them per call, each half of what is left at
// the shallowest level that still holds enough, so that when most of the grid
// runs dry at once the work spreads in fewer check periods.
constexpr uint32_t kDonateBurst = DSM_DONATE_BURST;

__device__ __forceinline__ bool donate_one(WaveHdr &S, const DfsArgs &a, uint32_t L0, uint32_t l, uint32_t n) {
  uint32_t lv = 0xFFFFFFFFu, mid = 0, keep_cn = 0;
  for (uint32_t k = L0; k <= l; k++) {
    const uint32_t p = S.pos[k], e = S.end[k];
    const uint32_t unf = e > p ? e - p : 0;
    if (k == l) {
      // entries one level above the leaf carry a leaf count each
      const uint32_t dmin = k + 2 == n ? a.donate_min_n2 : a.donate_min;
      if (unf >= dmin) {
        lv = k;
        mid = p + unf / 2;
      }
      break;
    }
    const uint32_t ci = S.ci[k], cn = S.cn[k];
    const uint32_t rem = cn - ci;
    // an entry two or more levels above the leaf roots a subtree worth a
    // hand-off; one level above, only a leaf scan
    const uint32_t need = (k + 3 <= n) ? 2u : 16u;
    if (rem + unf < need) continue;
    if (unf >= 2) {
      lv = k;
      mid = p + unf / 2;
      keep_cn = cn;
      break;
    }
    {
      // split the unconsumed chunk: keep [ci, ci+half), donate from the
      // source index of entry ci+half to the end of the range
      const uint32_t half = (rem + unf) / 2;
      const uint32_t kth = ci + half;
      if (half == 0 || kth >= cn) continue;
      uint64_t m = S.cmask[k];
      for (uint32_t i = 0; i < kth; i++) m &= m - 1;
      lv = k;
      mid = S.cbase[k] + (uint32_t)__builtin_ctzll(m);
      keep_cn = kth;
      break;
    }
  }
  if (lv == 0xFFFFFFFFu) return false;
  const uint32_t e = S.end[lv];
  uint32_t slot = 0, ok = 0;
  if (lane_id() == 0) {
    const uint32_t t = atomicAdd(&a.ctl->q_tail.v, 1u);
    if (t < a.qcap) {
      ok = 1;
      slot = t;
      atomicAdd(&a.ctl->work_left.v, 1);
    }
  }
  ok = bcast(ok);
  if (!ok) return false;
  slot = bcast(slot);
  const uint32_t lane = lane_id();
  uint32_t word = 0;
  if (lane == 0) word = S.plan;
  else if (lane == 1) word = lv;
  else if (lane == 2) word = mid;
  else if (lane == 3) word = e;
  else if (lane < 4u + kMaxQ) word = S.f[lane - 4];
  else if (lane == 4u + kMaxQ) word = S.tag;
  if (lane < (uint32_t)kTaskWords) reinterpret_cast<uint32_t *>(&a.queue[slot])[lane] = word;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (lane == 0) {
    st_release(&a.queue[slot].ready, a.epoch);
    // the donor keeps source indices below mid; after a split inside the
    // chunk, mid <= pos, so nothing is left to fetch at this level and a later
    // split there donates [mid', mid) only
    if (lv != l && keep_cn != S.cn[lv]) S.cn[lv] = keep_cn;
    S.end[lv] = mid;
  }
  wave_sync();
  return true;
}

__device__ void try_donate(WaveHdr &S, const DfsArgs &a, uint32_t L0, uint32_t l, uint32_t n) {
  // waves waiting for work = tickets taken - slots reserved
  uint32_t waiting = 0;
  if (lane_id() == 0) {
    const uint32_t h = ld_relaxed(&a.ctl->q_head.v), t = ld_relaxed(&a.ctl->q_tail.v);
    waiting = (h > t && t < a.qcap) ? h - t : 0u;
  }
  waiting = bcast(waiting);
  for (uint32_t k = 0; k < waiting && k < kDonateBurst; k++)
    if (!donate_one(S, a, L0, l, n)) break;
}

template <int WPB, bool TUP>
__global__ __launch_bounds__(WPB * 64) void k_dfs(DfsArgs a, uint32_t chunk_levels) {
  extern __shared__ __align__(16) uint8_t s_dyn[];
  const uint32_t lane = lane_id();
  WaveMem M;
  {
    uint8_t *base = s_dyn + (threadIdx.x >> 6) * wave_lds_bytes(chunk_levels);
    M.h = reinterpret_cast<WaveHdr *>(base);
    M.win = reinterpret_cast<uint32_t *>(base + (sizeof(WaveHdr) + 15) / 16 * 16);
    M.cand = M.win + kWin;
    M.cst = M.cand + chunk_levels * 64;
    M.cdg = M.cst + chunk_levels * 64;
    M.fbuf = M.cdg + chunk_levels * 64;
    M.fidx = reinterpret_cast<uint16_t *>(M.fbuf + kFCap);
  }
  WaveHdr &S = *M.h;
  if (lane == 0) S.plan = 0xFFFFFFFFu;
  wave_sync();
  // a seed table overflow leaves unwritten slots: match nothing, the host
  // reports the error (the table is sized so that it cannot happen)
  if (a.ctl->seed_overflow.v) return;
  const uint32_t nseeds = a.ctl->nseeds.v;
  // Seeds are claimed from a counter, never assigned by wave index: the grid
  // is sized to what the occupancy query says fits, but nothing guarantees
  // that every block is resident at once, and a seed reserved for a block
  // that has not started would keep the resident waves waiting for it.
  bool seeds_left = true;
  // Donated tasks are handed out by ticket: a wave out of work takes the next
  // ticket with one atomic add and then waits on the ready flag inside that
  // queue slot (a 128-byte line of its own), so publishing a task wakes exactly
  // one wave and idle waves never contend on a shared counter.
  uint32_t ticket = 0xFFFFFFFFu;
  uint32_t naps = 0;            // polls of the current ticket (backoff)
  uint64_t idle_since = 0;      // wall clock when the wave ran out of work (watchdog)
  uint64_t busy = 0;            // wall-clock ticks spent on tasks
  uint64_t first_t = 0, last_t = 0;  // start of the first task, end of the last
  for (;;) {
    // lane 0 claims work; every decision is then broadcast so that all
    // control flow below stays wave-uniform
    uint32_t got = 0, slot = 0, took = 0;
    if (lane == 0) {
      if (seeds_left) {
        const uint32_t i = atomicAdd(&a.ctl->seed_head.v, 1u);
        if (i < nseeds) {
          got = 2;
          slot = i;
        } else {
          seeds_left = false;
        }
      }
      if (!got) {
        if (ticket == 0xFFFFFFFFu) {
          ticket = atomicAdd(&a.ctl->q_head.v, 1u);
          took = 1;
        }
        // relaxed poll: an acquire here would invalidate the CU's L1 on every
        // poll of every waiting wave; the acquire fence follows a hit instead
        if (ticket < a.qcap && ld_relaxed(&a.queue[ticket].ready) == a.epoch) {
          got = 1;
          slot = ticket;
          ticket = 0xFFFFFFFFu;
        }
      }
    }
    got = bcast(got);
    slot = bcast(slot);
    if (bcast(took)) {
      naps = 0;
      idle_since = wall_clock64();
    }
    if (got == 0) {
      // waiting for the ticket's task: exponential backoff; the termination
      // counter, the abort flag and the watchdog every eighth poll
      if ((naps & 7) == 7) {
        int32_t left = 0;
        uint32_t ab = 0, expired = 0;
        if (lane == 0) {
          left = ld_relaxed(&a.ctl->work_left.v);
          ab = ld_relaxed(&a.ctl->abort.v);
          expired = wall_clock64() - idle_since > a.watchdog;
          if (expired && left > 0) st_release(&a.ctl->abort.v, 1u);
        }
        left = (int32_t)bcast((uint32_t)left);
        ab = bcast(ab);
        expired = bcast(expired);
        if (left <= 0 || ab || expired) break;
      }
      const uint32_t reps = (1u << (naps < 4 ? naps : 4u)) < a.poll_max ? (1u << (naps < 4 ? naps : 4u)) : a.poll_max;
      for (uint32_t r = 0; r < reps; r++) __builtin_amdgcn_s_sleep(8);
      naps++;
      continue;
    }
    // ---- load the task
    const uint64_t t_task = wall_clock64();
    if (first_t == 0) first_t = t_task;
    uint32_t plan, L0, lo, hi, tag, fw = 0;
    if (got == 1) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      const uint32_t word =
          lane < (uint32_t)kTaskWords ? reinterpret_cast<const uint32_t *>(&a.queue[slot])[lane] : 0u;
      plan = __builtin_amdgcn_readlane(word, 0);
      L0 = __builtin_amdgcn_readlane(word, 1);
      lo = __builtin_amdgcn_readlane(word, 2);
      hi = __builtin_amdgcn_readlane(word, 3);
      tag = __builtin_amdgcn_readlane(word, 4 + kMaxQ);
      fw = __shfl(word, (int)((lane + 4) & 63));
    } else {
      plan = a.seed_plan[slot];
      tag = a.seed_tag[slot];
      L0 = 2;
      lo = 0;
      hi = 0xFFFFFFFFu;
      fw = lane == 0 ? a.seed_s[slot] : a.seed_d[slot];
    }
    if (lane == 0) S.tag = tag;
    if (bcast(S.plan) != plan) {
      const Plan &P = a.plans[plan];
      if (lane < (uint32_t)kMaxQ) {
        S.sig[lane] = P.sig[lane];
        S.back[lane] = P.back[lane];
        S.inj[lane] = P.inj[lane];
        S.qv[lane] = P.qv[lane];
        S.lix[lane] = P.lix[lane];
      }
      if (lane == 0) {
        S.n = P.n;
        S.query = P.query;
        S.plan = plan;
      }
    }
    if (lane < L0) {
      S.f[lane] = fw;
      S.fs[lane] = a.g.start[fw];
      S.fd[lane] = a.g.deg[fw];
    }
    if (lane == 0) S.fstate = kFStale;
    wave_sync();
    const uint32_t n = bcast(S.n);
    uint64_t lcount = 0, mbytes = 0;
    uint32_t scanned = 0, loads = 0;  // per lane: source entries read, search loads
    if (L0 >= n) {
      // complete embedding (two-vertex query seeded by its only edge)
      if (lane == 0) lcount = 1;
      if (TUP) emit_tuples(S, a, n - 1, S.f[n - 1], lane == 0);
    } else {
      // a seed expands its partial match here; a donated task continues one
      // whose expansion the donor counted
      if (got == 2 && lane == 0) mbytes += list_bytes(S, S.back[L0]);
      init_level(S, a.g, L0, lo, hi);
      uint32_t l = L0;
      // highest position the leaf set F depends on (-1: none)
      int fdep = -1;
      if (n >= 4) {
        const uint32_t dep = (uint32_t)(S.back[n - 1] | S.inj[n - 1]) & ((1u << (n - 2)) - 1u);
        fdep = dep ? 31 - __builtin_clz(dep) : -1;
      }
      uint32_t tick = 0;
      for (;;) {
        if (l == n - 1) {
          // leaf: count every valid image in the remaining source range
          for (;;) {
            if (++tick >= a.donate_every) {
              tick = 0;
              try_donate(S, a, L0, l, n);
            }
            const uint32_t p = bcast(S.pos[l]), e = bcast(S.end[l]);
            if (p >= e) break;
            uint32_t w;
            const bool ok = test_chunk(M, a, l, p, e, w, loads);
            scanned += p + lane < e ? 1u : 0u;
            lcount += ok ? 1u : 0u;
            if (TUP) emit_tuples(S, a, l, w, ok);
            wave_sync();
            if (lane == 0) S.pos[l] = p + 64;
            wave_sync();
          }
          if (l == L0) break;
          l--;
          continue;
        }
        const uint32_t ci = bcast(S.ci[l]), cn = bcast(S.cn[l]);
        if (ci < cn) {
          const uint32_t c = (l - 2) * 64 + ci;
          const uint32_t w = bcast(M.cand[c]);
          const uint32_t ws = bcast(M.cst[c]);
          const uint32_t wd = bcast(M.cdg[c]);
          wave_sync();
          if (lane == 0) {
            S.ci[l] = ci + 1;
            S.f[l] = w;
            S.fs[l] = ws;
            S.fd[l] = wd;
            // F depends only on the positions of the leaf's backward and
            // same-label neighbours other than n-2: it stays valid while
            // deeper positions change
            if ((int)l <= fdep) S.fstate = kFStale;
          }
          wave_sync();
          l++;
          mbytes += lane == 0 ? list_bytes(S, S.back[l]) : 0ull;
          init_level(S, a.g, l, 0, 0xFFFFFFFFu);
          continue;
        }
        if (!TUP && a.leaf_sets && l + 2 == n && bcast(S.fstate) != kFLarge) {
          // one level above the leaf: each lane counts the leaf images of its
          // own candidate against the shared leaf set F, instead of a descent
          // per candidate
          bool large = false, descend = false;
          for (;;) {
            if (++tick >= a.donate_every) {
              tick = 0;
              try_donate(S, a, L0, l, n);
            }
            const uint32_t p = bcast(S.pos[l]), e = bcast(S.end[l]);
            if (p >= e) break;
            uint32_t w;
            const bool ok = test_chunk(M, a, l, p, e, w, loads);
            scanned += p + lane < e ? 1u : 0u;
            if (ballot(ok) == 0) {
              wave_sync();
              if (lane == 0) S.pos[l] = p + 64;
              wave_sync();
              continue;
            }
            if (bcast(S.fstate) == kFStale) compute_leaf_set(M, a, l + 1, scanned, loads);
            // backward lists of the leaf other than position n-2, read by a
            // merge for every candidate of this level
            const uint64_t fixed_leaf_bytes = list_bytes(S, S.back[l + 1] & ~(1u << l));
            if (bcast(S.fstate) == kFLarge) {
              // this chunk and the rest of the level take the descent path
              large = true;
              break;
            }
            bool defer = false, scan = false;
            LeafList ll{nullptr, 0, false};
            uint32_t wdeg = 0, lc = 0;
            if (ok) lc = leaf_count_lane(M, a, l + 1, w, defer, scan, ll, scanned, loads, wdeg);
            // inline work of the lanes up to this one; the first lane past the
            // budget descends, the lanes after it are tested again later
            uint32_t wk = 0;
            if (scan || defer) wk = ll.by_f ? S.fn * (33u - (uint32_t)__builtin_clz(ll.len | 1u)) : ll.len;
            uint32_t incl = wk;
            for (int o = 1; o < 64; o <<= 1) {
              const uint32_t y = __shfl_up(incl, o);
              if (lane >= (uint32_t)o) incl += y;
            }
            const uint64_t dm = ballot(wk > 0 && incl > a.leaf_budget);
            const uint32_t k0 = dm ? (uint32_t)__builtin_ctzll(dm) : 64u;
            if (lane >= k0) {
              lc = 0;
              scan = defer = false;
            } else if (ok) {
              mbytes += fixed_leaf_bytes + 4ull * wdeg;
            }
            lcount += lc;
            if (a.leaf_segments) lcount += leaf_scan_segments(M, a, l + 1, w, scan, ll, scanned);
            lcount += leaf_count_wave(M, a, l + 1, w, defer, ll, scanned, loads);
            if (dm) {
              // lane k0's candidate as a one-entry chunk at source index p + k0
              const uint32_t w0 = __shfl(w, (int)k0);
              wave_sync();
              if (lane == 0) {
                const uint32_t c = (l - 2) * 64;
                M.cand[c] = w0;
                M.cst[c] = a.g.start[w0];
                M.cdg[c] = a.g.deg[w0];
                S.cn[l] = 1;
                S.ci[l] = 0;
                S.cbase[l] = p + k0;
                S.cmask[l] = 1;
                S.pos[l] = p + k0 + 1;
              }
              wave_sync();
              descend = true;
              break;
            }
            wave_sync();
            if (lane == 0) S.pos[l] = p + 64;
            wave_sync();
          }
          if (descend) continue;
          if (!large) {
            if (l == L0) break;
            l--;
            continue;
          }
        }
        if (++tick >= a.donate_every) {
          tick = 0;
          try_donate(S, a, L0, l, n);
        }
        const uint32_t p = bcast(S.pos[l]), e = bcast(S.end[l]);
        if (p < e) {
          uint32_t w;
          const bool ok = test_chunk(M, a, l, p, e, w, loads);
          scanned += p + lane < e ? 1u : 0u;
          // prefetch the list of every survivor for the descent
          uint32_t ws = 0, wd = 0;
          if (ok) {
            ws = a.g.start[w];
            wd = a.g.deg[w];
          }
          const uint64_t m = ballot(ok);
          wave_sync();
          if (ok) {
            const uint32_t c = (l - 2) * 64 + mbcnt(m);
            M.cand[c] = w;
            M.cst[c] = ws;
            M.cdg[c] = wd;
          }
          if (lane == 0) {
            S.cn[l] = __popcll(m);
            S.ci[l] = 0;
            S.cbase[l] = p;
            S.cmask[l] = m;
            S.pos[l] = p + 64;
          }
          wave_sync();
          continue;
        }
        if (l == L0) break;
        l--;
      }
    }
    // wave sum of the lane counts
    uint64_t tot = lcount;
    uint64_t sc = scanned, ld = loads, mb = mbytes;
    for (int off = 32; off >= 1; off >>= 1) {
      tot += __shfl_xor(tot, off);
      sc += __shfl_xor(sc, off);
      ld += __shfl_xor(ld, off);
      mb += __shfl_xor(mb, off);
    }
    if (lane == 0) {
      const uint32_t tg = S.tag;
      if (tot)
        atomicAdd(&a.counts[((size_t)(tg >> 1) * a.nq + S.query) * 2 + (tg & 1)], (unsigned long long)tot);
      atomicAdd(&a.ctl->scanned.v, (unsigned long long)sc);
      atomicAdd(&a.ctl->loads.v, (unsigned long long)ld);
      atomicAdd(&a.ctl->mbytes.v, (unsigned long long)mb);
      atomicSub(&a.ctl->work_left.v, 1);
    }
    last_t = wall_clock64();
    busy += last_t - t_task;
    if (a.query_busy && lane == 0) atomicAdd(&a.query_busy[S.query], (unsigned long long)(last_t - t_task));
  }
  // time this wave spent on tasks (not waiting) and when its first task
  // started and its last ended, for the balance statistics
  const uint32_t slot_w = blockIdx.x * WPB + (threadIdx.x >> 6);
  if (a.wave_busy && lane == 0 && slot_w < a.wave_slots) {
    a.wave_busy[slot_w] = busy;
    a.wave_busy[a.wave_slots + slot_w] = first_t;
    a.wave_busy[2 * (size_t)a.wave_slots + slot_w] = last_t;
  }
}

// Copies every vertex's live list to its new slot (pool growth/compaction).
__global__ __launch_bounds__(256) void k_pool_move(const uint32_t *old_pool, uint32_t *new_pool,
                                                   const uint32_t *old_start, const uint32_t *new_start,
                                                   const uint32_t *deg, uint32_t n) {
  const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
  for (uint32_t v = blockIdx.x * 4 + wv; v < n; v += gridDim.x * 4) {
    const uint32_t d = deg[v];
    const uint32_t *src = old_pool + old_start[v];
    uint32_t *dst = new_pool + new_start[v];
    for (uint32_t j = lane; j < d; j += 64) dst[j] = src[j];
  }
}

template <int WPB>
int max_blocks(uint32_t chunk_levels) {
  int nb = 0;
  const size_t lds = (size_t)WPB * wave_lds_bytes(chunk_levels);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_dfs<WPB, false>, WPB * 64, lds) != hipSuccess) return 1;
  return nb > 0 ? nb : 1;
}

// ------------------------------------------------------ graph copy sync

// Bringing the second copy of the graph up to date before an overlapped
// update: instead of copying the whole graph, only what the previous update
// changed is copied from the up-to-date copy `x` to the stale one `y`: the
// touched vertices' slots, degrees and lists, the dirty vertices' label-sorted
// copies and bucket offsets, and the relabelled vertices' labels. Ranges of up
// to kSyncWave entries are copied by one wave; longer ones are queued
// (queue_big: a block each up to kBlockList entries, the whole grid beyond)
// for k_sync_big. A queue entry is the vertex with bit 31 set for a
// label-sorted copy.
constexpr uint32_t kSyncWave = 4096;

// mode 0: touched vertices (start, deg, cap and the pool slot); mode 1: dirty
// vertices (bucket offsets and the label-sorted copy of lists longer than
// kLabelSorted), except those whose bit is set in `skip`; mode 2: relabelled
// vertices (label).
__global__ __launch_bounds__(256) void k_sync_small(DevGraph x, DevGraph y, const uint32_t *list, const uint32_t *count,
                                                    uint32_t n_host, uint32_t mode, const uint32_t *skip, uint32_t *q,
                                                    uint32_t qcap, uint32_t *qctl) {
  const uint32_t lane = lane_id();
  const uint32_t total = count ? (*count < x.n ? *count : x.n) : n_host;
  for (uint32_t t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < total; t += (gridDim.x * blockDim.x) >> 6) {
    const uint32_t v = list[t];
    if (mode == 2) {
      if (lane == 0) y.lab[v] = x.lab[v];
      continue;
    }
    if (skip && ((skip[v >> 5] >> (v & 31)) & 1u)) continue;
    const uint32_t st = x.start[v], d = x.deg[v];
    if (mode == 0) {
      if (lane == 0) {
        y.start[v] = st;
        y.deg[v] = d;
        y.cap[v] = x.cap[v];
      }
    } else {
      if (d <= kLabelSorted) continue;
      const uint32_t nb = x.nbuckets + 1;
      if (lane < nb) y.loff[(size_t)v * nb + lane] = x.loff[(size_t)v * nb + lane];
    }
    const uint32_t *from = (mode == 0 ? x.pool : x.lpool) + st;
    uint32_t *to = (mode == 0 ? y.pool : y.lpool) + st;
    if (d <= kSyncWave) {
      copy_range(from, to, lane, 64, d);
    } else if (lane == 0) {
      queue_big(q, qcap, &qctl[0], &qctl[1], d, v | (mode << 31));
    }
  }
}

__global__ __launch_bounds__(256) void k_sync_big(DevGraph x, DevGraph y, const uint32_t *q, uint32_t qcap,
                                                  const uint32_t *qctl) {
  for_big_lists(q, qcap, qctl[0], qctl[1], [&](uint32_t e, uint32_t first, uint32_t stride) {
    const uint32_t v = e & 0x7FFFFFFFu;
    const bool sorted = e >> 31;
    const uint32_t st = x.start[v];
    copy_range((sorted ? x.lpool : x.pool) + st, (sorted ? y.lpool : y.pool) + st, first, stride, x.deg[v]);
  });
}

__global__ void k_sync_reset(uint32_t *qctl) {
  if (threadIdx.x < 2) qctl[threadIdx.x] = 0;
}

}  // namespace

hipError_t launch_apply_merge(const DevGraph &g, const BatchDev &b, unsigned long long scratch_cap,
                              uint32_t *scratch, ApplyCtl *actl, hipStream_t s) {
  if (b.ntouched == 0) return hipSuccess;
  const uint32_t blocks = (b.ntouched + kApplyWaves - 1) / kApplyWaves;
  hipLaunchKernelGGL(k_big_reset, dim3(1), dim3(64), 0, s, actl, 1, 0);
  hipLaunchKernelGGL(k_apply_merge, dim3(blocks < 4096 ? blocks : 4096), dim3(kApplyWaves * 64), 0, s, g, b,
                     scratch_cap, scratch, actl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_apply_merge_big, dim3(1024), dim3(256), 0, s, g, b, scratch, actl);
  return hipGetLastError();
}

hipError_t launch_apply_commit(const DevGraph &g, const BatchDev &b, const uint32_t *scratch,
                               unsigned long long pool_cap, ApplyCtl *actl, const DirtyArgs &d,
                               hipStream_t s, int only_failed) {
  if (b.ntouched == 0) return hipSuccess;
  const uint32_t blocks = (b.ntouched + kApplyWaves - 1) / kApplyWaves;
  hipLaunchKernelGGL(k_big_reset, dim3(1), dim3(64), 0, s, actl, 0, 1);
  hipLaunchKernelGGL(k_apply_commit, dim3(blocks < 4096 ? blocks : 4096), dim3(kApplyWaves * 64), 0, s, g,
                     b, scratch, pool_cap, actl, d, only_failed);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_apply_commit_big, dim3(1024), dim3(256), 0, s, g, b, scratch, actl);
  return hipGetLastError();
}

hipError_t launch_apply_commit(const DevGraph &g, const BatchDev &b, const uint32_t *scratch,
                               unsigned long long pool_cap, ApplyCtl *actl, const DirtyArgs &d,
                               hipStream_t s) {
  return launch_apply_commit(g, b, scratch, pool_cap, actl, d, s, 0);
}

hipError_t launch_apply_commit_failed(const DevGraph &g, const BatchDev &b, const uint32_t *scratch,
                                      unsigned long long pool_cap, ApplyCtl *actl, const DirtyArgs &d,
                                      hipStream_t s) {
  return launch_apply_commit(g, b, scratch, pool_cap, actl, d, s, 1);
}

hipError_t launch_dirty_compact(const DirtyArgs &d, uint32_t n, hipStream_t s) {
  const uint32_t nwords = (n + 31) / 32;
  const uint32_t blocks = (nwords + 255) / 256 < 4096 ? (nwords + 255) / 256 : 4096;
  hipLaunchKernelGGL(k_dirty_compact, dim3(blocks ? blocks : 1), dim3(256), 0, s, d, nwords);
  return hipGetLastError();
}

hipError_t launch_sync(const DevGraph &x, const DevGraph &y, const uint32_t *touched, uint32_t ntouched,
                       const uint32_t *dirty, const uint32_t *ndirty, const uint32_t *relabeled, uint32_t nrelabeled,
                       uint32_t *q, uint32_t qcap, uint32_t *qctl, hipStream_t s) {
  hipLaunchKernelGGL(k_sync_reset, dim3(1), dim3(64), 0, s, qctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint32_t *no_skip = nullptr;
  if (ntouched) {
    hipLaunchKernelGGL(k_sync_small, dim3((ntouched + 3) / 4 < 4096 ? (ntouched + 3) / 4 : 4096), dim3(256), 0, s, x, y,
                       touched, (const uint32_t *)nullptr, ntouched, 0u, no_skip, q, qcap, qctl);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (dirty) {
    hipLaunchKernelGGL(k_sync_small, dim3(4096), dim3(256), 0, s, x, y, dirty, ndirty, 0u, 1u, no_skip, q, qcap, qctl);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (nrelabeled) {
    hipLaunchKernelGGL(k_sync_small, dim3((nrelabeled + 3) / 4 < 4096 ? (nrelabeled + 3) / 4 : 4096), dim3(256), 0, s, x, y,
                       relabeled, (const uint32_t *)nullptr, nrelabeled, 2u, no_skip, q, qcap, qctl);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_sync_big, dim3(1024), dim3(256), 0, s, x, y, q, qcap, qctl);
  return hipGetLastError();
}

hipError_t launch_sync_sorted(const DevGraph &x, const DevGraph &y, const uint32_t *dirty, const uint32_t *ndirty,
                              const uint32_t *skip, uint32_t *q, uint32_t qcap, uint32_t *qctl, hipStream_t s) {
  hipLaunchKernelGGL(k_sync_reset, dim3(1), dim3(64), 0, s, qctl);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sync_small, dim3(4096), dim3(256), 0, s, x, y, dirty, ndirty, 0u, 1u, skip, q, qcap, qctl);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_sync_big, dim3(1024), dim3(256), 0, s, x, y, q, qcap, qctl);
  return hipGetLastError();
}

hipError_t launch_relabel(const DevGraph &g, const BatchDev &b, const FilterTables &ft,
                          const DirtyArgs &d, hipStream_t s) {
  if (b.nrelabel == 0) return hipSuccess;
  const uint32_t blocks = (b.nrelabel + 3) / 4;
  hipLaunchKernelGGL(k_relabel, dim3(blocks < 4096 ? blocks : 4096), dim3(256), 0, s, g, b, ft, d);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_relabel_big, dim3(1024), dim3(256), 0, s, g, b, d);
  return hipGetLastError();
}

hipError_t launch_filter_full(const DevGraph &g, const FilterTables &ft, hipStream_t s) {
  const uint32_t groups = (g.n + 63) / 64;
  const uint32_t blocks = (groups + 3) / 4;
  hipLaunchKernelGGL(k_filter_full, dim3(blocks < 8192 ? blocks : 8192), dim3(256), 0, s, g, ft);
  return hipGetLastError();
}

hipError_t launch_filter_dirty(const DevGraph &g, const FilterTables &ft, const DirtyArgs &d,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_filter_light, dim3(1024), dim3(256), 0, s, g, ft, d);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_filter_heavy, dim3(1024), dim3(256), 0, s, g, ft, d);
  return hipGetLastError();
}

hipError_t launch_lbuild(const DevGraph &g, const uint8_t *lmap, const uint32_t *list, const uint32_t *count,
                         const HugeSort &hs, hipStream_t s) {
  const uint32_t blocks = list ? 2048u : ((g.n + 3) / 4 < 8192 ? (g.n + 3) / 4 : 8192u);
  hipLaunchKernelGGL(k_huge_reset, dim3(1), dim3(64), 0, s, hs);
  hipLaunchKernelGGL(k_lbuild_pick, dim3(blocks), dim3(256), 0, s, g, list, count, hs);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lbuild, dim3(2048), dim3(256), 0, s, g, lmap, hs);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_lbuild_big, dim3(list ? 256u : 1024u), dim3(kLbuildWaves * 64), 0, s, g, lmap, list, count,
                     hs);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_lbuild_huge_count, dim3(512), dim3(kLbuildWaves * 64), 0, s, g, lmap, hs);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_lbuild_huge_scatter, dim3(512), dim3(kLbuildWaves * 64), 0, s, g, lmap, hs);
  return hipGetLastError();
}

hipError_t launch_seed(const SeedArgs &a, hipStream_t s) {
  if (a.nrec == 0 || a.nplans == 0) return hipSuccess;
  const uint32_t blocks = (a.nrec + 255) / 256;
  hipLaunchKernelGGL(k_seed, dim3(blocks < 4096 ? blocks : 4096), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_dfs(const DfsArgs &a, uint32_t blocks, uint32_t waves_per_block, uint32_t max_query_vertices,
                      hipStream_t s) {
  const uint32_t cl = max_query_vertices > 3 ? max_query_vertices - 3 : 1;
  const size_t per = wave_lds_bytes(cl);
  if (a.tuples) {
    hipLaunchKernelGGL((k_dfs<4, true>), dim3(blocks), dim3(256), 4 * per, s, a, cl);
    return hipGetLastError();
  }
  switch (waves_per_block) {
    case 1: hipLaunchKernelGGL((k_dfs<1, false>), dim3(blocks), dim3(64), per, s, a, cl); break;
    case 2: hipLaunchKernelGGL((k_dfs<2, false>), dim3(blocks), dim3(128), 2 * per, s, a, cl); break;
    case 8: hipLaunchKernelGGL((k_dfs<8, false>), dim3(blocks), dim3(512), 8 * per, s, a, cl); break;
    default: hipLaunchKernelGGL((k_dfs<4, false>), dim3(blocks), dim3(256), 4 * per, s, a, cl); break;
  }
  return hipGetLastError();
}

void round_node_params(RoundNodeArgs &st, uint32_t blocks, uint32_t waves_per_block,
                       uint32_t max_query_vertices, hipKernelNodeParams *reset, hipKernelNodeParams *seed0,
                       hipKernelNodeParams *seed1, hipKernelNodeParams *dfs) {
  st.no_actl = nullptr;
  st.no_tc = nullptr;
  st.no_counts = nullptr;
  st.no_ncounts = 0;
  st.reset_params[0] = &st.ctl;
  st.reset_params[1] = &st.no_actl;
  st.reset_params[2] = &st.no_tc;
  st.reset_params[3] = &st.no_counts;
  st.reset_params[4] = &st.no_ncounts;
  *reset = hipKernelNodeParams{};
  reset->func = reinterpret_cast<void *>(&k_ctl_reset);
  reset->gridDim = dim3(1);
  reset->blockDim = dim3(256);
  reset->sharedMemBytes = 0;
  reset->kernelParams = st.reset_params;
  reset->extra = nullptr;
  hipKernelNodeParams *sp[2] = {seed0, seed1};
  for (int i = 0; i < 2; i++) {
    st.seed_params[i][0] = &st.seed[i];
    const uint32_t b = st.seed[i].nrec ? (st.seed[i].nrec + 255) / 256 : 1u;
    *sp[i] = hipKernelNodeParams{};
    sp[i]->func = reinterpret_cast<void *>(&k_seed);
    sp[i]->gridDim = dim3(b < 4096 ? b : 4096);
    sp[i]->blockDim = dim3(256);
    sp[i]->sharedMemBytes = 0;
    sp[i]->kernelParams = st.seed_params[i];
    sp[i]->extra = nullptr;
  }
  st.chunk_levels = max_query_vertices > 3 ? max_query_vertices - 3 : 1;
  st.dfs_params[0] = &st.dfs;
  st.dfs_params[1] = &st.chunk_levels;
  const size_t per = wave_lds_bytes(st.chunk_levels);
  uint32_t wpb = st.dfs.tuples ? 4u : waves_per_block;
  void *fn;
  if (st.dfs.tuples) fn = reinterpret_cast<void *>(&k_dfs<4, true>);
  else if (wpb == 1) fn = reinterpret_cast<void *>(&k_dfs<1, false>);
  else if (wpb == 2) fn = reinterpret_cast<void *>(&k_dfs<2, false>);
  else if (wpb == 8) fn = reinterpret_cast<void *>(&k_dfs<8, false>);
  else {
    wpb = 4;
    fn = reinterpret_cast<void *>(&k_dfs<4, false>);
  }
  *dfs = hipKernelNodeParams{};
  dfs->func = fn;
  dfs->gridDim = dim3(blocks);
  dfs->blockDim = dim3(wpb * 64);
  dfs->sharedMemBytes = (unsigned int)(wpb * per);
  dfs->kernelParams = st.dfs_params;
  dfs->extra = nullptr;
}

uint32_t dfs_lds_bytes_per_wave(uint32_t max_query_vertices) {
  return wave_lds_bytes(max_query_vertices > 3 ? max_query_vertices - 3 : 1);
}

int dfs_max_blocks_per_cu(uint32_t waves_per_block, uint32_t max_query_vertices) {
  const uint32_t cl = max_query_vertices > 3 ? max_query_vertices - 3 : 1;
  switch (waves_per_block) {
    case 1: return max_blocks<1>(cl);
    case 2: return max_blocks<2>(cl);
    case 8: return max_blocks<8>(cl);
    default: return max_blocks<4>(cl);
  }
}

hipError_t launch_ctl_reset(DfsCtl *ctl, ApplyCtl *actl, unsigned long long *tuple_count,
                            unsigned long long *counts, uint32_t ncounts, hipStream_t s) {
  hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(256), 0, s, ctl, actl, tuple_count, counts, ncounts);
  return hipGetLastError();
}

hipError_t launch_scratch_reset(ApplyCtl *actl, hipStream_t s) {
  hipLaunchKernelGGL(k_scratch_reset, dim3(1), dim3(64), 0, s, actl);
  return hipGetLastError();
}

hipError_t launch_pool_move(const uint32_t *old_pool, uint32_t *new_pool, const uint32_t *old_start,
                            const uint32_t *new_start, const uint32_t *deg, uint32_t n, hipStream_t s) {
  const uint32_t blocks = (n + 3) / 4;
  hipLaunchKernelGGL(k_pool_move, dim3(blocks < 8192 ? blocks : 8192), dim3(256), 0, s, old_pool, new_pool,
                     old_start, new_start, deg, n);
  return hipGetLastError();
}

}  // namespace dsm
