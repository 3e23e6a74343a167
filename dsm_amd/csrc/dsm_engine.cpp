// Host side of the MI355X continuous subgraph matching engine: the C ABI of
// include/dsm.h over the kernels of dsm_kernels.hip.
//
// Per batch (dsm_apply_batch + dsm_match_batch):
//   1. delete records seed DFS roots in the pre-batch graph (negatives)
//   2. host: symmetrise, stable (src,dst) sort, final op per key, final label
//      per vertex (the update rule of cpuGraph.cpp:521-607)
//   3. device: merge/commit the per-vertex deltas in place, relabel, refresh
//      the candidate bitmaps of the dirty vertices
//   4. insert records seed DFS roots in the post-batch graph (positives)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dsm.h"
#include "dsm_internal.h"

using namespace dsm;

namespace {

thread_local std::string g_create_error;

}  // namespace

namespace dsm {
void set_global_error(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_create_error = buf;
}
}  // namespace dsm

namespace {

struct QueryHost {
  uint32_t n = 0;
  uint32_t lab[kMaxQ] = {};
  uint32_t adj[kMaxQ] = {};  // bitmask
  uint32_t order_variant = 0;  // rule of its matching orders (next_in_order)
};

constexpr uint32_t kOrderVariants = 4;

#ifndef DSM_DELTA_SYNC
#define DSM_DELTA_SYNC 1
#endif
// 1: an overlapped update brings the second copy of the graph up to date by
// copying only what the previous update changed (catch_up)
constexpr bool kDeltaSync = DSM_DELTA_SYNC;

// The next query vertex of a matching order, among the unplaced vertices with
// an earlier neighbour, by rule `variant` (ties: lowest id). An order never
// changes a count, only the work: the tuner picks the rule per query.
//   0: most earlier neighbours, then the rarest label in the data graph, then
//      the highest query degree;
//   1: most earlier neighbours, then the highest query degree, then the
//      rarest label;
//   2: the rarest label, then most earlier neighbours, then the highest degree;
//   3: most earlier neighbours, then the most frequent label, then the highest
//      query degree.
uint32_t next_in_order(const QueryHost &Q, uint32_t placed, const std::vector<uint64_t> &label_freq,
                       uint32_t variant) {
  int best = -1;
  int64_t bk[3] = {0, 0, 0};
  for (uint32_t w = 0; w < Q.n; w++) {
    if (placed >> w & 1) continue;
    const int64_t b = __builtin_popcount(Q.adj[w] & placed);
    if (b == 0) continue;
    const int64_t dg = __builtin_popcount(Q.adj[w]);
    const int64_t f = Q.lab[w] < label_freq.size() ? (int64_t)label_freq[Q.lab[w]] : 0;
    int64_t k[3];
    switch (variant) {
      case 1: k[0] = b; k[1] = dg; k[2] = -f; break;
      case 2: k[0] = -f; k[1] = b; k[2] = dg; break;
      case 3: k[0] = b; k[1] = f; k[2] = dg; break;
      default: k[0] = b; k[1] = -f; k[2] = dg; break;
    }
    if (best < 0 || k[0] > bk[0] || (k[0] == bk[0] && (k[1] > bk[1] || (k[1] == bk[1] && k[2] > bk[2])))) {
      best = (int)w;
      bk[0] = k[0];
      bk[1] = k[1];
      bk[2] = k[2];
    }
  }
  return (uint32_t)best;
}

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    reset();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e == hipSuccess) n = count;
    return e;
  }
  hipError_t ensure(size_t count) { return count <= n ? hipSuccess : alloc(count + count / 2); }
};

template <typename T>
struct PinnedBuf {
  T *p = nullptr;
  size_t n = 0;
  ~PinnedBuf() { reset(); }
  void reset() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t ensure(size_t count) {
    if (count <= n && p) return hipSuccess;
    reset();
    size_t c = count + count / 2 + 16;
    hipError_t e = hipHostMalloc(&p, c * sizeof(T), hipHostMallocDefault);
    if (e == hipSuccess) n = c;
    return e;
  }
};

uint8_t clamp_label(uint32_t l, const bool used[256]) { return (l < 255 && used[l]) ? (uint8_t)l : kNoLabel; }

}  // namespace

struct dsm_engine {
  dsm_options opt{};
  std::string err;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  uint32_t n = 0;
  std::vector<QueryHost> queries;
  bool label_used[256] = {};

  // candidate signatures and plans
  std::vector<SigReq> sigs;
  std::vector<uint32_t> sig_lo, sig_hi;  // 256
  std::vector<uint8_t> lmap;             // 256
  std::vector<std::vector<uint16_t>> qsig;  // per query, per query vertex
  std::vector<Plan> plans;
  std::vector<uint32_t> plan_key;
  uint32_t max_plans_per_key = 0;
  std::vector<uint64_t> label_freq;  // data vertices per label (orders the matching only)

  // device graph
  DevBuf<uint32_t> pool, start, deg, cap;
  DevBuf<uint8_t> lab;
  DevBuf<uint32_t> lpool, loff;  // label-sorted copies of long lists
  DevBuf<uint32_t> huge_list, huge_ctl, huge_hist;  // segmented builds of the longest lists
  DevBuf<uint32_t> big_list;                         // vertices whose copy a whole block builds
  DevBuf<uint32_t> mid_list;                         // vertices whose copy one wave builds
  uint32_t huge_segs = 0;
  uint32_t nbuckets = 1;
  unsigned long long pool_cap = 0;
  // filter
  DevBuf<uint64_t> bm;
  uint32_t words = 0;
  DevBuf<SigReq> d_sigs;
  DevBuf<uint32_t> d_sig_lo, d_sig_hi;
  DevBuf<uint8_t> d_lmap;
  DevBuf<Plan> d_plans;
  DevBuf<uint32_t> d_plan_key;
  // batch
  DevBuf<uint32_t> scratch;
  unsigned long long scratch_cap = 0;
  DevBuf<uint32_t> dirty_bits, dirty_list, heavy_list;
  DevBuf<uint32_t> bbuf;              // batch arrays (device)
  PinnedBuf<uint32_t> hbuf;           // batch arrays (pinned staging)
  DevBuf<uint32_t> ndeg, fail;
  DevBuf<unsigned long long> soff;
  DevBuf<ApplyCtl> actl;
  PinnedBuf<ApplyCtl> h_actl;
  // DFS
  DevBuf<DfsCtl> ctl;
  PinnedBuf<DfsCtl> h_ctl;
  DevBuf<Task> queue;
  uint32_t qcap = 0, epoch = 0;
  DevBuf<uint32_t> seed_plan, seed_s, seed_d, seed_tag;
  uint32_t seed_cap = 0;
  DevBuf<unsigned long long> counts;
  PinnedBuf<unsigned long long> h_counts;
  DevBuf<unsigned long long> tuple_count;
  uint32_t *tuples = nullptr;
  bool tuples_host = false;
  unsigned long long tuple_cap = 0;
  uint32_t dfs_blocks = 0, dfs_wpb = 4, max_qn = 2;
  uint32_t donate_min = 256, lane_work = 16, donate_every = 8, donate_min_n2 = 64, leaf_budget = 2048;
  bool leaf_sets = true;
  bool bit_probe_last = false;
  bool graph_rounds = true;
  PinnedBuf<uint32_t> pipe_h_rec;  // records of a pipelined round (dsm_process_batches)
  DevBuf<uint32_t> pipe_d_rec;
  bool leaf_segments = true;
  uint32_t poll_max = 16;
  // Second copy of the graph: a pipelined round matches on the engine's copy
  // while the next batch is applied to this one on stream2; then they swap.
  struct GraphCopy {
    DevBuf<uint32_t> pool, start, deg, cap, lpool, loff;
    DevBuf<uint8_t> lab;
    DevBuf<uint64_t> bm;
    DevBuf<ApplyCtl> actl;
    unsigned long long pool_cap = 0;
  } shadow;
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_pre = nullptr;  // all work on stream before a pipelined round
  bool overlap_apply = true;
  bool lpool_stale = false;  // the pool grew during the current update
  // The second copy of the graph differs from the current one only by the
  // last update, whose touched and relabelled vertices are kept here (its
  // dirty vertices stay on dirty_list): the next overlapped update copies just
  // those instead of the whole graph. Any other change clears it.
  bool shadow_delta_ok = false;
  bool grew = false;  // the pool grew during the update being applied
  std::vector<uint32_t> last_touched, last_relabel;
  DevBuf<uint32_t> d_last_touched, d_last_relabel, sync_q, sync_ctl;
  int cu_lds = 160 * 1024;   // LDS bytes per CU
  // per-wave task time of the last DFS launch (balance statistics)
  DevBuf<unsigned long long> wave_busy;
  PinnedBuf<unsigned long long> h_wave_busy;
  DevBuf<unsigned long long> query_busy;  // per query: DFS task ticks since the last read
  uint32_t wave_slots = 0, last_blocks = 0, last_wpb = 0;
  // the DFS round as a HIP graph (graph_round)
  hipGraph_t rg_graph = nullptr;
  hipGraphExec_t rg_exec = nullptr;
  hipGraphNode_t rg_reset = nullptr, rg_seed0 = nullptr, rg_seed1 = nullptr, rg_dfs = nullptr;
  uint32_t rg_key = 0;
  uint32_t watchdog_ms = 60000;
  int wall_khz = 100000;
  // records of the last batch (file order)
  std::vector<uint32_t> pos_s, pos_d, neg_s, neg_d;
  DevBuf<uint32_t> rec_neg, rec_pos;  // records on the device: src[k] | dst[k]
  bool batch_pending = false;
  hipEvent_t ev[12] = {};
  dsm_stats stats{};
  // cluster-encoding refinement (opt.cluster_filter): every query's position
  // rows follow the signature rows in bm; the DFS and the seeding read the
  // final positions
  bool cluster = false;
  uint32_t nsig_rows = 0, npos_rows = 0;
  std::vector<std::vector<uint32_t>> qref;  // per query, per vertex: the bitmap row the DFS uses
  std::vector<uint32_t> cstep_off;          // ops of pass step s: [cstep_off[s], cstep_off[s + 1])
  uint32_t cmax_step_ops = 0;
  DevBuf<ClusterOp> d_cops;
  DevBuf<uint32_t> d_crow, d_alive_src, cctl;
  uint32_t nalive_src = 0, nalive_rows = 0;
  DevBuf<uint64_t> calive;
  DevBuf<ClusterHeavy> cheavy;
  hipEvent_t ev_ref[2] = {};
  bool refine_timed = false;

  ~dsm_engine() {
    for (auto &x : ev_ref)
      if (x) (void)hipEventDestroy(x);
    if (rg_exec) (void)hipGraphExecDestroy(rg_exec);
    if (rg_graph) (void)hipGraphDestroy(rg_graph);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (tuples) {
      if (tuples_host) (void)hipHostFree(tuples);
      else (void)hipFree(tuples);
    }
    if (ev_pre) (void)hipEventDestroy(ev_pre);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int fail_msg(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }

  DevGraph dev_graph() {
    DevGraph g;
    g.n = n;
    g.pool = pool.p;
    g.start = start.p;
    g.deg = deg.p;
    g.cap = cap.p;
    g.lab = lab.p;
    g.lpool = lpool.p;
    g.loff = loff.p;
    g.nbuckets = nbuckets;
    return g;
  }
  FilterTables tables() {
    FilterTables t;
    t.sigs = d_sigs.p;
    t.sig_lo = d_sig_lo.p;
    t.sig_hi = d_sig_hi.p;
    t.lmap = d_lmap.p;
    t.nsig = (uint32_t)sigs.size();
    t.bm = bm.p;
    t.words = words;
    return t;
  }
  HugeSort huge_sort() {
    HugeSort h;
    h.list = huge_list.p;
    h.cap = (uint32_t)(huge_list.n / 2);
    h.ctl = huge_ctl.p;
    h.big = big_list.p;
    h.big_cap = (uint32_t)big_list.n;
    h.mid = mid_list.p;
    h.mid_cap = (uint32_t)mid_list.n;
    h.hist = huge_hist.p;
    h.segs = huge_segs;
    // the merge scratch (at least one entry per pool entry) is idle while the
    // label-sorted copies are rebuilt
    h.bucket_bytes = scratch_cap >= pool_cap ? reinterpret_cast<uint8_t *>(scratch.p) : nullptr;
    return h;
  }
  DirtyArgs dirty() {
    DirtyArgs d;
    d.bits = dirty_bits.p;
    d.list = dirty_list.p;
    d.heavy = heavy_list.p;
    d.list_cap = n;
    d.actl = actl.p;
    return d;
  }
};

#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess)                                                                     \
      return e->fail_msg(DSM_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                         __FILE__, __LINE__);                                                 \
  } while (0)

namespace {

// ------------------------------------------------------------ query set-up

// The refinement's passes for every query (dsm_cluster_plan.cpp), as device
// ops grouped by pass step: step s holds pass s of every query that has one.
// Rows: position p of query q is bitmap row nsig + qbase[q] + p; the alive row
// of query vertex u of q is abase[q] + u.
int build_cluster_ops(dsm_engine *e, const dsm_query_desc *qs, uint32_t nq) {
  std::vector<ClusterPlan> plans(nq);
  std::vector<uint32_t> qbase(nq), abase(nq);
  uint32_t npos = 0, nalive = 0, nsteps = 0;
  for (uint32_t q = 0; q < nq; q++) {
    if (build_cluster_plan(qs[q].num_vertices, qs[q].num_edges, qs[q].edge_src, qs[q].edge_dst, plans[q]))
      return e->fail_msg(DSM_ELIMIT, "query %u: its cluster plan exceeds the reference's limits (%u clusters, "
                         "%u small clusters per combine)", q, kMaxClusters, kMaxSmallClusters);
    qbase[q] = npos;
    abase[q] = nalive;
    npos += plans[q].positions;
    nalive += e->queries[q].n;
    nsteps = std::max<uint32_t>(nsteps, (uint32_t)plans[q].passes.size());
  }
  if ((uint64_t)e->nsig_rows + npos > 65535)
    return e->fail_msg(DSM_ELIMIT, "cluster positions of the queries exceed 65,535 bitmap rows");
  const uint32_t nsig = e->nsig_rows;
  auto row = [&](uint32_t q, uint32_t p) { return nsig + qbase[q] + p; };
  std::vector<std::vector<ClusterOp>> steps(nsteps);
  std::vector<uint32_t> crow;
  std::vector<bool> core_alive(nalive, false);
  for (uint32_t q = 0; q < nq; q++) {
    const ClusterPlan &P = plans[q];
    const QueryHost &Q = e->queries[q];
    for (size_t s = 0; s < P.passes.size(); s++) {
      const auto &pass = P.passes[s];
      ClusterOp op{};
      const ClusterPlan::Cluster &c =
          P.clusters[pass.kind == ClusterPlan::kCombine ? P.combines[pass.arg].big : pass.arg];
      const uint32_t core = c.members[0];
      op.alive = abase[q] + core;
      core_alive[op.alive] = true;
      op.core = row(q, c.pos);
      op.ntarget = (uint32_t)c.members.size() - 1;
      op.check_off = (uint32_t)crow.size();
      for (uint32_t t = 0; t < op.ntarget; t++) {
        op.tlab[t] = (uint8_t)Q.lab[c.members[t + 1]];
        op.tlix[t] = e->lmap[Q.lab[c.members[t + 1]]];
        bool seen = false;
        for (uint32_t k = 0; k < op.nlix; k++) seen = seen || op.lix[k] == op.tlix[t];
        if (!seen) op.lix[op.nlix++] = op.tlix[t];
      }
      if (pass.kind == ClusterPlan::kEncode) {
        // encodeKernel: a neighbour of a core candidate is a target's candidate
        op.kind = kOpEncode;
        for (uint32_t t = 0; t < op.ntarget; t++) {
          op.toff[t] = (uint16_t)crow.size();
          op.tcnt[t] = 1;
          crow.push_back(e->qsig[q][c.members[t + 1]]);
        }
      } else if (pass.kind == ClusterPlan::kMerge) {
        // mergeKernel: the core in both merged clusters; the targets are the
        // merged clusters' cores
        op.kind = kOpMerge;
        crow.push_back(row(q, c.left_pos));
        crow.push_back(row(q, c.right_pos));
        op.ncheck = 2;
        const uint32_t src[2] = {P.clusters[c.left].pos, P.clusters[c.right].pos};
        for (uint32_t t = 0; t < op.ntarget; t++) {
          op.toff[t] = (uint16_t)crow.size();
          op.tcnt[t] = 1;
          crow.push_back(row(q, src[t]));
        }
      } else {
        // combineMultipleClustersKernel: the core in every small cluster; a
        // target in every small cluster that holds it
        const ClusterPlan::Combine &cb = P.combines[pass.arg];
        op.kind = cb.keep ? kOpCombineKeep : kOpCombineNew;
        for (uint32_t sm : cb.small) crow.push_back(row(q, P.clusters[sm].pos));
        op.ncheck = (uint32_t)cb.small.size();
        for (uint32_t t = 0; t < op.ntarget; t++) {
          op.toff[t] = (uint16_t)crow.size();
          const uint32_t tu = c.members[t + 1];
          for (uint32_t sm : cb.small) {
            const auto &mm = P.clusters[sm].members;
            for (size_t x = 1; x < mm.size(); x++)
              if (mm[x] == tu) {
                crow.push_back(row(q, P.clusters[sm].pos + (uint32_t)x));
                break;
              }
          }
          op.tcnt[t] = (uint16_t)(crow.size() - op.toff[t]);
        }
      }
      if (crow.size() > 65535) return e->fail_msg(DSM_ELIMIT, "cluster plans too large");
      steps[s].push_back(op);
    }
    for (uint32_t u = 0; u < Q.n; u++) e->qref[q][u] = row(q, P.final_pos[u]);
  }
  std::vector<ClusterOp> ops;
  e->cstep_off.assign(1, 0);
  e->cmax_step_ops = 0;
  for (auto &st : steps) {
    ops.insert(ops.end(), st.begin(), st.end());
    e->cstep_off.push_back((uint32_t)ops.size());
    e->cmax_step_ops = std::max<uint32_t>(e->cmax_step_ops, (uint32_t)st.size());
  }
  // alive rows: every core's candidate list starts as its LDF/NLF row
  std::vector<uint32_t> asrc;
  for (uint32_t q = 0; q < nq; q++)
    for (uint32_t u = 0; u < e->queries[q].n; u++)
      if (core_alive[abase[q] + u]) {
        asrc.push_back(abase[q] + u);
        asrc.push_back(e->qsig[q][u]);
      }
  e->npos_rows = npos;
  e->nalive_rows = nalive;
  e->nalive_src = (uint32_t)asrc.size() / 2;
  HIPCHK(e->d_cops.alloc(ops.size()));
  HIPCHK(hipMemcpy(e->d_cops.p, ops.data(), ops.size() * sizeof(ClusterOp), hipMemcpyHostToDevice));
  HIPCHK(e->d_crow.alloc(crow.size()));
  if (!crow.empty()) HIPCHK(hipMemcpy(e->d_crow.p, crow.data(), crow.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(e->d_alive_src.alloc(asrc.size()));
  if (!asrc.empty()) HIPCHK(hipMemcpy(e->d_alive_src.p, asrc.data(), asrc.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(e->cctl.alloc(2 * (size_t)std::max<uint32_t>(nsteps, 1)));
  for (auto &x : e->ev_ref) HIPCHK(hipEventCreate(&x));
  return DSM_OK;
}

// The plans: every query edge in both orientations, each with its matching
// order (the query's order rule), grouped by the seed edge's label pair.
void build_plans(dsm_engine *e) {
  // plans: every query edge in both orientations
  struct PK {
    uint32_t key;
    Plan p;
  };
  std::vector<PK> pk;
  const uint32_t nq = (uint32_t)e->queries.size();
  for (uint32_t q = 0; q < nq; q++) {
    const QueryHost &Q = e->queries[q];
    for (uint32_t u = 0; u < Q.n; u++)
      for (uint32_t v = 0; v < Q.n; v++) {
        if (!(Q.adj[u] >> v & 1)) continue;
        Plan P{};
        P.n = (uint8_t)Q.n;
        P.query = (uint16_t)q;
        uint32_t order[kMaxQ];
        order[0] = u;
        order[1] = v;
        uint32_t placed = (1u << u) | (1u << v);
        for (uint32_t k = 2; k < Q.n; k++) {
          order[k] = next_in_order(Q, placed, e->label_freq, Q.order_variant);
          placed |= 1u << order[k];
        }
        for (uint32_t k = 0; k < Q.n; k++) {
          P.qv[k] = (uint8_t)order[k];
          P.sig[k] = (uint16_t)e->qref[q][order[k]];
          uint16_t back = 0, inj = 0;
          for (uint32_t j = 0; j < k; j++) {
            if (Q.adj[order[k]] >> order[j] & 1) back |= (uint16_t)(1u << j);
            if (Q.lab[order[j]] == Q.lab[order[k]]) inj |= (uint16_t)(1u << j);
          }
          P.back[k] = back;
          P.inj[k] = inj;
          P.lix[k] = e->lmap[Q.lab[order[k]]];
        }
        pk.push_back({(Q.lab[u] << 8) | Q.lab[v], P});
      }
  }
  std::stable_sort(pk.begin(), pk.end(), [](const PK &a, const PK &b) { return a.key < b.key; });
  e->plans.clear();
  e->plan_key.clear();
  e->max_plans_per_key = 0;
  for (size_t i = 0; i < pk.size();) {
    size_t j = i;
    while (j < pk.size() && pk[j].key == pk[i].key) j++;
    e->max_plans_per_key = std::max<uint32_t>(e->max_plans_per_key, (uint32_t)(j - i));
    i = j;
  }
  for (auto &x : pk) {
    e->plans.push_back(x.p);
    e->plan_key.push_back(x.key);
  }}

// label_freq[l]: data vertices of label l (l < 255), used only to order the
// matching: it never changes a count.
int build_queries(dsm_engine *e, const dsm_query_desc *qs, uint32_t nq, const std::vector<uint64_t> &label_freq) {
  if (nq == 0 || nq > DSM_MAX_QUERIES) return e->fail_msg(DSM_EINVAL, "need 1..%d queries", DSM_MAX_QUERIES);
  e->queries.resize(nq);
  for (uint32_t q = 0; q < nq; q++) {
    const dsm_query_desc &d = qs[q];
    QueryHost &Q = e->queries[q];
    if (d.num_vertices < 2 || d.num_vertices > (uint32_t)kMaxQ)
      return e->fail_msg(DSM_ELIMIT, "query %u: %u vertices (2..%d supported)", q, d.num_vertices, kMaxQ);
    Q.n = d.num_vertices;
    for (uint32_t u = 0; u < Q.n; u++) {
      if (d.labels[u] > DSM_MAX_LABEL)
        return e->fail_msg(DSM_ELIMIT, "query %u: label %u > %d", q, d.labels[u], DSM_MAX_LABEL);
      Q.lab[u] = d.labels[u];
      e->label_used[d.labels[u]] = true;
    }
    for (uint32_t i = 0; i < d.num_edges; i++) {
      uint32_t a = d.edge_src[i], b = d.edge_dst[i];
      if (a >= Q.n || b >= Q.n || a == b) return e->fail_msg(DSM_EINVAL, "query %u: bad edge %u-%u", q, a, b);
      Q.adj[a] |= 1u << b;
      Q.adj[b] |= 1u << a;
    }
    // connected (every later DFS position needs an earlier neighbour)
    uint32_t seen = 1, frontier = 1;
    while (frontier) {
      uint32_t nf = 0;
      for (uint32_t u = 0; u < Q.n; u++)
        if (frontier >> u & 1) nf |= Q.adj[u];
      frontier = nf & ~seen;
      seen |= nf;
    }
    if (seen != ((1u << Q.n) - 1)) return e->fail_msg(DSM_EINVAL, "query %u is not connected", q);
  }
  // relevant labels: those of query vertices with a neighbour (all of them)
  e->lmap.assign(256, 0xFF);
  uint32_t nrel = 0;
  for (uint32_t l = 0; l < 255; l++)
    if (e->label_used[l]) {
      if (nrel == (uint32_t)kMaxRel)
        return e->fail_msg(DSM_ELIMIT, "queries use more than %d distinct labels", kMaxRel);
      e->lmap[l] = (uint8_t)nrel++;
    }
  // signatures: (label, degree, neighbour-label counts) of every query vertex
  std::map<std::vector<uint64_t>, uint32_t> sigidx;
  std::vector<SigReq> tmp;
  e->qsig.assign(nq, std::vector<uint16_t>());
  for (uint32_t q = 0; q < nq; q++) {
    const QueryHost &Q = e->queries[q];
    e->qsig[q].resize(Q.n);
    for (uint32_t u = 0; u < Q.n; u++) {
      SigReq r{};
      r.label = Q.lab[u];
      r.deg = (uint32_t)__builtin_popcount(Q.adj[u]);
      uint8_t cnt[kMaxRel] = {};
      for (uint32_t w = 0; w < Q.n; w++)
        if (Q.adj[u] >> w & 1) cnt[e->lmap[Q.lab[w]]]++;
      for (int l = 0; l < kMaxRel; l++) r.nlc[l >> 3] |= (uint64_t)cnt[l] << ((l & 7) * 8);
      std::vector<uint64_t> key = {r.label, r.deg, r.nlc[0], r.nlc[1], r.nlc[2], r.nlc[3]};
      auto it = sigidx.find(key);
      if (it == sigidx.end()) {
        it = sigidx.emplace(key, (uint32_t)tmp.size()).first;
        tmp.push_back(r);
      }
      e->qsig[q][u] = (uint16_t)it->second;
    }
  }
  if (tmp.size() > 65535) return e->fail_msg(DSM_ELIMIT, "too many distinct query vertex signatures");
  // sort signatures by label; remap
  std::vector<uint32_t> ord(tmp.size());
  for (uint32_t i = 0; i < ord.size(); i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return tmp[a].label < tmp[b].label; });
  std::vector<uint32_t> newidx(tmp.size());
  e->sigs.resize(tmp.size());
  for (uint32_t i = 0; i < ord.size(); i++) {
    e->sigs[i] = tmp[ord[i]];
    newidx[ord[i]] = i;
  }
  for (auto &v : e->qsig)
    for (auto &s : v) s = (uint16_t)newidx[s];
  e->sig_lo.assign(256, 0);
  e->sig_hi.assign(256, 0);
  for (uint32_t i = 0; i < e->sigs.size(); i++) {
    uint32_t l = e->sigs[i].label;
    if (e->sig_hi[l] == 0) e->sig_lo[l] = i;
    e->sig_hi[l] = i + 1;
  }
  // cluster-encoding refinement: position rows after the signature rows
  e->nsig_rows = (uint32_t)e->sigs.size();
  e->qref.assign(nq, std::vector<uint32_t>());
  for (uint32_t q = 0; q < nq; q++)
    for (uint32_t u = 0; u < e->queries[q].n; u++) e->qref[q].push_back(e->qsig[q][u]);
  if (e->cluster) {
    if (int rc = build_cluster_ops(e, qs, nq)) return rc;
  }
  e->label_freq = label_freq;
  build_plans(e);

  return DSM_OK;
}

// ------------------------------------------------------------- graph set-up

// Segment slots for the segmented label sort: enough for every list of the
// pool to be split into kHugeSeg-entry segments, plus a margin.
int size_huge_sort(dsm_engine *e) {
  const uint32_t segs = (uint32_t)(e->pool_cap / kHugeSeg + 1024);
  if (segs > e->huge_segs) {
    HIPCHK(e->huge_hist.alloc((size_t)segs * (kMaxRel + 2)));
    e->huge_segs = segs;
  }
  if (!e->huge_list.p) {
    HIPCHK(e->huge_list.alloc(2 * 4096));
    HIPCHK(e->huge_ctl.alloc(4));
    // every vertex at most once (a dirty list holds each vertex once)
    HIPCHK(e->mid_list.alloc(e->n));
    // at most 2 m / 4,097 vertices have a degree above 4,096; a longer queue
    // makes k_lbuild_big look at every listed vertex instead
    HIPCHK(e->big_list.alloc((size_t)e->n / 64 + 4096));
  }
  return DSM_OK;
}

int upload_graph(dsm_engine *e, const dsm_graph_desc *g) {
  const uint32_t n = g->num_vertices;
  // sorted, deduplicated CSR
  std::vector<uint64_t> off(n + 1, 0);
  for (uint64_t i = 0; i < g->num_edges; i++) {
    uint32_t a = g->edge_src[i], b = g->edge_dst[i];
    if (a >= n || b >= n) return e->fail_msg(DSM_EINVAL, "edge %llu: endpoint out of range", (unsigned long long)i);
    off[a + 1]++;
    off[b + 1]++;
  }
  for (uint32_t v = 0; v < n; v++) off[v + 1] += off[v];
  std::vector<uint32_t> nb(off[n]);
  {
    std::vector<uint64_t> pos(off.begin(), off.end() - 1);
    for (uint64_t i = 0; i < g->num_edges; i++) {
      uint32_t a = g->edge_src[i], b = g->edge_dst[i];
      nb[pos[a]++] = b;
      nb[pos[b]++] = a;
    }
  }
  std::vector<uint32_t> dg(n), cp(n), st(n);
  unsigned long long total = 0;
  for (uint32_t v = 0; v < n; v++) {
    auto b = nb.begin() + off[v], en = nb.begin() + off[v + 1];
    std::sort(b, en);
    dg[v] = (uint32_t)(std::unique(b, en) - b);
    cp[v] = dg[v] + dg[v] / 4 + 4;
    st[v] = (uint32_t)total;
    total += cp[v];
  }
  const unsigned long long pcap = total + total / 2 + (1ull << 20);
  if (pcap >= (1ull << 32)) return e->fail_msg(DSM_ELIMIT, "graph too large for 32-bit pool offsets");
  std::vector<uint32_t> hp(total, 0xFFFFFFFFu);
  for (uint32_t v = 0; v < n; v++) std::copy(nb.begin() + off[v], nb.begin() + off[v] + dg[v], hp.begin() + st[v]);
  std::vector<uint8_t> hl(n);
  for (uint32_t v = 0; v < n; v++) hl[v] = clamp_label(g->labels[v], e->label_used);

  HIPCHK(e->pool.alloc(pcap));
  HIPCHK(e->start.alloc(n));
  HIPCHK(e->deg.alloc(n));
  HIPCHK(e->cap.alloc(n));
  HIPCHK(e->lab.alloc(n));
  HIPCHK(e->lpool.alloc(pcap));
  e->nbuckets = 1;
  for (int l = 0; l < 255; l++)
    if (e->lmap[l] != 0xFF) e->nbuckets = std::max<uint32_t>(e->nbuckets, e->lmap[l] + 2u);
  HIPCHK(e->loff.alloc((size_t)n * (e->nbuckets + 1)));
  e->pool_cap = pcap;
  if (int rc = size_huge_sort(e)) return rc;
  HIPCHK(hipMemcpy(e->pool.p, hp.data(), total * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->start.p, st.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->deg.p, dg.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->cap.p, cp.data(), (size_t)n * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->lab.p, hl.data(), n, hipMemcpyHostToDevice));
  // scratch holds one merged copy of every touched list
  e->scratch_cap = pcap + (1ull << 20);
  HIPCHK(e->scratch.alloc(e->scratch_cap));
  HIPCHK(e->actl.alloc(1));
  HIPCHK(e->h_actl.ensure(1));
  ApplyCtl a0{};
  a0.pool_used = total;
  HIPCHK(hipMemcpy(e->actl.p, &a0, sizeof a0, hipMemcpyHostToDevice));
  *e->h_actl.p = a0;
  HIPCHK(e->dirty_bits.alloc((n + 31) / 32));
  HIPCHK(hipMemset(e->dirty_bits.p, 0, ((n + 31) / 32) * 4));
  HIPCHK(e->dirty_list.alloc(n));
  HIPCHK(e->heavy_list.alloc(n));
  return DSM_OK;
}

int upload_tables(dsm_engine *e) {
  e->words = (e->n + 63) / 64;
  const size_t rows = (size_t)e->nsig_rows + e->npos_rows;
  HIPCHK(e->bm.alloc(rows * e->words));
  HIPCHK(hipMemset(e->bm.p, 0, rows * e->words * 8));
  if (e->cluster) HIPCHK(e->calive.alloc((size_t)e->nalive_rows * e->words));
  HIPCHK(e->d_sigs.alloc(e->sigs.size()));
  HIPCHK(hipMemcpy(e->d_sigs.p, e->sigs.data(), e->sigs.size() * sizeof(SigReq), hipMemcpyHostToDevice));
  HIPCHK(e->d_sig_lo.alloc(256));
  HIPCHK(e->d_sig_hi.alloc(256));
  HIPCHK(e->d_lmap.alloc(256));
  HIPCHK(hipMemcpy(e->d_sig_lo.p, e->sig_lo.data(), 256 * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->d_sig_hi.p, e->sig_hi.data(), 256 * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->d_lmap.p, e->lmap.data(), 256, hipMemcpyHostToDevice));
  HIPCHK(e->d_plans.alloc(e->plans.size()));
  HIPCHK(hipMemcpy(e->d_plans.p, e->plans.data(), e->plans.size() * sizeof(Plan), hipMemcpyHostToDevice));
  HIPCHK(e->d_plan_key.alloc(e->plan_key.size()));
  HIPCHK(hipMemcpy(e->d_plan_key.p, e->plan_key.data(), e->plan_key.size() * 4, hipMemcpyHostToDevice));
  const uint32_t nq = (uint32_t)e->queries.size();
  HIPCHK(e->counts.alloc((size_t)nq * 2 * kMaxSlots));
  HIPCHK(e->query_busy.alloc(nq));
  HIPCHK(hipMemset(e->query_busy.p, 0, (size_t)nq * 8));
  HIPCHK(e->h_counts.ensure((size_t)nq * 2 * kMaxSlots));
  HIPCHK(e->ctl.alloc(1));
  HIPCHK(e->h_ctl.ensure(1));
  e->qcap = e->opt.queue_capacity ? e->opt.queue_capacity : (1u << 20);
  HIPCHK(e->queue.alloc(e->qcap));
  HIPCHK(hipMemset(e->queue.p, 0, (size_t)e->qcap * sizeof(Task)));  // ready = 0: no epoch
  HIPCHK(e->tuple_count.alloc(1));
  if (e->opt.tuple_capacity) {
    e->tuple_cap = e->opt.tuple_capacity;
    size_t bytes = (size_t)e->tuple_cap * (kMaxQ + 2) * 4;
    if (e->opt.tuples_in_host_memory) {
      HIPCHK(hipHostMalloc((void **)&e->tuples, bytes, hipHostMallocMapped));
      e->tuples_host = true;
    } else {
      HIPCHK(hipMalloc((void **)&e->tuples, bytes));
    }
  }
  HIPCHK(e->ndeg.alloc(1024));
  HIPCHK(e->fail.alloc(1024));
  HIPCHK(e->soff.alloc(1024));
  return DSM_OK;
}

// ------------------------------------------------------------- DFS rounds

int ensure_seed_cap(dsm_engine *e, uint32_t nrec) {
  const uint64_t need = (uint64_t)nrec * std::max<uint32_t>(e->max_plans_per_key, 1);
  if (need > 0xFFFFFFF0ull) return e->fail_msg(DSM_ELIMIT, "seed table too large");
  if (need > e->seed_cap) {
    uint32_t c = (uint32_t)std::min<uint64_t>(need + need / 2 + 1024, 0xFFFFFFF0ull);
    HIPCHK(e->seed_plan.alloc(c));
    HIPCHK(e->seed_s.alloc(c));
    HIPCHK(e->seed_d.alloc(c));
    HIPCHK(e->seed_tag.alloc(c));
    e->seed_cap = c;
  }
  return DSM_OK;
}

// A DFS round: reset the control block, seed up to two record lists (each
// with its count tag, slot << 1 | sign), one persistent DFS launch over all
// their seeds, and a copy of the control block to the host. The seed table
// must hold the seeds of every list.
struct RoundPart {
  const uint32_t *rs, *rd;
  uint32_t n, tag;
};

SeedArgs seed_args(dsm_engine *e, const RoundPart &p) {
  SeedArgs sa;
  sa.g = e->dev_graph();
  sa.bm = e->bm.p;
  sa.words = e->words;
  sa.plans = e->d_plans.p;
  sa.plan_key = e->d_plan_key.p;
  sa.nplans = (uint32_t)e->plans.size();
  sa.rs = p.rs;
  sa.rd = p.rd;
  sa.nrec = p.n;
  sa.seed_plan = e->seed_plan.p;
  sa.seed_s = e->seed_s.p;
  sa.seed_d = e->seed_d.p;
  sa.seed_tag = e->seed_tag.p;
  sa.seed_cap = e->seed_cap;
  sa.tag = p.tag;
  sa.ctl = e->ctl.p;
  return sa;
}

DfsArgs dfs_args(dsm_engine *e) {
  DfsArgs da;
  da.g = e->dev_graph();
  da.bm = e->bm.p;
  da.words = e->words;
  da.plans = e->d_plans.p;
  da.seed_plan = e->seed_plan.p;
  da.seed_s = e->seed_s.p;
  da.seed_d = e->seed_d.p;
  da.seed_tag = e->seed_tag.p;
  da.queue = e->queue.p;
  da.qcap = e->qcap;
  da.epoch = ++e->epoch;
  da.donate_min = e->donate_min;
  da.lane_work = e->lane_work;
  da.leaf_budget = e->leaf_budget;
  da.bit_probe_last = e->bit_probe_last ? 1u : 0u;
  da.fshift = 0;
  while ((e->n > 0 ? (e->n - 1) >> da.fshift : 0u) >= 512u) da.fshift++;
  da.donate_every = e->donate_every;
  da.donate_min_n2 = e->donate_min_n2;
  da.leaf_segments = e->leaf_segments ? 1u : 0u;
  da.poll_max = e->poll_max;
  da.leaf_sets = e->leaf_sets ? 1u : 0u;
  da.nq = (uint32_t)e->queries.size();
  da.watchdog = (unsigned long long)e->wall_khz * e->watchdog_ms;
  da.ctl = e->ctl.p;
  da.counts = e->counts.p;
  da.tuples = e->tuples;
  da.tuple_cap = e->tuple_cap;
  da.tuple_count = e->tuple_count.p;
  da.wave_busy = e->wave_busy.p;
  da.query_busy = e->query_busy.p;
  da.wave_slots = e->wave_slots;
  return da;
}

// The grid: a round over a handful of records gets one block per CU, so that
// few waves pay the start-up and termination atomics.
uint32_t dfs_grid(dsm_engine *e, uint32_t nrec) {
  const uint64_t want =
      ((uint64_t)nrec * std::max<uint32_t>(e->max_plans_per_key, 1) * 8 + e->dfs_wpb - 1) / e->dfs_wpb;
  const uint32_t b = (uint32_t)std::min<uint64_t>(e->dfs_blocks, std::max<uint64_t>(want, (uint64_t)e->num_cu));
  e->last_blocks = b;
  e->last_wpb = e->tuples ? 4u : e->dfs_wpb;
  return b;
}

void drop_round_graph(dsm_engine *e) {
  if (e->rg_exec) (void)hipGraphExecDestroy(e->rg_exec);
  if (e->rg_graph) (void)hipGraphDestroy(e->rg_graph);
  e->rg_exec = nullptr;
  e->rg_graph = nullptr;
}

// The round as one HIP graph: ctl reset -> seed list 0 -> seed list 1 -> DFS ->
// control block to the host. Built once per launch shape (waves per block,
// tuples or not); every later round updates the kernel nodes' arguments and
// grids in the executable graph and launches it.
int graph_round(dsm_engine *e, const RoundPart *parts, int nparts, uint32_t nrec) {
  RoundNodeArgs st;
  st.ctl = e->ctl.p;
  const RoundPart none{nullptr, nullptr, 0, 0};
  st.seed[0] = seed_args(e, nparts > 0 ? parts[0] : none);
  st.seed[1] = seed_args(e, nparts > 1 ? parts[1] : none);
  st.dfs = dfs_args(e);
  hipKernelNodeParams pr, p0, p1, pd;
  round_node_params(st, dfs_grid(e, nrec), e->dfs_wpb, e->max_qn, &pr, &p0, &p1, &pd);
  const uint32_t key = (e->tuples ? 16u : 0u) | e->dfs_wpb;
  if (!e->rg_exec || e->rg_key != key) {
    drop_round_graph(e);
    HIPCHK(hipGraphCreate(&e->rg_graph, 0));
    HIPCHK(hipGraphAddKernelNode(&e->rg_reset, e->rg_graph, nullptr, 0, &pr));
    HIPCHK(hipGraphAddKernelNode(&e->rg_seed0, e->rg_graph, &e->rg_reset, 1, &p0));
    HIPCHK(hipGraphAddKernelNode(&e->rg_seed1, e->rg_graph, &e->rg_seed0, 1, &p1));
    HIPCHK(hipGraphAddKernelNode(&e->rg_dfs, e->rg_graph, &e->rg_seed1, 1, &pd));
    hipGraphNode_t cp, cw;
    HIPCHK(hipGraphAddMemcpyNode1D(&cp, e->rg_graph, &e->rg_dfs, 1, e->h_ctl.p, e->ctl.p, sizeof(DfsCtl),
                                   hipMemcpyDeviceToHost));
    HIPCHK(hipGraphAddMemcpyNode1D(&cw, e->rg_graph, &cp, 1, e->h_wave_busy.p, e->wave_busy.p,
                                   (size_t)e->wave_slots * 24, hipMemcpyDeviceToHost));
    HIPCHK(hipGraphInstantiate(&e->rg_exec, e->rg_graph, nullptr, nullptr, 0));
    e->rg_key = key;
  } else {
    HIPCHK(hipGraphExecKernelNodeSetParams(e->rg_exec, e->rg_seed0, &p0));
    HIPCHK(hipGraphExecKernelNodeSetParams(e->rg_exec, e->rg_seed1, &p1));
    HIPCHK(hipGraphExecKernelNodeSetParams(e->rg_exec, e->rg_dfs, &pd));
  }
  HIPCHK(hipGraphLaunch(e->rg_exec, e->stream));
  return DSM_OK;
}

// One round, eager or as a graph. ev0 / ev1 / ev2 (may be null) mark its start,
// the end of seeding (eager only; a graph round records ev1 at its start) and
// its end.
int dfs_round(dsm_engine *e, const RoundPart *parts, int nparts, hipEvent_t ev0, hipEvent_t ev1,
              hipEvent_t ev2) {
  uint32_t nrec = 0;
  for (int i = 0; i < nparts; i++) nrec += parts[i].n;
  if (ev0) HIPCHK(hipEventRecord(ev0, e->stream));
  if (e->graph_rounds) {
    if (ev1) HIPCHK(hipEventRecord(ev1, e->stream));
    int rc = graph_round(e, parts, nparts, nrec);
    if (rc) return rc;
  } else {
    HIPCHK(launch_ctl_reset(e->ctl.p, nullptr, nullptr, nullptr, 0, e->stream));
    for (int i = 0; i < nparts; i++)
      if (parts[i].n) HIPCHK(launch_seed(seed_args(e, parts[i]), e->stream));
    if (ev1) HIPCHK(hipEventRecord(ev1, e->stream));
    HIPCHK(launch_dfs(dfs_args(e), dfs_grid(e, nrec), e->dfs_wpb, e->max_qn, e->stream));
    HIPCHK(hipMemcpyAsync(e->h_ctl.p, e->ctl.p, sizeof(DfsCtl), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(e->h_wave_busy.p, e->wave_busy.p, (size_t)e->wave_slots * 24, hipMemcpyDeviceToHost,
                          e->stream));
  }
  if (ev2) HIPCHK(hipEventRecord(ev2, e->stream));
  return DSM_OK;
}

int run_round(dsm_engine *e, const uint32_t *rs, const uint32_t *rd, uint32_t nrec, uint32_t tag,
              hipEvent_t ev0, hipEvent_t ev1, hipEvent_t ev2) {
  const RoundPart p{rs, rd, nrec, tag};
  return dfs_round(e, &p, 1, ev0, ev1, ev2);
}

int check_round(dsm_engine *e, uint64_t *seeds) {
  const DfsCtl &c = *e->h_ctl.p;
  if (c.seed_overflow.v)
    return e->fail_msg(DSM_ELIMIT, "seed table overflow (%u seeds dropped)", c.seed_overflow.v);
  if (c.abort.v)
    return e->fail_msg(DSM_EDEVICE, "DFS watchdog: a wave waited more than %u ms for work (%d tasks left)",
                       e->watchdog_ms, c.work_left.v);
  *seeds = c.nseeds.v;
  e->stats.tasks_donated += std::min(c.q_tail.v, e->qcap);
  e->stats.dfs_entries += c.scanned.v;
  e->stats.dfs_search_loads += c.loads.v;
  e->stats.dfs_merge_bytes += c.mbytes.v;
  // balance: per-block task time of the launch
  {
    const uint32_t nb = e->last_blocks, wpb = e->last_wpb;
    if (nb && (uint64_t)nb * wpb <= e->wave_slots) {
      double tot = 0, mx = 0;
      for (uint32_t b = 0; b < nb; b++) {
        double s = 0;
        for (uint32_t w = 0; w < wpb; w++) s += (double)e->h_wave_busy.p[(size_t)b * wpb + w];
        tot += s;
        mx = std::max(mx, s);
      }
      const double to_ms = 1.0 / (double)e->wall_khz;
      e->stats.dfs_busy_ms += tot * to_ms;
      e->stats.dfs_waves += (uint64_t)nb * wpb;
      if (tot > 0) e->stats.dfs_block_imbalance += mx / (tot / nb);
      // head and tail idle: a wave without tasks counts as idle in the tail
      const unsigned long long *first = e->h_wave_busy.p + e->wave_slots;
      const unsigned long long *last = first + e->wave_slots;
      const uint32_t nw = nb * wpb;
      unsigned long long t0 = ~0ull, t1 = 0;
      for (uint32_t w = 0; w < nw; w++)
        if (first[w]) {
          t0 = std::min(t0, first[w]);
          t1 = std::max(t1, last[w]);
        }
      if (t1 > t0) {
        const double span = (double)(t1 - t0);
        double head = 0, tail = 0;
        for (uint32_t w = 0; w < nw; w++) {
          if (!first[w]) {
            tail += span;
            continue;
          }
          head += (double)(first[w] - t0);
          tail += (double)(t1 - last[w]);
        }
        e->stats.dfs_head_idle += head / (span * nw);
        e->stats.dfs_tail_idle += tail / (span * nw);
      }
    }
  }
  e->stats.dfs_launches += 1;
  return DSM_OK;
}

// ------------------------------------------------------------ apply rounds

struct HostDelta {
  std::vector<uint32_t> tv, ioff, doff, ins, del, rl_v, rl_l;
};

// The update rule of cpuGraph.cpp:521-607 reduced to its net effect: records
// symmetrised, stably sorted by (src,dst) (original records before reversed
// copies on ties), each directed key ends in the state of its LAST record, and
// each vertex's label becomes the label of the last insert record with that
// vertex as src.
void host_delta(const dsm_engine *e, const dsm_change *ch, uint32_t k, HostDelta &d) {
  struct R {
    uint32_t s, t, add, label, pos;
  };
  std::vector<R> r(2 * (size_t)k);
  for (uint32_t i = 0; i < k; i++) {
    r[i] = {ch[i].src, ch[i].dst, ch[i].add ? 1u : 0u, ch[i].label, i};
    r[k + i] = {ch[i].dst, ch[i].src, ch[i].add ? 1u : 0u, ch[i].label, k + i};
  }
  std::sort(r.begin(), r.end(), [](const R &a, const R &b) {
    if (a.s != b.s) return a.s < b.s;
    if (a.t != b.t) return a.t < b.t;
    return a.pos < b.pos;
  });
  d = HostDelta();
  d.ioff.push_back(0);
  d.doff.push_back(0);
  for (size_t i = 0; i < r.size();) {
    const uint32_t s = r[i].s;
    size_t j = i;
    int last_add = -1;
    while (j < r.size() && r[j].s == s) {
      size_t k2 = j;
      while (k2 + 1 < r.size() && r[k2 + 1].s == s && r[k2 + 1].t == r[j].t) k2++;
      if (r[k2].add) d.ins.push_back(r[j].t);
      else d.del.push_back(r[j].t);
      for (size_t x = j; x <= k2; x++)
        if (r[x].add) last_add = (int)x;
      j = k2 + 1;
    }
    d.tv.push_back(s);
    d.ioff.push_back((uint32_t)d.ins.size());
    d.doff.push_back((uint32_t)d.del.size());
    if (e->opt.relabel && last_add >= 0) {
      d.rl_v.push_back(s);
      d.rl_l.push_back(clamp_label(r[last_add].label, e->label_used));
    }
    i = j;
  }
}

// Uploads one apply round (a slice of every touched vertex's lists) and runs
// merge + commit. Grows the pool when a commit found no room.
int grow_pool(dsm_engine *e, unsigned long long extra);

int apply_round(dsm_engine *e, const HostDelta &d, const std::vector<uint32_t> &tv,
                const std::vector<uint32_t> &ioff, const std::vector<uint32_t> &doff,
                const std::vector<uint32_t> &ins, const std::vector<uint32_t> &del) {
  const uint32_t T = (uint32_t)tv.size();
  if (T == 0) return DSM_OK;
  (void)d;
  // layout: tv | ioff | doff | ins | del
  const size_t o_tv = 0, o_io = o_tv + T, o_do = o_io + T + 1, o_in = o_do + T + 1, o_de = o_in + ins.size();
  const size_t total = o_de + del.size();
  // device-only space after the uploaded part: prefix counts and big-list
  // indices of the hub lists that the whole grid copies
  const size_t o_gi = total, o_gd = o_gi + ins.size() + T, o_bm = o_gd + del.size() + T, o_bc = o_bm + T;
  const size_t dtotal = o_bc + T;
  // queued work must be done with the staging and device buffers before they are reused
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(e->hbuf.ensure(total));
  HIPCHK(e->bbuf.ensure(dtotal));
  uint32_t *h = e->hbuf.p;
  std::copy(tv.begin(), tv.end(), h + o_tv);
  std::copy(ioff.begin(), ioff.end(), h + o_io);
  std::copy(doff.begin(), doff.end(), h + o_do);
  std::copy(ins.begin(), ins.end(), h + o_in);
  std::copy(del.begin(), del.end(), h + o_de);
  HIPCHK(hipMemcpyAsync(e->bbuf.p, h, total * 4, hipMemcpyHostToDevice, e->stream));
  HIPCHK(e->ndeg.ensure(T));
  HIPCHK(e->fail.ensure(T));
  HIPCHK(e->soff.ensure(T));
  BatchDev b{};
  b.ntouched = T;
  b.tv = e->bbuf.p + o_tv;
  b.ioff = e->bbuf.p + o_io;
  b.doff = e->bbuf.p + o_do;
  b.ins = e->bbuf.p + o_in;
  b.del = e->bbuf.p + o_de;
  b.ndeg = e->ndeg.p;
  b.soff = e->soff.p;
  b.fail = e->fail.p;
  b.gicum = e->bbuf.p + o_gi;
  b.gdcum = e->bbuf.p + o_gd;
  b.big_merge = e->bbuf.p + o_bm;
  b.big_commit = e->bbuf.p + o_bc;
  // scratch must hold every touched list once plus the inserts: at most the
  // pool's entries plus this round's inserts (the stream is idle here, so a
  // larger scratch replaces the old one without copying)
  const unsigned long long need = e->pool_cap + ins.size() + (1ull << 20);
  if (need > e->scratch_cap) {
    HIPCHK(e->scratch.alloc(need + need / 4));
    e->scratch_cap = need + need / 4;
  }
  HIPCHK(launch_scratch_reset(e->actl.p, e->stream));
  HIPCHK(launch_apply_merge(e->dev_graph(), b, e->scratch_cap, e->scratch.p, e->actl.p, e->stream));
  HIPCHK(launch_apply_commit(e->dev_graph(), b, e->scratch.p, e->pool_cap, e->actl.p, e->dirty(), e->stream));
  HIPCHK(hipMemcpyAsync(e->h_actl.p, e->actl.p, sizeof(ApplyCtl), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  if (e->h_actl.p->merge_overflow)
    return e->fail_msg(DSM_ELIMIT, "update: %u merged lists found no room in the merge scratch",
                       e->h_actl.p->merge_overflow);
  for (int attempt = 0; e->h_actl.p->overflow && attempt < 4; attempt++) {
    int rc = grow_pool(e, e->pool_cap / 2 + (1ull << 22));
    if (rc) return rc;
    // the merge itself overflowing scratch is impossible by sizing; only commits retry
    HIPCHK(launch_apply_commit_failed(e->dev_graph(), b, e->scratch.p, e->pool_cap, e->actl.p, e->dirty(),
                                      e->stream));
    HIPCHK(hipMemcpyAsync(e->h_actl.p, e->actl.p, sizeof(ApplyCtl), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  if (e->h_actl.p->overflow) return e->fail_msg(DSM_ENOMEM, "neighbour pool could not be grown");
  return DSM_OK;
}

// Re-lays the pool with fresh slack (capacity deg*1.25+8 per vertex) into a
// new allocation of the live size plus `extra` entries.
int grow_pool(dsm_engine *e, unsigned long long extra) {
  const uint32_t n = e->n;
  std::vector<uint32_t> dg(n), st_new(n), cp_new(n);
  HIPCHK(hipMemcpy(dg.data(), e->deg.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  unsigned long long total = 0;
  for (uint32_t v = 0; v < n; v++) {
    cp_new[v] = dg[v] + dg[v] / 4 + 8;
    st_new[v] = (uint32_t)total;
    total += cp_new[v];
  }
  const unsigned long long ncap = total + extra;
  if (ncap >= (1ull << 32)) return e->fail_msg(DSM_ELIMIT, "neighbour pool exceeds 32-bit offsets");
  DevBuf<uint32_t> np, nst;
  HIPCHK(np.alloc(ncap));
  HIPCHK(nst.alloc(n));
  // on the engine's stream, ahead of the kernels that read the new starts
  HIPCHK(hipMemcpyAsync(nst.p, st_new.data(), (size_t)n * 4, hipMemcpyHostToDevice, e->stream));
  HIPCHK(launch_pool_move(e->pool.p, np.p, e->start.p, nst.p, e->deg.p, n, e->stream));
  HIPCHK(hipMemcpyAsync(e->cap.p, cp_new.data(), (size_t)n * 4, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  std::swap(e->pool.p, np.p);
  std::swap(e->pool.n, np.n);
  std::swap(e->start.p, nst.p);
  std::swap(e->start.n, nst.n);
  e->pool_cap = ncap;
  if (int rc = size_huge_sort(e)) return rc;
  // every label-sorted copy moves with the pool: the update that grew it
  // rebuilds them all once its commits are done (the rebuild uses the merge
  // scratch, which still holds the lists of the commits that failed)
  HIPCHK(e->lpool.alloc(ncap));
  e->lpool_stale = true;
  e->grew = true;
  // scratch must hold one copy of every list; the failed commits' lists move
  // with it
  if (e->scratch_cap < ncap + (1ull << 20)) {
    DevBuf<uint32_t> ns;
    HIPCHK(ns.alloc(ncap + (1ull << 20)));
    HIPCHK(hipMemcpyAsync(ns.p, e->scratch.p, e->scratch_cap * 4, hipMemcpyDeviceToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    std::swap(e->scratch.p, ns.p);
    std::swap(e->scratch.n, ns.n);
    e->scratch_cap = ncap + (1ull << 20);
  }
  ApplyCtl a = *e->h_actl.p;
  a.pool_used = total;
  a.overflow = 0;
  // ordered before the retried commits on the engine's stream; `a` must stay
  // alive until the copy is done
  HIPCHK(hipMemcpyAsync(e->actl.p, &a, sizeof a, hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  *e->h_actl.p = a;
  return DSM_OK;
}

int do_apply_delta(dsm_engine *e, const HostDelta &d) {
  const uint32_t T = (uint32_t)d.tv.size();
  // split into rounds of at most kMaxVertexChanges inserts and deletes per vertex
  uint32_t rounds = 1;
  for (uint32_t t = 0; t < T; t++) {
    uint32_t m = std::max(d.ioff[t + 1] - d.ioff[t], d.doff[t + 1] - d.doff[t]);
    rounds = std::max(rounds, (m + kMaxVertexChanges - 1) / kMaxVertexChanges);
  }
  for (uint32_t rd = 0; rd < rounds; rd++) {
    std::vector<uint32_t> tv, ioff{0}, doff{0}, ins, del;
    for (uint32_t t = 0; t < T; t++) {
      uint32_t i0 = d.ioff[t] + rd * kMaxVertexChanges, i1 = std::min(d.ioff[t + 1], i0 + kMaxVertexChanges);
      uint32_t d0 = d.doff[t] + rd * kMaxVertexChanges, d1 = std::min(d.doff[t + 1], d0 + kMaxVertexChanges);
      if (i0 >= i1 && d0 >= d1 && rd > 0) continue;
      tv.push_back(d.tv[t]);
      if (i0 < i1) ins.insert(ins.end(), d.ins.begin() + i0, d.ins.begin() + i1);
      if (d0 < d1) del.insert(del.end(), d.del.begin() + d0, d.del.begin() + d1);
      ioff.push_back((uint32_t)ins.size());
      doff.push_back((uint32_t)del.size());
    }
    int rc = apply_round(e, d, tv, ioff, doff, ins, del);
    if (rc) return rc;
  }
  e->stats.touched_vertices = T;
  e->stats.relocated_vertices = e->h_actl.p->relocated;
  // relabels
  BatchDev b{};
  const uint32_t R = (uint32_t)d.rl_v.size();
  if (R) {
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(e->hbuf.ensure(2 * (size_t)R));
    HIPCHK(e->bbuf.ensure(3 * (size_t)R));
    std::copy(d.rl_v.begin(), d.rl_v.end(), e->hbuf.p);
    std::copy(d.rl_l.begin(), d.rl_l.end(), e->hbuf.p + R);
    HIPCHK(hipMemcpyAsync(e->bbuf.p, e->hbuf.p, 2 * (size_t)R * 4, hipMemcpyHostToDevice, e->stream));
    b.nrelabel = R;
    b.rl_v = e->bbuf.p;
    b.rl_l = e->bbuf.p + R;
    b.rl_big = e->bbuf.p + 2 * (size_t)R;
    HIPCHK(launch_relabel(e->dev_graph(), b, e->tables(), e->dirty(), e->stream));
  }
  return DSM_OK;
}

int upload_records(dsm_engine *e, const std::vector<uint32_t> &s, const std::vector<uint32_t> &d,
                   DevBuf<uint32_t> &dev) {
  const size_t k = s.size();
  PinnedBuf<uint32_t> &h = e->hbuf;
  HIPCHK(hipStreamSynchronize(e->stream));
  HIPCHK(dev.ensure(2 * k + 2));
  HIPCHK(h.ensure(2 * k + 2));
  std::copy(s.begin(), s.end(), h.p);
  std::copy(d.begin(), d.end(), h.p + k);
  HIPCHK(hipMemcpyAsync(dev.p, h.p, 2 * k * 4 + 4, hipMemcpyHostToDevice, e->stream));
  return DSM_OK;
}


// The cluster-encoding refinement of every query on the current graph,
// enqueued on the engine's stream after the LDF/NLF rows are current: the
// position rows are cleared, each core's alive row loaded, then pass step s of
// every query runs as one k_cluster_pass launch plus one k_cluster_heavy
// launch for the chunks of candidates with long lists (filter.cu:936-1177 runs
// one kernel and one device synchronisation per cluster of one query).
int cluster_refine(dsm_engine *e) {
  if (!e->cluster) return DSM_OK;
  const uint32_t nsteps = (uint32_t)e->cstep_off.size() - 1;
  // heavy chunks of one pass: at most 2 x (list entries) / kClusterChunk
  const unsigned long long entries = std::max<unsigned long long>(e->h_actl.p->pool_used, 1);
  const size_t need = (size_t)e->cmax_step_ops * (2 * entries / kClusterChunk + 64);
  if (need > e->cheavy.n) HIPCHK(e->cheavy.alloc(need));
  HIPCHK(hipEventRecord(e->ev_ref[0], e->stream));
  HIPCHK(launch_cluster_init(e->bm.p, e->words, e->nsig_rows, e->npos_rows, e->calive.p, e->d_alive_src.p,
                             e->nalive_src, e->cctl.p, 2 * nsteps, e->stream));
  ClusterArgs a{};
  a.g = e->dev_graph();
  a.bm = e->bm.p;
  a.words = e->words;
  a.alive = e->calive.p;
  a.crow = e->d_crow.p;
  a.heavy = e->cheavy.p;
  a.heavy_cap = (uint32_t)std::min<size_t>(e->cheavy.n, 0xFFFFFFF0u);
  for (uint32_t st = 0; st < nsteps; st++) {
    a.ops = e->d_cops.p + e->cstep_off[st];
    a.nops = e->cstep_off[st + 1] - e->cstep_off[st];
    a.ctl = e->cctl.p + 2 * st;
    HIPCHK(launch_cluster_pass(a, (uint32_t)e->num_cu * 8, e->stream));
    HIPCHK(launch_cluster_heavy(a, (uint32_t)e->num_cu * 4, e->stream));
  }
  HIPCHK(hipEventRecord(e->ev_ref[1], e->stream));
  e->refine_timed = true;
  return DSM_OK;
}

// Adds the device time of the last refinement to the batch statistics (the
// refinement's stream must have been synchronised).
void take_refine_time(dsm_engine *e) {
  if (!e->refine_timed) return;
  e->refine_timed = false;
  float ms = 0;
  if (hipEventElapsedTime(&ms, e->ev_ref[0], e->ev_ref[1]) == hipSuccess) e->stats.ms_refine += ms;
}

// The label-sorted copies that a catch-up leaves to the update that follows
// it: those of the last update's dirty vertices (the shared dirty list still
// holds them until the next compaction), copied from `from` once the new
// update has marked its own dirty vertices, skipping those, whose copies it
// rebuilds from their lists anyway.
struct SortedSync {
  DevGraph from;
  const uint32_t *ndirty;
};

// One batch's update and candidate refresh from its host delta, enqueued on
// the engine's stream: reset the per-batch apply counters, apply (device
// merge/commit, which waits for the commit's status), relabel, make a
// catch-up's deferred label-sorted copies (sorted_sync), rebuild the dirty
// vertices' label-sorted copies, then refilter the dirty vertices (long lists
// take their neighbour-label counts from those copies). No wait at the end.
int apply_delta_and_refresh(dsm_engine *e, const HostDelta &d, const SortedSync *sorted_sync = nullptr) {
  int rc;
  e->shadow_delta_ok = false;
  e->grew = false;
  HIPCHK(launch_ctl_reset(nullptr, e->actl.p, nullptr, nullptr, 0, e->stream));
  if ((rc = do_apply_delta(e, d))) return rc;
  // the catch-up's deferred label-sorted copies: this update's dirty bits are
  // all set and the dirty list still holds the last update's vertices. After
  // a pool growth the slots moved and every copy is rebuilt instead.
  if (sorted_sync && !e->grew && !e->lpool_stale)
    HIPCHK(launch_sync_sorted(sorted_sync->from, e->dev_graph(), e->dirty_list.p, sorted_sync->ndirty,
                              e->dirty_bits.p, e->sync_q.p, (uint32_t)e->sync_q.n, e->sync_ctl.p, e->stream));
  HIPCHK(launch_dirty_compact(e->dirty(), e->n, e->stream));
  // the dirty list (touched vertices, relabelled ones and their neighbours) is
  // exactly the set whose label-sorted copies changed, unless the pool grew
  if (e->lpool_stale) {
    HIPCHK(launch_lbuild(e->dev_graph(), e->d_lmap.p, nullptr, nullptr, e->huge_sort(), e->stream));
    e->lpool_stale = false;
  } else {
    HIPCHK(launch_lbuild(e->dev_graph(), e->d_lmap.p, e->dirty_list.p, &e->actl.p->ndirty, e->huge_sort(),
                         e->stream));
  }
  HIPCHK(launch_filter_dirty(e->dev_graph(), e->tables(), e->dirty(), e->stream));
  return cluster_refine(e);
}

int apply_and_refresh(dsm_engine *e, const dsm_change *changes, uint32_t k) {
  HostDelta d;
  host_delta(e, changes, k, d);
  return apply_delta_and_refresh(e, d);
}

// Blocks per CU of a DFS shape: the occupancy limit and, with overlap_apply,
// at most what leaves 32 KB of the CU's LDS to the update kernels that run
// beside a pipelined round (k_apply_merge takes 16 KB per block). A shape
// that fills the LDS makes them wait for the round.
uint32_t dfs_bpc(const dsm_engine *e, uint32_t wpb) {
  uint32_t b = (uint32_t)std::max(1, dfs_max_blocks_per_cu(wpb, e->max_qn));
  if (e->overlap_apply) {
    const uint32_t per_block = wpb * dfs_lds_bytes_per_wave(e->max_qn);
    const int room = e->cu_lds - 32 * 1024;
    const uint32_t cap = room > (int)per_block ? (uint32_t)room / per_block : 1u;
    b = std::min(b, cap);
  }
  return b;
}

int create_impl(dsm_engine *e, const dsm_graph_desc *graph, const dsm_query_desc *queries,
                uint32_t num_queries) {
  int rc = 0;
  HIPCHK(hipSetDevice(e->opt.device));
  HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&e->ev_pre, hipEventDisableTiming));
  HIPCHK(hipDeviceGetAttribute(&e->num_cu, hipDeviceAttributeMultiprocessorCount, e->opt.device));
  HIPCHK(hipDeviceGetAttribute(&e->wall_khz, hipDeviceAttributeWallClockRate, e->opt.device));
  {
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, e->opt.device) ==
            hipSuccess &&
        lds > 0)
      e->cu_lds = lds;
  }
  e->overlap_apply = e->opt.overlap_apply != 0;
  HIPCHK(hipStreamCreateWithFlags(&e->stream2, hipStreamNonBlocking));
  if (e->wall_khz <= 0) e->wall_khz = 100000;
  if (e->opt.watchdog_ms) e->watchdog_ms = e->opt.watchdog_ms;
  for (auto &x : e->ev) HIPCHK(hipEventCreate(&x));
  e->cluster = e->opt.cluster_filter != 0;
  std::vector<uint64_t> label_freq(256, 0);
  for (uint32_t v = 0; graph->labels && v < graph->num_vertices; v++)
    if (graph->labels[v] < 255) label_freq[graph->labels[v]]++;
  if ((rc = build_queries(e, queries, num_queries, label_freq))) return rc;
  if ((rc = upload_graph(e, graph))) return rc;
  if ((rc = upload_tables(e))) return rc;
  uint32_t wpb = e->opt.waves_per_block;
  if (wpb != 1 && wpb != 2 && wpb != 4 && wpb != 8) wpb = 4;
  e->dfs_wpb = wpb;
  e->max_qn = 2;
  for (auto &q : e->queries) e->max_qn = std::max(e->max_qn, q.n);
  // keep a block's LDS (waves x per-wave stack) within 64 KB
  while (wpb > 1 && wpb * dfs_lds_bytes_per_wave(e->max_qn) > 65536) wpb /= 2;
  e->dfs_wpb = wpb;
  int per_cu = (int)dfs_bpc(e, wpb);
  if (e->opt.blocks_per_cu && (int)e->opt.blocks_per_cu < per_cu) per_cu = (int)e->opt.blocks_per_cu;
  e->dfs_blocks = (uint32_t)(per_cu * e->num_cu);
  // room for the largest grid any launch shape can have (the autotuner may
  // pick another shape later)
  {
    uint32_t most = 0;
    for (uint32_t w : {1u, 2u, 4u, 8u}) most = std::max<uint32_t>(most, (uint32_t)dfs_max_blocks_per_cu(w, e->max_qn) * w);
    e->wave_slots = most * (uint32_t)e->num_cu;
    // per wave: task time, first task start, last task end
    HIPCHK(e->wave_busy.alloc(3 * (size_t)e->wave_slots));
    HIPCHK(e->h_wave_busy.ensure(3 * (size_t)e->wave_slots));
  }
  e->donate_min = e->opt.donate_min ? e->opt.donate_min : 256;
  e->lane_work = e->opt.lane_work ? e->opt.lane_work : 16;
  e->leaf_budget = e->opt.leaf_budget ? e->opt.leaf_budget : 2048;
  e->donate_every = e->opt.donate_every ? e->opt.donate_every : 8;
  e->donate_min_n2 = e->opt.donate_min_n2 ? e->opt.donate_min_n2 : 64;
  e->leaf_sets = e->opt.leaf_sets != 0;
  e->bit_probe_last = e->opt.bit_probe_last != 0;
  e->graph_rounds = e->opt.graph_rounds != 0;
  e->leaf_segments = e->opt.leaf_segments != 0;
  e->poll_max = e->opt.poll_max ? std::min<uint32_t>(e->opt.poll_max, 16u) : 16u;
  HIPCHK(launch_filter_full(e->dev_graph(), e->tables(), e->stream));
  HIPCHK(launch_lbuild(e->dev_graph(), e->d_lmap.p, nullptr, nullptr, e->huge_sort(), e->stream));
  if ((rc = cluster_refine(e))) return rc;
  HIPCHK(hipStreamSynchronize(e->stream));
  e->refine_timed = false;
  return DSM_OK;
}

template <typename T>
void swap_buf(DevBuf<T> &a, DevBuf<T> &b) {
  std::swap(a.p, b.p);
  std::swap(a.n, b.n);
}

// Exchanges the engine's graph with the second copy.
void swap_graph(dsm_engine *e) {
  auto &c = e->shadow;
  swap_buf(e->pool, c.pool);
  swap_buf(e->start, c.start);
  swap_buf(e->deg, c.deg);
  swap_buf(e->cap, c.cap);
  swap_buf(e->lpool, c.lpool);
  swap_buf(e->loff, c.loff);
  swap_buf(e->lab, c.lab);
  swap_buf(e->bm, c.bm);
  swap_buf(e->actl, c.actl);
  std::swap(e->pool_cap, c.pool_cap);
}

template <typename T>
hipError_t copy_buf(DevBuf<T> &dst, const DevBuf<T> &src, size_t count, hipStream_t s) {
  if (dst.n < src.n) {
    hipError_t err = dst.alloc(src.n);
    if (err != hipSuccess) return err;
  }
  return count ? hipMemcpyAsync(dst.p, src.p, count * sizeof(T), hipMemcpyDeviceToDevice, s) : hipSuccess;
}

// Allocates the second copy of the graph at the current sizes; its contents
// are written by the first overlapped update.
int reserve_shadow(dsm_engine *e) {
  auto &c = e->shadow;
  auto like = [](auto &dst, const auto &src) -> hipError_t {
    return dst.n < src.n ? dst.alloc(src.n) : hipSuccess;
  };
  HIPCHK(like(c.pool, e->pool));
  HIPCHK(like(c.lpool, e->lpool));
  HIPCHK(like(c.start, e->start));
  HIPCHK(like(c.deg, e->deg));
  HIPCHK(like(c.cap, e->cap));
  HIPCHK(like(c.loff, e->loff));
  HIPCHK(like(c.lab, e->lab));
  HIPCHK(like(c.bm, e->bm));
  HIPCHK(like(c.actl, e->actl));
  return DSM_OK;
}

// Makes the engine's graph a copy of `c` on the engine's stream. The pool
// and the label-sorted copies are copied up to the pool's used end (the
// host's ApplyCtl mirror holds it after every update).
int copy_graph(dsm_engine *e, const dsm_engine::GraphCopy &c) {
  unsigned long long used = e->h_actl.p->pool_used;
  if (used == 0 || used > c.pool_cap) used = c.pool_cap;
  HIPCHK(copy_buf(e->pool, c.pool, used, e->stream));
  HIPCHK(copy_buf(e->lpool, c.lpool, used, e->stream));
  HIPCHK(copy_buf(e->start, c.start, c.start.n, e->stream));
  HIPCHK(copy_buf(e->deg, c.deg, c.deg.n, e->stream));
  HIPCHK(copy_buf(e->cap, c.cap, c.cap.n, e->stream));
  HIPCHK(copy_buf(e->loff, c.loff, c.loff.n, e->stream));
  HIPCHK(copy_buf(e->lab, c.lab, c.lab.n, e->stream));
  // the signature rows; the cluster position rows are recomputed by the update
  HIPCHK(copy_buf(e->bm, c.bm, (size_t)e->nsig_rows * e->words, e->stream));
  HIPCHK(copy_buf(e->actl, c.actl, 1, e->stream));
  e->pool_cap = c.pool_cap;
  return DSM_OK;
}

// Applies one batch to the second copy of the graph while a round that was
// enqueued on `stream` matches on the first: the copies swap, the new current
// one becomes a copy of the old one (after ev_pre, i.e. after every update
// enqueued before the round) and takes the batch, all on stream2. Returns
// once the update is done; the copies stay swapped, or with keep = false
// swap back (the update is dropped: the tuner's load).
DevGraph graph_of(const dsm_engine *e, const dsm_engine::GraphCopy &c) {
  DevGraph g;
  g.n = e->n;
  g.pool = c.pool.p;
  g.start = c.start.p;
  g.deg = c.deg.p;
  g.cap = c.cap.p;
  g.lab = c.lab.p;
  g.lpool = c.lpool.p;
  g.loff = c.loff.p;
  g.nbuckets = e->nbuckets;
  return g;
}

// Makes the engine's graph equal to `c` when it lags by exactly the last
// update (shadow_delta_ok): only that update's touched, dirty and relabelled
// vertices are copied (launch_sync), plus the signature rows and the apply
// status; otherwise the whole graph (copy_graph). With `deferred`, the dirty
// vertices' label-sorted copies are left to the update that follows: when the
// copy is by the last update only, *deferred describes them and *has_deferred
// is set (see SortedSync).
int catch_up(dsm_engine *e, const dsm_engine::GraphCopy &c, SortedSync *deferred = nullptr,
             bool *has_deferred = nullptr) {
  if (has_deferred) *has_deferred = false;
  if (!kDeltaSync || !e->shadow_delta_ok || e->pool_cap != c.pool_cap || e->n >= (1u << 31))
    return copy_graph(e, c);
  const uint32_t nt = (uint32_t)e->last_touched.size(), nr = (uint32_t)e->last_relabel.size();
  HIPCHK(e->d_last_touched.ensure(std::max<uint32_t>(nt, 1)));
  HIPCHK(e->d_last_relabel.ensure(std::max<uint32_t>(nr, 1)));
  if (nt)
    HIPCHK(hipMemcpyAsync(e->d_last_touched.p, e->last_touched.data(), (size_t)nt * 4, hipMemcpyHostToDevice,
                          e->stream));
  if (nr)
    HIPCHK(hipMemcpyAsync(e->d_last_relabel.p, e->last_relabel.data(), (size_t)nr * 4, hipMemcpyHostToDevice,
                          e->stream));
  // long lists queued: at most one per touched and one per dirty vertex
  HIPCHK(e->sync_q.ensure(2 * (size_t)e->n));
  HIPCHK(e->sync_ctl.ensure(2));
  const bool defer = deferred && has_deferred;
  HIPCHK(launch_sync(graph_of(e, c), e->dev_graph(), e->d_last_touched.p, nt, defer ? nullptr : e->dirty_list.p,
                     &c.actl.p->ndirty, e->d_last_relabel.p, nr, e->sync_q.p, (uint32_t)e->sync_q.n, e->sync_ctl.p,
                     e->stream));
  if (defer) {
    *deferred = SortedSync{graph_of(e, c), &c.actl.p->ndirty};
    *has_deferred = true;
  }
  HIPCHK(copy_buf(e->bm, c.bm, (size_t)e->nsig_rows * e->words, e->stream));
  HIPCHK(copy_buf(e->actl, c.actl, 1, e->stream));
  return DSM_OK;
}

int overlap_apply(dsm_engine *e, const HostDelta &d, bool stats, bool keep = true) {
  const ApplyCtl mirror = *e->h_actl.p;
  std::swap(e->stream, e->stream2);
  swap_graph(e);
  int rc = DSM_OK;
  auto run = [&]() -> int {
    HIPCHK(hipStreamWaitEvent(e->stream, e->ev_pre, 0));
    if (stats) HIPCHK(hipEventRecord(e->ev[0], e->stream));
    SortedSync ss;
    bool deferred = false;
    if (int r = catch_up(e, e->shadow, &ss, &deferred)) return r;
    if (int r = apply_delta_and_refresh(e, d, deferred ? &ss : nullptr)) return r;
    if (stats) HIPCHK(hipEventRecord(e->ev[2], e->stream));
    HIPCHK(hipMemcpyAsync(e->h_actl.p, e->actl.p, sizeof(ApplyCtl), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (stats) {
      e->stats.filter_entries += e->h_actl.p->filter_entries;
      e->stats.dirty_groups += e->h_actl.p->ndirty;
      e->stats.filter_bits += e->h_actl.p->filter_bits;
    }
    if (stats) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[2]));
      e->stats.ms_apply += ms;
      take_refine_time(e);
    }
    e->refine_timed = false;
    return DSM_OK;
  };
  rc = run();
  std::swap(e->stream, e->stream2);
  if (!keep) {
    swap_graph(e);
    *e->h_actl.p = mirror;
  } else if (rc == DSM_OK && !e->grew) {
    // the other copy now lags by exactly this update
    e->shadow_delta_ok = true;
    e->last_touched = d.tv;
    e->last_relabel = d.rl_v;
  }
  return rc;
}

}  // namespace

// =============================================================== C ABI

extern "C" {

void dsm_options_default(dsm_options *opt) {
  std::memset(opt, 0, sizeof *opt);
  opt->device = 0;
  opt->relabel = 1;
  opt->count_negatives = 1;
  opt->waves_per_block = 4;
  opt->leaf_sets = 1;
  opt->graph_rounds = 1;
  opt->leaf_segments = 1;
  opt->overlap_apply = 1;
}

const char *dsm_last_error(const dsm_engine *e) {
  if (!e) return g_create_error.c_str();
  return e->err.c_str();
}

int dsm_create(const dsm_graph_desc *graph, const dsm_query_desc *queries, uint32_t num_queries,
               const dsm_options *opt, dsm_engine **out) {
  if (!out || !graph || !queries) {
    g_create_error = "null argument";
    return DSM_EINVAL;
  }
  *out = nullptr;
  std::unique_ptr<dsm_engine> e(new dsm_engine());
  if (opt) e->opt = *opt;
  else dsm_options_default(&e->opt);
  auto fail = [&](int rc) {
    g_create_error = e->err;
    return rc;
  };
  if (graph->num_vertices == 0) {
    e->err = "empty data graph";
    return fail(DSM_EINVAL);
  }
  e->n = graph->num_vertices;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    e->err = "no HIP device available";
    return fail(DSM_EDEVICE);
  }
  if (e->opt.device < 0 || e->opt.device >= ndev) {
    e->err = "device ordinal out of range";
    return fail(DSM_EINVAL);
  }
  int rc = create_impl(e.get(), graph, queries, num_queries);
  if (rc) return fail(rc);
  *out = e.release();
  return DSM_OK;
}

void dsm_destroy(dsm_engine *e) {
  if (!e) return;
  (void)hipSetDevice(e->opt.device);
  (void)hipStreamSynchronize(e->stream);
  delete e;
}

int dsm_apply_batch(dsm_engine *e, const dsm_change *changes, uint32_t k) {
  if (!e || (!changes && k)) return DSM_EINVAL;
  for (uint32_t i = 0; i < k; i++)
    if (changes[i].src >= e->n || changes[i].dst >= e->n)
      return e->fail_msg(DSM_EINVAL, "change %u: endpoint out of range (%u,%u) for %u vertices", i,
                         changes[i].src, changes[i].dst, e->n);
  HIPCHK(hipSetDevice(e->opt.device));
  e->stats = dsm_stats{};
  e->pos_s.clear();
  e->pos_d.clear();
  e->neg_s.clear();
  e->neg_d.clear();
  for (uint32_t i = 0; i < k; i++) {
    if (changes[i].add) {
      e->pos_s.push_back(changes[i].src);
      e->pos_d.push_back(changes[i].dst);
    } else {
      e->neg_s.push_back(changes[i].src);
      e->neg_d.push_back(changes[i].dst);
    }
  }
  const uint32_t nq = (uint32_t)e->queries.size();
  HIPCHK(launch_ctl_reset(nullptr, nullptr, e->tuple_count.p, e->counts.p, nq * 2, e->stream));
  int rc;
  bool applied = false;
  // 1. negatives in the pre-batch graph; with overlap_apply the batch is
  // applied to the second copy of the graph meanwhile
  if (e->opt.count_negatives && !e->neg_s.empty()) {
    const uint32_t nn = (uint32_t)e->neg_s.size();
    if ((rc = ensure_seed_cap(e, nn))) return rc;
    DevBuf<uint32_t> &rb = e->rec_neg;
    if ((rc = upload_records(e, e->neg_s, e->neg_d, rb))) return rc;
    if (e->overlap_apply) {
      // the update refills the staging buffer the records came from
      HIPCHK(hipStreamSynchronize(e->stream));
      HIPCHK(hipEventRecord(e->ev_pre, e->stream));
    }
    if ((rc = run_round(e, rb.p, rb.p + nn, nn, 0, e->ev[6], e->ev[7], e->ev[8]))) return rc;
    if (e->overlap_apply) {
      HostDelta d;
      host_delta(e, changes, k, d);
      if ((rc = overlap_apply(e, d, true))) return rc;
      applied = true;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    if ((rc = check_round(e, &e->stats.seeds_neg))) return rc;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[7], e->ev[8]));
    e->stats.ms_dfs_neg = ms;
  }
  // 2-3. apply + relabel + filter (update, relabel, refilter and
  // label-sorted copies; already done beside the negatives with overlap)
  if (!applied) {
    HIPCHK(hipEventRecord(e->ev[0], e->stream));
    if ((rc = apply_and_refresh(e, changes, k))) return rc;
    HIPCHK(hipEventRecord(e->ev[2], e->stream));
  }
  HIPCHK(hipMemcpyAsync(e->h_actl.p, e->actl.p, sizeof(ApplyCtl), hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  if (!applied) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[2]));
    e->stats.ms_apply = ms;
    take_refine_time(e);
  }
  e->stats.ms_filter = 0;
  if (!applied) {
    e->stats.filter_entries = e->h_actl.p->filter_entries;
    e->stats.dirty_groups = e->h_actl.p->ndirty;
    e->stats.filter_bits = e->h_actl.p->filter_bits;
  }
  e->stats.relabeled_vertices = e->h_actl.p->relabeled;
  e->stats.pool_entries_used = e->h_actl.p->pool_used;
  e->stats.pool_entries_capacity = e->pool_cap;
  e->batch_pending = true;
  return DSM_OK;
}

int dsm_match_batch(dsm_engine *e, uint64_t *pos, uint64_t *neg) {
  if (!e) return DSM_EINVAL;
  HIPCHK(hipSetDevice(e->opt.device));
  int rc;
  const uint32_t nq = (uint32_t)e->queries.size();
  if (e->batch_pending && !e->pos_s.empty()) {
    const uint32_t np = (uint32_t)e->pos_s.size();
    if ((rc = ensure_seed_cap(e, np))) return rc;
    DevBuf<uint32_t> &rb = e->rec_pos;
    if ((rc = upload_records(e, e->pos_s, e->pos_d, rb))) return rc;
    if ((rc = run_round(e, rb.p, rb.p + np, np, 1, e->ev[3], e->ev[4], e->ev[5]))) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    if ((rc = check_round(e, &e->stats.seeds_pos))) return rc;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[3], e->ev[4]));
    e->stats.ms_seed = ms;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[4], e->ev[5]));
    e->stats.ms_dfs_pos = ms;
  }
  e->batch_pending = false;
  HIPCHK(hipMemcpyAsync(e->h_counts.p, e->counts.p, (size_t)nq * 2 * 8, hipMemcpyDeviceToHost, e->stream));
  unsigned long long tc = 0;
  HIPCHK(hipMemcpyAsync(&tc, e->tuple_count.p, 8, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  e->stats.tuples = tc;
  for (uint32_t q = 0; q < nq; q++) {
    if (pos) pos[q] = e->h_counts.p[2 * q + 1];
    if (neg) neg[q] = e->h_counts.p[2 * q + 0];
  }
  return DSM_OK;
}

int dsm_process_batch(dsm_engine *e, const dsm_change *changes, uint32_t k, uint64_t *pos, uint64_t *neg) {
  int rc = dsm_apply_batch(e, changes, k);
  if (rc) return rc;
  return dsm_match_batch(e, pos, neg);
}

int dsm_get_tuples(dsm_engine *e, uint32_t *out, uint64_t cap, uint64_t *num_out) {
  if (!e || !num_out) return DSM_EINVAL;
  const uint64_t found = e->stats.tuples;
  *num_out = found;
  const uint64_t have = std::min<uint64_t>(found, e->tuple_cap);
  const uint64_t copy = std::min<uint64_t>(have, cap);
  if (copy && out) {
    HIPCHK(hipSetDevice(e->opt.device));
    HIPCHK(hipMemcpy(out, e->tuples, copy * (kMaxQ + 2) * 4, hipMemcpyDefault));
  }
  if (found > e->tuple_cap) return e->fail_msg(DSM_ETRUNC, "%llu tuples found, capacity %llu",
                                               (unsigned long long)found, (unsigned long long)e->tuple_cap);
  return DSM_OK;
}

int dsm_get_stats(const dsm_engine *e, dsm_stats *out) {
  if (!e || !out) return DSM_EINVAL;
  *out = e->stats;
  return DSM_OK;
}

int dsm_get_graph(dsm_engine *e, uint64_t *offsets, uint32_t *neighbours, uint64_t cap, uint32_t *labels,
                  uint64_t *num_entries) {
  if (!e || !offsets || !num_entries) return DSM_EINVAL;
  HIPCHK(hipSetDevice(e->opt.device));
  HIPCHK(hipStreamSynchronize(e->stream));
  const uint32_t n = e->n;
  std::vector<uint32_t> st(n), dg(n);
  std::vector<uint8_t> lb(n);
  HIPCHK(hipMemcpy(st.data(), e->start.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(dg.data(), e->deg.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(lb.data(), e->lab.p, n, hipMemcpyDeviceToHost));
  uint64_t o = 0;
  for (uint32_t v = 0; v < n; v++) {
    offsets[v] = o;
    o += dg[v];
  }
  offsets[n] = o;
  *num_entries = o;
  if (labels)
    for (uint32_t v = 0; v < n; v++) labels[v] = lb[v];
  if (o > cap || !neighbours) return o > cap ? e->fail_msg(DSM_ETRUNC, "neighbour buffer too small") : DSM_OK;
  std::vector<uint32_t> hp((size_t)e->pool_cap);
  HIPCHK(hipMemcpy(hp.data(), e->pool.p, (size_t)e->pool_cap * 4, hipMemcpyDeviceToHost));
  for (uint32_t v = 0; v < n; v++) std::copy(hp.begin() + st[v], hp.begin() + st[v] + dg[v], neighbours + offsets[v]);
  return DSM_OK;
}

int dsm_get_candidates(dsm_engine *e, uint32_t q, uint32_t u, uint64_t *words) {
  if (!e || !words || q >= e->queries.size() || u >= e->queries[q].n) return DSM_EINVAL;
  HIPCHK(hipSetDevice(e->opt.device));
  HIPCHK(hipStreamSynchronize(e->stream));
  const uint32_t s = e->qsig[q][u];
  HIPCHK(hipMemcpy(words, e->bm.p + (size_t)s * e->words, (size_t)e->words * 8, hipMemcpyDeviceToHost));
  return DSM_OK;
}

uint32_t dsm_cluster_plan(uint32_t num_vertices, uint32_t num_edges, const uint32_t *edge_src,
                          const uint32_t *edge_dst, uint32_t *out, uint32_t cap) {
  if (!edge_src || !edge_dst || num_vertices > (uint32_t)kMaxQ) return 0;
  for (uint32_t i = 0; i < num_edges; i++)
    if (edge_src[i] >= num_vertices || edge_dst[i] >= num_vertices) return 0;
  ClusterPlan P;
  if (build_cluster_plan(num_vertices, num_edges, edge_src, edge_dst, P)) return 0;
  uint32_t w = 0;
  auto put = [&](uint32_t x) {
    if (out && w < cap) out[w] = x;
    w++;
  };
  put(num_vertices);
  uint32_t ncover = 0;
  for (auto &p : P.passes) ncover += p.kind == ClusterPlan::kEncode;
  put(ncover);
  for (auto &p : P.passes)
    if (p.kind == ClusterPlan::kEncode) put(P.clusters[p.arg].members[0]);
  put((uint32_t)P.clusters.size());
  for (auto &c : P.clusters) {
    put(c.valid ? 1 : 0);
    put((uint32_t)c.members.size());
    for (uint32_t u : c.members) put(u);
  }
  put(P.final_cluster);
  for (uint32_t u : P.final_pos) put(u);
  put((uint32_t)P.passes.size());
  for (auto &p : P.passes) {
    put(p.kind);
    put(p.arg);
  }
  return w;
}

int dsm_get_refined_candidates(dsm_engine *e, uint32_t q, uint32_t u, uint64_t *words) {
  if (!e || !words || q >= e->queries.size() || u >= e->queries[q].n) return DSM_EINVAL;
  HIPCHK(hipSetDevice(e->opt.device));
  HIPCHK(hipStreamSynchronize(e->stream));
  const uint32_t r = e->qref[q][u];
  HIPCHK(hipMemcpy(words, e->bm.p + (size_t)r * e->words, (size_t)e->words * 8, hipMemcpyDeviceToHost));
  return DSM_OK;
}

int dsm_get_query_time(dsm_engine *e, double *ms, uint32_t num_queries) {
  if (!e || !ms || num_queries != e->queries.size()) return DSM_EINVAL;
  HIPCHK(hipSetDevice(e->opt.device));
  HIPCHK(hipStreamSynchronize(e->stream));
  std::vector<unsigned long long> t(num_queries);
  HIPCHK(hipMemcpy(t.data(), e->query_busy.p, (size_t)num_queries * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemset(e->query_busy.p, 0, (size_t)num_queries * 8));
  for (uint32_t q = 0; q < num_queries; q++) ms[q] = (double)t[q] / (double)e->wall_khz;
  return DSM_OK;
}

int dsm_autotune(dsm_engine *e, const dsm_change *sample, uint32_t k, dsm_tune_report *rep) {
  if (!e || (!sample && k)) return DSM_EINVAL;
  if (e->batch_pending) return e->fail_msg(DSM_EINVAL, "a batch is applied but not matched yet");
  for (uint32_t i = 0; i < k; i++)
    if (sample[i].src >= e->n || sample[i].dst >= e->n)
      return e->fail_msg(DSM_EINVAL, "sample record %u: endpoint out of range", i);
  HIPCHK(hipSetDevice(e->opt.device));
  int rc;
  const uint32_t nq = (uint32_t)e->queries.size();
  std::vector<uint32_t> s(k), d(k);
  for (uint32_t i = 0; i < k; i++) {
    s[i] = sample[i].src;
    d[i] = sample[i].dst;
  }
  if ((rc = ensure_seed_cap(e, k))) return rc;
  DevBuf<uint32_t> &rb = e->rec_neg;
  if ((rc = upload_records(e, s, d, rb))) return rc;
  // the staging buffer is free again before an update (below) refills it
  HIPCHK(hipStreamSynchronize(e->stream));
  struct Cfg {
    uint32_t wpb, bpc, dmin, lw;
    bool ls, seg;
    uint32_t dev, lb;
    bool pl;
  };
  auto max_bpc = [&](uint32_t wpb) { return dfs_bpc(e, wpb); };
  // With overlap_apply, pipelined rounds run beside the next batch's update
  // (on the second copy of the graph), which competes for the CUs: each trial
  // round then runs beside the update of the first quarter of the sample (the
  // sample being a few batches), applied to the second copy and dropped, and
  // a trial's time is the makespan of both.
  HostDelta load;
  const bool loaded = e->overlap_apply && !e->tuple_cap && k >= 4;
  if (loaded) host_delta(e, sample, k / 4, load);
  // times one sample round under c and checks its counts against the first
  std::vector<unsigned long long> ref_counts;
  uint32_t trials = 0;
  auto timed = [&](const Cfg &c, float &ms) -> int {
    e->dfs_wpb = c.wpb;
    e->dfs_blocks = c.bpc * (uint32_t)e->num_cu;
    e->donate_min = c.dmin;
    e->lane_work = c.lw;
    e->leaf_budget = c.lb;
    e->leaf_sets = c.ls;
    e->leaf_segments = c.seg;
    e->donate_every = c.dev;
    e->bit_probe_last = c.pl;
    if (loaded) HIPCHK(hipEventRecord(e->ev_pre, e->stream));
    HIPCHK(launch_ctl_reset(nullptr, nullptr, e->tuple_count.p, e->counts.p, nq * 2, e->stream));
    int rc2 = run_round(e, rb.p, rb.p + k, k, 0, e->ev[9], e->ev[10], e->ev[11]);
    if (rc2) return rc2;
    if (loaded && (rc2 = overlap_apply(e, load, true, false))) return rc2;
    HIPCHK(hipMemcpyAsync(e->h_counts.p, e->counts.p, (size_t)nq * 2 * 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    uint64_t seeds = 0;
    if ((rc2 = check_round(e, &seeds))) return rc2;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[10], e->ev[11]));
    if (loaded) {
      // the makespan of the round and the update beside it: a shape that
      // leaves the update no room on the CUs delays it past the round
      float ua = 0;
      HIPCHK(hipEventElapsedTime(&ua, e->ev[10], e->ev[2]));
      ms = std::max(ms, ua);
    }
    std::vector<unsigned long long> cnt(e->h_counts.p, e->h_counts.p + 2 * (size_t)nq);
    if (ref_counts.empty()) ref_counts = cnt;
    else if (cnt != ref_counts)
      return e->fail_msg(DSM_EDEVICE, "autotune: launch configurations disagree on the counts");
    return DSM_OK;
  };
  Cfg best{e->dfs_wpb, e->dfs_blocks / (uint32_t)e->num_cu, e->donate_min, e->lane_work, e->leaf_sets,
           e->leaf_segments, e->donate_every, e->leaf_budget, e->bit_probe_last};
  // warm-up: clocks and caches settle before anything is compared
  float best_ms = 1e30f;
  for (int r = 0; r < 4; r++) {
    float t = 0;
    if ((rc = timed(best, t))) return rc;
    if (r >= 2) best_ms = std::min(best_ms, t);
  }
  const float ms0 = best_ms;
  // A candidate is timed in three pairs interleaved with the incumbent (so a
  // drift of the clocks hits both alike) and must beat it by 3% in every pair.
  auto consider = [&](const Cfg &c) -> int {
    float tb = 1e30f, tc = 1e30f;
    bool wins = true;
    for (int r = 0; r < 3; r++) {
      float t = 0, u = 0;
      int rc2 = timed(best, t);
      if (rc2) return rc2;
      tb = std::min(tb, t);
      if ((rc2 = timed(c, u))) return rc2;
      tc = std::min(tc, u);
      wins = wins && u < 0.97f * t;
    }
    trials++;
    if (wins) {
      best = c;
      best_ms = tc;
    } else {
      best_ms = std::min(best_ms, tb);
    }
    return DSM_OK;
  };
  // 1. waves per block (each at its occupancy limit); a tuple engine has one shape
  if (!e->tuple_cap) {
    const Cfg base = best;
    for (uint32_t w : {1u, 2u, 4u, 8u}) {
      if (w == base.wpb || w * dfs_lds_bytes_per_wave(e->max_qn) > 65536) continue;
      Cfg c = base;
      c.wpb = w;
      c.bpc = max_bpc(w);
      if ((rc = consider(c))) return rc;
    }
  }
  // 2. blocks per CU below the occupancy limit
  {
    const Cfg base = best;
    const uint32_t m = max_bpc(base.wpb);
    for (uint32_t b : {m, (m * 3 + 3) / 4, (m + 1) / 2}) {
      if (b == base.bpc || b == 0) continue;
      Cfg c = base;
      c.bpc = b;
      if ((rc = consider(c))) return rc;
    }
  }
  // 3. donation threshold and how often busy waves look for idle ones
  {
    const Cfg base = best;
    for (uint32_t v : {64u, 128u, 256u, 512u, 1024u}) {
      if (v == base.dmin) continue;
      Cfg c = base;
      c.dmin = v;
      if ((rc = consider(c))) return rc;
    }
    const Cfg base2 = best;
    for (uint32_t v : {4u, 8u, 16u, 32u}) {
      if (v == base2.dev) continue;
      Cfg c = base2;
      c.dev = v;
      if ((rc = consider(c))) return rc;
    }
  }
  // 4. leaf sets, 5. lane work
  if (!e->tuple_cap) {
    Cfg c = best;
    c.ls = !best.ls;
    if ((rc = consider(c))) return rc;
    if (best.ls) {
      Cfg c3 = best;
      c3.seg = !best.seg;
      if ((rc = consider(c3))) return rc;
      const Cfg base = best;
      for (uint32_t v : {8u, 16u, 24u, 48u, 64u}) {
        if (v == base.lw) continue;
        Cfg c2 = base;
        c2.lw = v;
        if ((rc = consider(c2))) return rc;
      }
      // 6. leaf budget per window (0xFFFFFFFF: none)
      const Cfg base6 = best;
      for (uint32_t v : {512u, 2048u, 8192u, 32768u, 0xFFFFFFFFu}) {
        if (v == base6.lb) continue;
        Cfg c6 = base6;
        c6.lb = v;
        if ((rc = consider(c6))) return rc;
      }
    }
  }
  // 7. candidate bit before or after the membership tests
  {
    Cfg c = best;
    c.pl = !best.pl;
    if ((rc = consider(c))) return rc;
  }
  // 8. matching order per query. The queries' plans are rebuilt with each
  // order rule in turn (all queries at once: each query's task time is
  // measured on its own), in the order 0..3 then 3..0 so that a drift of the
  // clocks hits every rule alike; a query keeps the rule with the least task
  // time if it beats its current rule by 5%. The new orders are kept only if
  // the sample round is then faster in three interleaved pairs.
  uint32_t order_changed = 0;
  {
    auto use_rules = [&](const std::vector<uint32_t> &rule) -> int {
      for (uint32_t q = 0; q < nq; q++) e->queries[q].order_variant = rule[q];
      build_plans(e);
      HIPCHK(hipMemcpy(e->d_plans.p, e->plans.data(), e->plans.size() * sizeof(Plan), hipMemcpyHostToDevice));
      return DSM_OK;
    };
    auto take_query_ms = [&](std::vector<double> *acc) -> int {
      std::vector<double> ms(nq);
      if (int r2 = dsm_get_query_time(e, ms.data(), nq)) return r2;
      if (acc)
        for (uint32_t q = 0; q < nq; q++) (*acc)[q] += ms[q];
      return DSM_OK;
    };
    std::vector<uint32_t> cur(nq);
    for (uint32_t q = 0; q < nq; q++) cur[q] = e->queries[q].order_variant;
    std::vector<std::vector<double>> qms(kOrderVariants, std::vector<double>(nq, 0.0));
    if ((rc = take_query_ms(nullptr))) return rc;
    for (uint32_t pass = 0; pass < 2; pass++)
      for (uint32_t i = 0; i < kOrderVariants; i++) {
        const uint32_t v = pass ? kOrderVariants - 1 - i : i;
        if ((rc = use_rules(std::vector<uint32_t>(nq, v)))) return rc;
        float t = 0;
        if ((rc = timed(best, t))) return rc;
        if ((rc = take_query_ms(&qms[v]))) return rc;
      }
    std::vector<uint32_t> pick = cur;
    for (uint32_t q = 0; q < nq; q++) {
      double bt = qms[cur[q]][q];
      for (uint32_t v = 0; v < kOrderVariants; v++)
        if (qms[v][q] < 0.95 * qms[cur[q]][q] && qms[v][q] < bt) {
          bt = qms[v][q];
          pick[q] = v;
        }
    }
    if (pick != cur) {
      float t_cur = 0, t_new = 0;
      for (int r = 0; r < 3; r++) {
        float t = 0;
        if ((rc = use_rules(cur)) || (rc = timed(best, t))) return rc;
        t_cur += t;
        if ((rc = use_rules(pick)) || (rc = timed(best, t))) return rc;
        t_new += t;
      }
      if (t_new < t_cur) {
        for (uint32_t q = 0; q < nq; q++) order_changed += pick[q] != cur[q];
      } else if ((rc = use_rules(cur))) {
        return rc;
      }
    } else if ((rc = use_rules(cur))) {
      return rc;
    }
    if ((rc = take_query_ms(nullptr))) return rc;
  }
  // the kept configuration, timed once more
  best_ms = 1e30f;
  for (int r = 0; r < 2; r++) {
    float t = 0;
    if ((rc = timed(best, t))) return rc;
    best_ms = std::min(best_ms, t);
  }
  if (rep) {
    rep->waves_per_block = best.wpb;
    rep->blocks_per_cu = best.bpc;
    rep->donate_min = best.dmin;
    rep->lane_work = best.lw;
    rep->leaf_sets = best.ls ? 1 : 0;
    rep->leaf_segments = best.seg ? 1 : 0;
    rep->donate_every = best.dev;
    rep->leaf_budget = best.lb;
    rep->bit_probe_last = best.pl ? 1 : 0;
    rep->order_changed = order_changed;
    rep->trials = trials;
    rep->ms_initial = ms0;
    rep->ms_best = best_ms;
  }
  return DSM_OK;
}

uint32_t dsm_num_vertices(const dsm_engine *e) { return e ? e->n : 0; }
uint32_t dsm_num_queries(const dsm_engine *e) { return e ? (uint32_t)e->queries.size() : 0; }

}  // extern "C"

// ============================================================ served stream
//
// A request is one update batch. The worker thread keeps a rolling pipeline
// over the submitted requests, in submission order, with the semantics of
// dsm_process_batch for each: request R's negatives are matched in the graph
// before R's update, its positives after it. The positives of the request in
// flight (P, already applied) and the negatives of the next one (R) see the
// same graph, so one DFS launch matches both, each seed tagged with its
// request's count slot. While that launch runs, the host computes R's delta;
// when it ends, P's counts are complete and are published at once, and R's
// update is applied. A request is admitted when it and the one in flight fit
// the memory budget together; one that does not fit even alone is rejected.

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// One update batch in the rolling pipeline (served stream and
// dsm_process_batches): its records, split into positive and negative seed
// lists, its count slot and, once published, its counts.
struct PipeItem {
  const dsm_change *rec = nullptr;
  uint32_t nrec = 0;
  std::vector<uint32_t> ps, pd, ns, nd;
  uint32_t slot = 0;
  std::vector<uint64_t> pos, neg;
};

void pipe_split(const dsm_engine *e, PipeItem &it) {
  for (uint32_t i = 0; i < it.nrec; i++) {
    const dsm_change &c = it.rec[i];
    if (c.add) {
      it.ps.push_back(c.src);
      it.pd.push_back(c.dst);
    } else if (e->opt.count_negatives) {
      it.ns.push_back(c.src);
      it.nd.push_back(c.dst);
    }
  }
}

struct ServerReq : PipeItem {
  uint64_t id = 0;
  std::vector<dsm_change> ch;
  uint64_t bytes = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
  uint32_t queue_depth = 0, fused = 0;
  int status = 1;  // 1 while pending or in flight
};

}  // namespace

struct dsm_server {
  dsm_engine *e = nullptr;
  uint64_t budget = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ServerReq *> pending;
  std::unordered_map<uint64_t, ServerReq *> reqs;
  uint64_t next_id = 1;
  bool stop = false;
  int failed = DSM_OK;  // after a device error every later request fails with it
  std::thread worker;
  PinnedBuf<uint32_t> h_rec;
  DevBuf<uint32_t> d_rec;
  std::string err;
};

namespace {

uint64_t request_bytes(const dsm_engine *e, uint32_t k) {
  // seed table (plan, src, dst, tag per seed) + records on the device + the
  // host-side delta arrays uploaded for the update
  return (uint64_t)k * ((uint64_t)std::max<uint32_t>(e->max_plans_per_key, 1) * 16 + 8 + 48);
}

void publish(dsm_server *sv, ServerReq *r, int status) {
  std::lock_guard<std::mutex> lk(sv->mu);
  r->status = status;
  r->t_done = now_s();
  sv->cv_done.notify_all();
}

// One DFS round over the current graph: the positives of P (if any) and the
// negatives of R (if any); then P's count slot is copied to the host. Waits
// for the round; while it runs, computes R's delta into *d, unless
// `have_delta` says that *d already holds it. With `stats`, the round's work
// counters and time go to the engine's statistics.
// One pipelined round: the positives of P and the negatives of R in one DFS
// launch on the current graph. With `applied` given and overlap_apply on, R's
// update is applied to the second copy of the graph meanwhile and *applied is
// set; otherwise the caller applies it after the round.
int pipe_round(dsm_engine *e, PinnedBuf<uint32_t> &h_rec, DevBuf<uint32_t> &d_rec, PipeItem *P, PipeItem *R,
               HostDelta *d, bool stats, bool *applied, bool have_delta = false) {
  const uint32_t nq = (uint32_t)e->queries.size();
  int rc;
  const uint32_t np = P ? (uint32_t)P->ps.size() : 0u, nn = R ? (uint32_t)R->ns.size() : 0u;
  const bool overlap = applied && R && d && e->overlap_apply;
  if (applied) *applied = false;
  if (overlap) HIPCHK(hipEventRecord(e->ev_pre, e->stream));
  if (R) HIPCHK(launch_ctl_reset(nullptr, nullptr, nullptr, e->counts.p + (size_t)R->slot * nq * 2, nq * 2,
                                 e->stream));
  if (np + nn) {
    if ((rc = ensure_seed_cap(e, np + nn))) return rc;
    HIPCHK(h_rec.ensure(2 * (size_t)(np + nn) + 2));
    HIPCHK(d_rec.ensure(2 * (size_t)(np + nn) + 2));
    uint32_t *h = h_rec.p;
    if (np) {
      std::copy(P->ps.begin(), P->ps.end(), h);
      std::copy(P->pd.begin(), P->pd.end(), h + np);
    }
    if (nn) {
      std::copy(R->ns.begin(), R->ns.end(), h + 2 * np);
      std::copy(R->nd.begin(), R->nd.end(), h + 2 * np + nn);
    }
    HIPCHK(hipMemcpyAsync(d_rec.p, h, 2 * (size_t)(np + nn) * 4, hipMemcpyHostToDevice, e->stream));
    RoundPart parts[2];
    int nparts = 0;
    if (np) parts[nparts++] = RoundPart{d_rec.p, d_rec.p + np, np, (P->slot << 1) | 1u};
    if (nn) parts[nparts++] = RoundPart{d_rec.p + 2 * np, d_rec.p + 2 * np + nn, nn, R->slot << 1};
    if ((rc = dfs_round(e, parts, nparts, stats ? e->ev[9] : nullptr, stats ? e->ev[10] : nullptr,
                        stats ? e->ev[11] : nullptr)))
      return rc;
  }
  if (P)
    HIPCHK(hipMemcpyAsync(e->h_counts.p + (size_t)P->slot * nq * 2, e->counts.p + (size_t)P->slot * nq * 2,
                          (size_t)nq * 2 * 8, hipMemcpyDeviceToHost, e->stream));
  if (R && d && !have_delta) host_delta(e, R->rec, R->nrec, *d);
  if (overlap) {
    if ((rc = overlap_apply(e, *d, stats))) return rc;
    *applied = true;
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  if (np + nn) {
    const DfsCtl &c = *e->h_ctl.p;
    if (c.seed_overflow.v) return e->fail_msg(DSM_ELIMIT, "seed table overflow in a pipelined round");
    if (c.abort.v)
      return e->fail_msg(DSM_EDEVICE, "DFS watchdog: a wave waited more than %u ms for work", e->watchdog_ms);
    if (stats) {
      uint64_t seeds = 0;
      if ((rc = check_round(e, &seeds))) return rc;
      e->stats.seeds_pos += seeds;
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e->ev[9], e->ev[11]));
      e->stats.ms_dfs_pos += ms;
    }
  }
  if (P) {
    P->pos.assign(nq, 0);
    P->neg.assign(nq, 0);
    const unsigned long long *c = e->h_counts.p + (size_t)P->slot * nq * 2;
    for (uint32_t q = 0; q < nq; q++) {
      P->pos[q] = c[2 * q + 1];
      P->neg[q] = c[2 * q + 0];
    }
  }
  return DSM_OK;
}

int serve_round(dsm_server *sv, ServerReq *P, ServerReq *R, HostDelta *d, bool *applied) {
  return pipe_round(sv->e, sv->h_rec, sv->d_rec, P, R, d, false, applied);
}

void server_loop(dsm_server *sv) {
  dsm_engine *e = sv->e;
  if (hipSetDevice(e->opt.device) != hipSuccess) sv->failed = DSM_EDEVICE;
  ServerReq *P = nullptr;  // applied, positives not yet matched
  for (;;) {
    ServerReq *R = nullptr;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      if (!P) sv->cv_work.wait(lk, [&] { return sv->stop || !sv->pending.empty(); });
      while (!sv->pending.empty()) {
        ServerReq *c = sv->pending.front();
        const int adm = sv->failed != DSM_OK ? sv->failed : dsm_admit(c->bytes, P ? P->bytes : 0, sv->budget);
        if (adm < 0) {
          // rejected without touching the graph
          sv->pending.pop_front();
          c->status = adm;
          c->t_start = c->t_done = now_s();
          sv->cv_done.notify_all();
          continue;
        }
        if (adm == 0) break;  // waits until the request in flight is done
        sv->pending.pop_front();
        R = c;
        R->queue_depth = (uint32_t)sv->pending.size();
        break;
      }
      if (!P && !R) {
        if (sv->stop && sv->pending.empty()) return;
        continue;
      }
    }
    if (R) {
      R->t_start = now_s();
      R->slot = P ? P->slot ^ 1u : 0u;
      R->fused = P ? 1u : 0u;
      pipe_split(e, *R);
    }
    HostDelta d;
    bool applied = false;
    int rc = serve_round(sv, P, R, R ? &d : nullptr, &applied);
    if (P) publish(sv, P, rc);
    if (rc == DSM_OK && R && !applied) rc = apply_delta_and_refresh(e, d);
    if (rc != DSM_OK) {
      std::lock_guard<std::mutex> lk(sv->mu);
      sv->failed = rc;
      sv->err = e->err;
    }
    if (R && rc != DSM_OK) {
      publish(sv, R, rc);
      R = nullptr;
    }
    P = R;
  }
}

}  // namespace

extern "C" {

int dsm_admit(uint64_t request_bytes, uint64_t in_flight_bytes, uint64_t budget) {
  if (request_bytes > budget) return DSM_EADMIT;
  return in_flight_bytes <= budget - request_bytes ? 1 : 0;
}

uint64_t dsm_estimate_request_bytes(const dsm_engine *e, uint32_t num_changes) {
  return e ? request_bytes(e, num_changes) : 0;
}

int dsm_server_create(dsm_engine *e, const dsm_server_options *opt, dsm_server **out) {
  if (!e || !out) return DSM_EINVAL;
  *out = nullptr;
  if (e->tuple_cap) return e->fail_msg(DSM_EINVAL, "served mode keeps counts only (tuple_capacity must be 0)");
  HIPCHK(hipSetDevice(e->opt.device));
  HIPCHK(hipStreamSynchronize(e->stream));
  // the second copy of the graph is part of the engine, not of the budget
  if (e->overlap_apply)
    if (int rc = reserve_shadow(e)) return rc;
  std::unique_ptr<dsm_server> sv(new dsm_server());
  sv->e = e;
  size_t freeb = 0, totalb = 0;
  HIPCHK(hipMemGetInfo(&freeb, &totalb));
  const double headroom = (opt && opt->hbm_headroom > 0) ? opt->hbm_headroom : 0.30;
  const double reserve = headroom * (double)totalb;
  if (opt && opt->budget_bytes) sv->budget = opt->budget_bytes;
  else sv->budget = (double)freeb > reserve ? (uint64_t)((double)freeb - reserve) : 0;
  dsm_server *p = sv.get();
  p->worker = std::thread(server_loop, p);
  *out = sv.release();
  return DSM_OK;
}

int dsm_server_submit(dsm_server *sv, const dsm_change *changes, uint32_t num_changes, uint64_t *id) {
  if (!sv || (!changes && num_changes) || !id) return DSM_EINVAL;
  const uint32_t n = sv->e->n;
  for (uint32_t i = 0; i < num_changes; i++)
    if (changes[i].src >= n || changes[i].dst >= n) return DSM_EINVAL;
  ServerReq *r = new ServerReq();
  r->ch.assign(changes, changes + num_changes);
  r->rec = r->ch.data();
  r->nrec = num_changes;
  r->bytes = request_bytes(sv->e, num_changes);
  r->t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    r->id = sv->next_id++;
    sv->reqs[r->id] = r;
    sv->pending.push_back(r);
    *id = r->id;
  }
  sv->cv_work.notify_one();
  return DSM_OK;
}

int dsm_server_wait(dsm_server *sv, uint64_t id, dsm_request_info *info, uint64_t *pos, uint64_t *neg) {
  if (!sv) return DSM_EINVAL;
  ServerReq *r;
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return DSM_EINVAL;
    r = it->second;
    sv->cv_done.wait(lk, [&] { return r->status != 1; });
    sv->reqs.erase(it);
  }
  const uint32_t nq = (uint32_t)sv->e->queries.size();
  if (info) {
    info->id = r->id;
    info->status = r->status;
    info->t_submit = r->t_submit;
    info->t_start = r->t_start;
    info->t_done = r->t_done;
    info->queue_depth = r->queue_depth;
    info->fused = r->fused;
    info->num_changes = (uint32_t)r->ch.size();
  }
  for (uint32_t q = 0; q < nq; q++) {
    if (pos) pos[q] = r->status == DSM_OK ? r->pos[q] : 0;
    if (neg) neg[q] = r->status == DSM_OK ? r->neg[q] : 0;
  }
  const int st = r->status;
  delete r;
  return st;
}

const char *dsm_server_error(const dsm_server *sv) { return sv ? sv->err.c_str() : ""; }

int dsm_process_batches(dsm_engine *e, const dsm_change *changes, const uint32_t *batch_offsets,
                        uint32_t num_batches, uint64_t *pos, uint64_t *neg, double *done_s) {
  if (!e || !batch_offsets || (!changes && batch_offsets[num_batches]) || !pos || !neg) return DSM_EINVAL;
  if (e->tuple_cap) return e->fail_msg(DSM_EINVAL, "pipelined batches keep counts only (tuple_capacity must be 0)");
  if (e->batch_pending) return e->fail_msg(DSM_EINVAL, "a batch is applied but not matched yet");
  for (uint32_t b = 0; b < num_batches; b++)
    if (batch_offsets[b + 1] < batch_offsets[b]) return e->fail_msg(DSM_EINVAL, "batch offsets decrease");
  for (uint32_t i = 0; i < batch_offsets[num_batches]; i++)
    if (changes[i].src >= e->n || changes[i].dst >= e->n)
      return e->fail_msg(DSM_EINVAL, "change %u: endpoint out of range (%u,%u) for %u vertices", i,
                         changes[i].src, changes[i].dst, e->n);
  HIPCHK(hipSetDevice(e->opt.device));
  e->stats = dsm_stats{};
  const uint32_t nq = (uint32_t)e->queries.size();
  std::vector<PipeItem> items(num_batches);
  for (uint32_t b = 0; b < num_batches; b++) {
    items[b].rec = changes + batch_offsets[b];
    items[b].nrec = batch_offsets[b + 1] - batch_offsets[b];
    pipe_split(e, items[b]);
  }
  auto publish_item = [&](uint32_t b) {
    for (uint32_t q = 0; q < nq; q++) {
      pos[(size_t)b * nq + q] = items[b].pos[q];
      neg[(size_t)b * nq + q] = items[b].neg[q];
    }
    if (done_s) done_s[b] = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  };
  int rc;
  bool apply_timed = false;
  auto take_apply_time = [&]() -> int {
    if (!apply_timed) return DSM_OK;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[2]));
    e->stats.ms_apply += ms;
    apply_timed = false;
    return DSM_OK;
  };
  // A batch's host delta (its records as per-vertex insert and delete lists)
  // depends on its records alone. A worker thread computes them in batch
  // order, ahead of the rounds, while this thread waits for the GPU; a round
  // then enqueues its update right behind its DFS launch instead of after
  // sorting its batch (at 32,768 records, a few ms per batch). The worker
  // stays at most kDeltasAhead batches ahead of the rounds, so the deltas
  // held at once stay bounded however many batches the call takes.
  constexpr uint32_t kDeltasAhead = 2;
  std::vector<HostDelta> deltas(num_batches);
  std::mutex dmu;
  std::condition_variable ready_cv, taken_cv;
  uint32_t dready = 0, dtaken = 0;
  bool dstop = false;
  std::thread dworker([&] {
    for (uint32_t b = 0; b < num_batches; b++) {
      {
        std::unique_lock<std::mutex> lk(dmu);
        taken_cv.wait(lk, [&] { return dstop || b < dtaken + kDeltasAhead; });
        if (dstop) return;
      }
      host_delta(e, items[b].rec, items[b].nrec, deltas[b]);
      {
        std::lock_guard<std::mutex> lk(dmu);
        dready = b + 1;
      }
      ready_cv.notify_one();
    }
  });
  struct JoinWorker {
    std::thread &t;
    std::mutex &mu;
    std::condition_variable &cv;
    bool &stop;
    ~JoinWorker() {
      {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
      }
      cv.notify_one();
      t.join();
    }
  } join_worker{dworker, dmu, taken_cv, dstop};
  PipeItem *P = nullptr;
  for (uint32_t b = 0; b < num_batches; b++) {
    PipeItem *R = &items[b];
    R->slot = P ? P->slot ^ 1u : 0u;
    {
      std::unique_lock<std::mutex> lk(dmu);
      ready_cv.wait(lk, [&] { return dready > b; });
      dtaken = b + 1;
    }
    taken_cv.notify_one();
    HostDelta d = std::move(deltas[b]);
    bool applied = false;
    if ((rc = pipe_round(e, e->pipe_h_rec, e->pipe_d_rec, P, R, &d, true, &applied, true))) return rc;
    if ((rc = take_apply_time())) return rc;
    if (P) publish_item(b - 1);
    if (!applied) {
      HIPCHK(hipEventRecord(e->ev[0], e->stream));
      if ((rc = apply_delta_and_refresh(e, d))) return rc;
      HIPCHK(hipEventRecord(e->ev[2], e->stream));
      apply_timed = true;
    }
    P = R;
  }
  if (P) {
    if ((rc = pipe_round(e, e->pipe_h_rec, e->pipe_d_rec, P, nullptr, nullptr, true, nullptr))) return rc;
    if ((rc = take_apply_time())) return rc;
    publish_item(num_batches - 1);
  }
  e->stats.pool_entries_used = e->h_actl.p->pool_used;
  e->stats.pool_entries_capacity = e->pool_cap;
  return DSM_OK;
}

void dsm_server_destroy(dsm_server *sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->worker.joinable()) sv->worker.join();
  for (auto &kv : sv->reqs) delete kv.second;
  delete sv;
}

}  // extern "C"
