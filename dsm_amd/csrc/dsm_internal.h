// Shared host/device definitions of the MI355X matching engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace dsm {

constexpr int kMaxQ = 16;       // query vertices per query (the reference's MAX_VQ, defs.h:18)
constexpr int kWave = 64;       // CDNA wavefront
constexpr int kMaxRel = 32;     // labels that can appear in an NLF requirement
constexpr int kMaxVertexChanges = 1024;  // per vertex per apply round (host splits beyond)
constexpr uint8_t kNoLabel = 255;  // device label of vertices whose label no query uses
#ifndef DSM_DONATE_BURST
#define DSM_DONATE_BURST 4  // ranges a busy wave may hand out per check
#endif
#ifndef DSM_LABEL_SORTED
#define DSM_LABEL_SORTED 64
#endif
constexpr uint32_t kLabelSorted = DSM_LABEL_SORTED;  // vertices of higher degree keep a label-sorted list copy
static_assert(kLabelSorted <= 64, "the heavy refilter reads the label-sorted copies of lists longer than 64");

// Matching plan of one (query, ordered query edge (u0,u1)) pair: positions 0
// and 1 hold the seeded edge, later positions follow a greedy order in which
// every vertex has an earlier neighbour.
struct alignas(16) Plan {
  uint8_t n;             // query vertices
  uint8_t pad0;
  uint16_t query;        // query index
  uint8_t qv[kMaxQ];     // query vertex at each position
  uint16_t sig[kMaxQ];   // candidate bitmap of each position
  uint16_t back[kMaxQ];  // earlier positions adjacent in the query
  uint16_t inj[kMaxQ];   // earlier positions with the same label (injectivity)
  uint8_t lix[kMaxQ];    // dense label index of each position (label-sorted lists)
  uint32_t pad1[3];
};
static_assert(sizeof(Plan) == 144, "plan layout");

// LDF/NLF requirement of one candidate bitmap: label, minimum degree and, for
// each of up to 32 relevant labels (dense index), the minimum number of
// neighbours with that label, one byte each (<= 15).
struct alignas(8) SigReq {
  uint32_t label;
  uint32_t deg;
  uint64_t nlc[4];
};

// A unit of DFS work: levels [0, level) are fixed in f; level `level` scans
// source-list indices [lo, hi). One 128-byte queue slot each; `ready` holds
// the launch epoch once the slot is published.
struct alignas(128) Task {
  uint32_t plan;
  uint32_t level;
  uint32_t lo, hi;
  uint32_t f[kMaxQ];
  uint32_t tag;          // count slot << 1 | sign (1 positive, 0 negative)
  uint32_t ready;
  uint32_t pad[10];
};
static_assert(sizeof(Task) == 128, "task slot layout");
constexpr int kTaskWords = 5 + kMaxQ;  // payload words (21)
constexpr uint32_t kMaxSlots = 64;     // count slots: requests served in one step

// Device-resident control block of one seed + DFS round. Every counter that
// waves update or poll concurrently has a 128-byte line of its own, so polling
// by idle waves does not serialise the atomics of busy ones.
template <typename T>
struct alignas(128) CtlLine {
  T v;
  uint32_t pad[(128 - sizeof(T)) / 4];
};
struct DfsCtl {
  CtlLine<uint32_t> nseeds;        // seeds written by the seed kernel
  CtlLine<uint32_t> seed_head;     // next seed to claim
  CtlLine<uint32_t> q_head;        // queue tickets taken by waves out of work
  CtlLine<uint32_t> q_tail;        // queue slots reserved by donors (= tasks donated)
  CtlLine<int32_t> work_left;      // seeds + donated tasks not yet finished
  CtlLine<uint32_t> seed_overflow; // seeds dropped for lack of capacity (must stay 0)
  CtlLine<uint32_t> abort;         // set when an idle wave outlived the watchdog
  CtlLine<unsigned long long> scanned;  // source-list entries tested
  CtlLine<unsigned long long> loads;    // 4-byte reads of membership searches
  CtlLine<unsigned long long> mbytes;   // 4 B x degree of every backward list of every expanded partial match
};

// Status of the apply kernels of one batch.
struct alignas(64) ApplyCtl {
  unsigned long long scratch_used;
  unsigned long long pool_used;     // bump pointer of the neighbour pool
  uint32_t overflow;                // commits that found no room (host grows the pool)
  uint32_t relocated;
  uint32_t ndirty;                  // vertices on the dirty list
  uint32_t nheavy;                  // dirty vertices re-filtered by a whole wave
  uint32_t relabeled;
  uint32_t nrelabel_big;            // relabelled vertices whose neighbours the whole grid marks
  uint32_t nmerge_big;              // touched vertices whose merge the whole grid copies
  uint32_t ncommit_big;             // touched vertices whose commit the whole grid copies
  uint32_t merge_overflow;          // merges that found no room in the scratch (must stay 0)
  uint32_t filter_bits;             // candidate bits the re-filter writes (work measure)
  unsigned long long filter_entries;  // list entries of the vertices re-filtered (work measure)
  // of the big lists above, those longer than kBlockList (split over the whole
  // grid; queued from the back of the same arrays, the rest from the front)
  uint32_t nrelabel_huge;
  uint32_t nmerge_huge;
  uint32_t ncommit_huge;
  uint32_t pad[13];
};
static_assert(sizeof(ApplyCtl) == 128, "apply status layout");

struct DevGraph {
  uint32_t n;
  uint32_t *pool;     // neighbour lists, sorted ascending, with slack
  uint32_t *start;    // first pool entry of each list
  uint32_t *deg;      // list length
  uint32_t *cap;      // slot capacity
  uint8_t *lab;       // vertex label (kNoLabel when no query uses it)
  // Vertices of degree > kLabelSorted also keep their list sorted by (dense
  // label, id) at lpool[start .. start + deg), with the bucket offsets of
  // label index i at loff[v * (nbuckets + 1) + i] (bucket nbuckets - 1 holds
  // labels no query uses).
  uint32_t *lpool;
  uint32_t *loff;
  uint32_t nbuckets;
};

struct FilterTables {
  const SigReq *sigs;       // [nsig]
  const uint32_t *sig_lo;   // [256] first sig of each label (sigs sorted by label)
  const uint32_t *sig_hi;   // [256] one past the last
  const uint8_t *lmap;      // [256] label -> dense relevant index or 0xFF
  uint32_t nsig;
  uint64_t *bm;             // [nsig][words]
  uint32_t words;           // ceil(n/64)
};

// Batch description uploaded per apply round.
struct BatchDev {
  uint32_t ntouched;
  const uint32_t *tv;       // touched vertices
  const uint32_t *ioff;     // [ntouched+1] into ins
  const uint32_t *doff;     // [ntouched+1] into del
  const uint32_t *ins;
  const uint32_t *del;
  uint32_t *ndeg;           // [ntouched] new degree (merge -> commit)
  unsigned long long *soff; // [ntouched] scratch offset
  uint32_t *fail;           // [ntouched] commit found no room
  // lists longer than kApplyWave: their merge and commit copies are split over
  // the whole grid. icum/dcum of touched vertex t live at ioff[t] + t and
  // doff[t] + t; big_merge / big_commit list the vertices (indices t).
  uint32_t *gicum;          // [ins + ntouched]
  uint32_t *gdcum;          // [del + ntouched]
  uint32_t *big_merge;      // [ntouched]
  uint32_t *big_commit;     // [ntouched]
  uint32_t nrelabel;
  const uint32_t *rl_v;
  const uint32_t *rl_l;     // device label (already clamped)
  uint32_t *rl_big;         // [nrelabel] relabelled vertices of degree > kRelabelWave
};

struct DfsArgs {
  DevGraph g;
  const uint64_t *bm;
  uint32_t words;
  const Plan *plans;
  const uint32_t *seed_plan;
  const uint32_t *seed_s;
  const uint32_t *seed_d;
  Task *queue;
  uint32_t qcap;
  uint32_t epoch;
  uint32_t donate_min;
  uint32_t lane_work;            // longest leaf list one lane scans alone
  uint32_t leaf_budget;          // inline leaf-count work of a window before a candidate descends
  uint32_t donate_every;         // chunks between checks for waiting waves
  uint32_t donate_min_n2;        // smallest range split off one level above the leaf
  uint32_t leaf_segments;        // leaf lists scanned as one segmented array
  uint32_t poll_max;             // longest sleep between polls of a waiting wave (s_sleep 8 units)
  uint32_t leaf_sets;            // fused leaf level on/off
  uint32_t bit_probe_last;       // candidate bit probed after the membership tests
  uint32_t fshift;               // ids >> fshift < 512: the leaf set's bucket of an id
  uint32_t nq;                   // queries (count layout)
  const uint32_t *seed_tag;      // per seed: count slot << 1 | sign
  unsigned long long watchdog;   // wall-clock ticks an idle wave may wait
  DfsCtl *ctl;
  unsigned long long *counts;    // [slot][nq][2]: index (slot*nq + q)*2 + sign
  uint32_t *tuples;              // [tuple_cap][kMaxQ+2] or null
  unsigned long long tuple_cap;
  unsigned long long *tuple_count;  // tuples found this batch
  unsigned long long *wave_busy;    // per wave of the grid: wall-clock ticks on tasks (or null)
  uint32_t wave_slots;              // entries of wave_busy
  unsigned long long *query_busy;   // per query: wall-clock ticks of its tasks (or null)
};

struct SeedArgs {
  DevGraph g;
  const uint64_t *bm;
  uint32_t words;
  const Plan *plans;
  const uint32_t *plan_key;   // sorted (label0 << 8 | label1) of each plan
  uint32_t nplans;
  const uint32_t *rs;
  const uint32_t *rd;
  uint32_t nrec;
  uint32_t *seed_plan;
  uint32_t *seed_s;
  uint32_t *seed_d;
  uint32_t *seed_tag;
  uint32_t seed_cap;
  uint32_t tag;               // written to every seed of these records
  DfsCtl *ctl;
};

struct DirtyArgs {
  uint32_t *bits;      // [ceil(n/32)] dirty flags
  uint32_t *list;      // dirty vertices
  uint32_t *heavy;     // dirty vertices with a long list
  uint32_t list_cap;
  ApplyCtl *actl;
};

// Message of the last failed call that has no engine (dsm_last_error(NULL)).
void set_global_error(const char *fmt, ...);

// Sorted-set intersection (dsm_intersect.hip): k_intersect writes chunk c's
// members of B at hits[c * 64 ..] and their number at cnt[c].
uint32_t intersect_chunks(uint32_t na);
hipError_t launch_intersect(const uint32_t *A, uint32_t na, const uint32_t *B, uint32_t nb, uint32_t *hits,
                            uint32_t *cnt, hipStream_t s);

// ---- cluster-encoding refinement (dsm_cluster_plan.cpp, dsm_cluster.hip) ---

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMaxClusters = 150;       // MAX_CLUSTERS, defs.h:22
constexpr uint32_t kMaxClusterLayers = 100;  // MAX_LAYERS, defs.h:23
constexpr uint32_t kMaxSmallClusters = 32;   // per combine (filter.cu:1056-1061)

// Host plan of one query's refinement: clusters with their bit positions, the
// combines, the passes in execution order and each query vertex's final
// position (its refined candidate set).
struct ClusterPlan {
  struct Cluster {
    std::vector<uint32_t> members;  // core first, then its neighbours
    bool valid = true;
    bool merged = false;            // made by a merge of clusters left and right
    uint32_t left = kNone, right = kNone;
    uint32_t left_pos = kNone, right_pos = kNone;  // the core's positions in left / right
    uint32_t pos = 0;               // first bit position
  };
  struct Combine {
    uint32_t big = kNone;           // the cluster that receives the combination
    bool keep = false;              // big existed (AND into it) or was created (OR)
    std::vector<uint32_t> small;
  };
  enum { kEncode = 0, kMerge = 1, kCombine = 2 };
  struct Pass {
    uint32_t kind, arg;             // encode / merge: cluster; combine: index into combines
  };
  std::vector<Cluster> clusters;
  std::vector<Combine> combines;
  std::vector<Pass> passes;
  std::vector<uint32_t> final_pos;  // per query vertex
  uint32_t final_cluster = 0;
  uint32_t positions = 0;
};

// Returns 0, or -1 when the plan exceeds one of the reference's limits.
int build_cluster_plan(uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev, ClusterPlan &out);

// One pass of one query on the device. Bitmap rows index `bm` (signature rows
// first, then every query's position rows); `alive` rows are the candidate
// lists the passes prune (d_u_candidate_vs_ of the reference).
enum : uint32_t { kOpEncode = 0, kOpMerge = 1, kOpCombineNew = 2, kOpCombineKeep = 3 };
struct alignas(16) ClusterOp {
  uint32_t kind;
  uint32_t alive;       // alive row of the cluster's core
  uint32_t core;        // bm row of the core position; targets at core+1 .. core+ntarget
  uint32_t ntarget;
  uint32_t check_off;   // a core candidate survives when all rows crow[check_off ..+ncheck) hold it
  uint32_t ncheck;
  uint32_t pad[2];
  uint16_t toff[kMaxQ]; // target t: the AND of rows crow[toff[t] .. +tcnt[t]) at the neighbour
  uint16_t tcnt[kMaxQ];
  uint8_t tlab[kMaxQ];  // device label of target t's query vertex (every condition implies it)
  uint8_t tlix[kMaxQ];  // its dense label index (the bucket of the label-sorted lists)
  uint8_t lix[kMaxQ];   // the distinct tlix values, nlix of them
  uint32_t nlix;
  uint32_t pad2[3];
};

struct ClusterHeavy {
  // entries [lo, lo + kClusterChunk) of candidate v's label bucket li (in the
  // label-sorted copy) for pass op
  uint32_t op, v, lo, li;
};
constexpr uint32_t kClusterChunk = 4096;

struct ClusterArgs {
  DevGraph g;
  uint64_t *bm;
  uint32_t words;
  uint64_t *alive;
  const ClusterOp *ops;
  uint32_t nops;
  const uint32_t *crow;
  ClusterHeavy *heavy;
  uint32_t heavy_cap;
  uint32_t *ctl;        // [2]: heavy entries queued, heavy entries claimed
};

// Zeroes bm rows [first_row, first_row + nrows), copies the listed (alive row,
// bm row) pairs into the alive rows and zeroes nctl counters.
hipError_t launch_cluster_init(uint64_t *bm, uint32_t words, uint32_t first_row, uint32_t nrows, uint64_t *alive,
                               const uint32_t *alive_src, uint32_t nalive_src, uint32_t *ctl, uint32_t nctl,
                               hipStream_t s);
hipError_t launch_cluster_pass(const ClusterArgs &a, uint32_t max_blocks, hipStream_t s);
hipError_t launch_cluster_heavy(const ClusterArgs &a, uint32_t blocks, hipStream_t s);

// ---- launchers (dsm_kernels.hip) -------------------------------------------
hipError_t launch_apply_merge(const DevGraph &g, const BatchDev &b, unsigned long long scratch_cap,
                              uint32_t *scratch, ApplyCtl *actl, hipStream_t s);
hipError_t launch_apply_commit(const DevGraph &g, const BatchDev &b, const uint32_t *scratch,
                               unsigned long long pool_cap, ApplyCtl *actl, const DirtyArgs &d,
                               hipStream_t s);
hipError_t launch_apply_commit_failed(const DevGraph &g, const BatchDev &b, const uint32_t *scratch,
                                      unsigned long long pool_cap, ApplyCtl *actl, const DirtyArgs &d,
                                      hipStream_t s);
// The dirty list (d.list, its length added to d.actl->ndirty) from the dirty
// bits of the update's n vertices.
hipError_t launch_dirty_compact(const DirtyArgs &d, uint32_t n, hipStream_t s);
// Copies what the last update changed from graph copy x to the stale copy y:
// the touched vertices' slots and lists, the dirty vertices' label-sorted
// copies (ndirty on the device; none when dirty is null), the relabelled
// vertices' labels. q (qcap entries) and qctl[2] are scratch for the long
// lists.
hipError_t launch_sync(const DevGraph &x, const DevGraph &y, const uint32_t *touched, uint32_t ntouched,
                       const uint32_t *dirty, const uint32_t *ndirty, const uint32_t *relabeled, uint32_t nrelabeled,
                       uint32_t *q, uint32_t qcap, uint32_t *qctl, hipStream_t s);
// The label-sorted copies and bucket offsets of the listed (dirty) vertices
// from x to y, except those whose bit is set in skip.
hipError_t launch_sync_sorted(const DevGraph &x, const DevGraph &y, const uint32_t *dirty, const uint32_t *ndirty,
                              const uint32_t *skip, uint32_t *q, uint32_t qcap, uint32_t *qctl, hipStream_t s);
hipError_t launch_relabel(const DevGraph &g, const BatchDev &b, const FilterTables &ft,
                          const DirtyArgs &d, hipStream_t s);
hipError_t launch_filter_full(const DevGraph &g, const FilterTables &ft, hipStream_t s);
// Label-sorted copies of lists longer than kHugeList entries are built in
// segments of kHugeSeg entries spread over the grid: the vertices are queued
// here with their first segment slot, and per-segment bucket counts go to hist.
constexpr uint32_t kHugeList = 65536;
constexpr uint32_t kHugeSeg = 65536;
struct HugeSort {
  uint32_t *list;      // [cap][2]: vertex, first segment slot
  uint32_t cap;
  uint32_t *ctl;       // [4]: queued vertices, segment slots taken, vertices on big, on mid
  uint32_t *big;       // [big_cap]: vertices of degree > 4,096 for the block-per-list build
  uint32_t big_cap;
  uint32_t *mid;       // [mid_cap]: vertices of degree 65..4,096 for the wave-per-list build
  uint32_t mid_cap;
  uint32_t *hist;      // [segs][kMaxRel + 2]
  uint32_t segs;
  // one byte per pool entry (indexed like the pool): the bucket of each entry,
  // written by the counting pass so that the scatter pass reads it back
  // instead of gathering the neighbour's label again (null: gather twice)
  uint8_t *bucket_bytes;
};

// Rebuilds the label-sorted copies of the listed vertices (list + count on the
// device), or of every vertex when list is null.
hipError_t launch_lbuild(const DevGraph &g, const uint8_t *lmap, const uint32_t *list, const uint32_t *count,
                         const HugeSort &hs, hipStream_t s);
hipError_t launch_filter_dirty(const DevGraph &g, const FilterTables &ft, const DirtyArgs &d,
                               hipStream_t s);
hipError_t launch_seed(const SeedArgs &a, hipStream_t s);
hipError_t launch_dfs(const DfsArgs &a, uint32_t blocks, uint32_t waves_per_block, uint32_t max_query_vertices,
                      hipStream_t s);
// Resets the DFS control block (ctl) and/or the per-batch state (actl's
// counters, the tuple counter and the nq*2 match counts); null pointers are skipped.
hipError_t launch_ctl_reset(DfsCtl *ctl, ApplyCtl *actl, unsigned long long *tuple_count,
                            unsigned long long *counts, uint32_t ncounts, hipStream_t s);
hipError_t launch_scratch_reset(ApplyCtl *actl, hipStream_t s);
hipError_t launch_pool_move(const uint32_t *old_pool, uint32_t *new_pool, const uint32_t *old_start,
                            const uint32_t *new_start, const uint32_t *deg, uint32_t n,
                            hipStream_t s);
int dfs_max_blocks_per_cu(uint32_t waves_per_block, uint32_t max_query_vertices);

// Kernel-node parameters of one DFS round for a HIP graph (ctl reset, two seed
// lists, the DFS), the same launches launch_ctl_reset / launch_seed /
// launch_dfs make. The node's kernelParams point into `st`, which must stay
// valid until the node is added or updated (the values are copied then). A
// seed list with no records still gets one block, which does nothing.
struct RoundNodeArgs {
  DfsCtl *ctl;
  ApplyCtl *no_actl;
  unsigned long long *no_tc;
  unsigned long long *no_counts;
  uint32_t no_ncounts;
  SeedArgs seed[2];
  DfsArgs dfs;
  uint32_t chunk_levels;
  void *reset_params[5];
  void *seed_params[2][1];
  void *dfs_params[2];
};
void round_node_params(RoundNodeArgs &st, uint32_t blocks, uint32_t waves_per_block,
                       uint32_t max_query_vertices, hipKernelNodeParams *reset, hipKernelNodeParams *seed0,
                       hipKernelNodeParams *seed1, hipKernelNodeParams *dfs);
uint32_t dfs_lds_bytes_per_wave(uint32_t max_query_vertices);

}  // namespace dsm
