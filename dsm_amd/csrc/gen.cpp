// Seeded synthetic workload generator: data graphs (uniform G(n,m) or R-MAT),
// vertex labels (uniform or Zipf), update streams and queries extracted from
// the data graph by randomized walk, plus a writer for the DSM text format that
// the CDSM command line reads (io.cu:122-252 of the reference defines it).
//
// Every routine is deterministic in its seed (splitmix64 / xoshiro256**), so a
// workload is identified by its parameters alone.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (auto &w : s) {
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      w = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? (uint64_t)(((__uint128_t)next() * n) >> 64) : 0; }
};

// Open-addressing set of undirected edge keys.
struct EdgeSet {
  std::vector<uint64_t> t;
  uint64_t mask;
  explicit EdgeSet(uint64_t expect) {
    uint64_t c = 16;
    while (c < expect * 2) c <<= 1;
    t.assign(c, ~0ull);
    mask = c - 1;
  }
  static uint64_t h(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
    return k;
  }
  bool insert(uint64_t k) {
    uint64_t i = h(k) & mask;
    while (t[i] != ~0ull) {
      if (t[i] == k) return false;
      i = (i + 1) & mask;
    }
    t[i] = k;
    return true;
  }
};

inline uint64_t ekey(uint32_t a, uint32_t b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)a << 32) | b;
}

}  // namespace

static int g_progress = 0;  // > 0: long generators report progress on stderr

extern "C" {

void gen_set_progress(int on) { g_progress = on; }

// G(n,m): m distinct undirected edges, no self loops, uniform over pairs.
int64_t gen_edges_uniform(uint32_t n, uint64_t m, uint64_t seed, uint32_t *eu, uint32_t *ev) {
  if (n < 2 || m > (uint64_t)n * (n - 1) / 2) return -1;
  Rng r(seed);
  EdgeSet set(m);
  uint64_t k = 0;
  while (k < m) {
    uint32_t a = (uint32_t)r.below(n), b = (uint32_t)r.below(n);
    if (a == b || !set.insert(ekey(a, b))) continue;
    eu[k] = a; ev[k] = b; k++;
  }
  return (int64_t)k;
}

// R-MAT with quadrant probabilities (a,b,c,1-a-b-c) over 2^scale ids, ids >= n
// rejected, self loops and duplicates dropped until m distinct edges exist;
// vertex ids are then relabelled by a seeded random permutation so hubs do not
// cluster at low ids.
int64_t gen_edges_rmat(uint32_t n, uint64_t m, double a, double b, double c, uint64_t seed,
                       uint32_t *eu, uint32_t *ev) {
  if (n < 2) return -1;
  int scale = 0;
  while ((1ull << scale) < n) scale++;
  Rng r(seed);
  EdgeSet set(m);
  uint64_t k = 0, tries = 0;
  const double ab = a + b, abc = a + b + c;
  while (k < m) {
    if (++tries > m * 64 + 1000000) return -2;
    uint32_t u = 0, v = 0;
    for (int i = 0; i < scale; i++) {
      double x = r.uniform();
      uint32_t bu = 0, bv = 0;
      if (x < a) {
      } else if (x < ab) {
        bv = 1;
      } else if (x < abc) {
        bu = 1;
      } else {
        bu = 1; bv = 1;
      }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if (u >= n || v >= n || u == v || !set.insert(ekey(u, v))) continue;
    eu[k] = u; ev[k] = v; k++;
    if (g_progress && (k & ((1ull << 25) - 1)) == 0) {
      fprintf(stderr, "R-MAT: %llu of %llu edges\n", (unsigned long long)k, (unsigned long long)m);
      fflush(stderr);
    }
  }
  std::vector<uint32_t> perm(n);
  for (uint32_t i = 0; i < n; i++) perm[i] = i;
  for (uint32_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[r.below(i + 1)]);
  for (uint64_t i = 0; i < m; i++) { eu[i] = perm[eu[i]]; ev[i] = perm[ev[i]]; }
  return (int64_t)k;
}

// R-MAT as above, generated in parallel with the same result for any thread
// count. Candidate edges come in chunks of 2^20 draws, chunk c from its own
// generator seeded by (seed, c); the sequence is chunk 0, 1, 2, ... in order.
// Rejected draws (ids >= n, self loops) are dropped, and of the rest the first
// m distinct undirected edges in sequence order are kept. Duplicates are found
// by key-sharded hash sets: every thread scans a round's candidates in order
// and inserts those of its own shards, so "first in sequence order" holds
// without locks. Ids are then relabelled by a seeded permutation.
int64_t gen_edges_rmat_chunked(uint32_t n, uint64_t m, double a, double b, double c, uint64_t seed,
                               uint32_t threads, uint32_t *eu, uint32_t *ev) {
  if (n < 2) return -1;
  int scale = 0;
  while ((1ull << scale) < n) scale++;
  if (threads == 0) threads = 1;
  if (threads > 64) threads = 64;
  const uint64_t kChunk = 1ull << 20;
  const uint32_t kShards = 256;
  const double ab = a + b, abc = a + b + c;
  // shard tables: open addressing, capacity ~2.2x the expected keys per shard
  const uint64_t per = (uint64_t)((double)m / kShards * 2.2) + 1024;
  std::vector<std::vector<uint64_t>> tab(kShards, std::vector<uint64_t>());
  std::vector<uint64_t> used(kShards, 0);
  auto shard_of = [](uint64_t h) { return (uint32_t)(h >> 56); };
  auto slot_of = [per](uint64_t h) { return (uint64_t)(((__uint128_t)(h << 8) * per) >> 64); };
  std::vector<uint64_t> cand;
  std::vector<uint8_t> first;
  uint64_t k = 0, next_chunk = 0;
  int fail = 0;
  while (k < m) {
    uint64_t need = m - k;
    uint64_t nch = (uint64_t)((double)need * 1.3 / (double)kChunk) + 1;
    if (next_chunk + nch > (m * 64) / kChunk + 64) return -2;
    cand.assign(nch * kChunk, ~0ull);
    first.assign(nch * kChunk, 0);
    // 1. draws, chunk by chunk (chunks split over the threads)
    {
      std::vector<std::thread> th;
      for (uint32_t t = 0; t < threads; t++)
        th.emplace_back([&, t]() {
          for (uint64_t ci = t; ci < nch; ci += threads) {
            Rng r(seed * 0x9E3779B97F4A7C15ull + (next_chunk + ci) * 0xD1B54A32D192ED03ull + 0x5851F42D4C957F2Dull);
            uint64_t *out = cand.data() + ci * kChunk;
            for (uint64_t i = 0; i < kChunk; i++) {
              uint32_t u = 0, v = 0;
              for (int l = 0; l < scale; l++) {
                double x = r.uniform();
                uint32_t bu = 0, bv = 0;
                if (x < a) {
                } else if (x < ab) {
                  bv = 1;
                } else if (x < abc) {
                  bu = 1;
                } else {
                  bu = 1; bv = 1;
                }
                u = (u << 1) | bu;
                v = (v << 1) | bv;
              }
              out[i] = (u >= n || v >= n || u == v) ? ~0ull : ekey(u, v);
            }
          }
        });
      for (auto &x : th) x.join();
    }
    next_chunk += nch;
    // 2. first occurrences, thread t owning shards t, t+threads, ...
    {
      std::vector<std::thread> th;
      std::vector<int> bad(threads, 0);
      for (uint32_t t = 0; t < threads; t++)
        th.emplace_back([&, t]() {
          for (uint32_t s = t; s < kShards; s += threads)
            if (tab[s].empty()) tab[s].assign(per, ~0ull);
          for (uint64_t i = 0; i < cand.size(); i++) {
            uint64_t key = cand[i];
            if (key == ~0ull) continue;
            uint64_t h = EdgeSet::h(key);
            uint32_t s = shard_of(h);
            if (s % threads != t) continue;
            std::vector<uint64_t> &T = tab[s];
            uint64_t j = slot_of(h);
            while (T[j] != ~0ull && T[j] != key) j = (j + 1 == per) ? 0 : j + 1;
            if (T[j] == key) continue;
            if (++used[s] * 10 > per * 9) { bad[t] = 1; return; }
            T[j] = key;
            first[i] = 1;
          }
        });
      for (auto &x : th) x.join();
      for (int x : bad) fail |= x;
    }
    if (fail) return -3;
    // 3. the first occurrences in sequence order, up to m edges
    for (uint64_t i = 0; i < cand.size() && k < m; i++)
      if (first[i]) {
        eu[k] = (uint32_t)(cand[i] >> 32);
        ev[k] = (uint32_t)cand[i];
        k++;
      }
    if (g_progress) {
      fprintf(stderr, "R-MAT (chunked): %llu of %llu edges after %llu chunks\n", (unsigned long long)k,
              (unsigned long long)m, (unsigned long long)next_chunk);
      fflush(stderr);
    }
  }
  Rng r(seed ^ 0xA0761D6478BD642Full);
  std::vector<uint32_t> perm(n);
  for (uint32_t i = 0; i < n; i++) perm[i] = i;
  for (uint32_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[r.below(i + 1)]);
  {
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < threads; t++)
      th.emplace_back([&, t]() {
        uint64_t lo = m * t / threads, hi = m * (t + 1) / threads;
        for (uint64_t i = lo; i < hi; i++) { eu[i] = perm[eu[i]]; ev[i] = perm[ev[i]]; }
      });
    for (auto &x : th) x.join();
  }
  return (int64_t)k;
}

// Labels in [0,nl): uniform when s == 0, else Zipf P(l) ~ 1/(l+1)^s.
void gen_labels(uint32_t n, uint32_t nl, double s, uint64_t seed, uint32_t *lab) {
  Rng r(seed);
  std::vector<double> cdf(nl);
  double acc = 0;
  for (uint32_t l = 0; l < nl; l++) { acc += (s == 0.0) ? 1.0 : 1.0 / std::pow((double)(l + 1), s); cdf[l] = acc; }
  for (auto &x : cdf) x /= acc;
  for (uint32_t v = 0; v < n; v++) {
    double x = r.uniform();
    lab[v] = (uint32_t)(std::lower_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
    if (lab[v] >= nl) lab[v] = nl - 1;
  }
}

// Seeded Fisher-Yates permutation of [0,n).
void gen_permutation(uint64_t n, uint64_t seed, uint64_t *out) {
  Rng r(seed);
  for (uint64_t i = 0; i < n; i++) out[i] = i;
  for (uint64_t i = n; i > 1; i--) std::swap(out[i - 1], out[r.below(i)]);
}

// Sorted CSR of an undirected edge list (both directions stored).
void gen_csr(uint32_t n, uint64_t m, const uint32_t *eu, const uint32_t *ev, uint64_t *off, uint32_t *nbr) {
  std::vector<uint64_t> cnt(n + 1, 0);
  for (uint64_t i = 0; i < m; i++) { cnt[eu[i]]++; cnt[ev[i]]++; }
  off[0] = 0;
  for (uint32_t v = 0; v < n; v++) off[v + 1] = off[v] + cnt[v];
  std::vector<uint64_t> pos(off, off + n);
  for (uint64_t i = 0; i < m; i++) { nbr[pos[eu[i]]++] = ev[i]; nbr[pos[ev[i]]++] = eu[i]; }
  for (uint32_t v = 0; v < n; v++) std::sort(nbr + off[v], nbr + off[v + 1]);
}

// Finds one injective, non-induced embedding of the pattern (k vertices, pm
// edges pu[i]-pv[i]) in the data graph by randomized depth-first walk: a random
// start vertex, then for each pattern vertex (in BFS order from pattern vertex
// 0) a random neighbour of its mapped parent that closes every pattern edge to
// already-mapped vertices. Start vertices with degree above max_start_deg are
// skipped, and the walk only visits vertices with label >= min_label (so that
// queries can be drawn over rarer labels); each start gets `budget` extension
// steps. Writes the data vertex of
// each pattern vertex to out. Returns 0 on success, -1 when no start succeeds.
int gen_find_pattern(uint32_t n, const uint64_t *off, const uint32_t *nbr, const uint32_t *lab,
                     uint32_t min_label, uint32_t k, uint32_t pm, const uint32_t *pu, const uint32_t *pv,
                     uint64_t seed, uint32_t max_start_deg, uint32_t max_walk_deg, uint64_t budget,
                     uint32_t attempts, uint32_t *out) {
  if (k == 0 || k > 32) return -1;
  std::vector<uint32_t> adj(k, 0);
  for (uint32_t i = 0; i < pm; i++) { adj[pu[i]] |= 1u << pv[i]; adj[pv[i]] |= 1u << pu[i]; }
  // BFS order of the pattern and each vertex's parent
  std::vector<uint32_t> order, parent(k, 0), pos(k, ~0u);
  order.push_back(0);
  pos[0] = 0;
  for (size_t h = 0; h < order.size(); h++)
    for (uint32_t w = 0; w < k; w++)
      if ((adj[order[h]] >> w & 1) && pos[w] == ~0u) { pos[w] = (uint32_t)order.size(); parent[w] = order[h]; order.push_back(w); }
  if (order.size() != k) return -1;  // disconnected pattern
  Rng r(seed);
  std::vector<uint32_t> f(k), iter(k), cnt(k);
  std::vector<std::vector<uint32_t>> cand(k);
  for (uint32_t at = 0; at < attempts; at++) {
    uint32_t v0 = (uint32_t)r.below(n);
    uint64_t d0 = off[v0 + 1] - off[v0];
    if (d0 == 0 || d0 > max_start_deg || lab[v0] < min_label) continue;
    f[order[0]] = v0;
    uint64_t steps = 0;
    int lvl = 1;
    bool fresh = true;
    while (lvl > 0 && lvl < (int)k && steps < budget) {
      uint32_t u = order[lvl];
      if (fresh) {
        uint32_t p = f[parent[u]];
        auto &c = cand[lvl];
        c.assign(nbr + off[p], nbr + off[p + 1]);
        for (size_t i = c.size(); i > 1; i--) std::swap(c[i - 1], c[r.below(i)]);
        iter[lvl] = 0;
        fresh = false;
      }
      bool advanced = false;
      auto &c = cand[lvl];
      while (iter[lvl] < c.size()) {
        uint32_t w = c[iter[lvl]++];
        steps++;
        if (off[w + 1] - off[w] > max_walk_deg || lab[w] < min_label) continue;
        bool ok = true;
        for (int j = 0; j < lvl && ok; j++)
          if (f[order[j]] == w) ok = false;
        for (int j = 0; j < lvl && ok; j++) {
          uint32_t x = order[j];
          if (!(adj[u] >> x & 1)) continue;
          uint32_t y = f[x];
          if (!std::binary_search(nbr + off[w], nbr + off[w + 1], y)) ok = false;
        }
        if (!ok) continue;
        f[u] = w;
        advanced = true;
        break;
      }
      if (advanced) { lvl++; fresh = true; }
      else { lvl--; fresh = false; }
    }
    if (lvl == (int)k) {
      for (uint32_t i = 0; i < k; i++) out[i] = f[i];
      return 0;
    }
  }
  return -1;
}

// Pilot of a query's enumeration cost over given edge records: for every record
// (s,d) and ordered query edge (u,v) with matching labels, counts the nodes of
// the backtracking tree that extends u->s, v->d (every valid partial embedding
// above the last level, plus one per 64 source entries scanned at the last
// level), with the usual candidate rule (label, adjacency to every mapped query
// neighbour, injectivity) on the given graph. Stops as soon as the total
// exceeds `budget` and then returns budget + 1.
namespace {
struct Pilot {
  const uint64_t *off;
  const uint32_t *nbr;
  const uint32_t *lab;
  const uint32_t *qlab;
  uint32_t k;
  uint32_t adj[32];
  uint32_t order[32];
  uint32_t f[32];
  uint64_t nodes, budget;
  bool adjacent(uint32_t a, uint32_t b) const {
    return std::binary_search(nbr + off[a], nbr + off[a + 1], b);
  }
  void rec(uint32_t p) {
    if (p == k || nodes > budget) return;
    const uint32_t u = order[p];
    uint32_t src = 0;
    uint64_t sdeg = ~0ull;
    for (uint32_t j = 0; j < p; j++)
      if (adj[u] >> order[j] & 1) {
        const uint64_t dg = off[f[order[j]] + 1] - off[f[order[j]]];
        if (dg < sdeg) { sdeg = dg; src = order[j]; }
      }
    const uint32_t x = f[src];
    if (p + 1 == k) {
      // the last level is counted in 64-entry scan chunks, as a wave64
      // enumerator counts leaves with one ballot per chunk
      nodes += (off[x + 1] - off[x] + 63) / 64;
      return;
    }
    for (uint64_t i = off[x]; i < off[x + 1] && nodes <= budget; i++) {
      const uint32_t w = nbr[i];
      if (lab[w] != qlab[u]) continue;
      bool ok = true;
      for (uint32_t j = 0; j < p && ok; j++) {
        const uint32_t qj = order[j];
        if (f[qj] == w) ok = false;
        else if (qj != src && (adj[u] >> qj & 1) && !adjacent(w, f[qj])) ok = false;
      }
      if (!ok) continue;
      nodes++;
      f[u] = w;
      rec(p + 1);
    }
  }
};
}  // namespace

uint64_t gen_pilot_nodes(uint32_t n, const uint64_t *off, const uint32_t *nbr, const uint32_t *lab, uint32_t k,
                         uint32_t pm, const uint32_t *pu, const uint32_t *pv, const uint32_t *qlab, uint64_t nrec,
                         const uint32_t *rs, const uint32_t *rd, uint64_t budget) {
  if (k < 2 || k > 32) return budget + 1;
  Pilot P;
  P.off = off; P.nbr = nbr; P.lab = lab; P.qlab = qlab; P.k = k; P.nodes = 0; P.budget = budget;
  for (uint32_t i = 0; i < 32; i++) P.adj[i] = 0;
  for (uint32_t i = 0; i < pm; i++) { P.adj[pu[i]] |= 1u << pv[i]; P.adj[pv[i]] |= 1u << pu[i]; }
  for (uint64_t r = 0; r < nrec && P.nodes <= budget; r++) {
    const uint32_t s = rs[r], d = rd[r];
    if (s >= n || d >= n || s == d) continue;
    for (uint32_t u = 0; u < k && P.nodes <= budget; u++) {
      if (qlab[u] != lab[s]) continue;
      for (uint32_t v = 0; v < k; v++) {
        if (!(P.adj[u] >> v & 1) || qlab[v] != lab[d]) continue;
        uint32_t placed = (1u << u) | (1u << v);
        P.order[0] = u; P.order[1] = v;
        for (uint32_t p = 2; p < k; p++) {
          int best = -1, bb = -1;
          for (uint32_t w = 0; w < k; w++) {
            if (placed >> w & 1) continue;
            int b = __builtin_popcount(P.adj[w] & placed);
            if (b > bb) { bb = b; best = (int)w; }
          }
          P.order[p] = (uint32_t)best;
          placed |= 1u << best;
        }
        P.f[u] = s; P.f[v] = d;
        P.nodes++;
        P.rec(2);
      }
    }
  }
  return P.nodes > budget ? budget + 1 : P.nodes;
}

// Writes a data graph + update batches (or, with nb == UINT32_MAX, a query
// graph without the '-' terminator) in the DSM text format. Degrees on the
// v-lines are the incident e-line counts. Returns 0 or -1 on I/O failure.
int gen_write_dsm(const char *path, uint32_t n, const uint32_t *lab, uint64_t m, const uint32_t *eu,
                  const uint32_t *ev, uint32_t nb, const uint32_t *bstart, const uint8_t *add,
                  const uint32_t *src, const uint32_t *dst, const uint32_t *clab) {
  FILE *f = std::fopen(path, "w");
  if (!f) return -1;
  std::vector<uint32_t> deg(n, 0);
  for (uint64_t i = 0; i < m; i++) { deg[eu[i]]++; deg[ev[i]]++; }
  std::vector<char> buf(1 << 20);
  std::setvbuf(f, buf.data(), _IOFBF, buf.size());
  std::fprintf(f, "t %u %llu\n", n, (unsigned long long)m);
  for (uint32_t v = 0; v < n; v++) std::fprintf(f, "v %u %u %u\n", v, lab[v], deg[v]);
  for (uint64_t i = 0; i < m; i++) std::fprintf(f, "e %u %u\n", eu[i], ev[i]);
  if (nb != UINT32_MAX) {
    std::fprintf(f, "-\n");
    for (uint32_t b = 0; b < nb; b++) {
      std::fprintf(f, "b %u %u\n", b, bstart[b + 1] - bstart[b]);
      for (uint32_t i = bstart[b]; i < bstart[b + 1]; i++)
        std::fprintf(f, "%c %u %u %u\n", add[i] ? 'a' : 'd', src[i], dst[i], clab[i]);
    }
  }
  int rc = std::ferror(f) ? -1 : 0;
  if (std::fclose(f) != 0) rc = -1;
  return rc;
}

}  // extern "C"
