/*
 * CPU oracle for the continuous subgraph matching path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline), never as the product path.
 *
 * It restates, in plain C, the semantics of the reference's matching path:
 *   - DSM text parsing ............ io.cu:122-252 (readGraphToCPUDynamic)
 *   - batch update rule ........... cpuGraph.cpp:521-607 (updateHostGraph)
 *   - LDF/NLF candidate filter .... filter.cu:394-566 (oneRoundFilterBidirectionKernel)
 *   - edge seeding ................ main.cu:268-329
 *   - DFS candidate rule .......... gpu_match.cu:96-325 (get_new_v), :417-792 (match)
 * plus an independent brute-force enumerator used to pin the DFS.
 *
 * Parity pins (tests/test_oracle.py):
 *   parse ........ fixture closed form (11 V / 13 E / 1 batch; query 5 V / 5 E)
 *   update ....... hand-derived known answers (relabel quirk, ordering, idempotence)
 *                  + invariants (symmetric, sorted, deduplicated, deg = list length)
 *   LDF/NLF ...... the example's hand-derived candidate sets + soundness against
 *                  brute-force embedding projections
 *   counts ....... brute force (no filter, no order, no intersection) on random
 *                  small graphs; the example's known answer (4, 8 tuples)
 *   tuples ....... brute force multiset equality
 * No function here is "parity unpinned".
 *
 * Decisions where the reference leaves behaviour undefined (stated once here):
 *   - ties in the (src,dst) sort of updateHostGraph use a stable order (the
 *     reference's std::sort leaves them unspecified): original records first,
 *     in file order, then the reversed copies, in file order;
 *   - a seed whose data edge (s,d) is absent from the graph it is evaluated in
 *     contributes 0 (the reference never checks it, which makes its count depend
 *     on the candidate filter);
 *   - negative counts: delete records seed in the pre-batch graph, symmetric to
 *     the insert rule (the reference defines none).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define OQ_MAXV 32

/* ------------------------------------------------------------------ graph */

typedef struct {
    uint32_t n;
    uint32_t *lab;
    uint32_t *deg;
    uint32_t *cap;
    uint32_t **adj; /* sorted, deduplicated */
} og_t;

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

static void adj_reserve(og_t *g, uint32_t v, uint32_t need) {
    if (g->cap[v] >= need) return;
    uint32_t c = g->cap[v] ? g->cap[v] : 4;
    while (c < need) c *= 2;
    g->adj[v] = (uint32_t *)realloc(g->adj[v], (size_t)c * 4);
    g->cap[v] = c;
}

/* lower bound in sorted array */
static uint32_t lb(const uint32_t *a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static int has_edge(const og_t *g, uint32_t a, uint32_t b) {
    uint32_t p = lb(g->adj[a], g->deg[a], b);
    return p < g->deg[a] && g->adj[a][p] == b;
}

static void set_insert(og_t *g, uint32_t v, uint32_t x) {
    uint32_t p = lb(g->adj[v], g->deg[v], x);
    if (p < g->deg[v] && g->adj[v][p] == x) return;
    adj_reserve(g, v, g->deg[v] + 1);
    memmove(g->adj[v] + p + 1, g->adj[v] + p, (size_t)(g->deg[v] - p) * 4);
    g->adj[v][p] = x;
    g->deg[v]++;
}

static void set_erase(og_t *g, uint32_t v, uint32_t x) {
    uint32_t p = lb(g->adj[v], g->deg[v], x);
    if (!(p < g->deg[v] && g->adj[v][p] == x)) return;
    memmove(g->adj[v] + p, g->adj[v] + p + 1, (size_t)(g->deg[v] - p - 1) * 4);
    g->deg[v]--;
}

/* Build from an undirected edge list; each edge is stored both ways, then the
 * lists are sorted and deduplicated (the state updateHostGraph produces,
 * cpuGraph.cpp:576-598). */
og_t *og_new(uint32_t n, const uint32_t *labels, uint64_t m,
             const uint32_t *eu, const uint32_t *ev) {
    og_t *g = (og_t *)calloc(1, sizeof(og_t));
    g->n = n;
    g->lab = (uint32_t *)malloc((size_t)n * 4 + 4);
    g->deg = (uint32_t *)calloc((size_t)n + 1, 4);
    g->cap = (uint32_t *)calloc((size_t)n + 1, 4);
    g->adj = (uint32_t **)calloc((size_t)n + 1, sizeof(uint32_t *));
    memcpy(g->lab, labels, (size_t)n * 4);
    uint32_t *cnt = (uint32_t *)calloc((size_t)n + 1, 4);
    for (uint64_t i = 0; i < m; i++) { cnt[eu[i]]++; cnt[ev[i]]++; }
    for (uint32_t v = 0; v < n; v++) adj_reserve(g, v, cnt[v] ? cnt[v] : 1);
    for (uint64_t i = 0; i < m; i++) {
        g->adj[eu[i]][g->deg[eu[i]]++] = ev[i];
        g->adj[ev[i]][g->deg[ev[i]]++] = eu[i];
    }
    /* lists are independent: sorted on the host's threads */
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint32_t v = 0; v < n; v++) {
        qsort(g->adj[v], g->deg[v], 4, cmp_u32);
        uint32_t w = 0;
        for (uint32_t i = 0; i < g->deg[v]; i++)
            if (w == 0 || g->adj[v][w - 1] != g->adj[v][i]) g->adj[v][w++] = g->adj[v][i];
        g->deg[v] = w;
    }
    free(cnt);
    return g;
}

void og_free(og_t *g) {
    if (!g) return;
    for (uint32_t v = 0; v < g->n; v++) free(g->adj[v]);
    free(g->adj); free(g->deg); free(g->cap); free(g->lab); free(g);
}

uint64_t og_num_entries(const og_t *g) {
    uint64_t s = 0;
    for (uint32_t v = 0; v < g->n; v++) s += g->deg[v];
    return s;
}

void og_export(const og_t *g, uint64_t *offsets, uint32_t *nbrs, uint32_t *labels) {
    uint64_t o = 0;
    for (uint32_t v = 0; v < g->n; v++) {
        offsets[v] = o;
        memcpy(nbrs + o, g->adj[v], (size_t)g->deg[v] * 4);
        o += g->deg[v];
        labels[v] = g->lab[v];
    }
    offsets[g->n] = o;
}

/* ------------------------------------------------------------ update rule */

typedef struct { uint32_t src, dst, label, add, pos; } rec_t;

static int cmp_rec(const void *a, const void *b) {
    const rec_t *x = (const rec_t *)a, *y = (const rec_t *)b;
    if (x->src != y->src) return (x->src > y->src) - (x->src < y->src);
    if (x->dst != y->dst) return (x->dst > y->dst) - (x->dst < y->dst);
    return (x->pos > y->pos) - (x->pos < y->pos); /* stable */
}

/* updateHostGraph (cpuGraph.cpp:521-607): every record is symmetrised
 * (:527-533), the 2k records are sorted by (src,dst) (:536-542) and applied in
 * that order: an add inserts dst into src's set and, when relabel != 0, sets
 * label[src] = record label (:562-566); a delete erases (:567-571).
 * Returns 0, or -1 when an endpoint is out of range. */
int og_apply(og_t *g, uint32_t k, const uint8_t *add, const uint32_t *src,
             const uint32_t *dst, const uint32_t *label, int relabel) {
    for (uint32_t i = 0; i < k; i++)
        if (src[i] >= g->n || dst[i] >= g->n) return -1;
    rec_t *r = (rec_t *)malloc(sizeof(rec_t) * (size_t)(2 * k + 1));
    for (uint32_t i = 0; i < k; i++) {
        r[i].src = src[i]; r[i].dst = dst[i]; r[i].label = label[i]; r[i].add = add[i]; r[i].pos = i;
        r[k + i].src = dst[i]; r[k + i].dst = src[i]; r[k + i].label = label[i]; r[k + i].add = add[i]; r[k + i].pos = k + i;
    }
    qsort(r, 2 * (size_t)k, sizeof(rec_t), cmp_rec);
    for (uint32_t i = 0; i < 2 * k; i++) {
        if (r[i].add) {
            if (relabel) g->lab[r[i].src] = r[i].label;
            set_insert(g, r[i].src, r[i].dst);
        } else {
            set_erase(g, r[i].src, r[i].dst);
        }
    }
    free(r);
    return 0;
}

/* ------------------------------------------------------------------ query */

typedef struct {
    uint32_t n;
    uint32_t lab[OQ_MAXV];
    uint32_t adjm[OQ_MAXV];          /* adjacency bitmask */
    uint32_t deg[OQ_MAXV];
    uint32_t nvlq;                    /* max query label + 1 (io.cu:336) */
} oq_t;

/* The query is the ORIGINAL query graph (hq_backup): main.cu:297 tests edges
 * against it. Duplicate edge lines collapse; self loops are rejected. */
oq_t *oq_new(uint32_t n, const uint32_t *labels, uint32_t m, const uint32_t *eu, const uint32_t *ev) {
    if (n == 0 || n > OQ_MAXV) return NULL;
    oq_t *q = (oq_t *)calloc(1, sizeof(oq_t));
    q->n = n;
    for (uint32_t i = 0; i < n; i++) {
        q->lab[i] = labels[i];
        if (labels[i] + 1 > q->nvlq) q->nvlq = labels[i] + 1;
    }
    for (uint32_t i = 0; i < m; i++) {
        if (eu[i] >= n || ev[i] >= n || eu[i] == ev[i]) { free(q); return NULL; }
        q->adjm[eu[i]] |= 1u << ev[i];
        q->adjm[ev[i]] |= 1u << eu[i];
    }
    for (uint32_t i = 0; i < n; i++) q->deg[i] = (uint32_t)__builtin_popcount(q->adjm[i]);
    return q;
}

void oq_free(oq_t *q) { free(q); }

/* ------------------------------------------------------------- LDF / NLF */

/* v in C(u) iff L(v)=L(u), deg(v) >= deg_Q(u), and for every label l < NUM_VLQ
 * the count of v's neighbours with label l is >= u's count (filter.cu:475-547;
 * vertices whose label >= NUM_VLQ are skipped, :475; neighbours with such
 * labels are not counted, :499). out: q->n rows of g->n bytes. */
void og_ldf_nlf(const og_t *g, const oq_t *q, uint8_t *out) {
    uint32_t nl = q->nvlq;
    uint32_t qnlc[OQ_MAXV][OQ_MAXV + 1];
    memset(qnlc, 0, sizeof(qnlc));
    uint32_t *dn = (uint32_t *)calloc(nl + 1, 4);
    for (uint32_t u = 0; u < q->n; u++)
        for (uint32_t w = 0; w < q->n; w++)
            if (q->adjm[u] >> w & 1) qnlc[u][q->lab[w]]++;
    memset(out, 0, (size_t)q->n * g->n);
    for (uint32_t v = 0; v < g->n; v++) {
        if (g->lab[v] >= nl) continue;
        memset(dn, 0, (nl + 1) * 4);
        for (uint32_t i = 0; i < g->deg[v]; i++) {
            uint32_t l = g->lab[g->adj[v][i]];
            if (l < nl) dn[l]++;
        }
        for (uint32_t u = 0; u < q->n; u++) {
            if (q->lab[u] != g->lab[v] || q->deg[u] > g->deg[v]) continue;
            int ok = 1;
            for (uint32_t l = 0; l < nl; l++)
                if (dn[l] < qnlc[u][l]) { ok = 0; break; }
            if (ok) out[(size_t)u * g->n + v] = 1;
        }
    }
    free(dn);
}

/* ------------------------------------------------------------------- DFS */

typedef struct {
    const og_t *g;
    const oq_t *q;
    const uint8_t *cand;      /* q->n x g->n */
    uint32_t order[OQ_MAXV];  /* query vertex at each position */
    uint32_t back[OQ_MAXV];   /* bitmask of earlier positions adjacent */
    uint32_t f[OQ_MAXV];      /* data vertex at each position */
    uint64_t count;
    uint32_t *tup;            /* tuple sink (query-vertex order), may be NULL */
    uint64_t tup_cap, tup_n;
} dfs_t;

/* Count-neutral order: positions 0,1 are the seeded query edge; then the
 * unplaced vertex with the most placed neighbours, lowest id on ties. (The
 * reference's computeMatchingOrder, order.cpp:450-644, chooses differently; the
 * order never changes a count.) */
static void make_order(dfs_t *s, uint32_t u, uint32_t v) {
    const oq_t *q = s->q;
    uint32_t placed = (1u << u) | (1u << v);
    s->order[0] = u; s->order[1] = v;
    for (uint32_t p = 2; p < q->n; p++) {
        int best = -1, bestc = -1;
        for (uint32_t w = 0; w < q->n; w++) {
            if (placed >> w & 1) continue;
            int c = __builtin_popcount(q->adjm[w] & placed);
            if (c > bestc) { bestc = c; best = (int)w; }
        }
        s->order[p] = (uint32_t)best;
        placed |= 1u << best;
    }
    for (uint32_t p = 0; p < q->n; p++) {
        s->back[p] = 0;
        for (uint32_t j = 0; j < p; j++)
            if (q->adjm[s->order[p]] >> s->order[j] & 1) s->back[p] |= 1u << j;
    }
}

static void emit(dfs_t *s) {
    if (s->tup && s->tup_n < s->tup_cap) {
        uint32_t *t = s->tup + s->tup_n * s->q->n;
        for (uint32_t p = 0; p < s->q->n; p++) t[s->order[p]] = s->f[p];
    }
    s->tup_n++;
}

static void dfs(dfs_t *s, uint32_t p) {
    const og_t *g = s->g;
    uint32_t n = s->q->n;
    if (p == n) { s->count++; if (s->tup) emit(s); return; }
    uint32_t u = s->order[p];
    uint32_t bm = s->back[p];
    /* backward neighbour with the smallest list drives the scan (the reference
     * intersects in ascending degree, gpu_match.cu:189-247). */
    int src = -1;
    for (uint32_t j = 0; j < p; j++)
        if (bm >> j & 1) {
            if (src < 0 || g->deg[s->f[j]] < g->deg[s->f[src]]) src = (int)j;
        }
    if (src < 0) return; /* disconnected query: rejected at create time */
    const uint32_t *list = g->adj[s->f[src]];
    uint32_t len = g->deg[s->f[src]];
    const uint8_t *cu = s->cand + (size_t)u * g->n;
    for (uint32_t i = 0; i < len; i++) {
        uint32_t w = list[i];
        if (!cu[w]) continue;
        int ok = 1;
        for (uint32_t j = 0; j < p && ok; j++) if (s->f[j] == w) ok = 0; /* injectivity */
        for (uint32_t j = 0; j < p && ok; j++)
            if ((bm >> j & 1) && (int)j != src && !has_edge(g, s->f[j], w)) ok = 0;
        if (!ok) continue;
        s->f[p] = w;
        dfs(s, p + 1);
    }
}

/* Seeding (main.cu:268-329) for one record (s,d): every ordered query edge
 * (u,v) with s in C(u) and d in C(v) seeds f(u)=s, f(v)=d; the DFS counts the
 * embeddings that extend it. cand must be og_ldf_nlf's output for g. */
static void seed_record(dfs_t *st, uint32_t s, uint32_t d) {
    const og_t *g = st->g;
    const oq_t *q = st->q;
    if (s >= g->n || d >= g->n || s == d || !has_edge(g, s, d)) return;
    for (uint32_t u = 0; u < q->n; u++) {
        if (!st->cand[(size_t)u * g->n + s]) continue;
        for (uint32_t v = 0; v < q->n; v++) {
            if (!(q->adjm[u] >> v & 1)) continue;
            if (!st->cand[(size_t)v * g->n + d]) continue;
            make_order(st, u, v);
            st->f[0] = s; st->f[1] = d;
            dfs(st, 2);
        }
    }
}

/* Per-record embedding counts for k records evaluated in the current graph. */
void og_count_records(const og_t *g, const oq_t *q, const uint8_t *cand, uint32_t k,
                      const uint32_t *src, const uint32_t *dst, uint64_t *out) {
    dfs_t st; memset(&st, 0, sizeof(st));
    st.g = g; st.q = q; st.cand = cand;
    for (uint32_t i = 0; i < k; i++) {
        st.count = 0;
        seed_record(&st, src[i], dst[i]);
        out[i] = st.count;
    }
}

/* Tuples (f(u_0) .. f(u_{n-1})) of every seeded embedding, in seed order.
 * Returns the number of tuples found (may exceed cap; only cap are written). */
uint64_t og_tuples_records(const og_t *g, const oq_t *q, const uint8_t *cand, uint32_t k,
                           const uint32_t *src, const uint32_t *dst,
                           uint32_t *out, uint64_t cap) {
    dfs_t st; memset(&st, 0, sizeof(st));
    st.g = g; st.q = q; st.cand = cand;
    st.tup = out; st.tup_cap = cap;
    static uint32_t dummy[OQ_MAXV];
    if (!out) { st.tup = dummy; st.tup_cap = 0; }
    for (uint32_t i = 0; i < k; i++) seed_record(&st, src[i], dst[i]);
    return st.tup_n;
}

/* ------------------------------------------------------------ brute force */

/* Independent enumerator: for record (s,d) and each ordered query edge (u,v),
 * fixes f(u)=s, f(v)=d and tries EVERY data vertex of the right label for each
 * remaining query vertex (ascending query id), checking each query edge among
 * assigned vertices with a direct adjacency test. No candidate filter, no
 * adjacency-list scan, no matching order. */
typedef struct {
    const og_t *g; const oq_t *q;
    uint32_t f[OQ_MAXV]; uint32_t assigned; uint64_t count;
    uint32_t **bylab; uint32_t *nbylab;
    uint32_t *sink; uint64_t sink_cap; /* optional tuple sink (query-vertex order) */
} bf_t;

static void bf_rec(bf_t *b, uint32_t u) {
    const oq_t *q = b->q;
    while (u < q->n && (b->assigned >> u & 1)) u++;
    if (u == q->n) {
        if (b->sink && b->count < b->sink_cap) memcpy(b->sink + b->count * q->n, b->f, q->n * 4);
        b->count++;
        return;
    }
    uint32_t l = q->lab[u];
    for (uint32_t i = 0; i < b->nbylab[l]; i++) {
        uint32_t w = b->bylab[l][i];
        int ok = 1;
        for (uint32_t x = 0; x < q->n && ok; x++)
            if ((b->assigned >> x & 1) && b->f[x] == w) ok = 0;
        for (uint32_t x = 0; x < q->n && ok; x++)
            if ((b->assigned >> x & 1) && (q->adjm[u] >> x & 1) && !has_edge(b->g, w, b->f[x])) ok = 0;
        if (!ok) continue;
        b->f[u] = w; b->assigned |= 1u << u;
        bf_rec(b, u + 1);
        b->assigned &= ~(1u << u);
    }
}

uint64_t og_brute_record(const og_t *g, const oq_t *q, uint32_t s, uint32_t d) {
    if (s >= g->n || d >= g->n || s == d || !has_edge(g, s, d)) return 0;
    bf_t b; memset(&b, 0, sizeof(b));
    b.g = g; b.q = q;
    uint32_t nl = q->nvlq;
    b.bylab = (uint32_t **)calloc(nl, sizeof(uint32_t *));
    b.nbylab = (uint32_t *)calloc(nl, 4);
    for (uint32_t v = 0; v < g->n; v++) if (g->lab[v] < nl) b.nbylab[g->lab[v]]++;
    for (uint32_t l = 0; l < nl; l++) { b.bylab[l] = (uint32_t *)malloc((size_t)b.nbylab[l] * 4 + 4); b.nbylab[l] = 0; }
    for (uint32_t v = 0; v < g->n; v++) if (g->lab[v] < nl) b.bylab[g->lab[v]][b.nbylab[g->lab[v]]++] = v;
    for (uint32_t u = 0; u < q->n; u++) {
        if (q->lab[u] != g->lab[s]) continue;
        for (uint32_t v = 0; v < q->n; v++) {
            if (!(q->adjm[u] >> v & 1) || q->lab[v] != g->lab[d]) continue;
            b.assigned = (1u << u) | (1u << v);
            b.f[u] = s; b.f[v] = d;
            bf_rec(&b, 0);
        }
    }
    for (uint32_t l = 0; l < nl; l++) free(b.bylab[l]);
    free(b.bylab); free(b.nbylab);
    return b.count;
}

/* Total number of embeddings of q in g (every injective, label-preserving map
 * that sends every query edge onto a data edge), by the same brute force. */
uint64_t og_brute_total(const og_t *g, const oq_t *q, uint32_t *sink, uint64_t sink_cap) {
    bf_t b; memset(&b, 0, sizeof(b));
    b.g = g; b.q = q; b.sink = sink; b.sink_cap = sink_cap;
    uint32_t nl = q->nvlq;
    b.bylab = (uint32_t **)calloc(nl, sizeof(uint32_t *));
    b.nbylab = (uint32_t *)calloc(nl, 4);
    for (uint32_t v = 0; v < g->n; v++) if (g->lab[v] < nl) b.nbylab[g->lab[v]]++;
    for (uint32_t l = 0; l < nl; l++) { b.bylab[l] = (uint32_t *)malloc((size_t)b.nbylab[l] * 4 + 4); b.nbylab[l] = 0; }
    for (uint32_t v = 0; v < g->n; v++) if (g->lab[v] < nl) b.bylab[g->lab[v]][b.nbylab[g->lab[v]]++] = v;
    b.assigned = 0;
    bf_rec(&b, 0);
    for (uint32_t l = 0; l < nl; l++) free(b.bylab[l]);
    free(b.bylab); free(b.nbylab);
    return b.count;
}

/* ------------------------------------------------- cluster-encoding filter */

/* Restatement of the reference's cluster-encoding candidate refinement
 * (clusterFilter, filter.cu:1525-1654): the query is reduced to a BFS spanning
 * tree (decycle, decycle.cpp:6-135), a minimum vertex cover of the tree is
 * chosen (getVertexCover / dfs_cvc, filter.cu:23-116), each cover vertex and
 * its tree neighbours form a cluster, clusters sharing a vertex are merged
 * layer by layer and clusters with the same core are combined
 * (clustering, filter.cu:131-392; encodingMeta::init, structure.cu:107-135).
 * Every cluster member owns one bit position; encodeKernel, mergeKernel and
 * combineMultipleClustersKernel (filter.cu:615-934), run in the order of
 * encode() (filter.cu:936-1177), set and clear those bits per data vertex, and
 * a query vertex's refined candidates are the bits of its position in the
 * last valid cluster that holds it (filter.cu:1589-1609; compact,
 * filter.cu:1692-1750).
 *
 * The plan uses the query's adjacency in edge-line order (readGraphToCPU-
 * Dynamic appends neighbours as the e lines come, io.cu:176-196), which fixes
 * the BFS tree and the cover's order. One deviation: the reference updates a
 * vertex's own core bit with a plain read-modify-write of its 32-bit word
 * (filter.cu:765, :876-884) while other warps may atomically update other bits
 * of the same word; a lost update there would make its result depend on
 * timing. Here every bit update takes effect. */

#define OC_MAXC 150 /* MAX_CLUSTERS, defs.h:22 */
#define OC_MAXL 100 /* MAX_LAYERS, defs.h:23 */
#define OC_MAXSMALL 32 /* small clusters per combine (filter.cu:1056-1061) */

enum { OC_ENCODE = 0, OC_MERGE = 1, OC_COMBINE = 2 };

typedef struct {
    uint32_t n;
    uint32_t tdeg[OQ_MAXV], tadj[OQ_MAXV][OQ_MAXV]; /* spanning tree, edge-line order */
    uint32_t ncover, cover[OQ_MAXV];
    uint32_t nc;                                    /* clusters, valid or not */
    uint32_t csz[OC_MAXC], cu[OC_MAXC][OQ_MAXV + 1];
    uint8_t valid[OC_MAXC];
    uint32_t coff[OC_MAXC + 1];                     /* first bit position of each cluster */
    uint32_t nlayers, per_layer[OC_MAXL + 2], layer_off[OC_MAXL + 2];
    int checkpoint[OC_MAXL + 2];
    uint32_t nmerge, mleft[OC_MAXC], mright[OC_MAXC];
    uint32_t ncomb, cbig[OC_MAXC], ctype[OC_MAXC], cnsmall[OC_MAXC], csmall[OC_MAXC][OC_MAXSMALL];
    uint32_t final_cluster;
    uint32_t enc_pos[OQ_MAXV];
    /* the device passes in execution order */
    uint32_t nops, op_kind[3 * OC_MAXC], op_arg[3 * OC_MAXC], op_arg2[3 * OC_MAXC];
} ocp_t;

/* decycle.cpp:6-135: root = the first vertex of highest degree; BFS in
 * adjacency order keeps the edge to each newly reached vertex. */
static void oc_decycle(ocp_t *p, uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev) {
    uint32_t deg[OQ_MAXV] = {0}, adj[OQ_MAXV][2 * OQ_MAXV], eid[OQ_MAXV][2 * OQ_MAXV];
    uint8_t keep[2 * OQ_MAXV * OQ_MAXV];
    memset(keep, 0, sizeof(keep));
    for (uint32_t i = 0; i < m; i++) {
        adj[eu[i]][deg[eu[i]]] = ev[i]; eid[eu[i]][deg[eu[i]]++] = i;
        adj[ev[i]][deg[ev[i]]] = eu[i]; eid[ev[i]][deg[ev[i]]++] = i;
    }
    uint32_t root = 0;
    for (uint32_t u = 1; u < n; u++) if (deg[u] > deg[root]) root = u;
    uint32_t qu[OQ_MAXV], qh = 0, qt = 0, seen = 1u << root;
    qu[qt++] = root;
    while (qh < qt) {
        uint32_t t = qu[qh++];
        for (uint32_t k = 0; k < deg[t]; k++) {
            uint32_t w = adj[t][k];
            if (seen >> w & 1) continue;
            seen |= 1u << w;
            qu[qt++] = w;
            keep[eid[t][k]] = 1;
        }
    }
    p->n = n;
    for (uint32_t u = 0; u < n; u++) {
        p->tdeg[u] = 0;
        for (uint32_t k = 0; k < deg[u]; k++)
            if (keep[eid[u][k]]) p->tadj[u][p->tdeg[u]++] = adj[u][k];
    }
}

static void oc_cvc(const ocp_t *p, uint32_t v, uint32_t father, int f[][2]) {
    f[v][0] = 0; f[v][1] = 1;
    for (uint32_t k = 0; k < p->tdeg[v]; k++) {
        uint32_t c = p->tadj[v][k];
        if (c == father) continue;
        oc_cvc(p, c, v, f);
        f[v][0] += f[c][1];
        f[v][1] += f[c][0] < f[c][1] ? f[c][0] : f[c][1];
    }
}

/* filter.cu:23-116: tree DP, then a top-down pass with an explicit stack
 * (children pushed in adjacency order, so popped last first). */
static void oc_cover(ocp_t *p) {
    int f[OQ_MAXV][2];
    uint32_t root = 0;
    for (uint32_t u = 1; u < p->n; u++) if (p->tdeg[u] > p->tdeg[root]) root = u;
    oc_cvc(p, root, 0xFFFFFFFFu, f);
    uint32_t st[OQ_MAXV], sf[OQ_MAXV], ssel[OQ_MAXV], sp = 0;
    st[sp] = root; sf[sp] = 0xFFFFFFFFu; ssel[sp] = 0; sp++;
    p->ncover = 0;
    while (sp) {
        sp--;
        uint32_t cur = st[sp], fa = sf[sp], fsel = ssel[sp], child_sel;
        if (f[cur][1] <= f[cur][0]) {
            p->cover[p->ncover++] = cur;
            child_sel = 1;
        } else if (fsel || fa == 0xFFFFFFFFu) {
            child_sel = 0;
        } else {
            p->cover[p->ncover++] = cur;
            child_sel = 1;
        }
        for (uint32_t k = 0; k < p->tdeg[cur]; k++) {
            uint32_t c = p->tadj[cur][k];
            if (c == fa) continue;
            st[sp] = c; sf[sp] = cur; ssel[sp] = child_sel; sp++;
        }
    }
}

/* filter.cu:131-392: returns 0, or -1 when a limit of the reference
 * (MAX_CLUSTERS, MAX_LAYERS, 32 small clusters per combine) is exceeded. */
static int oc_clustering(ocp_t *p) {
    uint32_t nc = p->ncover;
    if (nc > OC_MAXC) return -1;
    memset(p->valid, 1, sizeof(p->valid));
    for (uint32_t c = 0; c < nc; c++) {
        uint32_t core = p->cover[c];
        p->csz[c] = p->tdeg[core] + 1;
        p->cu[c][0] = core;
        for (uint32_t k = 0; k < p->tdeg[core]; k++) p->cu[c][k + 1] = p->tadj[core][k];
    }
    for (uint32_t l = 0; l < OC_MAXL + 2; l++) { p->per_layer[l] = 0; p->layer_off[l] = 0; p->checkpoint[l] = -1; }
    p->per_layer[0] = nc;
    p->nmerge = p->ncomb = 0;
    p->final_cluster = 0;
    uint32_t layer = 0, i = 0, actual;
    do {
        uint32_t nnew = 0;
        for (; i < nc; i++) {
            if (!p->valid[i]) continue;
            uint32_t core_i = p->cu[i][0];
            for (uint32_t j = i + 1; j < nc; j++) {
                if (!p->valid[j]) continue;
                uint32_t core_j = p->cu[j][0];
                for (uint32_t a = 1; a < p->csz[i]; a++)
                    for (uint32_t b = 0; b < p->csz[j]; b++) {
                        if (p->cu[i][a] != p->cu[j][b]) continue;
                        uint32_t x = nc + nnew++;
                        if (x >= OC_MAXC) return -1;
                        p->final_cluster = x;
                        p->valid[x] = 1;
                        p->cu[x][0] = p->cu[i][a];
                        p->cu[x][1] = core_i;
                        if (b && core_i != core_j) { p->cu[x][2] = core_j; p->csz[x] = 3; }
                        else p->csz[x] = 2;
                        p->mleft[p->nmerge] = i;
                        p->mright[p->nmerge] = j;
                        p->nmerge++;
                    }
            }
        }
        layer++;
        if (layer > OC_MAXL) return -1;
        actual = nnew;
        p->checkpoint[layer] = (int)(nc + nnew);
        for (uint32_t o = nc; o < nc + nnew; o++) {
            if (!p->valid[o]) continue;
            uint32_t core = p->cu[o][0];
            uint32_t members[OC_MAXC], nm = 0, nbr = 0; /* nbr: set of non-core vertices */
            members[nm++] = o;
            for (uint32_t k = 1; k < p->csz[o]; k++) nbr |= 1u << p->cu[o][k];
            for (uint32_t x = o + 1; x < nc + nnew; x++) {
                if (!p->valid[x] || p->cu[x][0] != core) continue;
                members[nm++] = x;
                for (uint32_t k = 1; k < p->csz[x]; k++) nbr |= 1u << p->cu[x][k];
            }
            if (nm == 1) continue;
            uint32_t want = (uint32_t)__builtin_popcount(nbr) + 1, big = 0xFFFFFFFFu;
            for (uint32_t k = 0; k < nm; k++) /* members are in ascending order */
                if (p->csz[members[k]] == want) { big = members[k]; break; }
            uint32_t ci = p->ncomb;
            p->cnsmall[ci] = 0;
            if (big != 0xFFFFFFFFu) {
                p->ctype[ci] = 1;
                for (uint32_t k = 0; k < nm; k++)
                    if (members[k] != big) {
                        if (p->cnsmall[ci] >= OC_MAXSMALL) return -1;
                        p->csmall[ci][p->cnsmall[ci]++] = members[k];
                    }
            } else {
                p->ctype[ci] = 0;
                big = nc + nnew++;
                actual++;
                if (big >= OC_MAXC) return -1;
                p->valid[big] = 1;
                p->cu[big][0] = core;
                p->csz[big] = 1;
                for (uint32_t u = 0; u < p->n; u++) if (nbr >> u & 1) p->cu[big][p->csz[big]++] = u;
                for (uint32_t k = 0; k < nm; k++) {
                    if (p->cnsmall[ci] >= OC_MAXSMALL) return -1;
                    p->csmall[ci][p->cnsmall[ci]++] = members[k];
                }
            }
            p->final_cluster = big;
            p->cbig[ci] = big;
            actual -= p->cnsmall[ci];
            for (uint32_t k = 0; k < p->cnsmall[ci]; k++) p->valid[p->csmall[ci][k]] = 0;
            p->ncomb++;
        }
        nc += nnew;
        p->per_layer[layer] = nnew;
    } while (actual > 1);
    p->nc = nc;
    for (uint32_t l = 1; l <= layer; l++) p->layer_off[l] = p->layer_off[l - 1] + p->per_layer[l - 1];
    p->layer_off[layer + 1] = nc;
    p->nlayers = layer + 1;
    p->coff[0] = 0;
    for (uint32_t c = 0; c < nc; c++) p->coff[c + 1] = p->coff[c] + p->csz[c];
    return 0;
}

/* encode(), filter.cu:936-1177: the order in which the passes run, including
 * where each layer's combines fall (its combine checkpoint, or a later one, or
 * the final pass). */
static void oc_schedule(ocp_t *p) {
    p->nops = 0;
    for (uint32_t c = 0; c < p->per_layer[0]; c++) {
        p->op_kind[p->nops] = OC_ENCODE; p->op_arg[p->nops++] = c;
    }
    uint32_t layer = 1, sum = p->per_layer[0], cptr = 0, mptr = 0;
    for (uint32_t ci = p->per_layer[0]; ci < p->nc; ci++) {
        if (ci == sum + p->per_layer[layer]) { sum += p->per_layer[layer]; layer++; }
        if ((int)ci == p->checkpoint[layer]) {
            uint32_t next = p->layer_off[layer + 1];
            while (cptr < p->ncomb && p->cbig[cptr] < next) {
                p->op_kind[p->nops] = OC_COMBINE; p->op_arg[p->nops++] = cptr++;
            }
            if (ci != next) { ci = next - 1; continue; }
        }
        p->op_kind[p->nops] = OC_MERGE; p->op_arg[p->nops] = ci; p->op_arg2[p->nops++] = mptr++;
    }
    if (p->checkpoint[p->nlayers - 1] == (int)p->nc)
        while (cptr < p->ncomb) { p->op_kind[p->nops] = OC_COMBINE; p->op_arg[p->nops++] = cptr++; }
    /* filter.cu:1589-1609: each vertex's first position, scanning the valid
     * clusters down from the final one */
    uint32_t seen = 0, cnt = 0;
    for (int c = (int)p->final_cluster; c >= 0; c--) {
        if (!p->valid[c]) continue;
        if (cnt == p->n) break;
        for (uint32_t k = 0; k < p->csz[c]; k++) {
            uint32_t u = p->cu[c][k];
            if (!(seen >> u & 1)) { seen |= 1u << u; p->enc_pos[u] = p->coff[c] + k; cnt++; }
        }
    }
}

static int oc_plan(ocp_t *p, uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev) {
    if (n < 2 || n > 16) return -1;
    memset(p, 0, sizeof(*p));
    oc_decycle(p, n, m, eu, ev);
    oc_cover(p);
    if (oc_clustering(p)) return -1;
    oc_schedule(p);
    return 0;
}

/* The plan as flat words for the tests: n, ncover, cover[ncover], nc, then per
 * cluster (valid, size, members...), final cluster, enc_pos[n], nops, then
 * (kind, arg) per pass. Returns the word count (0 on a limit error). */
uint32_t oq_cluster_plan(uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev, uint32_t *out,
                         uint32_t cap) {
    ocp_t *p = (ocp_t *)calloc(1, sizeof(ocp_t));
    uint32_t w = 0;
#define PUT(x) do { if (w < cap) out[w] = (x); w++; } while (0)
    if (oc_plan(p, n, m, eu, ev) == 0) {
        PUT(p->n); PUT(p->ncover);
        for (uint32_t k = 0; k < p->ncover; k++) PUT(p->cover[k]);
        PUT(p->nc);
        for (uint32_t c = 0; c < p->nc; c++) {
            PUT(p->valid[c]); PUT(p->csz[c]);
            for (uint32_t k = 0; k < p->csz[c]; k++) PUT(p->cu[c][k]);
        }
        PUT(p->final_cluster);
        for (uint32_t u = 0; u < p->n; u++) PUT(p->enc_pos[u]);
        PUT(p->nops);
        for (uint32_t k = 0; k < p->nops; k++) { PUT(p->op_kind[k]); PUT(p->op_arg[k]); }
    }
#undef PUT
    free(p);
    return w;
}

typedef struct {
    const og_t *g;
    uint64_t words;
    uint64_t *bits; /* [positions][words] */
} oc_enc_t;

static inline int oc_get(const oc_enc_t *e, uint32_t pos, uint32_t v) {
    return (int)(e->bits[(uint64_t)pos * e->words + (v >> 6)] >> (v & 63) & 1);
}
static inline void oc_set(oc_enc_t *e, uint32_t pos, uint32_t v) {
    e->bits[(uint64_t)pos * e->words + (v >> 6)] |= 1ull << (v & 63);
}
static inline void oc_clr(oc_enc_t *e, uint32_t pos, uint32_t v) {
    e->bits[(uint64_t)pos * e->words + (v >> 6)] &= ~(1ull << (v & 63));
}

/* First position at or after `from` that holds query vertex u (the scans of
 * filter.cu:1104-1110 over query_us_compact_). */
static uint32_t oc_find(const ocp_t *p, uint32_t from_cluster, uint32_t u) {
    for (uint32_t c = from_cluster; c < p->nc; c++)
        for (uint32_t k = 0; k < p->csz[c]; k++)
            if (p->cu[c][k] == u) return p->coff[c] + k;
    return 0xFFFFFFFFu;
}

/* Refined candidates of the query (edge lines eu/ev in file order) in graph
 * g, from the LDF/NLF candidates `ldf` (qn rows of g->n bytes, og_ldf_nlf of
 * the same query): out receives qn rows of g->n bytes. Returns 0, or -1 when
 * the plan exceeds a limit of the reference. */
int og_cluster_filter(const og_t *g, uint32_t qn, uint32_t qm, const uint32_t *eu, const uint32_t *ev,
                      const uint8_t *ldf, uint8_t *out) {
    ocp_t *p = (ocp_t *)calloc(1, sizeof(ocp_t));
    if (oc_plan(p, qn, qm, eu, ev)) { free(p); return -1; }
    uint32_t n = g->n;
    oc_enc_t e;
    e.g = g;
    e.words = ((uint64_t)n + 63) / 64;
    e.bits = (uint64_t *)calloc((size_t)p->coff[p->nc] * e.words + 1, 8);
    /* candidate lists (d_u_candidate_vs_): LDF/NLF, entries removed by the
     * merge and combine passes */
    uint8_t *alive = (uint8_t *)malloc((size_t)qn * n + 1);
    memcpy(alive, ldf, (size_t)qn * n);
    uint32_t *snap = (uint32_t *)malloc((size_t)n * 4 + 4);
    for (uint32_t k = 0; k < p->nops; k++) {
        uint32_t kind = p->op_kind[k], arg = p->op_arg[k];
        uint32_t c = kind == OC_COMBINE ? p->cbig[arg] : arg;
        uint32_t core = p->cu[c][0], pos = p->coff[c], ns = 0;
        uint8_t *al = alive + (size_t)core * n;
        /* the list as the pass starts: a pass removes entries but still
         * visits the neighbours of every entry it started with */
        for (uint32_t v = 0; v < n; v++) if (al[v]) snap[ns++] = v;
        if (kind == OC_ENCODE) { /* encodeKernel, filter.cu:615-717 */
            for (uint32_t s = 0; s < ns; s++) {
                uint32_t v = snap[s];
                oc_set(&e, pos, v);
                for (uint32_t a = 0; a < g->deg[v]; a++) {
                    uint32_t w = g->adj[v][a];
                    for (uint32_t i = 1; i < p->csz[c]; i++)
                        if (ldf[(size_t)p->cu[c][i] * n + w]) oc_set(&e, pos + i, w);
                }
            }
        } else if (kind == OC_MERGE) { /* mergeKernel, filter.cu:719-803; filter.cu:1101-1127 */
            uint32_t left = p->mleft[p->op_arg2[k]], right = p->mright[p->op_arg2[k]];
            uint32_t lpos = oc_find(p, left, core), rpos = oc_find(p, right, core);
            uint32_t lcore = p->coff[left], rcore = p->coff[right];
            for (uint32_t s = 0; s < ns; s++) {
                uint32_t v = snap[s];
                if (oc_get(&e, lpos, v) && oc_get(&e, rpos, v)) oc_set(&e, pos, v);
                else al[v] = 0;
                for (uint32_t a = 0; a < g->deg[v]; a++) {
                    uint32_t w = g->adj[v][a];
                    if (oc_get(&e, lcore, w)) oc_set(&e, pos + 1, w);
                    if (rcore != rpos && oc_get(&e, rcore, w)) oc_set(&e, pos + 2, w);
                }
            }
        } else { /* combineMultipleClustersKernel, filter.cu:805-934 */
            uint32_t type = p->ctype[arg], nsm = p->cnsmall[arg];
            const uint32_t *sm = p->csmall[arg];
            /* per target position of the big cluster: the position of the same
             * query vertex in each small cluster that holds it (non-core part) */
            uint32_t tpos[OQ_MAXV + 1][OC_MAXSMALL], ntp[OQ_MAXV + 1];
            for (uint32_t i = 1; i < p->csz[c]; i++) {
                ntp[i] = 0;
                for (uint32_t j = 0; j < nsm; j++)
                    for (uint32_t x = 1; x < p->csz[sm[j]]; x++)
                        if (p->cu[sm[j]][x] == p->cu[c][i]) { tpos[i][ntp[i]++] = p->coff[sm[j]] + x; break; }
            }
            for (uint32_t s = 0; s < ns; s++) {
                uint32_t v = snap[s];
                int all = 1;
                for (uint32_t j = 0; j < nsm; j++) all &= oc_get(&e, p->coff[sm[j]], v);
                if (!all) { oc_clr(&e, pos, v); al[v] = 0; }
                else if (type == 0) oc_set(&e, pos, v);
                for (uint32_t a = 0; a < g->deg[v]; a++) {
                    uint32_t w = g->adj[v][a];
                    for (uint32_t i = 1; i < p->csz[c]; i++) {
                        int fin = 1;
                        for (uint32_t t = 0; t < ntp[i]; t++) fin &= oc_get(&e, tpos[i][t], w);
                        if (type == 1 && !fin) oc_clr(&e, pos + i, w);
                        if (type == 0 && fin) oc_set(&e, pos + i, w);
                    }
                }
            }
        }
    }
    for (uint32_t u = 0; u < qn; u++)
        for (uint32_t v = 0; v < n; v++) out[(size_t)u * n + v] = (uint8_t)oc_get(&e, p->enc_pos[u], v);
    free(snap); free(alive); free(e.bits); free(p);
    return 0;
}

/* ---------------------------------------------------------- DSM text file */

/* Token parser following readGraphToCPUDynamic (io.cu:122-252):
 *   t n m ; v id label deg (x n) ; e a b (x m) ; '-' ; b bid k ; (a|d) s d l (x k)
 * Query files end after the e lines. Stores the edge lines as read (no dedup).
 * Returns NULL on malformed input. */
typedef struct {
    uint32_t n, m, nb;
    uint32_t *lab, *deg, *eu, *ev;
    uint32_t *bstart;            /* nb+1 offsets into the change arrays */
    uint8_t *cadd; uint32_t *csrc, *cdst, *clab;
    uint64_t nchanges;
} of_t;

static int next_tok(FILE *f, char *buf, int cap) {
    int c, i = 0;
    do { c = fgetc(f); } while (c == ' ' || c == '\n' || c == '\t' || c == '\r');
    if (c == EOF) return 0;
    while (c != EOF && c != ' ' && c != '\n' && c != '\t' && c != '\r') {
        if (i < cap - 1) buf[i++] = (char)c;
        c = fgetc(f);
    }
    buf[i] = 0;
    return 1;
}

static int tok_u32(FILE *f, uint32_t *out) {
    char b[32];
    if (!next_tok(f, b, 32)) return 0;
    char *e; unsigned long v = strtoul(b, &e, 10);
    if (*e || e == b) return 0;
    *out = (uint32_t)v;
    return 1;
}

void of_free(of_t *o) {
    if (!o) return;
    free(o->lab); free(o->deg); free(o->eu); free(o->ev); free(o->bstart);
    free(o->cadd); free(o->csrc); free(o->cdst); free(o->clab); free(o);
}

of_t *of_read(const char *path, int is_query) {
    FILE *f = fopen(path, "r");
    if (!f) return NULL;
    of_t *o = (of_t *)calloc(1, sizeof(of_t));
    char t[32];
    if (!next_tok(f, t, 32) || strcmp(t, "t") || !tok_u32(f, &o->n) || !tok_u32(f, &o->m)) goto bad;
    o->lab = (uint32_t *)calloc((size_t)o->n + 1, 4);
    o->deg = (uint32_t *)calloc((size_t)o->n + 1, 4);
    o->eu = (uint32_t *)malloc((size_t)o->m * 4 + 4);
    o->ev = (uint32_t *)malloc((size_t)o->m * 4 + 4);
    uint32_t nv = 0, ne = 0;
    int ended = 0;
    while (next_tok(f, t, 32)) {
        if (!strcmp(t, "v")) {
            uint32_t id, l, d;
            if (!tok_u32(f, &id) || !tok_u32(f, &l) || !tok_u32(f, &d) || id != nv || id >= o->n) goto bad;
            o->lab[id] = l; o->deg[id] = d; nv++;
        } else if (!strcmp(t, "e")) {
            uint32_t a, b;
            if (!tok_u32(f, &a) || !tok_u32(f, &b) || ne >= o->m || a >= o->n || b >= o->n) goto bad;
            o->eu[ne] = a; o->ev[ne] = b; ne++;
        } else if (!strcmp(t, "-")) {
            ended = 1; break;
        } else goto bad;
    }
    if (nv != o->n || ne != o->m) goto bad;
    o->bstart = (uint32_t *)calloc(1, 4);
    if (!is_query && ended) {
        uint64_t capc = 1024;
        o->cadd = (uint8_t *)malloc(capc); o->csrc = (uint32_t *)malloc(capc * 4);
        o->cdst = (uint32_t *)malloc(capc * 4); o->clab = (uint32_t *)malloc(capc * 4);
        while (next_tok(f, t, 32)) {
            uint32_t bid, k;
            if (strcmp(t, "b") || !tok_u32(f, &bid) || !tok_u32(f, &k) || bid != o->nb) goto bad;
            for (uint32_t i = 0; i < k; i++) {
                uint32_t s, d, l;
                if (!next_tok(f, t, 32) || (strcmp(t, "a") && strcmp(t, "d"))) goto bad;
                if (!tok_u32(f, &s) || !tok_u32(f, &d) || !tok_u32(f, &l)) goto bad;
                if (s >= o->n || d >= o->n) goto bad;
                if (o->nchanges == capc) {
                    capc *= 2;
                    o->cadd = (uint8_t *)realloc(o->cadd, capc); o->csrc = (uint32_t *)realloc(o->csrc, capc * 4);
                    o->cdst = (uint32_t *)realloc(o->cdst, capc * 4); o->clab = (uint32_t *)realloc(o->clab, capc * 4);
                }
                o->cadd[o->nchanges] = t[0] == 'a'; o->csrc[o->nchanges] = s;
                o->cdst[o->nchanges] = d; o->clab[o->nchanges] = l; o->nchanges++;
            }
            o->nb++;
            o->bstart = (uint32_t *)realloc(o->bstart, (size_t)(o->nb + 1) * 4);
            o->bstart[o->nb] = (uint32_t)o->nchanges;
        }
    }
    fclose(f);
    return o;
bad:
    fclose(f);
    of_free(o);
    return NULL;
}

/* accessors for ctypes */
uint32_t of_n(const of_t *o) { return o->n; }
uint32_t of_m(const of_t *o) { return o->m; }
uint32_t of_nb(const of_t *o) { return o->nb; }
uint64_t of_nchanges(const of_t *o) { return o->nchanges; }
void of_copy(const of_t *o, uint32_t *lab, uint32_t *deg, uint32_t *eu, uint32_t *ev) {
    memcpy(lab, o->lab, (size_t)o->n * 4); memcpy(deg, o->deg, (size_t)o->n * 4);
    memcpy(eu, o->eu, (size_t)o->m * 4); memcpy(ev, o->ev, (size_t)o->m * 4);
}
void of_copy_changes(const of_t *o, uint32_t *bstart, uint8_t *add, uint32_t *src, uint32_t *dst, uint32_t *lab) {
    memcpy(bstart, o->bstart, (size_t)(o->nb + 1) * 4);
    memcpy(add, o->cadd, o->nchanges); memcpy(src, o->csrc, o->nchanges * 4);
    memcpy(dst, o->cdst, o->nchanges * 4); memcpy(lab, o->clab, o->nchanges * 4);
}
