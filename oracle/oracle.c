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
