// DSM text format reader (the reference's readGraphToCPUDynamic, io.cu:122-252):
//
//   t <n> <m>
//   v <id> <label> <deg>        n lines, ids 0..n-1 in order
//   e <a> <b> [<elabel>]        m lines, each undirected edge once; the
//                               optional edge label is kept but never matched
//   -                           data files only; ends the graph
//   b <bid> <k>                 bid = 0,1,2,... in order
//   a|d <src> <dst> <label>     k lines
//
// Tokens are whitespace separated, as with ifstream >> (io.cu:162-246). Unlike
// the reference, which exits on a bad token and builds a corrupt CSR when a
// v-line degree disagrees with the e-lines, every problem is reported as
// DSM_EFORMAT with a message.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dsm.h"

struct dsm_file {
  bool is_query = false;
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<uint32_t> labels, deg, eu, ev;
  std::vector<uint32_t> elab;  // edge labels of the e lines (0 where absent)
  bool has_elab = false;
  std::vector<uint32_t> bstart;
  std::vector<dsm_change> changes;
};

namespace {

thread_local std::string g_file_error;

struct Reader {
  FILE *f;
  std::vector<char> buf;
  size_t pos = 0, len = 0;
  bool eof = false;
  std::string held;  // a token read ahead and given back
  bool has_held = false;
  explicit Reader(FILE *fp) : f(fp), buf(1 << 20) {}
  void unget(const std::string &t) {
    held = t;
    has_held = true;
  }
  int getc() {
    if (pos == len) {
      if (eof) return EOF;
      len = fread(buf.data(), 1, buf.size(), f);
      pos = 0;
      if (len == 0) {
        eof = true;
        return EOF;
      }
    }
    return (unsigned char)buf[pos++];
  }
  // next whitespace-separated token; false at end of input
  bool token(std::string &t) {
    if (has_held) {
      t = held;
      has_held = false;
      return true;
    }
    t.clear();
    int c;
    do {
      c = getc();
    } while (c == ' ' || c == '\n' || c == '\t' || c == '\r');
    if (c == EOF) return false;
    while (c != EOF && c != ' ' && c != '\n' && c != '\t' && c != '\r') {
      t.push_back((char)c);
      c = getc();
    }
    return true;
  }
  bool u32(uint32_t &out) {
    std::string t;
    return token(t) && parse_u32(t, out);
  }
  static bool parse_u32(const std::string &t, uint32_t &out) {
    if (t.empty() || t.size() > 10) return false;
    uint64_t v = 0;
    for (char ch : t) {
      if (ch < '0' || ch > '9') return false;
      v = v * 10 + (uint64_t)(ch - '0');
    }
    if (v > 0xFFFFFFFFull) return false;
    out = (uint32_t)v;
    return true;
  }
};

int bad(const char *path, const char *what) {
  g_file_error = std::string(path) + ": " + what;
  return DSM_EFORMAT;
}

}  // namespace

extern "C" {

const char *dsm_file_error(void) { return g_file_error.c_str(); }

int dsm_file_read(const char *path, int is_query, dsm_file **out) {
  if (!path || !out) return DSM_EINVAL;
  *out = nullptr;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) {
    g_file_error = std::string("cannot open ") + path;
    return DSM_EIO;
  }
  Reader r(fp);
  dsm_file *f = new dsm_file();
  f->is_query = is_query != 0;
  auto fail = [&](const char *what) {
    std::fclose(fp);
    delete f;
    return bad(path, what);
  };
  std::string t;
  uint32_t n, m32;
  if (!r.token(t) || t != "t" || !r.u32(n) || !r.u32(m32)) return fail("expected 't <n> <m>'");
  f->n = n;
  f->m = m32;
  f->labels.resize(n);
  f->deg.resize(n);
  f->eu.reserve(m32);
  f->ev.reserve(m32);
  uint32_t nv = 0;
  bool ended = false;
  while (r.token(t)) {
    if (t == "v") {
      uint32_t id, l, d;
      if (!r.u32(id) || !r.u32(l) || !r.u32(d)) return fail("bad v line");
      if (id != nv || id >= n) return fail("v lines must list ids 0..n-1 in order");
      f->labels[id] = l;
      f->deg[id] = d;
      nv++;
    } else if (t == "e") {
      uint32_t a, b;
      if (!r.u32(a) || !r.u32(b)) return fail("bad e line");
      if (a >= n || b >= n) return fail("edge endpoint out of range");
      if (f->eu.size() >= m32) return fail("more e lines than declared");
      f->eu.push_back(a);
      f->ev.push_back(b);
      // an edge label, when the next token is a number rather than a line type
      uint32_t el = 0;
      if (r.token(t)) {
        if (Reader::parse_u32(t, el)) f->has_elab = true;
        else r.unget(t);
      }
      f->elab.push_back(el);
    } else if (t == "-") {
      ended = true;
      break;
    } else {
      return fail("unexpected token in graph section");
    }
  }
  if (nv != n) return fail("fewer v lines than declared");
  if (f->eu.size() != m32) return fail("fewer e lines than declared");
  {
    std::vector<uint32_t> cnt(n, 0);
    for (size_t i = 0; i < f->eu.size(); i++) {
      cnt[f->eu[i]]++;
      cnt[f->ev[i]]++;
    }
    for (uint32_t v = 0; v < n; v++)
      if (cnt[v] != f->deg[v]) return fail("v line degree differs from the vertex's e lines");
  }
  f->bstart.push_back(0);
  if (!f->is_query && ended) {
    uint32_t bid, k;
    while (r.token(t)) {
      if (t != "b" || !r.u32(bid) || !r.u32(k)) return fail("expected 'b <bid> <k>'");
      if (bid != f->bstart.size() - 1) return fail("batch ids must be 0,1,2,... in order");
      for (uint32_t i = 0; i < k; i++) {
        dsm_change c;
        if (!r.token(t) || (t != "a" && t != "d")) return fail("expected 'a' or 'd'");
        c.add = t == "a";
        if (!r.u32(c.src) || !r.u32(c.dst) || !r.u32(c.label)) return fail("bad change line");
        if (c.src >= n || c.dst >= n) return fail("change endpoint out of range");
        f->changes.push_back(c);
      }
      f->bstart.push_back((uint32_t)f->changes.size());
    }
  } else if (f->is_query && ended) {
    return fail("query files end after their e lines");
  }
  std::fclose(fp);
  *out = f;
  return DSM_OK;
}

void dsm_file_free(dsm_file *f) { delete f; }

int dsm_file_graph(const dsm_file *f, dsm_graph_desc *g) {
  if (!f || !g) return DSM_EINVAL;
  g->num_vertices = f->n;
  g->labels = f->labels.data();
  g->num_edges = f->m;
  g->edge_src = f->eu.data();
  g->edge_dst = f->ev.data();
  return DSM_OK;
}

int dsm_file_query(const dsm_file *f, dsm_query_desc *q) {
  if (!f || !q) return DSM_EINVAL;
  q->num_vertices = f->n;
  q->labels = f->labels.data();
  q->num_edges = (uint32_t)f->m;
  q->edge_src = f->eu.data();
  q->edge_dst = f->ev.data();
  return DSM_OK;
}

int dsm_file_edge_labels(const dsm_file *f, const uint32_t **labels) {
  if (!f || !labels) return DSM_EINVAL;
  *labels = f->elab.data();
  return f->has_elab ? 1 : 0;
}

uint32_t dsm_file_num_batches(const dsm_file *f) { return f ? (uint32_t)f->bstart.size() - 1 : 0; }

int dsm_file_batch(const dsm_file *f, uint32_t b, const dsm_change **changes, uint32_t *num) {
  if (!f || !changes || !num || b + 1 >= f->bstart.size()) return DSM_EINVAL;
  *changes = f->changes.data() + f->bstart[b];
  *num = f->bstart[b + 1] - f->bstart[b];
  return DSM_OK;
}

}  // extern "C"
