// CDSM command line: drop-in for the reference's `CDSM -q <query> -d <data>
// [--gpu N]` (main.cu:37-45). For every update batch of the data file it
// prints the positive match count as `res: <N>` (the reference's intended
// output, gpu_match.cu:1338/1429), one line per batch and query.
//
// Extensions (not in the reference): -q may be repeated (standing queries
// matched together), --neg prints negative counts, --no-relabel ignores the
// label field of insert records, --stats prints per-batch device timings.
// Without --stats the batches are matched pipelined (dsm_process_batches:
// one DFS launch per graph state); the counts are the same either way.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dsm.h"

static void usage() {
  std::fprintf(stderr,
               "usage: CDSM -q <query file> [-q <query file> ...] -d <data file> [--gpu N]\n"
               "            [--neg] [--no-relabel] [--stats]\n");
}

int main(int argc, char **argv) {
  std::vector<std::string> qpaths;
  std::string dpath;
  int gpu = 0;
  bool neg = false, relabel = true, stats = false;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "-q" && i + 1 < argc) qpaths.push_back(argv[++i]);
    else if (a == "-d" && i + 1 < argc) dpath = argv[++i];
    else if (a == "--gpu" && i + 1 < argc) gpu = std::atoi(argv[++i]);
    else if (a == "--neg") neg = true;
    else if (a == "--no-relabel") relabel = false;
    else if (a == "--stats") stats = true;
    else {
      usage();
      return 1;
    }
  }
  if (qpaths.empty() || dpath.empty()) {
    usage();
    return 1;
  }
  std::vector<dsm_file *> qf(qpaths.size(), nullptr);
  std::vector<dsm_query_desc> qd(qpaths.size());
  for (size_t i = 0; i < qpaths.size(); i++) {
    if (dsm_file_read(qpaths[i].c_str(), 1, &qf[i]) != DSM_OK) {
      std::fprintf(stderr, "error: %s\n", dsm_file_error());
      return 1;
    }
    dsm_file_query(qf[i], &qd[i]);
  }
  auto t0 = std::chrono::steady_clock::now();
  dsm_file *df = nullptr;
  if (dsm_file_read(dpath.c_str(), 0, &df) != DSM_OK) {
    std::fprintf(stderr, "error: %s\n", dsm_file_error());
    return 1;
  }
  dsm_graph_desc gd;
  dsm_file_graph(df, &gd);
  auto t1 = std::chrono::steady_clock::now();
  std::printf("read (seconds): %.6f\n", std::chrono::duration<double>(t1 - t0).count());
  dsm_options opt;
  dsm_options_default(&opt);
  opt.device = gpu;
  opt.relabel = relabel ? 1 : 0;
  opt.count_negatives = neg ? 1 : 0;
  dsm_engine *e = nullptr;
  int rc = dsm_create(&gd, qd.data(), (uint32_t)qd.size(), &opt, &e);
  if (rc != DSM_OK) {
    std::fprintf(stderr, "error: %s\n", dsm_last_error(nullptr));
    return 1;
  }
  auto t2 = std::chrono::steady_clock::now();
  std::printf("device set-up (seconds): %.6f\n", std::chrono::duration<double>(t2 - t1).count());
  const uint32_t nq = (uint32_t)qd.size();
  std::vector<uint64_t> pos(nq), ng(nq);
  const uint32_t nb = dsm_file_num_batches(df);
  auto print_counts = [&](const uint64_t *p, const uint64_t *n) {
    for (uint32_t q = 0; q < nq; q++) {
      if (nq > 1) std::printf("query %u ", q);
      std::printf("res: %llu\n", (unsigned long long)p[q]);
      if (neg) std::printf("neg: %llu\n", (unsigned long long)n[q]);
    }
  };
  if (!stats && nb > 0) {
    std::vector<dsm_change> all;
    std::vector<uint32_t> off(nb + 1, 0);
    for (uint32_t b = 0; b < nb; b++) {
      const dsm_change *ch;
      uint32_t k;
      dsm_file_batch(df, b, &ch, &k);
      all.insert(all.end(), ch, ch + k);
      off[b + 1] = off[b] + k;
    }
    std::vector<uint64_t> P((size_t)nb * nq), N((size_t)nb * nq);
    rc = dsm_process_batches(e, all.data(), off.data(), nb, P.data(), N.data(), nullptr);
    if (rc != DSM_OK) {
      std::fprintf(stderr, "error: %s\n", dsm_last_error(e));
      dsm_destroy(e);
      return 1;
    }
    for (uint32_t b = 0; b < nb; b++) print_counts(P.data() + (size_t)b * nq, N.data() + (size_t)b * nq);
  }
  for (uint32_t b = 0; stats && b < nb; b++) {
    const dsm_change *ch;
    uint32_t k;
    dsm_file_batch(df, b, &ch, &k);
    auto s0 = std::chrono::steady_clock::now();
    rc = dsm_process_batch(e, ch, k, pos.data(), ng.data());
    auto s1 = std::chrono::steady_clock::now();
    if (rc != DSM_OK) {
      std::fprintf(stderr, "error: batch %u: %s\n", b, dsm_last_error(e));
      dsm_destroy(e);
      return 1;
    }
    print_counts(pos.data(), ng.data());
    {
      dsm_stats st;
      dsm_get_stats(e, &st);
      std::printf("batch %u: %u changes, %.3f ms wall; apply %.3f filter %.3f seed %.3f dfs+ %.3f dfs- %.3f ms; "
                  "seeds %llu/%llu donated %llu\n",
                  b, k, std::chrono::duration<double, std::milli>(s1 - s0).count(), st.ms_apply, st.ms_filter,
                  st.ms_seed, st.ms_dfs_pos, st.ms_dfs_neg, (unsigned long long)st.seeds_pos,
                  (unsigned long long)st.seeds_neg, (unsigned long long)st.tasks_donated);
    }
  }
  dsm_destroy(e);
  dsm_file_free(df);
  for (auto *f : qf) dsm_file_free(f);
  return 0;
}
