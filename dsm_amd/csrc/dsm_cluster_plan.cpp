// Host planning of the cluster-encoding candidate refinement (the reference's
// clusterFilter, filter.cu:1525-1654) for one query. Pure host code: the
// passes it plans run as k_cluster_pass / k_cluster_heavy (dsm_cluster.hip).
//
//   spanning tree     decycle.cpp:6-135   BFS from the first vertex of highest
//                                         degree, neighbours in edge-line order
//   vertex cover      filter.cu:23-116    tree DP + top-down pass with a stack
//   clusters          filter.cu:131-392   cover vertex + tree neighbours; merge
//                                         clusters sharing a vertex layer by
//                                         layer; combine same-core clusters
//   bit positions     structure.cu:107-135
//   pass order        filter.cu:936-1177  (encode(), including where each
//                                         layer's combines run)
//   final positions   filter.cu:1589-1609 (enc_pos_of_u_)
#include <algorithm>
#include <cstdint>
#include <set>
#include <vector>

#include "dsm_internal.h"

namespace dsm {

namespace {

struct Builder {
  uint32_t n = 0;
  std::vector<std::vector<uint32_t>> tree;  // adjacency of the spanning tree, edge-line order
  std::vector<uint32_t> cover;
  ClusterPlan *P = nullptr;

  // decycle.cpp:6-135
  void spanning_tree(uint32_t m, const uint32_t *eu, const uint32_t *ev) {
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> adj(n);  // (neighbour, edge line)
    for (uint32_t i = 0; i < m; i++) {
      adj[eu[i]].push_back({ev[i], i});
      adj[ev[i]].push_back({eu[i], i});
    }
    uint32_t root = 0;
    for (uint32_t u = 1; u < n; u++)
      if (adj[u].size() > adj[root].size()) root = u;
    std::vector<bool> kept(m, false), seen(n, false);
    std::vector<uint32_t> fifo{root};
    seen[root] = true;
    for (size_t h = 0; h < fifo.size(); h++)
      for (auto [w, line] : adj[fifo[h]])
        if (!seen[w]) {
          seen[w] = true;
          kept[line] = true;
          fifo.push_back(w);
        }
    tree.assign(n, {});
    for (uint32_t u = 0; u < n; u++)
      for (auto [w, line] : adj[u])
        if (kept[line]) tree[u].push_back(w);
  }

  // filter.cu:23-116
  void vertex_cover() {
    uint32_t root = 0;
    for (uint32_t u = 1; u < n; u++)
      if (tree[u].size() > tree[root].size()) root = u;
    std::vector<int> without(n), with(n);  // cover size of the subtree without / with the vertex
    // post-order by an explicit DFS
    std::vector<std::pair<uint32_t, uint32_t>> order, st{{root, kNone}};
    while (!st.empty()) {
      auto [v, parent] = st.back();
      st.pop_back();
      order.push_back({v, parent});
      for (uint32_t c : tree[v])
        if (c != parent) st.push_back({c, v});
    }
    for (size_t i = order.size(); i-- > 0;) {
      auto [v, parent] = order[i];
      without[v] = 0;
      with[v] = 1;
      for (uint32_t c : tree[v])
        if (c != parent) {
          without[v] += with[c];
          with[v] += std::min(without[c], with[c]);
        }
    }
    struct Item {
      uint32_t v, parent;
      bool parent_in;
    };
    std::vector<Item> stack{{root, kNone, false}};
    cover.clear();
    while (!stack.empty()) {
      Item it = stack.back();
      stack.pop_back();
      bool take;
      if (with[it.v] <= without[it.v]) take = true;
      else take = !(it.parent_in || it.parent == kNone);
      if (take) cover.push_back(it.v);
      for (uint32_t c : tree[it.v])
        if (c != it.parent) stack.push_back({c, it.v, take});
    }
  }

  // filter.cu:131-392
  int clusters() {
    auto &C = P->clusters;
    C.clear();
    for (uint32_t core : cover) {
      ClusterPlan::Cluster c;
      c.members.push_back(core);
      c.members.insert(c.members.end(), tree[core].begin(), tree[core].end());
      C.push_back(c);
    }
    std::vector<uint32_t> per_layer{(uint32_t)C.size()};
    std::vector<int> checkpoint{-1};
    uint32_t i = 0, actual = 0;
    do {
      const uint32_t nc = (uint32_t)C.size();
      // merges: every vertex a non-core member of cluster i shares with cluster j > i
      for (; i < nc; i++) {
        if (!C[i].valid) continue;
        for (uint32_t j = i + 1; j < nc; j++) {
          if (!C[j].valid) continue;
          for (size_t a = 1; a < C[i].members.size(); a++)
            for (size_t b = 0; b < C[j].members.size(); b++) {
              if (C[i].members[a] != C[j].members[b]) continue;
              ClusterPlan::Cluster x;
              x.members = {C[i].members[a], C[i].members[0]};
              if (b > 0 && C[i].members[0] != C[j].members[0]) x.members.push_back(C[j].members[0]);
              x.left = i;
              x.right = j;
              x.merged = true;
              C.push_back(x);
              P->final_cluster = (uint32_t)C.size() - 1;
            }
        }
      }
      const uint32_t merged_end = (uint32_t)C.size();
      checkpoint.push_back((int)merged_end);
      actual = merged_end - nc;
      // combines: clusters of this layer with the same core
      for (uint32_t o = nc; o < C.size(); o++) {
        if (!C[o].valid) continue;
        const uint32_t core = C[o].members[0];
        std::vector<uint32_t> same{o};
        std::set<uint32_t> nbrs(C[o].members.begin() + 1, C[o].members.end());
        for (uint32_t x = o + 1; x < C.size(); x++)
          if (C[x].valid && C[x].members[0] == core) {
            same.push_back(x);
            nbrs.insert(C[x].members.begin() + 1, C[x].members.end());
          }
        if (same.size() < 2) continue;
        ClusterPlan::Combine cb;
        uint32_t big = kNone;
        for (uint32_t x : same)
          if (C[x].members.size() == nbrs.size() + 1) {
            big = x;
            break;
          }
        if (big != kNone) {
          cb.keep = true;
          for (uint32_t x : same)
            if (x != big) cb.small.push_back(x);
        } else {
          ClusterPlan::Cluster nw;
          nw.members.push_back(core);
          nw.members.insert(nw.members.end(), nbrs.begin(), nbrs.end());
          C.push_back(nw);
          big = (uint32_t)C.size() - 1;
          actual++;
          cb.keep = false;
          cb.small = same;
        }
        if (cb.small.size() > kMaxSmallClusters) return -1;
        cb.big = big;
        P->final_cluster = big;
        actual -= (uint32_t)cb.small.size();
        for (uint32_t x : cb.small) C[x].valid = false;
        P->combines.push_back(cb);
      }
      per_layer.push_back((uint32_t)C.size() - nc);
      if (C.size() > kMaxClusters || per_layer.size() > kMaxClusterLayers) return -1;
    } while (actual > 1);
    // bit positions: clusters in index order, members in order
    uint32_t pos = 0;
    for (auto &c : C) {
      c.pos = pos;
      pos += (uint32_t)c.members.size();
    }
    P->positions = pos;
    schedule(per_layer, checkpoint);
    return 0;
  }

  uint32_t position_of(uint32_t from, uint32_t u) const {
    for (uint32_t c = from; c < P->clusters.size(); c++) {
      const auto &m = P->clusters[c].members;
      for (size_t k = 0; k < m.size(); k++)
        if (m[k] == u) return P->clusters[c].pos + (uint32_t)k;
    }
    return kNone;
  }

  // encode(), filter.cu:936-1177. A layer's combines run when the pass loop
  // reaches the layer's first combined cluster; a layer without one never
  // meets its checkpoint (the loop moves to the next layer first), so its
  // combines run at a later checkpoint or in the final sweep.
  void schedule(const std::vector<uint32_t> &per_layer, const std::vector<int> &checkpoint) {
    auto &C = P->clusters;
    const uint32_t nc = (uint32_t)C.size();
    std::vector<uint32_t> layer_off(per_layer.size() + 1, 0);
    for (size_t l = 0; l < per_layer.size(); l++) layer_off[l + 1] = layer_off[l] + per_layer[l];
    auto per = [&](size_t l) { return l < per_layer.size() ? per_layer[l] : 0u; };
    auto cp = [&](size_t l) { return l < checkpoint.size() ? checkpoint[l] : -1; };
    auto off = [&](size_t l) { return l < layer_off.size() ? layer_off[l] : nc; };
    P->passes.clear();
    for (uint32_t c = 0; c < per_layer[0]; c++) P->passes.push_back({ClusterPlan::kEncode, c});
    size_t layer = 1, comb = 0;
    uint32_t sum = per_layer[0];
    for (uint32_t ci = per_layer[0]; ci < nc; ci++) {
      if (ci == sum + per(layer)) sum += per(layer++);
      if ((int)ci == cp(layer)) {
        while (comb < P->combines.size() && P->combines[comb].big < off(layer + 1))
          P->passes.push_back({ClusterPlan::kCombine, (uint32_t)comb++});
        if (ci != off(layer + 1)) {
          ci = off(layer + 1) - 1;
          continue;
        }
      }
      P->passes.push_back({ClusterPlan::kMerge, ci});
    }
    if (cp(per_layer.size() - 1) == (int)nc)
      while (comb < P->combines.size()) P->passes.push_back({ClusterPlan::kCombine, (uint32_t)comb++});
    // each vertex: its first position in the valid clusters, from the final one down
    P->final_pos.assign(n, kNone);
    uint32_t found = 0;
    for (int c = (int)P->final_cluster; c >= 0 && found < n; c--) {
      if (!C[c].valid) continue;
      for (size_t k = 0; k < C[c].members.size(); k++) {
        uint32_t u = C[c].members[k];
        if (P->final_pos[u] == kNone) {
          P->final_pos[u] = C[c].pos + (uint32_t)k;
          found++;
        }
      }
    }
    // merge passes: where the connection vertex sits in the two merged clusters
    for (auto &c : C)
      if (c.merged) {
        c.left_pos = position_of(c.left, c.members[0]);
        c.right_pos = position_of(c.right, c.members[0]);
      }
  }
};

}  // namespace

int build_cluster_plan(uint32_t n, uint32_t m, const uint32_t *eu, const uint32_t *ev, ClusterPlan &out) {
  if (n < 2 || n > (uint32_t)kMaxQ) return -1;
  out = ClusterPlan();
  Builder b;
  b.n = n;
  b.P = &out;
  b.spanning_tree(m, eu, ev);
  b.vertex_cover();
  return b.clusters();
}

}  // namespace dsm
