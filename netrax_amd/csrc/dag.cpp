// Op-DAG compiler: which CLVs exist and what each is computed from.
//
// NetRAX's DP (src/likelihood/ImprovedLoglikelihood.cpp:194-408) keeps, for
// every node, one CLV per distinct displayed subtree below it, tagged with
// the reticulation choices under which it exists
// (ReticulationConfigSet, src/graph/ReticulationConfigSet.cpp).  Two
// configurations share a CLV at a node exactly when they leave the same
// alive edges below it: inactive reticulation edges are dropped and dead
// nodes -- nodes without an active alive child -- vanish
// (src/helper/NetworkFunctions.cpp:228-279).  This compiler reproduces that
// set directly: it walks the 2^r configurations (bit r set = second parent,
// src/helper/ReticulationConfigHelper.cpp:188-203), keys every alive inner
// node by the bitset of alive edges below it, and interns (node, key) as one
// CLV op whose children are the alive children's ops (or tips).  Unary
// nodes stay as one-child ops, as the reference keeps them with a fake
// all-ones sibling (ImprovedLoglikelihood.cpp:115-152).
//
// Root classes are the root variants: configurations with the same alive
// edge set.  Each class is scored at its displayed-tree root, the first
// node with two alive children below the network root
// (ReticulationConfigHelper.cpp:251-272).
#include <algorithm>
#include <stdexcept>
#include <string>
#include <unordered_map>

#include "engine.hpp"

namespace nrx {

namespace {

struct KeyHash {
  size_t operator()(const std::string &s) const { return std::hash<std::string>()(s); }
};

}  // namespace

Dag compile_dag(const nrx_network_desc &net) {
  const int N = net.n_nodes, E = net.n_edges, R = net.n_retic, T = net.n_tips;
  if (N < 2 || T < 2 || E != N - 1 + R || net.root < T || net.root >= N)
    throw std::invalid_argument("inconsistent network sizes");
  if (R < 0 || R > 20)
    throw std::invalid_argument("between 0 and 20 reticulations are supported");
  std::vector<int> indeg(N, 0), outdeg(N, 0);
  for (int e = 0; e < E; ++e) {
    const int p = net.edge_parent[e], c = net.edge_child[e];
    if (p < 0 || p >= N || c < 0 || c >= N || p == c)
      throw std::invalid_argument("edge endpoint out of range");
    if (p < T) throw std::invalid_argument("a tip has a child");
    indeg[c]++;
    outdeg[p]++;
  }
  std::vector<int> retic_of(N, -1);
  for (int r = 0; r < R; ++r) {
    const int e1 = net.retic_first_edge[r], e2 = net.retic_second_edge[r];
    if (e1 < 0 || e1 >= E || e2 < 0 || e2 >= E || e1 == e2 ||
        net.edge_child[e1] != net.edge_child[e2])
      throw std::invalid_argument("bad reticulation edges");
    retic_of[net.edge_child[e1]] = r;
  }
  for (int v = 0; v < N; ++v) {
    const int want_in = v == net.root ? 0 : (retic_of[v] >= 0 ? 2 : 1);
    if (indeg[v] != want_in) throw std::invalid_argument("bad in-degree");
    if (v < T ? outdeg[v] != 0 : (outdeg[v] < 1 || outdeg[v] > 2))
      throw std::invalid_argument("bad out-degree");
  }
  // children CSR and topological order (parents first)
  std::vector<int> kstart(N + 1, 0), kedge(E);
  for (int e = 0; e < E; ++e) kstart[net.edge_parent[e] + 1]++;
  for (int v = 0; v < N; ++v) kstart[v + 1] += kstart[v];
  {
    std::vector<int> fill(N, 0);
    for (int e = 0; e < E; ++e) {
      const int p = net.edge_parent[e];
      kedge[kstart[p] + fill[p]++] = e;
    }
  }
  std::vector<int> topo;
  topo.reserve(N);
  {
    std::vector<int> rem = indeg;
    std::vector<int> stack{net.root};
    while (!stack.empty()) {
      const int v = stack.back();
      stack.pop_back();
      topo.push_back(v);
      for (int k = kstart[v]; k < kstart[v + 1]; ++k) {
        const int c = net.edge_child[kedge[k]];
        if (--rem[c] == 0) stack.push_back(c);
      }
    }
    if (static_cast<int>(topo.size()) != N) throw std::invalid_argument("network has a cycle");
  }

  const int W = (E + 63) / 64;
  Dag dag;
  dag.n_retic = R;
  std::vector<std::unordered_map<std::string, int, KeyHash>> interned(N);
  std::unordered_map<std::string, int, KeyHash> class_of_key;
  std::vector<uint64_t> key(static_cast<size_t>(N) * W);
  std::vector<char> act(E), alive(N);
  std::vector<int> slot_of(N, -1), level_of_slot;
  const uint32_t n_cfg = 1u << R;
  for (uint32_t t = 0; t < n_cfg; ++t) {
    std::fill(act.begin(), act.end(), 1);
    for (int r = 0; r < R; ++r) act[((t >> r) & 1u) ? net.retic_first_edge[r]
                                                    : net.retic_second_edge[r]] = 0;
    std::fill(key.begin(), key.end(), 0);
    for (int i = N - 1; i >= 0; --i) {
      const int v = topo[i];
      if (v < T) {
        alive[v] = 1;
        continue;
      }
      alive[v] = 0;
      uint64_t *kv = &key[static_cast<size_t>(v) * W];
      Variant var;
      var.node = v;
      int lvl = 0;
      for (int k = kstart[v]; k < kstart[v + 1]; ++k) {
        const int e = kedge[k], c = net.edge_child[e];
        if (!act[e] || !alive[c]) continue;
        alive[v] = 1;
        kv[e / 64] |= 1ull << (e % 64);
        const uint64_t *kc = &key[static_cast<size_t>(c) * W];
        for (int w = 0; w < W; ++w) kv[w] |= kc[w];
        var.kid[var.nkids] = c < T ? tip_code(c) : slot_of[c];
        var.edge[var.nkids] = e;
        if (c >= T) lvl = std::max(lvl, level_of_slot[slot_of[c]]);
        var.nkids++;
      }
      if (!alive[v]) {
        slot_of[v] = -1;
        continue;
      }
      std::string ks(reinterpret_cast<const char *>(kv), sizeof(uint64_t) * W);
      auto it = interned[v].find(ks);
      if (it == interned[v].end()) {
        var.level = lvl + 1;
        const int slot = static_cast<int>(dag.var.size());
        dag.var.push_back(var);
        level_of_slot.push_back(var.level);
        it = interned[v].emplace(std::move(ks), slot).first;
      }
      slot_of[v] = it->second;
    }
    // displayed-tree root: first node with two alive active children
    int u = net.root;
    for (;;) {
      int n_alive = 0, only = -1;
      for (int k = kstart[u]; k < kstart[u + 1]; ++k) {
        const int e = kedge[k], c = net.edge_child[e];
        if (act[e] && alive[c]) {
          ++n_alive;
          only = c;
        }
      }
      if (n_alive == 1 && only >= T)
        u = only;
      else
        break;
    }
    if (u < T) throw std::invalid_argument("displayed tree has a single tip");
    const uint64_t *kr = &key[static_cast<size_t>(net.root) * W];
    std::string ks(reinterpret_cast<const char *>(kr), sizeof(uint64_t) * W);
    auto cit = class_of_key.find(ks);
    if (cit == class_of_key.end()) {
      RootClass rc;
      rc.dtroot_slot = slot_of[u];
      rc.edges.assign(kr, kr + W);
      dag.classes.push_back(std::move(rc));
      cit = class_of_key.emplace(std::move(ks), static_cast<int>(dag.classes.size()) - 1).first;
    }
    RootClass &rc = dag.classes[cit->second];
    if (rc.dtroot_slot != slot_of[u])
      throw std::logic_error("root class with two displayed-tree roots");
    rc.configs.push_back(t);
  }
  return dag;
}

}  // namespace nrx
