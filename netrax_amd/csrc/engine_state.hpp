// Engine state shared by the host orchestration (engine.cpp) and the
// kernel launchers (kernels.hip).
#pragma once

#include <memory>
#include <vector>

#include "engine.hpp"

struct nrx_engine {};

namespace nrx {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void alloc(size_t n) {
    if (n <= bytes && p) return;
    release();
    if (n == 0) n = 16;
    NRX_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

struct PartInfo {
  int n_sites = 0;        // patterns
  int64_t site_off = 0;   // first padded site
  int tile_begin = 0, tile_end = 0;
  bool model_set = false;
  EigenModel model;
  std::vector<double> rates;  // [C]
  std::vector<double> subst;  // exchangeabilities as given
  double alpha = 0.0;         // Gamma shape of the rates; NaN after nrx_set_category_rates
};

// Result block offsets (doubles) inside the device/host result buffers.
struct ResultLayout {
  int total, per_part, root_logl, tree_logl, alpha, dbrlen, part_dbrlen, dprob, size;
};

struct Engine : nrx_engine {
  nrx_options opt{};
  int N = 0, E = 0, R = 0, T = 0, root = 0;
  std::vector<int32_t> edge_parent, edge_child, retic_first, retic_second;
  int P = 0, S = 0, C = 0;
  std::vector<PartInfo> parts;
  DevLayout L{};
  Dag dag;

  // host parameters
  std::vector<double> brlen;  // clamped
  std::vector<double> probs;  // first-parent probabilities
  std::vector<double> class_logprob, class_dlogprob;  // [nc], [nc][R]
  std::vector<char> class_use;

  // schedule (rebuilt when the set of used classes changes)
  std::vector<char> needed;                 // [n_var]
  std::vector<std::vector<int>> level_ops;  // variant slots per level (needed only)
  std::vector<int> root_slots;              // distinct displayed-tree-root CLVs
  std::vector<int> class_root_index;        // [nc]
  // plain form: ops in level order, cut into launches with bounded LDS tables
  std::vector<DevOp> ops_all;
  std::vector<LevelChunk> chunks;
  // plain form reverse sweep: runs of ops without a common child, per chunk
  std::vector<LevelChunk> adj_runs;
  std::vector<int> adj_run_begin;
  // batched form: groups of independent ops in a depth-first-biased order
  std::vector<DevOp> ops_dfs;
  std::vector<int32_t> group_begin;  // [n_groups + 1] into ops_dfs
  std::vector<int32_t> edge_use_begin, edge_use_list;  // gradient rows per edge
  int n_uses = 0;
  int pw = 1;  // site-sum partials per tile (one per wave)
  // gradient partials [site block][row][wave]: waves per block, blocks per tile
  int gpw = 1, gbpt = 1;
  // FP32 form shape (nrx_autotune / nrx_set_tuning): ops per group of the
  // persistent kernels and sites per workgroup (4-state data; 32 for 20)
  int tune_group = 4;
  int tune_tile = TILE;
  int tune_lanes = 1;  // lanes per (site, category) row of the 4-state reverse sweep
  int schedule_version = 0;  // bumped by every schedule rebuild
  int pass_slots = 0;        // LDS pass slots the FP32 4-state reverse sweep uses

  // branch derivatives: sumtables of one branch for every displayed tree
  // containing it (valid while nothing but that branch's length changes)
  bool deriv_valid = false;
  int deriv_edge = -1;
  std::vector<int> deriv_class;       // used classes whose alive edges contain the edge
  std::vector<int> deriv_path;        // per such class: sumtable row, or -1 (edge above the root)
  int deriv_n_paths = 0;
  std::vector<double> deriv_base_logl;  // [P][nc] tree lnLs when the sumtables were built
  DevBuf d_steps, d_path_begin, d_st, d_st_partials, d_st_sums;
  // device Newton-Raphson: state, class table, tree lnLs; its captured
  // sequence, valid while the sumtable buffers keep their addresses
  DevBuf d_nr_state, d_nr_cls;
  DevBuf d_mg_g, d_mg_r;  // model gradient partials
  // Memory budget of the FP64 form: with n_seg > 1 the CLV, adjoint and
  // scaler rows live in n_phys physical rows.  The level chunks are cut into
  // n_seg segments; a CLV consumed in a later segment than its own keeps a
  // fixed row (the boundary rows), the others share the rows of a pool that
  // each segment reuses.  The forward pass runs the segments in order; the
  // reverse sweep recomputes a segment's pooled CLVs before its adjoint.
  uint64_t mem_budget = 0;  // bytes of CLV + adjoint + scaler rows, 0 = all resident
  struct BudgetPlan {
    int n_seg = 0;  // 0 or 1: every CLV resident
    int n_phys = 0, n_boundary = 0;
    std::vector<int> seg_chunk;         // [n_seg + 1] chunk ranges of the segments
    std::vector<int> rec_begin;         // [n_chunks + 1] ranges of ops_rec per chunk
    std::vector<int32_t> root_phys;     // per segment, its roots' physical rows ...
    std::vector<int32_t> root_index;    // ... and their root indices
    std::vector<int> seg_root_begin;    // [n_seg + 1] ranges of the two above
  } bp;
  DevBuf d_ops_b, d_ops_rec, d_root_phys, d_root_index;
  hipGraphExec_t nr_graph = nullptr;
  int nr_graph_key[4] = {-1, -1, -1, -1};  // variant, mode, max_iters, paths
  const void *nr_graph_bufs[3] = {nullptr, nullptr, nullptr};
  NewtonState *h_nr_state = nullptr;  // pinned
  char *h_stage = nullptr;             // pinned staging for per-branch uploads
  size_t h_stage_bytes = 0, h_stage_off = 0;
  std::vector<double> h_st_sums;

  // dirty tracking
  bool model_dirty = true;
  std::vector<char> edge_dirty;  // P needs recompute
  std::vector<char> clv_dirty;   // [n_var]
  bool probs_dirty = true;
  bool cache_valid = false;
  int cache_variant = -1;
  double cached_total = 0.0;
  std::vector<double> cached_per_part;

  // device memory
  hipStream_t stream = nullptr;
  DevBuf d_tips, d_weights, d_tile_part, d_part_tile_begin;
  DevBuf d_V, d_Vinv, d_eval, d_rates, d_freqs, d_brlen;
  DevBuf d_P, d_dP, d_d2P, d_Pf, d_dPf;
  DevBuf d_clv, d_scl, d_adj, d_adj_exp, d_Qf;
  DevBuf d_Pa;  // 20-state FP32 form: matrix-core operand tables
  DevBuf d_ops, d_ops_dirty, d_Qd;
  DevBuf d_adj_walk;  // the level chunks top level first (the reverse sweep's walk)
  DevBuf d_fwd_walk;  // the level chunks lowest level first (the forward pass's walk)
  DevBuf d_dirty_walk;  // op ranges of d_ops_dirty per level (incremental forward)
  // FP32 form, incremental evaluation: the invalid ops, their groups and
  // their descriptor
  DevBuf d_ops_inc, d_group_begin_inc, d_netdesc_inc;
  DevBuf d_ops_dfs, d_group_begin, d_edge_use_begin, d_edge_use_list;
  DevBuf d_netdesc, d_work;  // FP32 form: this engine's NetDesc and (0, tile) work table
  DevBuf d_classes, d_class_dlogprob, d_root_slots;
  DevBuf d_root_partials, d_beta, d_grad_partials, d_root_sums, d_use_sums;
  DevBuf d_site_seed;  // plain form: adjoint seed factor w / (C Ltilde) [root][site]
  DevBuf d_result;
  double *h_result = nullptr;  // pinned
  double *h_brlen = nullptr;   // pinned staging of the branch lengths
  hipEvent_t ev_brlen = nullptr;  // the last copy out of h_brlen
  ResultLayout rl{};
  size_t clv_elems_per_slot = 0;  // n_sites_pad * C * S

  // FP32 form: instantiated graphs of one evaluation, [variant][gradient],
  // valid for schedule `graph_version`
  hipGraphExec_t graph_exec[2][2][2] = {};  // [variant][gradient][P matrices updated in the graph]
  int graph_version = -1;

  // timing
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  // event pairs around pruning (kind 0) and adjoint (kind 1) launches
  std::vector<hipEvent_t> ev_k[2];
  int n_k[2] = {0, 0};
  nrx_timing last_timing{};

  ~Engine();
};

}  // namespace nrx
