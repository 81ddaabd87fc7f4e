// Host orchestration and C ABI of the network-likelihood engine.
//
// The call sequence mirrors NetRAX's computeLoglikelihoodImproved
// (src/likelihood/ImprovedLoglikelihood.cpp:646-671): invalidate (when not
// incremental), update P matrices of invalid edges, recompute invalid CLVs
// level by level, evaluate every displayed-tree root, mix per partition,
// sum partitions, and fail on -infinity.  A result is reused when nothing
// changed since the last call (reuseOldDisplayedTreesCheck,
// src/graph/AnnotatedNetwork.cpp:359-394).
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <stdexcept>
#include <string>
#include <thread>

#include "engine_state.hpp"
#include "schedule.hpp"

namespace nrx {

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

// a call this engine's configuration does not provide (NRX_ERR_UNSUPPORTED)
struct unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <typename F>
int guarded(F &&f) {
  try {
    return f();
  } catch (const unsupported &ex) {
    return fail(NRX_ERR_UNSUPPORTED, ex.what());
  } catch (const std::invalid_argument &ex) {
    return fail(NRX_ERR_ARG, ex.what());
  } catch (const std::runtime_error &ex) {
    const std::string m = ex.what();
    return fail(m.rfind("HIP error", 0) == 0 ? NRX_ERR_DEVICE : NRX_ERR_STATE, m);
  } catch (const std::exception &ex) {
    return fail(NRX_ERR_STATE, ex.what());
  }
}

// Host-to-device copy of a (pageable, possibly short-lived) host vector;
// waits for completion so the source may be freed right after.
template <typename T>
void upload(DevBuf &b, const std::vector<T> &v, hipStream_t s) {
  b.alloc(std::max<size_t>(sizeof(T) * v.size(), 16));
  if (!v.empty()) {
    NRX_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
    NRX_HIP(hipStreamSynchronize(s));
  }
}

}  // namespace

static void destroy_graphs(hipGraphExec_t (&g)[2][2]) {
  for (auto &row : g)
    for (auto &x : row) {
      if (x) (void)hipGraphExecDestroy(x);
      x = nullptr;
    }
}

static void destroy_graphs(hipGraphExec_t (&g)[2][2][2]) {
  for (auto &plane : g) destroy_graphs(plane);
}

// Captures the work `enqueue` puts on `stream` into an instantiated graph.
template <typename F>
static hipGraphExec_t capture_graph(hipStream_t stream, F &&enqueue) {
  hipGraph_t g = nullptr;
  NRX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  try {
    enqueue();
  } catch (...) {
    (void)hipStreamEndCapture(stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  NRX_HIP(hipStreamEndCapture(stream, &g));
  hipGraphExec_t exec = nullptr;
  const hipError_t err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  NRX_HIP(err);
  return exec;
}

Engine::~Engine() {
  if (stream) (void)hipStreamSynchronize(stream);
  destroy_graphs(graph_exec);
  if (nr_graph) (void)hipGraphExecDestroy(nr_graph);
  if (h_nr_state) (void)hipHostFree(h_nr_state);
  if (h_stage) (void)hipHostFree(h_stage);
  if (h_result) (void)hipHostFree(h_result);
  if (h_brlen) (void)hipHostFree(h_brlen);
  if (ev_brlen) (void)hipEventDestroy(ev_brlen);
  if (ev_begin) (void)hipEventDestroy(ev_begin);
  if (ev_end) (void)hipEventDestroy(ev_end);
  for (auto &v : ev_k)
    for (auto ev : v) (void)hipEventDestroy(ev);
  if (stream) (void)hipStreamDestroy(stream);
}

// ------------------------------------------------------------ host logic

static void compute_class_probs(Engine &e) {
  const int nc = static_cast<int>(e.dag.classes.size()), R = e.R;
  e.class_logprob.assign(nc, 0.0);
  e.class_dlogprob.assign(static_cast<size_t>(std::max(nc, 1)) * std::max(R, 1), 0.0);
  e.class_use.assign(nc, 0);
  for (int k = 0; k < nc; ++k) {
    const auto &cfgs = e.dag.classes[k].configs;
    // log P of one configuration: sum over reticulations of log p or
    // log(1 - p) (ReticulationConfigSet.cpp:70-88); several member
    // configurations are summed in probability space (:98-113)
    std::vector<double> lps(cfgs.size());
    double mx = -INFINITY;
    for (size_t m = 0; m < cfgs.size(); ++m) {
      double lp = 0.0;
      for (int r = 0; r < R; ++r)
        lp += ((cfgs[m] >> r) & 1u) ? std::log(1.0 - e.probs[r]) : std::log(e.probs[r]);
      lps[m] = lp;
      mx = std::max(mx, lp);
    }
    double s = 0.0;
    std::vector<double> d(R, 0.0);
    for (size_t m = 0; m < cfgs.size(); ++m) {
      const double w = std::exp(lps[m] - mx);
      s += w;
      for (int r = 0; r < R; ++r)
        d[r] += w * (((cfgs[m] >> r) & 1u) ? -1.0 / (1.0 - e.probs[r]) : 1.0 / e.probs[r]);
    }
    e.class_logprob[k] = cfgs.size() == 1 ? lps[0] : mx + std::log(s);
    for (int r = 0; r < R; ++r) e.class_dlogprob[static_cast<size_t>(k) * R + r] = d[r] / s;
    e.class_use[k] = e.class_logprob[k] >= e.opt.min_tree_logprob;
  }
}

static void upload_netdesc(Engine &e);
static void stream_sync(Engine &e);

// Bytes of one CLV row with its adjoint and scaler rows (FP64 form).
static size_t budget_row_bytes(const Engine &e) {
  return sizeof(double) * e.clv_elems_per_slot * 2 + sizeof(int32_t) * static_cast<size_t>(e.L.n_sites_pad);
}

static void resize_rows(Engine &e, size_t rows) {
  rows = std::max<size_t>(rows, 1);
  e.d_clv.release();
  e.d_clv.alloc(sizeof(double) * e.clv_elems_per_slot * rows);
  e.d_scl.release();
  e.d_scl.alloc(sizeof(int32_t) * static_cast<size_t>(e.L.n_sites_pad) * rows);
  e.d_adj.release();
  e.d_adj.alloc(sizeof(double) * e.clv_elems_per_slot * rows);
  std::fill(e.clv_dirty.begin(), e.clv_dirty.end(), 1);
  e.cache_valid = false;
  e.deriv_valid = false;
}

// Segments of the level schedule whose rows fit the memory budget: the
// fewest segments (least recomputation) that fit, cut at level boundaries
// with about equal op counts.  The reverse sweep then recomputes every
// segment but the last one, at most one more forward pass.
static void plan_budget(Engine &e) {
  auto &bp = e.bp;
  const int nv = static_cast<int>(e.dag.var.size());
  const bool was_on = bp.n_seg > 1;
  const size_t row = budget_row_bytes(e);
  if (e.mem_budget == 0 || e.opt.precision != NRX_FP64 ||
      static_cast<size_t>(std::max(nv, 1)) * row <= e.mem_budget) {
    bp = {};
    if (was_on) resize_rows(e, nv);
    return;
  }
  const int nch = static_cast<int>(e.chunks.size());
  const int total = static_cast<int>(e.ops_all.size());
  std::vector<int> seg_of(nv, -1), last(nv, -1), cuts;
  size_t smallest = SIZE_MAX;
  bool found = false;
  for (int K = 2; K <= nch && !found; ++K) {
    cuts.assign(1, 0);
    int acc = 0;
    for (int c = 0; c < nch; ++c) {
      acc += e.chunks[c].op_end - e.chunks[c].op_begin;
      const int k = static_cast<int>(cuts.size());
      if (k < K && c + 1 < nch && static_cast<long>(acc) * K >= static_cast<long>(k) * total) cuts.push_back(c + 1);
    }
    cuts.push_back(nch);
    const int ns = static_cast<int>(cuts.size()) - 1;
    std::fill(seg_of.begin(), seg_of.end(), -1);
    std::fill(last.begin(), last.end(), -1);
    for (int j = 0; j < ns; ++j)
      for (int c = cuts[j]; c < cuts[j + 1]; ++c)
        for (int i = e.chunks[c].op_begin; i < e.chunks[c].op_end; ++i) {
          const DevOp &op = e.ops_all[i];
          seg_of[op.out] = j;
          for (int k = 0; k < op.nk; ++k)
            if (op.kid[k] >= 0) last[op.kid[k]] = std::max(last[op.kid[k]], j);
        }
    int nb = 0;
    std::vector<int> internal(ns, 0);
    for (int s = 0; s < nv; ++s) {
      if (seg_of[s] < 0) continue;
      if (last[s] > seg_of[s])
        ++nb;
      else
        ++internal[seg_of[s]];
    }
    const size_t phys = static_cast<size_t>(nb) + *std::max_element(internal.begin(), internal.end());
    smallest = std::min(smallest, phys * row);
    if (phys * row <= e.mem_budget) {
      found = true;
      bp = {};
      bp.n_seg = ns;
      bp.n_boundary = nb;
      bp.n_phys = static_cast<int>(phys);
      bp.seg_chunk = cuts;
    }
  }
  if (!found)
    throw std::invalid_argument("memory budget of " + std::to_string(e.mem_budget) +
                                " bytes is below the smallest segment plan (" +
                                std::to_string(std::min(smallest, static_cast<size_t>(std::max(nv, 1)) * row)) +
                                " bytes)");
  // physical rows: boundary CLVs first, then each segment's pool
  std::vector<int> phys(nv, -1), pool_next(bp.n_seg, bp.n_boundary);
  int nb = 0;
  for (int s = 0; s < nv; ++s)
    if (seg_of[s] >= 0 && last[s] > seg_of[s]) phys[s] = nb++;
  for (int s = 0; s < nv; ++s)
    if (seg_of[s] >= 0 && phys[s] < 0) phys[s] = pool_next[seg_of[s]]++;
  std::vector<DevOp> ops = e.ops_all, rec;
  bp.rec_begin.assign(1, 0);
  for (int c = 0; c < nch; ++c) {
    for (int i = e.chunks[c].op_begin; i < e.chunks[c].op_end; ++i) {
      DevOp &op = ops[i];
      const bool pooled = phys[op.out] >= bp.n_boundary;
      op.out = phys[op.out];
      for (int k = 0; k < op.nk; ++k)
        if (op.kid[k] >= 0) op.kid[k] = phys[op.kid[k]];
      if (pooled) rec.push_back(op);
    }
    bp.rec_begin.push_back(static_cast<int>(rec.size()));
  }
  bp.seg_root_begin.assign(1, 0);
  for (int j = 0; j < bp.n_seg; ++j) {
    for (size_t r = 0; r < e.root_slots.size(); ++r)
      if (seg_of[e.root_slots[r]] == j) {
        bp.root_phys.push_back(phys[e.root_slots[r]]);
        bp.root_index.push_back(static_cast<int32_t>(r));
      }
    bp.seg_root_begin.push_back(static_cast<int>(bp.root_phys.size()));
  }
  upload(e.d_ops_b, ops, e.stream);
  upload(e.d_ops_rec, rec, e.stream);
  upload(e.d_root_phys, bp.root_phys, e.stream);
  upload(e.d_root_index, bp.root_index, e.stream);
  resize_rows(e, bp.n_phys);
}

// Builds the level schedule, the batched order, root CLV set and gradient
// rows for the classes that pass the log-probability cut, and uploads them.
static void build_schedule(Engine &e) {
  const int nv = static_cast<int>(e.dag.var.size());
  const int nc = static_cast<int>(e.dag.classes.size());
  e.needed.assign(nv, 0);
  for (int k = 0; k < nc; ++k)
    if (e.class_use[k]) e.needed[e.dag.classes[k].dtroot_slot] = 1;
  for (int s = nv - 1; s >= 0; --s)  // parents have larger slot ids
    if (e.needed[s])
      for (int k = 0; k < e.dag.var[s].nkids; ++k)
        if (e.dag.var[s].kid[k] >= 0) e.needed[e.dag.var[s].kid[k]] = 1;
  int max_level = 0;
  for (int s = 0; s < nv; ++s)
    if (e.needed[s]) max_level = std::max(max_level, e.dag.var[s].level);
  e.level_ops.assign(max_level + 1, {});
  for (int s = 0; s < nv; ++s)
    if (e.needed[s]) e.level_ops[e.dag.var[s].level].push_back(s);

  // root CLVs
  e.root_slots.clear();
  e.class_root_index.assign(nc, -1);
  for (int k = 0; k < nc; ++k) {
    if (!e.class_use[k]) continue;
    const int s = e.dag.classes[k].dtroot_slot;
    auto it = std::find(e.root_slots.begin(), e.root_slots.end(), s);
    if (it == e.root_slots.end()) {
      e.root_slots.push_back(s);
      it = e.root_slots.end() - 1;
    }
    e.class_root_index[k] = static_cast<int>(it - e.root_slots.begin());
  }
  std::vector<int> root_of_slot(nv, -1);
  for (size_t r = 0; r < e.root_slots.size(); ++r) root_of_slot[e.root_slots[r]] = static_cast<int>(r);

  auto make_op = [&](int s) {
    const Variant &v = e.dag.var[s];
    DevOp op{};
    op.out = s;
    op.nk = v.nkids;
    for (int k = 0; k < 2; ++k) {
      op.kid[k] = k < v.nkids ? v.kid[k] : NO_CHILD;
      op.edge[k] = k < v.nkids ? v.edge[k] : 0;
      op.gslot[k] = 0;  // numbered below, once the adjoint order is known
    }
    op.root = root_of_slot[s];
    op.flags = 0;
    return op;
  };

  // plain form: level chunks (one per pruning level), ops in slot order;
  // the 4-state passes walk them in one launch, 20 states launch per level
  e.ops_all.clear();
  e.chunks.clear();
  for (int l = 1; l <= max_level; ++l) {
    const auto &lv = e.level_ops[l];
    if (lv.empty()) continue;
    LevelChunk ch{};
    ch.level = l;
    ch.op_begin = static_cast<int>(e.ops_all.size());
    for (int s : lv) e.ops_all.push_back(make_op(s));
    ch.op_end = static_cast<int>(e.ops_all.size());
    e.chunks.push_back(ch);
  }
  {
    // adjoint: chunks backwards, each chunk's runs in order, ops in index
    // order (the runs are contiguous ranges of the chunk)
    split_level_runs(e.ops_all, e.chunks, e.adj_runs, e.adj_run_begin);
    std::vector<int> order;
    for (auto it = e.chunks.rbegin(); it != e.chunks.rend(); ++it)
      for (int i = it->op_begin; i < it->op_end; ++i) order.push_back(i);
    set_adjoint_flags(e.ops_all, order, nv);
  }

  // batched form (FP32): groups of up to tune_group independent ops in
  // depth-first priority, then the LDS pass slots of the 4-state sweep
  {
    std::vector<DevOp> proto(nv);
    for (int v = 0; v < nv; ++v)
      if (e.needed[v]) proto[v] = make_op(v);
    build_batched_order(e.dag, e.root_slots, proto, e.tune_group, e.ops_dfs, e.group_begin);
    e.pass_slots = 0;
    if (e.S == 4 && e.opt.precision == NRX_FP32)
      e.pass_slots = adjoint_pass_slots(e.C, e.tune_group, e.tune_tile, e.tune_lanes, e.E,
                                        e.L.n_sites_pad / e.tune_tile);
    set_pass_slots(e.ops_dfs, e.group_begin, nv, e.pass_slots);
  }

  // gradient rows, one per (op, child), numbered in the order this engine's
  // adjoint walks the ops (plain form: levels top down; FP32 form: groups
  // backwards), so that a workgroup's consecutive ops write consecutive rows
  {
    std::vector<int> walk;  // indices into the form's op list
    std::vector<DevOp> &ops = e.opt.precision == NRX_FP32 ? e.ops_dfs : e.ops_all;
    if (e.opt.precision == NRX_FP32) {
      for (int g = static_cast<int>(e.group_begin.size()) - 2; g >= 0; --g)
        for (int i = e.group_begin[g]; i < e.group_begin[g + 1]; ++i) walk.push_back(i);
    } else {
      for (auto it = e.chunks.rbegin(); it != e.chunks.rend(); ++it)
        for (int i = it->op_begin; i < it->op_end; ++i) walk.push_back(i);
    }
    std::vector<std::array<int32_t, 2>> gslot(nv, {0, 0});
    std::vector<std::vector<int>> by_edge(e.E);
    e.n_uses = 0;
    for (int i : walk) {
      const DevOp &op = ops[i];
      for (int k = 0; k < op.nk; ++k) {
        gslot[op.out][k] = e.n_uses;
        by_edge[op.edge[k]].push_back(e.n_uses);
        e.n_uses++;
      }
    }
    for (std::vector<DevOp> *list : {&e.ops_all, &e.ops_dfs})
      for (DevOp &op : *list)
        for (int k = 0; k < op.nk; ++k) op.gslot[k] = gslot[op.out][k];
    e.edge_use_begin.assign(e.E + 1, 0);
    e.edge_use_list.clear();
    for (int ed = 0; ed < e.E; ++ed) {
      e.edge_use_begin[ed] = static_cast<int>(e.edge_use_list.size());
      for (int u : by_edge[ed]) e.edge_use_list.push_back(u);
    }
    e.edge_use_begin[e.E] = static_cast<int>(e.edge_use_list.size());
  }

  std::vector<DevClass> dc(nc);
  for (int k = 0; k < nc; ++k) {
    dc[k].logprob = e.class_logprob[k];
    dc[k].root_index = e.class_root_index[k];
    dc[k].use = e.class_use[k];
  }
  upload(e.d_ops, e.ops_all, e.stream);
  {
    std::vector<LevelChunk> walk(e.chunks.rbegin(), e.chunks.rend());
    upload(e.d_adj_walk, walk, e.stream);
    upload(e.d_fwd_walk, e.chunks, e.stream);
  }
  upload(e.d_ops_dfs, e.ops_dfs, e.stream);
  upload(e.d_group_begin, e.group_begin, e.stream);
  upload(e.d_edge_use_begin, e.edge_use_begin, e.stream);
  upload(e.d_edge_use_list, e.edge_use_list, e.stream);
  upload(e.d_classes, dc, e.stream);
  upload(e.d_class_dlogprob, e.class_dlogprob, e.stream);
  std::vector<int32_t> rs(e.root_slots.begin(), e.root_slots.end());
  upload(e.d_root_slots, rs, e.stream);
  // gradient partials per site block: plain form 64-site tiles of C waves;
  // FP32 form its reverse-sweep workgroups (tune_tile sites of tune_lanes
  // lanes per row, or 32 sites for 20-state data)
  if (e.opt.precision != NRX_FP32) {
    e.gpw = TILE * e.C / 64;
    e.gbpt = 1;
  } else if (e.S == AA_STATES) {
    e.gpw = e.C;
    e.gbpt = TILE / AA_TILE;
  } else {
    e.gpw = e.tune_tile * e.C * e.tune_lanes / 64;
    e.gbpt = TILE / e.tune_tile;
  }
  if (e.gpw * e.gbpt < e.pw) throw std::logic_error("gradient partial layout does not cover the tile");
  const size_t n_wt = static_cast<size_t>(e.L.n_tiles) * e.pw;
  const size_t n_gwt = static_cast<size_t>(e.L.n_tiles) * e.gpw * e.gbpt;
  e.d_root_partials.alloc(sizeof(double) * std::max<size_t>(1, rs.size()) * n_wt);
  e.d_beta.alloc(sizeof(double) * std::max<size_t>(1, rs.size()) * e.P);
  e.d_grad_partials.alloc(sizeof(double) * std::max(1, e.n_uses) * n_gwt);
  e.d_site_seed.alloc(sizeof(double) * std::max<size_t>(1, rs.size()) * e.L.n_sites_pad);
  e.d_root_sums.alloc(sizeof(double) * std::max<size_t>(1, rs.size()) * e.P);
  e.d_use_sums.alloc(sizeof(double) * std::max(1, e.n_uses) * static_cast<size_t>(e.P));
  stream_sync(e);
  upload_netdesc(e);
  plan_budget(e);
  e.schedule_version++;
}


// Device descriptor of this engine for the persistent kernels (FP32 form).
NetDesc make_netdesc(const Engine &e) {
  NetDesc d{};
  d.L = e.L;
  d.ops = e.d_ops_dfs.as<DevOp>();
  d.group_begin = e.d_group_begin.as<int32_t>();
  d.n_groups = static_cast<int32_t>(e.group_begin.size()) - 1;
  d.n_root = static_cast<int32_t>(e.root_slots.size());
  d.n_uses = e.n_uses;
  d.n_pass = e.pass_slots;
  d.clv = e.d_clv.as<float>();
  d.scl = e.d_scl.as<int32_t>();
  d.adj = e.d_adj.as<float>();
  d.adj_exp = e.d_adj_exp.as<int32_t>();
  d.tips = e.d_tips.as<uint8_t>();
  d.Pf = e.d_Pf.as<float>();
  d.Qf = e.d_Qf.as<float>();
  d.Pa = e.d_Pa.as<float>();
  d.tile_part = e.d_tile_part.as<int32_t>();
  d.freqs = e.d_freqs.as<double>();
  d.weights = e.d_weights.as<double>();
  d.beta = e.d_beta.as<double>();
  d.root_partials = e.d_root_partials.as<double>();
  d.grad_partials = e.d_grad_partials.as<double>();
  // every buffer the persistent kernels read must exist before its address
  // is handed to the device
  const bool aa = e.S == AA_STATES;
  const void *need[] = {d.ops, d.group_begin, d.clv, d.scl, d.adj, d.adj_exp, d.tips,
                        aa ? static_cast<const void *>(d.Pa) : d.Pf, d.Qf,
                        d.tile_part, d.freqs, d.weights, d.beta, d.root_partials, d.grad_partials};
  for (const void *p : need)
    if (!p) throw std::logic_error("device descriptor built before its buffers were allocated");
  return d;
}

static void upload_netdesc(Engine &e) {
  if (e.opt.precision != NRX_FP32) return;
  std::vector<NetDesc> nd{make_netdesc(e)};
  upload(e.d_netdesc, nd, e.stream);
  // work items: tiles of tune_tile sites, or of 32 sites for 20-state data
  const int n_work = static_cast<int>(e.L.n_sites_pad / (e.S == AA_STATES ? AA_TILE : e.tune_tile));
  std::vector<WorkItem> work(n_work);
  for (int t = 0; t < n_work; ++t) work[t] = WorkItem{0, t};
  upload(e.d_work, work, e.stream);
}

static void upload_classes_only(Engine &e) {
  const int nc = static_cast<int>(e.dag.classes.size());
  std::vector<DevClass> dc(nc);
  for (int k = 0; k < nc; ++k) {
    dc[k].logprob = e.class_logprob[k];
    dc[k].root_index = e.class_root_index[k];
    dc[k].use = e.class_use[k];
  }
  upload(e.d_classes, dc, e.stream);
  upload(e.d_class_dlogprob, e.class_dlogprob, e.stream);
  stream_sync(e);
}

static void set_probs(Engine &e, const double *p) {
  for (int r = 0; r < e.R; ++r) {
    if (!(p[r] > 0.0 && p[r] < 1.0))
      throw std::invalid_argument("reticulation probabilities must lie in (0, 1)");
    e.probs[r] = p[r];
  }
  const std::vector<char> old_use = e.class_use;
  compute_class_probs(e);
  if (old_use != e.class_use) {
    e.deriv_valid = false;
    build_schedule(e);
    std::fill(e.clv_dirty.begin(), e.clv_dirty.end(), 1);
  } else {
    upload_classes_only(e);
  }
  e.probs_dirty = true;
}

static void mark_clv_dirty_for_edges(Engine &e) {
  const int nv = static_cast<int>(e.dag.var.size());
  for (int s = 0; s < nv; ++s) {
    const Variant &v = e.dag.var[s];
    bool d = e.clv_dirty[s];
    for (int k = 0; k < v.nkids && !d; ++k) {
      if (e.edge_dirty[v.edge[k]]) d = true;
      if (v.kid[k] >= 0 && e.clv_dirty[v.kid[k]]) d = true;
    }
    e.clv_dirty[s] = d;
  }
}

static void check_ready(const Engine &e) {
  for (int p = 0; p < e.P; ++p)
    if (!e.parts[p].model_set) throw std::runtime_error("model of a partition is not set");
}

static void upload_models(Engine &e) {
  const int S = e.S, C = e.C, P = e.P;
  std::vector<double> V(P * S * S), Vi(P * S * S), ev(P * S), rates(P * C), fr(P * S);
  for (int p = 0; p < P; ++p) {
    const PartInfo &pi = e.parts[p];
    std::copy(pi.model.V.begin(), pi.model.V.end(), V.begin() + p * S * S);
    std::copy(pi.model.Vinv.begin(), pi.model.Vinv.end(), Vi.begin() + p * S * S);
    std::copy(pi.model.eval.begin(), pi.model.eval.end(), ev.begin() + p * S);
    std::copy(pi.rates.begin(), pi.rates.end(), rates.begin() + p * C);
    std::copy(pi.model.freqs.begin(), pi.model.freqs.end(), fr.begin() + p * S);
  }
  {
    // r_c Q per partition and category: dP/dt = Q P gives P' x = r_c Q (P x)
    std::vector<double> Qd(static_cast<size_t>(P) * C * S * S);
    std::vector<float> Qf(Qd.size());
    for (int p = 0; p < P; ++p) {
      const EigenModel &m = e.parts[p].model;
      for (int c = 0; c < C; ++c)
        for (int i = 0; i < S; ++i)
          for (int j = 0; j < S; ++j) {
            double q = 0.0;
            for (int k = 0; k < S; ++k) q += m.V[i * S + k] * m.eval[k] * m.Vinv[k * S + j];
            const size_t idx = ((static_cast<size_t>(p) * C + c) * S + i) * S + j;
            Qd[idx] = q * e.parts[p].rates[c];
            Qf[idx] = static_cast<float>(Qd[idx]);
          }
    }
    upload(e.d_Qd, Qd, e.stream);
    if (e.opt.precision == NRX_FP32) upload(e.d_Qf, Qf, e.stream);
  }
  upload(e.d_V, V, e.stream);
  upload(e.d_Vinv, Vi, e.stream);
  upload(e.d_eval, ev, e.stream);
  upload(e.d_rates, rates, e.stream);
  upload(e.d_freqs, fr, e.stream);
}

// P matrices for changed edges / models; marks dependent CLVs invalid.
// Branch lengths to the device through the engine's pinned staging buffer,
// asynchronously on stream s (no host round trip per evaluation).  The
// buffer is rewritten only once the previous copy out of it has finished.
static void upload_brlen(Engine &e, hipStream_t s) {
  NRX_HIP(hipEventSynchronize(e.ev_brlen));
  std::copy(e.brlen.begin(), e.brlen.end(), e.h_brlen);
  NRX_HIP(hipMemcpyAsync(e.d_brlen.p, e.h_brlen, sizeof(double) * e.E, hipMemcpyHostToDevice, s));
  NRX_HIP(hipEventRecord(e.ev_brlen, s));
}

// Host side of a P-matrix update: marks the CLVs the changed edges and
// models touch, and (do_update) stages the models and branch lengths on the
// stream.  Returns whether the P kernels (enqueue_pmatrices) must run.
static bool stage_pmatrices(Engine &e, bool do_update) {
  const bool any_edge = std::any_of(e.edge_dirty.begin(), e.edge_dirty.end(), [](char c) { return c; });
  if (!(e.model_dirty || any_edge)) return false;
  if (e.model_dirty) std::fill(e.clv_dirty.begin(), e.clv_dirty.end(), 1);
  mark_clv_dirty_for_edges(e);
  if (!do_update) return false;
  if (e.model_dirty) upload_models(e);
  upload_brlen(e, e.stream);
  e.model_dirty = false;
  std::fill(e.edge_dirty.begin(), e.edge_dirty.end(), 0);
  return true;
}

static void enqueue_pmatrices(Engine &e) {
  launch_pmatrices(e);
  if (e.opt.precision == NRX_FP32 && e.S == AA_STATES) launch_mfma_tables(e);
}

static void record_kernel_begin(Engine &e, int kind) {
  if (!e.opt.timing) return;
  auto &evs = e.ev_k[kind];
  if (static_cast<int>(evs.size()) < 2 * (e.n_k[kind] + 1)) {
    for (int i = 0; i < 2; ++i) {
      hipEvent_t ev;
      NRX_HIP(hipEventCreate(&ev));
      evs.push_back(ev);
    }
  }
  NRX_HIP(hipEventRecord(evs[2 * e.n_k[kind]], e.stream));
}

static void record_kernel_end(Engine &e, int kind) {
  if (!e.opt.timing) return;
  NRX_HIP(hipEventRecord(e.ev_k[kind][2 * e.n_k[kind] + 1], e.stream));
  e.n_k[kind]++;
}

// Waits for the engine's stream.  Every copy staged through the pinned
// staging memory has then been read, so the staging starts over.
static void stream_sync(Engine &e) {
  NRX_HIP(hipStreamSynchronize(e.stream));
  e.h_stage_off = 0;
}

// Copies v to device buffer b through the engine's pinned staging memory,
// after the copies staged since the last stream_sync, without waiting.
template <typename T>
static void stage_upload(Engine &e, DevBuf &b, const std::vector<T> &v) {
  size_t &off = e.h_stage_off;
  const size_t bytes = sizeof(T) * v.size();
  b.alloc(std::max<size_t>(bytes, 16));
  if (bytes == 0) return;
  if (off + bytes > e.h_stage_bytes) {
    const size_t cap = std::max<size_t>(2 * (off + bytes), 1 << 16);
    stream_sync(e);  // earlier copies may still read the old staging
    if (e.h_stage) NRX_HIP(hipHostFree(e.h_stage));
    e.h_stage = nullptr;
    e.h_stage_bytes = 0;
    NRX_HIP(hipHostMalloc(reinterpret_cast<void **>(&e.h_stage), cap, hipHostMallocDefault));
    e.h_stage_bytes = cap;
  }
  std::memcpy(e.h_stage + off, v.data(), bytes);
  NRX_HIP(hipMemcpyAsync(b.p, e.h_stage + off, bytes, hipMemcpyHostToDevice, e.stream));
  off += (bytes + 255) & ~size_t(255);
}

static void forward(Engine &e, bool incremental) {
  if (!incremental) std::fill(e.clv_dirty.begin(), e.clv_dirty.end(), 1);
  bool all = true;
  for (const DevOp &op : e.ops_all)
    if (!e.clv_dirty[op.out]) {
      all = false;
      break;
    }
  // 4 states: one launch walking the levels (a workgroup per tile); 20
  // states: one launch per level (one op per workgroup, so a level fills
  // the GPU)
  const bool walk = e.S != AA_STATES;
  if (all) {
    if (walk) {
      record_kernel_begin(e, 0);
      launch_forward_walk(e, e.d_ops.as<DevOp>(), e.d_fwd_walk.as<LevelChunk>(), static_cast<int>(e.chunks.size()));
      record_kernel_end(e, 0);
    } else {
      for (const LevelChunk &ch : e.chunks) {
        record_kernel_begin(e, 0);
        launch_forward_level(e, e.d_ops.as<DevOp>() + ch.op_begin, ch.op_end - ch.op_begin, ch);
        record_kernel_end(e, 0);
      }
    }
  } else {
    // only the invalid CLVs, level by level
    std::vector<DevOp> dirty;
    std::vector<LevelChunk> lv;
    for (const LevelChunk &ch : e.chunks) {
      LevelChunk d = ch;
      d.op_begin = static_cast<int>(dirty.size());
      for (int i = ch.op_begin; i < ch.op_end; ++i)
        if (e.clv_dirty[e.ops_all[i].out]) dirty.push_back(e.ops_all[i]);
      d.op_end = static_cast<int>(dirty.size());
      if (d.op_end > d.op_begin) lv.push_back(d);
    }
    if (!dirty.empty()) {
      stage_upload(e, e.d_ops_dirty, dirty);
      if (walk) {
        stage_upload(e, e.d_dirty_walk, lv);
        record_kernel_begin(e, 0);
        launch_forward_walk(e, e.d_ops_dirty.as<DevOp>(), e.d_dirty_walk.as<LevelChunk>(), static_cast<int>(lv.size()));
        record_kernel_end(e, 0);
      } else {
        for (const LevelChunk &d : lv) {
          record_kernel_begin(e, 0);
          launch_forward_level(e, e.d_ops_dirty.as<DevOp>() + d.op_begin, d.op_end - d.op_begin, d);
          record_kernel_end(e, 0);
        }
      }
    }
  }
  for (const DevOp &op : e.ops_all) e.clv_dirty[op.out] = 0;
}

static void begin_timing(Engine &e) {
  e.n_k[0] = e.n_k[1] = 0;
  if (e.opt.timing) NRX_HIP(hipEventRecord(e.ev_begin, e.stream));
}

static void end_timing(Engine &e) {
  if (!e.opt.timing) return;
  NRX_HIP(hipEventRecord(e.ev_end, e.stream));
  NRX_HIP(hipEventSynchronize(e.ev_end));
  float ms = 0.f;
  NRX_HIP(hipEventElapsedTime(&ms, e.ev_begin, e.ev_end));
  nrx_timing t{};
  t.total_ms = ms;
  double acc[2] = {0.0, 0.0};
  for (int k = 0; k < 2; ++k)
    for (int i = 0; i < e.n_k[k]; ++i) {
      NRX_HIP(hipEventElapsedTime(&ms, e.ev_k[k][2 * i], e.ev_k[k][2 * i + 1]));
      acc[k] += ms;
    }
  t.fwd_ms = acc[0];
  t.fwd_launches = e.n_k[0];
  t.adj_ms = acc[1];
  t.adj_launches = e.n_k[1];
  e.last_timing = t;
}

static void fetch_results(Engine &e) {
  NRX_HIP(hipMemcpyAsync(e.h_result, e.d_result.p, sizeof(double) * e.rl.size,
                         hipMemcpyDeviceToHost, e.stream));
  stream_sync(e);
}

// Reverse sweep of level chunks c_hi down to c_lo (plain form).  20
// states: one launch per run of ops without a common child, one op per
// workgroup, so that a level fills the GPU.  4 states: one launch for all
// of them, a workgroup per tile walking the levels and their ops in order
// (the order the first-delivery flags were set in); per-level launches
// cost a fixed time each against little work per op.
static void adjoint_chunks(Engine &e, int c_hi, int c_lo, const DevOp *ops) {
  if (e.S == AA_STATES) {
    for (int c = c_hi; c >= c_lo; --c)
      for (int r = e.adj_run_begin[c]; r < e.adj_run_begin[c + 1]; ++r) {
        record_kernel_begin(e, 1);
        launch_adjoint_level(e, e.adj_runs[r], ops);
        record_kernel_end(e, 1);
      }
  } else if (c_hi >= c_lo) {
    record_kernel_begin(e, 1);
    launch_adjoint_levels(e, ops, c_hi, c_lo);
    record_kernel_end(e, 1);
  }
}

// Forward pass, root sums and (with `gradient`) reverse sweep of the FP64
// form under a memory budget; the mixture runs between the two.
static void budget_forward(Engine &e, bool gradient) {
  const auto &bp = e.bp;
  const DevOp *ops = e.d_ops_b.as<DevOp>();
  for (int j = 0; j < bp.n_seg; ++j) {
    for (int c = bp.seg_chunk[j]; c < bp.seg_chunk[j + 1]; ++c) {
      const LevelChunk &ch = e.chunks[c];
      record_kernel_begin(e, 0);
      launch_forward_level(e, ops + ch.op_begin, ch.op_end - ch.op_begin, ch);
      record_kernel_end(e, 0);
    }
    const int r0 = bp.seg_root_begin[j], nr = bp.seg_root_begin[j + 1] - r0;
    launch_root_partials(e, gradient, e.d_root_phys.as<int32_t>() + r0, e.d_root_index.as<int32_t>() + r0, nr);
  }
}

static void budget_adjoint(Engine &e) {
  const auto &bp = e.bp;
  const DevOp *ops = e.d_ops_b.as<DevOp>(), *rec = e.d_ops_rec.as<DevOp>();
  for (int j = bp.n_seg - 1; j >= 0; --j) {
    if (j < bp.n_seg - 1)  // the last segment's pool still holds its CLVs
      for (int c = bp.seg_chunk[j]; c < bp.seg_chunk[j + 1]; ++c) {
        const int n = bp.rec_begin[c + 1] - bp.rec_begin[c];
        if (n == 0) continue;
        record_kernel_begin(e, 0);
        launch_forward_level(e, rec + bp.rec_begin[c], n, e.chunks[c]);
        record_kernel_end(e, 0);
      }
    adjoint_chunks(e, bp.seg_chunk[j + 1] - 1, bp.seg_chunk[j], ops);
  }
}

// FP32 form, incremental evaluation: the ops whose CLVs are invalid, in the
// forward order and groups of the full list (a subset of a group of
// independent ops is independent), as a device descriptor of their own.
// A group is flagged independent of the one before it when it reads none of
// that group's outputs, as the kernels' early-load rule requires.  Returns
// the number of groups (0: nothing to recompute), or -1 when every op is
// invalid (the full list, and its graph, is used instead); `desc` points to
// the staged descriptor.
static int stage_invalid_ops(Engine &e, const NetDesc *&desc) {
  const int ng = static_cast<int>(e.group_begin.size()) - 1;
  std::vector<DevOp> ops;
  std::vector<int32_t> gb(1, 0);
  std::vector<int32_t> prev, cur;
  for (int g = 0; g < ng; ++g) {
    cur.clear();
    const size_t first = ops.size();
    for (int i = e.group_begin[g]; i < e.group_begin[g + 1]; ++i) {
      const DevOp &op = e.ops_dfs[i];
      if (!e.clv_dirty[op.out]) continue;
      ops.push_back(op);
      cur.push_back(op.out);
    }
    if (cur.empty()) continue;
    bool indep = !prev.empty();
    for (size_t i = first; i < ops.size() && indep; ++i)
      for (int k = 0; k < ops[i].nk; ++k)
        if (std::find(prev.begin(), prev.end(), ops[i].kid[k]) != prev.end()) indep = false;
    ops[first].flags = indep ? (ops[first].flags | OP_GROUP_INDEP) : (ops[first].flags & ~OP_GROUP_INDEP);
    gb.push_back(static_cast<int32_t>(ops.size()));
    prev.swap(cur);
  }
  if (ops.size() == e.ops_dfs.size()) return -1;
  const int n = static_cast<int>(gb.size()) - 1;
  if (n == 0) return 0;
  stage_upload(e, e.d_ops_inc, ops);
  stage_upload(e, e.d_group_begin_inc, gb);
  NetDesc d = make_netdesc(e);
  d.ops = e.d_ops_inc.as<DevOp>();
  d.group_begin = e.d_group_begin_inc.as<int32_t>();
  d.n_groups = n;
  stage_upload(e, e.d_netdesc_inc, std::vector<NetDesc>{d});
  desc = e.d_netdesc_inc.as<NetDesc>();
  return n;
}

// `fetch` false: the results stay on the device (for a caller that reads
// them there, such as the device Newton sweep) and no host wait is made;
// the host copy is then stale, so the cache is marked invalid.
static int evaluate(Engine &e, int variant, bool incremental, bool update_pm, bool gradient, bool fetch = true) {
  check_ready(e);
  if (variant != NRX_AVERAGE_DISPLAYED_TREES && variant != NRX_BEST_DISPLAYED_TREE)
    throw std::invalid_argument("unknown likelihood variant");
  begin_timing(e);
  const bool pm = stage_pmatrices(e, update_pm);
  // an FP32 evaluation replayed from a graph runs the P kernels inside it
  const bool pm_in_graph = pm && e.opt.precision == NRX_FP32 && e.opt.use_graph;
  if (pm && !pm_in_graph) enqueue_pmatrices(e);
  const bool any_dirty = !incremental || std::any_of(e.ops_all.begin(), e.ops_all.end(),
                                                     [&](const DevOp &o) { return e.clv_dirty[o.out]; });
  if (!gradient && !any_dirty && !e.probs_dirty && e.cache_valid && e.cache_variant == variant) {
    if (pm_in_graph) enqueue_pmatrices(e);  // staged above: the P matrices must follow
    end_timing(e);
    return NRX_OK;  // nothing changed since the last evaluation
  }
  bool results_copied = false;
  if (e.opt.precision == NRX_FP32) {
    // batched form: the op list in one persistent launch, root site sums
    // fused into it.  A full evaluation recomputes every CLV (one captured
    // graph with use_graph); an incremental one with some CLVs still valid
    // runs the invalid ops only, through a descriptor of their own, eagerly
    const bool aa32 = e.S == AA_STATES;
    const NetDesc *inc = nullptr;
    int n_inc_groups = -1;
    if (incremental) n_inc_groups = stage_invalid_ops(e, inc);
    const bool partial = n_inc_groups >= 0;
    auto enqueue = [&e, aa32, variant, gradient, pm_in_graph, inc, n_inc_groups](bool timed) {
      if (pm_in_graph) enqueue_pmatrices(e);
      if (timed) record_kernel_begin(e, 0);
      if (n_inc_groups == 0) {
        // no CLV to recompute (a probability change, or a repeated call)
      } else if (aa32) {
        launch_forward_mfma(e, inc);
      } else {
        launch_forward_persistent(e, inc);
      }
      if (timed) record_kernel_end(e, 0);
      launch_mixture(e, variant, e.pw);
      if (gradient) {
        if (timed) record_kernel_begin(e, 1);
        if (aa32)
          launch_adjoint_mfma(e);
        else
          launch_adjoint_persistent(e);
        if (timed) record_kernel_end(e, 1);
        launch_grad_reduce(e);
      }
    };
    if (e.opt.use_graph && !partial) {
      // one graph launch per evaluation; re-captured after a schedule change
      if (e.graph_version != e.schedule_version) {
        destroy_graphs(e.graph_exec);
        e.graph_version = e.schedule_version;
      }
      hipGraphExec_t &ge = e.graph_exec[variant][gradient ? 1 : 0][pm_in_graph ? 1 : 0];
      if (!ge)
        ge = capture_graph(e.stream, [&]() {
          enqueue(false);
          // the result block comes back inside the graph (a copy node)
          NRX_HIP(hipMemcpyAsync(e.h_result, e.d_result.p, sizeof(double) * e.rl.size, hipMemcpyDeviceToHost,
                                 e.stream));
        });
      NRX_HIP(hipGraphLaunch(ge, e.stream));
      results_copied = true;
    } else {
      enqueue(true);
    }
    std::fill(e.clv_dirty.begin(), e.clv_dirty.end(), 0);
  } else if (e.bp.n_seg > 1) {
    // memory budget: segments forward, then the reverse sweep recomputing
    budget_forward(e, gradient);
    launch_mixture(e, variant, e.pw);
    if (gradient) {
      budget_adjoint(e);
      launch_grad_reduce(e);
    }
    std::fill(e.clv_dirty.begin(), e.clv_dirty.end(), 0);
  } else {
    forward(e, incremental);
    launch_root_partials(e, gradient);
    launch_mixture(e, variant, e.pw);
    if (gradient) {
      e.d_adj.alloc(sizeof(double) * e.clv_elems_per_slot * e.dag.var.size());
      adjoint_chunks(e, static_cast<int>(e.chunks.size()) - 1, 0, e.d_ops.as<DevOp>());
      launch_grad_reduce(e);
    }
  }
  end_timing(e);
  if (!fetch && !results_copied) {
    e.probs_dirty = false;
    e.cache_valid = false;
    return NRX_OK;
  }
  if (results_copied)
    stream_sync(e);
  else
    fetch_results(e);
  e.probs_dirty = false;
  e.cache_valid = true;
  e.cache_variant = variant;
  e.cached_total = e.h_result[e.rl.total];
  e.cached_per_part.assign(e.h_result + e.rl.per_part, e.h_result + e.rl.per_part + e.P);
  if (!(e.cached_total > -INFINITY) || std::isnan(e.cached_total)) {
    e.cache_valid = false;
    return fail(NRX_ERR_NEGINF, "Invalid network likelihood: negative infinity");
  }
  return NRX_OK;
}


// Root-to-branch paths of branch `edge` for every used displayed tree that
// contains it (alive edge set, as isActiveBranch + the target variant
// existing, LikelihoodDerivatives.cpp:313-343).  cls: those classes;
// path: per class its path index, or -1 when the tree's root lies below the
// branch (the branch is on the dead path above the displayed-tree root and
// contributes derivative 0); steps / begin: the paths' steps.
static void make_deriv_paths(const Engine &e, int edge, std::vector<int> &cls, std::vector<int> &path,
                             std::vector<PathStep> &steps, std::vector<int32_t> &begin) {
  const int nc = static_cast<int>(e.dag.classes.size());
  cls.clear();
  path.clear();
  steps.clear();
  begin.assign(1, 0);
  for (int k = 0; k < nc; ++k) {
    if (!e.class_use[k]) continue;
    const auto &ed = e.dag.classes[k].edges;
    if (!((ed[edge / 64] >> (edge % 64)) & 1ull)) continue;
    cls.push_back(k);
    // depth-first search below the displayed-tree root for the op holding `edge`
    std::vector<PathStep> p;
    std::function<bool(int)> find = [&](int slot) -> bool {
      const Variant &v = e.dag.var[slot];
      for (int c = 0; c < v.nkids; ++c) {
        PathStep ps{};
        ps.out = slot;
        ps.child = v.kid[c];
        ps.edge_child = v.edge[c];
        ps.sib = v.nkids == 2 ? v.kid[1 - c] : NO_CHILD;
        ps.edge_sib = v.nkids == 2 ? v.edge[1 - c] : 0;
        ps.nk = v.nkids;
        p.push_back(ps);
        if (v.edge[c] == edge) return true;
        if (v.kid[c] >= 0 && find(v.kid[c])) return true;
        p.pop_back();
      }
      return false;
    };
    if (find(e.dag.classes[k].dtroot_slot)) {
      // the kernels take each step's own scaler from the previous step's
      // child: the path must run down through the children it names
      for (size_t i = 1; i < p.size(); ++i)
        if (p[i].out != p[i - 1].child) throw std::logic_error("derivative path is not a chain of ops");
      path.push_back(static_cast<int>(begin.size()) - 1);
      steps.insert(steps.end(), p.begin(), p.end());
      begin.push_back(static_cast<int32_t>(steps.size()));
    } else {
      path.push_back(-1);
    }
  }
}

static int max_path_steps(const std::vector<int32_t> &begin) {
  int m = 0;
  for (size_t i = 1; i < begin.size(); ++i) m = std::max(m, begin[i] - begin[i - 1]);
  return m;
}

// Branch derivatives and model gradients read the CLVs of every op after
// the evaluation; under a memory budget only some of them survive it.
static void require_resident(const Engine &e) {
  if (e.bp.n_seg > 1) throw unsupported("this call needs every CLV resident; the memory budget is active");
}

// Sumtables of branch `edge` (the device-side computePartitionSumtables).
static void build_sumtables(Engine &e, int edge) {
  require_resident(e);
  const int nc = static_cast<int>(e.dag.classes.size());
  std::vector<PathStep> steps;
  std::vector<int32_t> begin;
  make_deriv_paths(e, edge, e.deriv_class, e.deriv_path, steps, begin);
  const int np = static_cast<int>(begin.size()) - 1;
  e.deriv_n_paths = np;
  if (np > 0) {
    stage_upload(e, e.d_steps, steps);
    stage_upload(e, e.d_path_begin, begin);
    e.d_st.alloc(sizeof(double) * e.clv_elems_per_slot * np);
    e.d_st_partials.alloc(sizeof(double) * np * 3 * static_cast<size_t>(e.L.n_tiles) * e.pw);
    e.d_st_sums.alloc(sizeof(double) * np * 3 * e.P);
    e.h_st_sums.assign(static_cast<size_t>(np) * 3 * e.P, 0.0);
    launch_sumtables(e, np, max_path_steps(begin));
  }
  e.deriv_base_logl.assign(e.h_result + e.rl.tree_logl, e.h_result + e.rl.tree_logl + e.P * nc);
  e.deriv_edge = edge;
  e.deriv_valid = true;
}

// Per partition: each tree's lnL at the evaluated length and its first two
// derivatives from the path sums (sums[path][3][P]: log rho, rho'/rho,
// rho''/rho - (rho'/rho)^2), then the mixture of computePartitionLhData
// (LikelihoodDerivatives.cpp:30-188) or the true derivative.
static void mix_brlen_derivatives(const Engine &e, int variant, int mode, const std::vector<double> &base_logl,
                                  const std::vector<int> &cls, const std::vector<int> &path,
                                  const double *sums, double *d1, double *d2, double *part_d1,
                                  double *part_d2) {
  const int nc = static_cast<int>(e.dag.classes.size());
  double D1 = 0.0, D2 = 0.0;
  std::vector<double> ll(nc), l1(nc), l2(nc);
  std::vector<char> has(nc, 0);
  for (int k : cls) has[k] = 1;
  for (int p = 0; p < e.P; ++p) {
    for (int k = 0; k < nc; ++k) {
      ll[k] = base_logl[static_cast<size_t>(p) * nc + k];
      l1[k] = l2[k] = 0.0;
    }
    for (size_t i = 0; i < cls.size(); ++i) {
      const int k = cls[i], j = path[i];
      if (j < 0) continue;
      const double *v = sums + static_cast<size_t>(j) * 3 * e.P;
      ll[k] += v[0 * e.P + p];
      l1[k] = v[1 * e.P + p];
      l2[k] = v[2 * e.P + p];
    }
    double pd1 = 0.0, pd2 = 0.0;
    std::vector<int> members;
    for (int k = 0; k < nc; ++k)
      if (e.class_use[k] && (mode == NRX_DERIV_TRUE || has[k])) members.push_back(k);
    if (members.size() == 1 && mode == NRX_DERIV_REFERENCE) {
      pd1 = l1[members[0]];  // single-tree mode (LikelihoodDerivatives.cpp:145-151)
      pd2 = l2[members[0]];
    } else if (!members.empty()) {
      if (variant == NRX_AVERAGE_DISPLAYED_TREES) {
        double mx = -INFINITY;
        for (int k : members) mx = std::max(mx, e.class_logprob[k] + ll[k]);
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int k : members) {
          const double wgt = std::exp(e.class_logprob[k] + ll[k] - mx);
          s0 += wgt;
          s1 += wgt * l1[k];
          s2 += wgt * (l1[k] * l1[k] + l2[k]);
        }
        pd1 = s1 / s0;
        pd2 = (s2 * s0 - s1 * s1) / (s0 * s0);
      } else if (mode == NRX_DERIV_REFERENCE) {
        // the reference picks the tree maximising tree_logl * tree_prob
        // (LikelihoodDerivatives.cpp:161-167)
        double best = -INFINITY;
        for (int k : members) {
          const double score = ll[k] * std::exp(e.class_logprob[k]);
          if (score > best) {
            best = score;
            pd1 = l1[k];
            pd2 = l2[k];
          }
        }
      } else {
        double best = -INFINITY;
        for (int k : members) {
          const double score = e.class_logprob[k] + ll[k];
          if (score > best) {
            best = score;
            pd1 = l1[k];
            pd2 = l2[k];
          }
        }
      }
    }
    if (part_d1) part_d1[p] = pd1;
    if (part_d2) part_d2[p] = pd2;
    D1 += pd1;
    D2 += pd2;
  }
  *d1 = D1;
  *d2 = D2;
}


// ------------------------------------------------------------ batches

struct Batch {
  std::vector<Engine *> engines;
  std::vector<int> versions;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  DevBuf d_nets, d_work, d_small, d_offsets, d_out;
  double *h_out = nullptr;
  std::vector<int64_t> offsets;
  int n_work = 0, C = 0, max_pm = 0, max_root_rows = 0, max_use_rows = 0, max_edges = 0,
      max_res = 0;
  size_t lds_fwd = 0, lds_adj = 0;
  bool use_graph = false;
  int G = 4, TS = TILE, H = 1;  // shape of the persistent kernels, shared by all engines
  hipGraphExec_t graph_exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [variant][gradient]
  ~Batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    destroy_graphs(graph_exec);
    if (h_out) (void)hipHostFree(h_out);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

static SmallDesc make_smalldesc(const Engine &e) {
  SmallDesc d{};
  d.L = e.L;
  d.V = e.d_V.as<double>();
  d.Vinv = e.d_Vinv.as<double>();
  d.ev = e.d_eval.as<double>();
  d.rates = e.d_rates.as<double>();
  d.brlen = e.d_brlen.as<double>();
  d.P = e.d_P.as<double>();
  d.dP = e.d_dP.as<double>();
  d.d2P = e.d_d2P.as<double>();
  d.Pf = e.d_Pf.as<float>();
  d.n_root = static_cast<int32_t>(e.root_slots.size());
  d.nc = static_cast<int32_t>(e.dag.classes.size());
  d.R = e.R;
  d.n_uses = e.n_uses;
  d.part_tile_begin = e.d_part_tile_begin.as<int32_t>();
  d.classes = e.d_classes.as<DevClass>();
  d.class_dlogprob = e.d_class_dlogprob.as<double>();
  d.root_partials = e.d_root_partials.as<double>();
  d.root_sums = e.d_root_sums.as<double>();
  d.grad_partials = e.d_grad_partials.as<double>();
  d.use_sums = e.d_use_sums.as<double>();
  d.beta = e.d_beta.as<double>();
  d.edge_use_begin = e.d_edge_use_begin.as<int32_t>();
  d.edge_use_list = e.d_edge_use_list.as<int32_t>();
  d.res = e.d_result.as<double>();
  d.rl_total = e.rl.total;
  d.rl_per_part = e.rl.per_part;
  d.rl_root_logl = e.rl.root_logl;
  d.rl_tree_logl = e.rl.tree_logl;
  d.rl_alpha = e.rl.alpha;
  d.rl_dbrlen = e.rl.dbrlen;
  d.rl_part_dbrlen = e.rl.part_dbrlen;
  d.rl_dprob = e.rl.dprob;
  d.pw = e.pw;
  d.gpw = e.gpw;
  d.gbpt = e.gbpt;
  const void *need[] = {d.V, d.Vinv, d.ev, d.rates, d.brlen, d.P, d.dP, d.d2P, d.Pf,
                        d.part_tile_begin, d.classes, d.class_dlogprob, d.root_partials,
                        d.root_sums, d.grad_partials, d.use_sums, d.beta, d.edge_use_begin,
                        d.edge_use_list, d.res};
  for (const void *p : need)
    if (!p) throw std::logic_error("device descriptor built before its buffers were allocated");
  return d;
}

// Descriptors and the (network, tile) work table; tiles of networks with
// longer op lists come first so the longest workgroups start earliest.
static void build_batch(Batch &b) {
  const int n = static_cast<int>(b.engines.size());
  destroy_graphs(b.graph_exec);  // they hold the old tables' addresses
  std::vector<NetDesc> nets(n);
  std::vector<SmallDesc> small(n);
  std::vector<std::pair<int, int>> order;  // (cost, engine)
  b.offsets.assign(n + 1, 0);
  b.max_pm = b.max_root_rows = b.max_use_rows = b.max_edges = b.max_res = 0;
  b.lds_fwd = b.lds_adj = 0;
  b.versions.resize(n);
  b.G = b.engines[0]->tune_group;
  b.TS = b.engines[0]->tune_tile;
  b.H = b.engines[0]->tune_lanes;
  // One batch launch holds far more workgroups than one engine's, so each
  // engine keeps only as many LDS pass slots as leave the launch's
  // workgroups per CU unchanged (a slot given up moves its deliveries back
  // to HBM); its own graphs then re-capture.
  int64_t total_work = 0;
  for (const Engine *e : b.engines) total_work += e->L.n_sites_pad / b.TS;
  for (Engine *e : b.engines) {
    if (e->S != 4 || e->tune_group != b.G || e->tune_tile != b.TS || e->tune_lanes != b.H) continue;
    const int k = adjoint_pass_slots(e->C, b.G, b.TS, b.H, e->E, total_work);
    if (k >= e->pass_slots) continue;
    e->pass_slots = k;
    set_pass_slots(e->ops_dfs, e->group_begin, static_cast<int>(e->dag.var.size()), k);
    upload(e->d_ops_dfs, e->ops_dfs, e->stream);
    upload_netdesc(*e);
    NRX_HIP(hipStreamSynchronize(e->stream));
    e->schedule_version++;
  }
  for (int i = 0; i < n; ++i) {
    const Engine &e = *b.engines[i];
    if (e.tune_group != b.G || e.tune_tile != b.TS || e.tune_lanes != b.H)
      throw std::invalid_argument("batch engines must share the FP32 kernel shape (nrx_set_tuning)");
    nets[i] = make_netdesc(e);
    small[i] = make_smalldesc(e);
    b.versions[i] = e.schedule_version;
    order.push_back({static_cast<int>(e.ops_dfs.size()), i});
    b.offsets[i + 1] = b.offsets[i] + e.rl.size;
    b.max_pm = std::max(b.max_pm, e.P * e.E * e.C * e.S * e.S);
    b.max_root_rows = std::max(b.max_root_rows, static_cast<int>(e.root_slots.size()) * e.P);
    b.max_use_rows = std::max(b.max_use_rows, e.n_uses * e.P);
    b.max_edges = std::max(b.max_edges, e.E);
    b.max_res = std::max(b.max_res, e.rl.size);
    b.lds_fwd = std::max(b.lds_fwd, sizeof(float) * e.E * e.C * 16);
    b.lds_adj = std::max(b.lds_adj, adjoint_persistent_lds(e.E, e.C, b.TS, e.pass_slots));
  }
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
                     return x.first > y.first;
                   });
  std::vector<WorkItem> work;
  for (const auto &o : order)
    for (int t = 0; t < b.engines[o.second]->L.n_sites_pad / b.TS; ++t) work.push_back({o.second, t});
  b.n_work = static_cast<int>(work.size());
  upload(b.d_nets, nets, b.stream);
  upload(b.d_small, small, b.stream);
  upload(b.d_work, work, b.stream);
  upload(b.d_offsets, b.offsets, b.stream);
  b.d_out.alloc(sizeof(double) * b.offsets[n]);
  if (b.h_out) NRX_HIP(hipHostFree(b.h_out));
  NRX_HIP(hipHostMalloc(reinterpret_cast<void **>(&b.h_out), sizeof(double) * b.offsets[n],
                        hipHostMallocDefault));
}

}  // namespace nrx

extern "C" {

int nrx_batch_create(nrx_engine *const *engines, int32_t n_engines, nrx_batch **out) {
  using namespace nrx;
  return guarded([&]() -> int {
    if (!engines || n_engines <= 0 || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    auto b = std::make_unique<Batch>();
    for (int i = 0; i < n_engines; ++i) {
      Engine *e = static_cast<Engine *>(engines[i]);
      if (!e) throw std::invalid_argument("null engine in batch");
      // an engine listed twice would have its tiles worked by two
      // workgroups at once, racing on its adjoint rows
      if (std::find(b->engines.begin(), b->engines.end(), e) != b->engines.end())
        throw std::invalid_argument("engine " + std::to_string(i) + " appears twice in the batch");
      if (e->opt.precision != NRX_FP32 || e->S != 4)
        throw std::invalid_argument("batches take FP32-form engines with 4-state data");
      if (i > 0 && (e->C != b->engines[0]->C || e->opt.device != b->engines[0]->opt.device))
        throw std::invalid_argument("batch engines must share rate categories and device");
      b->engines.push_back(e);
    }
    b->C = b->engines[0]->C;
    b->use_graph = b->engines[0]->opt.use_graph != 0;
    NRX_HIP(hipSetDevice(b->engines[0]->opt.device));
    NRX_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    NRX_HIP(hipEventCreate(&b->ev0));
    NRX_HIP(hipEventCreate(&b->ev1));
    build_batch(*b);
    *out = reinterpret_cast<nrx_batch *>(b.release());
    return NRX_OK;
  });
}

int nrx_batch_destroy(nrx_batch *batch) {
  using namespace nrx;
  return guarded([&]() -> int {
    delete reinterpret_cast<Batch *>(batch);
    return NRX_OK;
  });
}

int nrx_batch_evaluate(nrx_batch *batch, int32_t variant, double *lnl, double *dbrlen,
                       double *dprob, double *device_ms) {
  using namespace nrx;
  return guarded([&]() -> int {
    if (!batch || !lnl) throw std::invalid_argument("null argument");
    if (variant != NRX_AVERAGE_DISPLAYED_TREES && variant != NRX_BEST_DISPLAYED_TREE)
      throw std::invalid_argument("unknown likelihood variant");
    Batch &b = *reinterpret_cast<Batch *>(batch);
    const int n = static_cast<int>(b.engines.size());
    bool rebuild = false;
    for (int i = 0; i < n; ++i) {
      Engine &e = *b.engines[i];
      check_ready(e);
      // models and branch lengths of every engine, as update_pmatrices does
      if (e.model_dirty) upload_models(e);
      upload_brlen(e, b.stream);  // ordered before the batch's kernels
      e.model_dirty = false;
      std::fill(e.edge_dirty.begin(), e.edge_dirty.end(), 0);
      rebuild |= b.versions[i] != e.schedule_version;
    }
    if (rebuild) build_batch(b);
    const bool grad = dbrlen != nullptr;
    auto enqueue = [&]() {
      const SmallDesc *sd = b.d_small.as<SmallDesc>();
      launch_pmatrices_multi(b.stream, sd, n, b.max_pm);
      launch_forward_persistent_multi(b.stream, b.C, b.G, b.TS, b.d_nets.as<NetDesc>(),
                                      b.d_work.as<WorkItem>(), b.n_work, b.lds_fwd);
      launch_root_reduce_multi(b.stream, sd, n, b.max_root_rows);
      launch_mixture_multi(b.stream, sd, n, variant);
      if (grad) {
        launch_adjoint_persistent_multi(b.stream, b.C, b.G, b.TS, b.H, b.d_nets.as<NetDesc>(),
                                        b.d_work.as<WorkItem>(), b.n_work, b.lds_adj);
        launch_grad_reduce_multi(b.stream, sd, n, b.max_use_rows, b.max_edges);
      }
      launch_gather_results(b.stream, sd, n, b.d_offsets.as<int64_t>(), b.d_out.as<double>(),
                            b.max_res);
    };
    hipGraphExec_t *ge = nullptr;
    if (b.use_graph) {
      ge = &b.graph_exec[variant][grad ? 1 : 0];
      if (!*ge) *ge = capture_graph(b.stream, enqueue);
    }
    NRX_HIP(hipEventRecord(b.ev0, b.stream));
    if (ge)
      NRX_HIP(hipGraphLaunch(*ge, b.stream));
    else
      enqueue();
    NRX_HIP(hipEventRecord(b.ev1, b.stream));
    NRX_HIP(hipMemcpyAsync(b.h_out, b.d_out.p, sizeof(double) * b.offsets[n],
                           hipMemcpyDeviceToHost, b.stream));
    NRX_HIP(hipStreamSynchronize(b.stream));
    if (device_ms) {
      float ms = 0.f;
      NRX_HIP(hipEventElapsedTime(&ms, b.ev0, b.ev1));
      *device_ms = ms;
    }
    int64_t eo = 0, ro = 0;
    int bad = -1;
    for (int i = 0; i < n; ++i) {
      Engine &e = *b.engines[i];
      const double *r = b.h_out + b.offsets[i];
      lnl[i] = r[e.rl.total];
      if (!(lnl[i] > -INFINITY) && bad < 0) bad = i;
      if (grad) {
        std::copy(r + e.rl.dbrlen, r + e.rl.dbrlen + e.E, dbrlen + eo);
        if (dprob && e.R > 0) std::copy(r + e.rl.dprob, r + e.rl.dprob + e.R, dprob + ro);
      }
      eo += e.E;
      ro += e.R;
      // CLVs are now valid; the engine's own cached result is not
      std::fill(e.clv_dirty.begin(), e.clv_dirty.end(), 0);
      e.cache_valid = false;
      e.deriv_valid = false;
    }
    if (bad >= 0)
      return fail(NRX_ERR_NEGINF, "Invalid network likelihood: negative infinity (engine " +
                                      std::to_string(bad) + ")");
    return NRX_OK;
  });
}

}  // extern "C"

namespace nrx {

}  // namespace nrx

using namespace nrx;

extern "C" {

void nrx_default_options(nrx_options *opt) {
  if (!opt) return;
  std::memset(opt, 0, sizeof(*opt));
  opt->precision = NRX_FP64;
  opt->device = 0;
  opt->min_tree_logprob = std::log(1e-6);
  opt->brlen_min = 1e-6;
  opt->brlen_max = 100.0;
  opt->use_graph = 0;
  opt->timing = 0;
}

const char *nrx_last_error(void) { return g_last_error.c_str(); }

int nrx_create(const nrx_network_desc *net, const nrx_partition_desc *parts, int32_t n_parts,
               const nrx_options *opt, nrx_engine **out) {
  return guarded([&]() -> int {
    if (!net || !parts || n_parts <= 0 || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    auto e = std::make_unique<Engine>();
    if (opt)
      e->opt = *opt;
    else
      nrx_default_options(&e->opt);
    if (e->opt.precision != NRX_FP64 && e->opt.precision != NRX_FP32)
      throw std::invalid_argument("unknown precision");
    e->N = net->n_nodes;
    e->E = net->n_edges;
    e->R = net->n_retic;
    e->T = net->n_tips;
    e->root = net->root;
    e->edge_parent.assign(net->edge_parent, net->edge_parent + net->n_edges);
    e->edge_child.assign(net->edge_child, net->edge_child + net->n_edges);
    if (e->R > 0) {
      e->retic_first.assign(net->retic_first_edge, net->retic_first_edge + e->R);
      e->retic_second.assign(net->retic_second_edge, net->retic_second_edge + e->R);
    }
    nrx_network_desc nd = *net;
    nd.edge_parent = e->edge_parent.data();
    nd.edge_child = e->edge_child.data();
    nd.retic_first_edge = e->retic_first.data();
    nd.retic_second_edge = e->retic_second.data();
    e->dag = compile_dag(nd);

    if (n_parts > 64) throw std::invalid_argument("at most 64 partitions per engine");
    e->P = n_parts;
    e->S = parts[0].states;
    e->C = parts[0].rate_cats;
    if (e->S != 4 && e->S != 20) throw std::invalid_argument("states must be 4 or 20");
    if (e->C != 1 && e->C != 2 && e->C != 4 && e->C != 8)
      throw std::invalid_argument("rate categories must be 1, 2, 4 or 8");
    const bool aa32 = e->opt.precision == NRX_FP32 && e->S == AA_STATES;
    // P (and Q) of one partition are staged in LDS (64 KiB per workgroup)
    if (e->opt.precision == NRX_FP32 && !aa32 &&
        sizeof(float) * (static_cast<size_t>(e->E) + 1) * e->C * 16 > 65536)
      return fail(NRX_ERR_UNSUPPORTED, "network too large for the FP32 form's LDS tables");
    // site-sum partials per 64-site tile: one per wave (the 20-state FP32
    // form runs two 32-site workgroups of C waves per tile)
    e->pw = aa32 ? 2 * e->C : TILE * e->C / 64;
    e->parts.resize(n_parts);
    int64_t off = 0;
    int tile = 0;
    for (int p = 0; p < n_parts; ++p) {
      if (parts[p].states != e->S || parts[p].rate_cats != e->C)
        throw std::invalid_argument("all partitions must share states and rate categories");
      if (parts[p].n_sites <= 0 || !parts[p].tip_masks || !parts[p].weights)
        throw std::invalid_argument("empty partition");
      PartInfo &pi = e->parts[p];
      pi.n_sites = parts[p].n_sites;
      pi.site_off = off;
      const int nt = (pi.n_sites + TILE - 1) / TILE;
      pi.tile_begin = tile;
      pi.tile_end = tile + nt;
      tile += nt;
      off += static_cast<int64_t>(nt) * TILE;
    }
    DevLayout &L = e->L;
    L.n_sites_pad = off;
    L.n_tiles = tile;
    L.n_tiles_grid = (tile + 7) / 8 * 8;
    L.n_parts = n_parts;
    L.n_edges = e->E;
    L.S = e->S;
    L.C = e->C;

    NRX_HIP(hipSetDevice(e->opt.device));
    NRX_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    NRX_HIP(hipEventCreate(&e->ev_begin));
    NRX_HIP(hipEventCreate(&e->ev_end));
    NRX_HIP(hipEventCreateWithFlags(&e->ev_brlen, hipEventDisableTiming));

    // tips, weights, tile -> partition
    const int64_t n = L.n_sites_pad;
    const uint32_t full = e->S == 4 ? 0xFu : ((1u << 20) - 1u);
    std::vector<double> w(n, 0.0);
    std::vector<int32_t> tile_part(tile), ptb(n_parts + 1);
    for (int p = 0; p < n_parts; ++p) {
      const PartInfo &pi = e->parts[p];
      for (int s = 0; s < pi.n_sites; ++s) w[pi.site_off + s] = parts[p].weights[s];
      for (int t = pi.tile_begin; t < pi.tile_end; ++t) tile_part[t] = p;
      ptb[p] = pi.tile_begin;
    }
    ptb[n_parts] = tile;
    if (e->S == 4) {
      std::vector<uint8_t> tips(static_cast<size_t>(e->T) * n, static_cast<uint8_t>(full));
      for (int p = 0; p < n_parts; ++p) {
        const PartInfo &pi = e->parts[p];
        for (int t = 0; t < e->T; ++t)
          for (int s = 0; s < pi.n_sites; ++s) {
            const uint32_t m = parts[p].tip_masks[static_cast<size_t>(t) * pi.n_sites + s] & full;
            if (m == 0) throw std::invalid_argument("tip state mask is empty");
            tips[static_cast<size_t>(t) * n + pi.site_off + s] = static_cast<uint8_t>(m);
          }
      }
      upload(e->d_tips, tips, e->stream);
    } else {
      std::vector<uint32_t> tips(static_cast<size_t>(e->T) * n, full);
      for (int p = 0; p < n_parts; ++p) {
        const PartInfo &pi = e->parts[p];
        for (int t = 0; t < e->T; ++t)
          for (int s = 0; s < pi.n_sites; ++s) {
            const uint32_t m = parts[p].tip_masks[static_cast<size_t>(t) * pi.n_sites + s] & full;
            if (m == 0) throw std::invalid_argument("tip state mask is empty");
            tips[static_cast<size_t>(t) * n + pi.site_off + s] = m;
          }
      }
      upload(e->d_tips, tips, e->stream);
    }
    upload(e->d_weights, w, e->stream);
    upload(e->d_tile_part, tile_part, e->stream);
    upload(e->d_part_tile_begin, ptb, e->stream);

    // model and branch-length buffers have fixed sizes: allocate them now so
    // their addresses, which the device descriptors hold, never change
    const size_t ss = static_cast<size_t>(n_parts) * e->S * e->S;
    e->d_V.alloc(sizeof(double) * ss);
    e->d_Vinv.alloc(sizeof(double) * ss);
    e->d_eval.alloc(sizeof(double) * n_parts * e->S);
    e->d_rates.alloc(sizeof(double) * n_parts * e->C);
    e->d_freqs.alloc(sizeof(double) * n_parts * e->S);
    e->d_brlen.alloc(sizeof(double) * e->E);
    NRX_HIP(hipHostMalloc(reinterpret_cast<void **>(&e->h_brlen), sizeof(double) * std::max(1, e->E),
                          hipHostMallocDefault));
    e->d_Qd.alloc(sizeof(double) * ss * e->C);
    if (e->opt.precision == NRX_FP32) e->d_Qf.alloc(sizeof(float) * ss * e->C);
    const size_t pm = static_cast<size_t>(n_parts) * e->E * e->C * e->S * e->S;
    e->d_P.alloc(sizeof(double) * pm);
    e->d_dP.alloc(sizeof(double) * pm);
    e->d_d2P.alloc(sizeof(double) * pm);
    e->clv_elems_per_slot = static_cast<size_t>(n) * e->C * e->S;
    const size_t nv = e->dag.var.size();
    const size_t real_b = e->opt.precision == NRX_FP32 ? sizeof(float) : sizeof(double);
    e->d_clv.alloc(real_b * e->clv_elems_per_slot * std::max<size_t>(nv, 1));
    e->d_scl.alloc(sizeof(int32_t) * static_cast<size_t>(n) * std::max<size_t>(nv, 1));
    if (e->opt.precision == NRX_FP32) {
      e->d_Pf.alloc(sizeof(float) * pm);
      e->d_adj.alloc(sizeof(float) * e->clv_elems_per_slot * std::max<size_t>(nv, 1));
      // adjoint exponents: per site (4 states) or per (site, category) (20 states)
      const size_t per_slot = aa32 ? static_cast<size_t>(n) * e->C : static_cast<size_t>(n);
      e->d_adj_exp.alloc(sizeof(int32_t) * per_slot * std::max<size_t>(nv, 1));
      if (aa32)
        e->d_Pa.alloc(sizeof(float) * static_cast<size_t>(n_parts) * e->E * e->C * AA_KINDS * AA_TABLE);
    }

    const int nc = static_cast<int>(e->dag.classes.size());
    ResultLayout &rl = e->rl;
    int o = 0;
    rl.total = o++;
    rl.per_part = o;
    o += n_parts;
    rl.root_logl = o;
    o += n_parts * nc;
    rl.tree_logl = o;
    o += n_parts * nc;
    rl.alpha = o;
    o += n_parts * nc;
    rl.dbrlen = o;
    o += e->E;
    rl.part_dbrlen = o;
    o += n_parts * e->E;
    rl.dprob = o;
    o += std::max(e->R, 1);
    rl.size = o;
    e->d_result.alloc(sizeof(double) * rl.size);
    NRX_HIP(hipHostMalloc(reinterpret_cast<void **>(&e->h_result), sizeof(double) * rl.size,
                          hipHostMallocDefault));

    e->brlen.assign(e->E, 0.1);
    e->probs.assign(e->R, 0.5);
    e->edge_dirty.assign(e->E, 1);
    e->clv_dirty.assign(nv, 1);
    compute_class_probs(*e);
    build_schedule(*e);
    *out = e.release();
    return NRX_OK;
  });
}

int nrx_destroy(nrx_engine *eng) {
  return guarded([&]() -> int {
    delete static_cast<Engine *>(eng);
    return NRX_OK;
  });
}

int nrx_set_model(nrx_engine *eng, int32_t partition, const double *subst_rates,
                  const double *freqs, double alpha) {
  return guarded([&]() -> int {
    Engine &e = *static_cast<Engine *>(eng);
    if (!eng || partition < 0 || partition >= e.P || !subst_rates || !freqs)
      throw std::invalid_argument("bad model arguments");
    PartInfo &pi = e.parts[partition];
    build_model(e.S, subst_rates, freqs, pi.model);
    pi.rates = gamma_mean_rates(alpha, e.C);
    pi.subst.assign(subst_rates, subst_rates + e.S * (e.S - 1) / 2);
    pi.alpha = alpha;
    pi.model_set = true;
    e.model_dirty = true;
    e.deriv_valid = false;
    return NRX_OK;
  });
}

int nrx_set_category_rates(nrx_engine *eng, int32_t partition, const double *rates) {
  return guarded([&]() -> int {
    Engine &e = *static_cast<Engine *>(eng);
    if (!eng || partition < 0 || partition >= e.P || !rates)
      throw std::invalid_argument("bad arguments");
    PartInfo &pi = e.parts[partition];
    if (!pi.model_set) throw std::runtime_error("set the model before the category rates");
    for (int c = 0; c < e.C; ++c)
      if (!(rates[c] >= 0.0)) throw std::invalid_argument("category rates must be >= 0");
    pi.rates.assign(rates, rates + e.C);
    pi.alpha = std::nan("");
    e.model_dirty = true;
    e.deriv_valid = false;
    return NRX_OK;
  });
}

int nrx_set_branch_lengths(nrx_engine *eng, const double *brlen) {
  return guarded([&]() -> int {
    Engine &e = *static_cast<Engine *>(eng);
    if (!eng || !brlen) throw std::invalid_argument("null argument");
    for (int i = 0; i < e.E; ++i) {
      if (std::isnan(brlen[i])) throw std::invalid_argument("branch length is NaN");
      // clamp as createNetworkPllTreeinfo does (RaxmlWrapper.cpp:652-663)
      const double b = std::min(std::max(brlen[i], e.opt.brlen_min), e.opt.brlen_max);
      if (b != e.brlen[i]) {
        e.brlen[i] = b;
        e.edge_dirty[i] = 1;
        // sumtables of a branch stay valid while only that branch moves
        if (i != e.deriv_edge) e.deriv_valid = false;
      }
    }
    return NRX_OK;
  });
}

int nrx_set_reticulation_probs(nrx_engine *eng, const double *p) {
  return guarded([&]() -> int {
    Engine &e = *static_cast<Engine *>(eng);
    if (!eng || (e.R > 0 && !p)) throw std::invalid_argument("null argument");
    set_probs(e, p);
    return NRX_OK;
  });
}

int nrx_loglikelihood(nrx_engine *eng, int32_t variant, int32_t incremental,
                      int32_t update_pmatrices, double *total, double *per_partition) {
  return guarded([&]() -> int {
    if (!eng || !total) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    const int rc = evaluate(e, variant, incremental != 0, update_pmatrices != 0, false);
    if (rc != NRX_OK) return rc;
    *total = e.cached_total;
    if (per_partition)
      std::copy(e.cached_per_part.begin(), e.cached_per_part.end(), per_partition);
    return NRX_OK;
  });
}

int nrx_gradient(nrx_engine *eng, int32_t variant, double *total, double *dbrlen, double *dprob,
                 double *per_partition_dbrlen) {
  return guarded([&]() -> int {
    if (!eng || !total || !dbrlen) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    const int rc = evaluate(e, variant, true, true, true);
    if (rc != NRX_OK) return rc;
    *total = e.cached_total;
    std::copy(e.h_result + e.rl.dbrlen, e.h_result + e.rl.dbrlen + e.E, dbrlen);
    if (dprob && e.R > 0) std::copy(e.h_result + e.rl.dprob, e.h_result + e.rl.dprob + e.R, dprob);
    if (per_partition_dbrlen)
      std::copy(e.h_result + e.rl.part_dbrlen, e.h_result + e.rl.part_dbrlen + e.P * e.E,
                per_partition_dbrlen);
    return NRX_OK;
  });
}

int nrx_brlen_derivatives(nrx_engine *eng, int32_t variant, int32_t edge, int32_t mode,
                          double brlen, double *d1, double *d2, double *part_d1,
                          double *part_d2) {
  return guarded([&]() -> int {
    if (!eng || !d1 || !d2) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    if (edge < 0 || edge >= e.E) throw std::invalid_argument("edge index out of range");
    if (mode != NRX_DERIV_REFERENCE && mode != NRX_DERIV_TRUE)
      throw std::invalid_argument("unknown derivative mode");
    if (variant != NRX_AVERAGE_DISPLAYED_TREES && variant != NRX_BEST_DISPLAYED_TREE)
      throw std::invalid_argument("unknown likelihood variant");
    if (!(brlen >= 0.0)) throw std::invalid_argument("branch length must be >= 0");
    // the same clamp as nrx_set_branch_lengths (RaxmlWrapper.cpp:652-663):
    // derivatives at a length the engine would never evaluate are not given
    brlen = std::min(std::max(brlen, e.opt.brlen_min), e.opt.brlen_max);
    if (e.opt.precision == NRX_FP32 && e.S == AA_STATES)
      return fail(NRX_ERR_UNSUPPORTED,
                  "per-branch sumtables of 20-state data need the FP64 form (the FP32 form "
                  "returns the gradient of every branch through nrx_gradient)");
    if (!e.deriv_valid || e.deriv_edge != edge) {
      const int rc = evaluate(e, variant, true, true, false);
      if (rc != NRX_OK) return rc;
      build_sumtables(e, edge);
    }
    const int np = e.deriv_n_paths;
    if (np > 0) {
      launch_sumtable_eval(e, np, brlen);
      NRX_HIP(hipMemcpyAsync(e.h_st_sums.data(), e.d_st_sums.p, sizeof(double) * np * 3 * e.P,
                             hipMemcpyDeviceToHost, e.stream));
      stream_sync(e);
    }
    mix_brlen_derivatives(e, variant, mode, e.deriv_base_logl, e.deriv_class, e.deriv_path, e.h_st_sums.data(),
                          d1, d2, part_d1, part_d2);
    return NRX_OK;
  });
}

int nrx_brlen_derivatives_noalloc(nrx_engine *eng, int32_t variant, int32_t edge, int32_t mode, double brlen,
                                  double *d1, double *d2, double *part_d1, double *part_d2) {
  return guarded([&]() -> int {
    if (!eng || !d1 || !d2) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    if (edge < 0 || edge >= e.E) throw std::invalid_argument("edge index out of range");
    if (mode != NRX_DERIV_REFERENCE && mode != NRX_DERIV_TRUE)
      throw std::invalid_argument("unknown derivative mode");
    if (variant != NRX_AVERAGE_DISPLAYED_TREES && variant != NRX_BEST_DISPLAYED_TREE)
      throw std::invalid_argument("unknown likelihood variant");
    if (!(brlen >= 0.0)) throw std::invalid_argument("branch length must be >= 0");
    brlen = std::min(std::max(brlen, e.opt.brlen_min), e.opt.brlen_max);
    if (e.opt.precision == NRX_FP32 && e.S == AA_STATES)
      return fail(NRX_ERR_UNSUPPORTED,
                  "per-branch derivatives of 20-state data need the FP64 form (the FP32 form "
                  "returns the gradient of every branch through nrx_gradient)");
    std::vector<int> cls, path;
    std::vector<PathStep> steps;
    std::vector<int32_t> begin;
    make_deriv_paths(e, edge, cls, path, steps, begin);
    const int np = static_cast<int>(begin.size()) - 1;
    const int nc = static_cast<int>(e.dag.classes.size());
    // layout of the fixed scratch: path steps and bounds, per-wave partials,
    // per-(path, partition) sums
    require_resident(e);
    auto aligned = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t o_steps = 0;
    const size_t o_begin = aligned(o_steps + sizeof(PathStep) * steps.size());
    const size_t o_part = aligned(o_begin + sizeof(int32_t) * begin.size());
    const size_t o_sums = aligned(o_part + sizeof(double) * np * 3 * static_cast<size_t>(e.L.n_tiles) * e.pw);
    std::vector<double> sums(static_cast<size_t>(np) * 3 * e.P, 0.0);
    const size_t need = o_sums + sizeof(double) * sums.size();
    if (need > NOALLOC_SCRATCH_BYTES)
      throw std::invalid_argument("branch derivatives of this engine need " + std::to_string(need) +
                                  " bytes of scratch, more than the no-alloc limit of " +
                                  std::to_string(NOALLOC_SCRATCH_BYTES));
    // the CLVs, P matrices and tree lnLs of the last evaluation are used as
    // they are: refreshing them could allocate
    const bool stale = !e.cache_valid || e.cache_variant != variant || e.model_dirty || e.probs_dirty ||
                       std::any_of(e.edge_dirty.begin(), e.edge_dirty.end(), [](char c) { return c; }) ||
                       std::any_of(e.clv_dirty.begin(), e.clv_dirty.end(), [](char c) { return c; });
    if (stale)
      return fail(NRX_ERR_STATE,
                  "parameters changed since the last evaluation of this variant: call nrx_loglikelihood "
                  "(or nrx_gradient) first");
    if (np > 0) {
      NRX_HIP(hipSetDevice(e.opt.device));
      char *base = static_cast<char *>(noalloc_scratch());
      NRX_HIP(hipMemcpyAsync(base + o_steps, steps.data(), sizeof(PathStep) * steps.size(),
                             hipMemcpyHostToDevice, e.stream));
      NRX_HIP(hipMemcpyAsync(base + o_begin, begin.data(), sizeof(int32_t) * begin.size(),
                             hipMemcpyHostToDevice, e.stream));
      launch_deriv_fused(e, reinterpret_cast<const PathStep *>(base + o_steps),
                         reinterpret_cast<const int32_t *>(base + o_begin), np, max_path_steps(begin), brlen,
                         reinterpret_cast<double *>(base + o_part), reinterpret_cast<double *>(base + o_sums));
      NRX_HIP(hipMemcpyAsync(sums.data(), base + o_sums, sizeof(double) * sums.size(), hipMemcpyDeviceToHost,
                             e.stream));
      stream_sync(e);
    }
    const std::vector<double> base_logl(e.h_result + e.rl.tree_logl, e.h_result + e.rl.tree_logl + e.P * nc);
    mix_brlen_derivatives(e, variant, mode, base_logl, cls, path, sums.data(), d1, d2, part_d1, part_d2);
    return NRX_OK;
  });
}

int nrx_get_branch_lengths(nrx_engine *eng, double *brlen) {
  return guarded([&]() -> int {
    if (!eng || !brlen) throw std::invalid_argument("null argument");
    const Engine &e = *static_cast<Engine *>(eng);
    std::copy(e.brlen.begin(), e.brlen.end(), brlen);
    return NRX_OK;
  });
}

// The device Newton sweep of one engine over `edges` (see
// nrx_optimize_branches); throws on bad arguments.
static int optimize_branches_impl(Engine &e, int variant, int mode, const int32_t *edges, int n_edges,
                                  int max_iters, double *logl) {
  {
    if (n_edges > 0 && !edges) throw std::invalid_argument("null argument");
    if (mode != NRX_DERIV_REFERENCE && mode != NRX_DERIV_TRUE)
      throw std::invalid_argument("unknown derivative mode");
    if (variant != NRX_AVERAGE_DISPLAYED_TREES && variant != NRX_BEST_DISPLAYED_TREE)
      throw std::invalid_argument("unknown likelihood variant");
    if (max_iters < 1 || max_iters > 256) throw std::invalid_argument("max_iters must lie in [1, 256]");
    if (e.opt.precision == NRX_FP32 && e.S == AA_STATES)
      return fail(NRX_ERR_UNSUPPORTED, "per-branch sumtables of 20-state data need the FP64 form");
    std::vector<int> cand(edges, edges + std::max(0, n_edges));
    for (int ed : cand)
      if (ed < 0 || ed >= e.E) throw std::invalid_argument("edge index out of range");
    std::sort(cand.begin(), cand.end());  // optimize_branches takes the candidates in index order
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    const int nc = static_cast<int>(e.dag.classes.size());
    // sumtable buffers for every class at once: their addresses then stay
    // fixed for the captured Newton sequence
    e.d_st.alloc(sizeof(double) * e.clv_elems_per_slot * std::max(nc, 1));
    e.d_st_partials.alloc(sizeof(double) * std::max(nc, 1) * 3 * static_cast<size_t>(e.L.n_tiles) * e.pw);
    e.d_st_sums.alloc(sizeof(double) * std::max(nc, 1) * 3 * e.P);
    e.d_nr_state.alloc(sizeof(NewtonState));
    e.d_nr_cls.alloc(sizeof(DerivClassDev) * std::max(nc, 1));
    if (!e.h_nr_state)
      NRX_HIP(hipHostMalloc(reinterpret_cast<void **>(&e.h_nr_state), sizeof(NewtonState), hipHostMallocDefault));
    const void *bufs[3] = {e.d_st.p, e.d_st_partials.p, e.d_st_sums.p};
    static_assert(sizeof(e.nr_graph_bufs) == sizeof(bufs), "graph key buffers");
    // Newton rounds per graph launch: most branches converge within one
    // launch; a branch that has not is launched again
    const int rounds = std::min<int>(max_iters, 6);
    const int key[4] = {variant, mode, rounds, nc};
    if (e.nr_graph && (!std::equal(key, key + 4, e.nr_graph_key) || !std::equal(bufs, bufs + 3, e.nr_graph_bufs))) {
      (void)hipGraphExecDestroy(e.nr_graph);
      e.nr_graph = nullptr;
    }
    int rc = evaluate(e, variant, true, true, false);
    if (rc != NRX_OK) return rc;
    const double start = e.cached_total;
    const double tol = e.opt.brlen_min > 0.0 ? e.opt.brlen_min / 10.0 : 1e-7;
    std::vector<DerivClassDev> cls(nc);
    const double *base_logl = e.d_result.as<double>() + e.rl.tree_logl;  // tree lnLs, read on the device
    for (int ed : cand) {
      // the branch's evaluation leaves its results on the device: the
      // Newton kernels read the tree lnLs there, so the only host wait of
      // a branch is the one that reads its Newton state
      rc = evaluate(e, variant, true, true, false, /*fetch=*/false);
      if (rc != NRX_OK) return rc;
      build_sumtables(e, ed);
      e.deriv_valid = false;  // its host-side tree lnLs are stale (results stayed on the device)
      const int np = e.deriv_n_paths;
      if (np == 0) continue;  // the branch is in no displayed tree's path: derivative 0, length kept
      std::vector<char> has(nc, 0);
      std::vector<int> path_of(nc, -1);
      for (size_t i = 0; i < e.deriv_class.size(); ++i) {
        has[e.deriv_class[i]] = 1;
        path_of[e.deriv_class[i]] = e.deriv_path[i];
      }
      for (int k = 0; k < nc; ++k) cls[k] = DerivClassDev{e.class_logprob[k], e.class_use[k], has[k], path_of[k], 0};
      stage_upload(e, e.d_nr_cls, cls);
      NewtonState &st = *e.h_nr_state;
      st = NewtonState{};
      st.lo = e.opt.brlen_min;
      st.hi = e.opt.brlen_max;
      st.x = st.x0 = std::min(std::max(e.brlen[ed], st.lo), st.hi);
      st.a = st.lo;
      st.b = st.hi;
      st.tol = tol;
      st.n_paths = np;
      st.max_iters = max_iters;
      NRX_HIP(hipMemcpyAsync(e.d_nr_state.p, e.h_nr_state, sizeof(NewtonState), hipMemcpyHostToDevice, e.stream));
      if (!e.nr_graph) {
        e.nr_graph = capture_graph(e.stream, [&]() {
          enqueue_newton(e, variant, mode, rounds, nc, e.d_nr_state.as<NewtonState>(),
                         e.d_nr_cls.as<DerivClassDev>(), base_logl);
        });
        std::copy(key, key + 4, e.nr_graph_key);
        std::copy(bufs, bufs + 3, e.nr_graph_bufs);
      }
      do {
        NRX_HIP(hipGraphLaunch(e.nr_graph, e.stream));
        NRX_HIP(hipMemcpyAsync(e.h_nr_state, e.d_nr_state.p, sizeof(NewtonState), hipMemcpyDeviceToHost,
                               e.stream));
        stream_sync(e);  // one wait per launch (usually one per branch)
      } while (!st.done);
      const double x = st.accepted ? st.x : st.x0;
      if (x != e.brlen[ed]) {
        e.brlen[ed] = x;
        e.edge_dirty[ed] = 1;  // the sumtables of this branch stay valid
      }
    }
    rc = evaluate(e, variant, true, true, false);
    if (rc != NRX_OK) return rc;
    if (logl) *logl = e.cached_total;
    // optimize_branches' final check (BranchLengthOptimization.cpp:422-477)
    if (e.cached_total < start && std::fabs(e.cached_total - start) >= 1e-3)
      throw std::runtime_error("Overall loglikelihood got worse");
    return NRX_OK;
  }
}

int nrx_optimize_branches(nrx_engine *eng, int32_t variant, int32_t mode, const int32_t *edges, int32_t n_edges,
                          int32_t max_iters, double *logl) {
  return guarded([&]() -> int {
    if (!eng) throw std::invalid_argument("null argument");
    return optimize_branches_impl(*static_cast<Engine *>(eng), variant, mode, edges, n_edges, max_iters, logl);
  });
}

int nrx_model_gradient(nrx_engine *eng, int32_t variant, double *total, double *d_subst, double *d_freqs,
                       double *d_alpha, double *d_rates) {
  return guarded([&]() -> int {
    if (!eng || !total || !d_subst || !d_freqs) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    if (e.opt.precision != NRX_FP64 || e.S != 4)
      return fail(NRX_ERR_UNSUPPORTED, "model gradients are provided for the FP64 form with 4-state data");
    require_resident(e);
    const int rc = evaluate(e, variant, true, true, true);  // CLVs, adjoints, mixture weights
    if (rc != NRX_OK) return rc;
    *total = e.cached_total;
    const int S = e.S, C = e.C, P = e.P, nr = S * (S - 1) / 2;
    const int n_ops = static_cast<int>(e.ops_all.size());
    const int n_root = static_cast<int>(e.root_slots.size());
    // slices per partition so that the grid fills the GPU
    int n_split = std::max(1, (2048 + n_ops * P - 1) / std::max(1, n_ops * P));
    for (const PartInfo &pi : e.parts) n_split = std::min(n_split, std::max(1, pi.tile_end - pi.tile_begin));
    const int nch = P * n_split;
    const size_t ng = static_cast<size_t>(std::max(1, e.n_uses)) * nch * C * S * S;
    const size_t nrr = static_cast<size_t>(std::max(1, n_root)) * nch * S;
    e.d_mg_g.alloc(sizeof(double) * ng);
    e.d_mg_r.alloc(sizeof(double) * nrr);
    launch_model_grad(e, n_split, e.d_mg_g.as<double>(), e.d_mg_r.as<double>());
    std::vector<double> G(ng), Rr(nrr);
    NRX_HIP(hipMemcpyAsync(G.data(), e.d_mg_g.p, sizeof(double) * ng, hipMemcpyDeviceToHost, e.stream));
    NRX_HIP(hipMemcpyAsync(Rr.data(), e.d_mg_r.p, sizeof(double) * nrr, hipMemcpyDeviceToHost, e.stream));
    stream_sync(e);
    for (int p = 0; p < P; ++p) {
      const PartInfo &pi = e.parts[p];
      const EigenModel &m = pi.model;
      const double *pf = m.freqs.data();
      // Q and its unnormalised form R (Q = R / mu), as build_model forms them
      std::vector<double> R(S * S, 0.0);
      for (int i = 0, k = 0; i < S; ++i)
        for (int j = i + 1; j < S; ++j, ++k) {
          R[i * S + j] = pi.subst[k] * pf[j];
          R[j * S + i] = pi.subst[k] * pf[i];
        }
      double mu = 0.0;
      for (int i = 0; i < S; ++i) {
        double rowsum = 0.0;
        for (int j = 0; j < S; ++j)
          if (j != i) rowsum += R[i * S + j];
        R[i * S + i] = -rowsum;
        mu += pf[i] * rowsum;
      }
      std::vector<double> Q(S * S);
      for (int q = 0; q < S * S; ++q) Q[q] = R[q] / mu;
      // dJ/dQ through P_{e,c} = exp(Q r_c t_e) = V exp(L s) V^-1:
      // dJ/dQ = V^-T ((V^T G V^-T) o Psi) V^T, Psi_kl = (e^{l_k s} - e^{l_l s}) / (l_k - l_l)
      // (s e^{l_k s} for equal eigenvalues); dJ/dr_c = sum_e t_e tr(G^T Q P)
      std::vector<double> dQ(S * S, 0.0), dr(C, 0.0), Gm(S * S), H(S * S), T1(S * S), Ps(S * S), QP(S * S);
      for (int ed = 0; ed < e.E; ++ed) {
        for (int c = 0; c < C; ++c) {
          std::fill(Gm.begin(), Gm.end(), 0.0);
          for (int i = e.edge_use_begin[ed]; i < e.edge_use_begin[ed + 1]; ++i) {
            const int u = e.edge_use_list[i];
            for (int sl = 0; sl < n_split; ++sl) {
              const double *g = &G[((static_cast<size_t>(u) * nch + p * n_split + sl) * C + c) * S * S];
              for (int q = 0; q < S * S; ++q) Gm[q] += g[q];
            }
          }
          const double s = pi.rates[c] * e.brlen[ed];
          // H = V^T G Vinv^T
          for (int a = 0; a < S; ++a)
            for (int j = 0; j < S; ++j) {
              double acc = 0.0;
              for (int i = 0; i < S; ++i) acc += m.V[i * S + a] * Gm[i * S + j];
              T1[a * S + j] = acc;
            }
          for (int a = 0; a < S; ++a)
            for (int b = 0; b < S; ++b) {
              double acc = 0.0;
              for (int j = 0; j < S; ++j) acc += T1[a * S + j] * m.Vinv[b * S + j];
              H[a * S + b] = acc;
            }
          for (int a = 0; a < S; ++a)
            for (int b = 0; b < S; ++b) {
              const double la = m.eval[a], lb = m.eval[b];
              const double psi = std::fabs(la - lb) > 1e-10 * std::max(1.0, std::fabs(la))
                                     ? (std::exp(la * s) - std::exp(lb * s)) / (la - lb)
                                     : s * std::exp(0.5 * (la + lb) * s);
              H[a * S + b] *= psi;
            }
          // dQ += Vinv^T H V^T
          for (int i = 0; i < S; ++i)
            for (int b = 0; b < S; ++b) {
              double acc = 0.0;
              for (int a = 0; a < S; ++a) acc += m.Vinv[a * S + i] * H[a * S + b];
              T1[i * S + b] = acc;
            }
          for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j) {
              double acc = 0.0;
              for (int b = 0; b < S; ++b) acc += T1[i * S + b] * m.V[j * S + b];
              dQ[i * S + j] += acc;
            }
          // tr(G^T Q P) with P = V exp(L s) Vinv
          for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j) {
              double acc = 0.0;
              for (int k = 0; k < S; ++k) acc += m.V[i * S + k] * std::exp(m.eval[k] * s) * m.Vinv[k * S + j];
              Ps[i * S + j] = acc;
            }
          double tr = 0.0;
          for (int i = 0; i < S; ++i)
            for (int j = 0; j < S; ++j) {
              double acc = 0.0;
              for (int k = 0; k < S; ++k) acc += Q[i * S + k] * Ps[k * S + j];
              tr += Gm[i * S + j] * acc;
            }
          dr[c] += e.brlen[ed] * tr;
        }
      }
      // chain to exchangeabilities and frequencies: Q = R / mu,
      // R_ij = rho_ij pi_j (i != j), R_ii = -sum_{j != i} R_ij,
      // mu = sum_i pi_i sum_{j != i} rho_ij pi_j
      double dmu = 0.0;
      for (int q = 0; q < S * S; ++q) dmu -= dQ[q] * R[q] / (mu * mu);
      std::vector<double> D(S * S, 0.0);  // dJ/dR_ij (i != j) with the diagonal's dependence folded in
      for (int i = 0; i < S; ++i)
        for (int j = 0; j < S; ++j)
          if (i != j) D[i * S + j] = (dQ[i * S + j] - dQ[i * S + i]) / mu;
      for (int i = 0, k = 0; i < S; ++i)
        for (int j = i + 1; j < S; ++j, ++k)
          d_subst[static_cast<size_t>(p) * nr + k] =
              D[i * S + j] * pf[j] + D[j * S + i] * pf[i] + dmu * 2.0 * pf[i] * pf[j];
      for (int kq = 0; kq < S; ++kq) {
        double g = 0.0, dmu_dpi = 0.0;
        for (int i = 0; i < S; ++i) {
          if (i == kq) continue;
          const double rho = pi.subst[i < kq ? (i * (2 * S - i - 1)) / 2 + (kq - i - 1)
                                             : (kq * (2 * S - kq - 1)) / 2 + (i - kq - 1)];
          g += D[i * S + kq] * rho;
          dmu_dpi += 2.0 * rho * pf[i];
        }
        g += dmu * dmu_dpi;
        // the frequencies' direct share at the displayed-tree roots
        for (int r = 0; r < n_root; ++r)
          for (int sl = 0; sl < n_split; ++sl)
            g += Rr[(static_cast<size_t>(r) * nch + p * n_split + sl) * S + kq];
        d_freqs[static_cast<size_t>(p) * S + kq] = g;
      }
      if (d_rates)
        for (int c = 0; c < C; ++c) d_rates[static_cast<size_t>(p) * C + c] = dr[c];
      if (d_alpha) {
        // dlnL/dalpha = sum_c dJ/dr_c dr_c/dalpha; the mean-rate function's
        // derivative by a central difference (it is smooth in alpha)
        double ga = 0.0;
        if (C > 1 && std::isfinite(pi.alpha)) {
          const double h = 1e-5 * pi.alpha;
          const std::vector<double> up = gamma_mean_rates(pi.alpha + h, C), dn = gamma_mean_rates(pi.alpha - h, C);
          for (int c = 0; c < C; ++c) ga += dr[c] * (up[c] - dn[c]) / (2.0 * h);
        } else if (C > 1) {
          ga = std::nan("");  // rates were set explicitly: no alpha
        }
        d_alpha[p] = ga;
      }
    }
    return NRX_OK;
  });
}

// First and second derivatives of lnL in each reticulation probability,
// from the mixture weights of the last evaluation.  A class's probability
// is multilinear in the probabilities, so each partition's likelihood (sum
// of class weights, AVERAGE) or winning class (BEST) gives
// d2 lnL_p / dp_r^2 = -(d lnL_p / dp_r)^2: the second derivative is
// -sum_p (d1_{p,r})^2, with d1_{p,r} = sum_k alpha_{p,k} dlogP_k/dp_r.
static void prob_derivatives(const Engine &e, double *d1, double *d2) {
  const int nc = static_cast<int>(e.dag.classes.size()), R = e.R;
  const double *alpha = e.h_result + e.rl.alpha;
  for (int r = 0; r < R; ++r) {
    double s1 = 0.0, s2 = 0.0;
    for (int p = 0; p < e.P; ++p) {
      double g = 0.0;
      for (int k = 0; k < nc; ++k) g += alpha[p * nc + k] * e.class_dlogprob[static_cast<size_t>(k) * R + r];
      s1 += g;
      s2 -= g * g;
    }
    if (d1) d1[r] = s1;
    if (d2) d2[r] = s2;
  }
}

int nrx_prob_derivatives(nrx_engine *eng, int32_t variant, double *total, double *d1, double *d2) {
  return guarded([&]() -> int {
    if (!eng || !total) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    const int rc = evaluate(e, variant, true, true, false);
    if (rc != NRX_OK) return rc;
    *total = e.cached_total;
    prob_derivatives(e, d1, d2);
    return NRX_OK;
  });
}

// Rounds of per-reticulation Newton searches over `retics` (all when
// null) of one engine (see nrx_optimize_reticulations).
static int optimize_reticulations_impl(Engine &e, int variant, int max_iters, const int32_t *retics, int n_retics,
                                       double brprob_min, double brprob_max, double tolerance, double *logl) {
  {
    std::vector<int> list;
    if (retics) {
      for (int i = 0; i < n_retics; ++i) {
        if (retics[i] < 0 || retics[i] >= e.R) throw std::invalid_argument("reticulation index out of range");
        list.push_back(retics[i]);
      }
    } else {
      for (int r = 0; r < e.R; ++r) list.push_back(r);
    }
    if (!(brprob_min > 0.0 && brprob_max < 1.0 && brprob_min < brprob_max) || !(tolerance > 0.0))
      throw std::invalid_argument("bad reticulation probability bounds or tolerance");
    int rc = evaluate(e, variant, true, true, false);
    if (rc != NRX_OK) return rc;
    std::vector<double> probs = e.probs, d1(e.R), d2(e.R);
    // lnL, d1 and d2 of reticulation r at probability x
    auto probe = [&](int r, double x, double &f, double &g, double &h) -> int {
      probs[r] = x;
      set_probs(e, probs.data());
      const int c = evaluate(e, variant, true, true, false);
      if (c != NRX_OK) return c;
      prob_derivatives(e, d1.data(), d2.data());
      f = e.cached_total;
      g = d1[r];
      h = d2[r];
      return NRX_OK;
    };
    double act = e.cached_total;
    for (int iter = 0; iter < max_iters; ++iter) {
      double loop = act;
      for (int r : list) {
        // safeguarded Newton on d lnL / dp_r = 0 inside [lo, hi]: lnL is
        // concave in p_r when the used classes stay the same, so the slope's
        // sign brackets the maximum; a step that leaves the bracket bisects
        // it, and a bound is tried directly when the slope points past it
        double x = probs[r], f, g, h;
        rc = probe(r, x, f, g, h);
        if (rc != NRX_OK) return rc;
        double best_x = x, best_f = f, lo = brprob_min, hi = brprob_max;
        bool tried_lo = false, tried_hi = false;
        for (int it = 0; it < 100; ++it) {
          if (g > 0.0)
            lo = x;
          else if (g < 0.0)
            hi = x;
          else
            break;
          if (hi - lo <= tolerance) break;
          double xn = h < 0.0 ? x - g / h : std::nan("");
          if (!(xn > lo && xn < hi)) {
            if (g > 0.0 && hi == brprob_max && !tried_hi) {
              xn = brprob_max;
              tried_hi = true;
            } else if (g < 0.0 && lo == brprob_min && !tried_lo) {
              xn = brprob_min;
              tried_lo = true;
            } else {
              xn = 0.5 * (lo + hi);
            }
          }
          const double step = std::fabs(xn - x);
          x = xn;
          rc = probe(r, x, f, g, h);
          if (rc != NRX_OK) return rc;
          if (f > best_f) {
            best_f = f;
            best_x = x;
          }
          if (step <= tolerance) break;
        }
        if (probs[r] != best_x) {
          probs[r] = best_x;
          set_probs(e, probs.data());
          rc = evaluate(e, variant, true, true, false);
          if (rc != NRX_OK) return rc;
        }
        loop = e.cached_total;
      }
      const bool same = loop == act;
      act = loop;
      if (same) break;
    }
    if (logl) *logl = act;
    return NRX_OK;
  }
}

int nrx_optimize_reticulations(nrx_engine *eng, int32_t variant, int32_t max_iters, double brprob_min,
                               double brprob_max, double tolerance, double *logl) {
  return guarded([&]() -> int {
    if (!eng) throw std::invalid_argument("null argument");
    return optimize_reticulations_impl(*static_cast<Engine *>(eng), variant, max_iters, nullptr, 0, brprob_min,
                                       brprob_max, tolerance, logl);
  });
}

int nrx_get_reticulation_probs(nrx_engine *eng, double *probs) {
  return guarded([&]() -> int {
    if (!eng || !probs) throw std::invalid_argument("null argument");
    const Engine &e = *static_cast<Engine *>(eng);
    std::copy(e.probs.begin(), e.probs.end(), probs);
    return NRX_OK;
  });
}

int nrx_tree_logls(nrx_engine *eng, double *tree_logl, double *class_logprob) {
  return guarded([&]() -> int {
    if (!eng) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    if (!e.cache_valid) throw std::runtime_error("no evaluation yet");
    const int nc = static_cast<int>(e.dag.classes.size());
    if (tree_logl)
      std::copy(e.h_result + e.rl.tree_logl, e.h_result + e.rl.tree_logl + e.P * nc, tree_logl);
    if (class_logprob) std::copy(e.class_logprob.begin(), e.class_logprob.end(), class_logprob);
    return NRX_OK;
  });
}

int nrx_get_stats(nrx_engine *eng, nrx_stats *out) {
  return guarded([&]() -> int {
    if (!eng || !out) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    std::memset(out, 0, sizeof(*out));
    out->n_variants = static_cast<int>(e.dag.var.size());
    out->n_ops_needed = static_cast<int>(e.ops_all.size());
    out->n_levels = static_cast<int>(e.level_ops.size()) - 1;
    out->n_classes = static_cast<int>(e.dag.classes.size());
    out->n_root_clvs = static_cast<int>(e.root_slots.size());
    out->n_uses = e.n_uses;
    for (const DevOp &op : e.ops_all)
      for (int k = 0; k < op.nk; ++k) out->n_inner_uses += op.kid[k] >= 0;
    out->n_sites_padded = e.L.n_sites_pad;
    // algorithmic bytes over the real (unpadded) sites
    double sites = 0.0;
    for (const auto &p : e.parts) sites += p.n_sites;
    const double b = e.opt.precision == NRX_FP64 ? 8.0 : 4.0;
    const double row = sites * e.C * e.S * b;  // one CLV over all sites
    const double scl = 4.0 * sites;            // one scaler row
    const double tip_b = (e.S == 4 ? 1.0 : 4.0) * sites;
    double fwd_ops = 0.0;
    for (const DevOp &op : e.ops_all) {
      for (int k = 0; k < op.nk; ++k) fwd_ops += op.kid[k] >= 0 ? row + scl : tip_b;
      fwd_ops += row + scl;  // parent CLV and scaler written
    }
    const double root = e.root_slots.size() * (row + scl + 8.0 * sites);
    const double pm = static_cast<double>(e.P) * e.E * e.C * e.S * e.S * b;
    // push-form adjoint, over the op list and flags of the engine's form:
    // per op read its scaler, its adjoint (when accumulated) and its children
    // with their scalers; per inner child write the delivered adjoint, and
    // read it back first when another consumer delivered before.  A root op
    // reads the seed factor (plain form, 8 bytes per site) or its own CLV
    // (FP32 form).  Adjoints handed over through LDS (FP32, 4 states) move no
    // HBM bytes.  Adjoint exponents: per site, or per (site, category) for
    // 20-state FP32.
    const bool fp32 = e.opt.precision == NRX_FP32;
    const double aexp = fp32 ? (e.S == AA_STATES ? scl * e.C : scl) : 0.0;
    const bool pass = fp32 && e.S == 4;
    out->n_lds_handoffs = 0;
    if (pass)
      for (const DevOp &op : e.ops_dfs)
        out->n_lds_handoffs += ((op.flags & OP_PASS_KID0) != 0) + ((op.flags & OP_PASS_KID1) != 0);
    double adj = 0.0;
    for (const DevOp &op : fp32 ? e.ops_dfs : e.ops_all) {
      adj += scl;
      if ((op.flags & OP_HAS_CONSUMERS) && !(pass && (op.flags & OP_A_PASS))) adj += row + aexp;
      if (op.root >= 0) adj += fp32 ? row : 8.0 * sites;
      for (int k = 0; k < op.nk; ++k) {
        if (op.kid[k] < 0) {
          adj += tip_b;
          continue;
        }
        adj += row + scl;  // child CLV and scaler
        if (pass && (op.flags & (k == 0 ? OP_PASS_KID0 : OP_PASS_KID1))) continue;
        adj += row + aexp;  // delivered adjoint
        if (!(op.flags & (k == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1))) adj += row + aexp;
      }
    }
    adj += pm;
    out->fwd_op_bytes = fwd_ops;
    out->root_bytes = root;
    out->adj_bytes = adj;
    out->fwd_bytes = pm + fwd_ops + root;
    out->grad_bytes = out->fwd_bytes + adj;
    out->clv_bytes = static_cast<double>(e.d_clv.bytes + e.d_scl.bytes + e.d_adj.bytes);
    return NRX_OK;
  });
}

int nrx_last_timing(nrx_engine *eng, nrx_timing *out) {
  return guarded([&]() -> int {
    if (!eng || !out) throw std::invalid_argument("null argument");
    *out = static_cast<Engine *>(eng)->last_timing;
    return NRX_OK;
  });
}

int nrx_set_memory_budget(nrx_engine *eng, uint64_t bytes) {
  return guarded([&]() -> int {
    if (!eng) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    if (bytes != 0 && e.opt.precision != NRX_FP64)
      return fail(NRX_ERR_UNSUPPORTED, "memory budgets are provided for the FP64 form");
    const uint64_t old = e.mem_budget;
    e.mem_budget = bytes;
    try {
      plan_budget(e);
    } catch (...) {
      e.mem_budget = old;  // the previous plan stays in force
      throw;
    }
    return NRX_OK;
  });
}

int nrx_memory_plan(nrx_engine *eng, nrx_memory_plan_info *out) {
  return guarded([&]() -> int {
    if (!eng || !out) throw std::invalid_argument("null argument");
    const Engine &e = *static_cast<Engine *>(eng);
    out->segments = std::max(1, e.bp.n_seg);
    const int nv = static_cast<int>(e.dag.var.size());
    out->rows = e.bp.n_seg > 1 ? e.bp.n_phys : nv;
    out->boundary_rows = e.bp.n_seg > 1 ? e.bp.n_boundary : nv;
    out->resident_rows = nv;
    out->row_bytes = static_cast<double>(budget_row_bytes(e));
    return NRX_OK;
  });
}

int nrx_batch_optimize(nrx_batch *batch, int32_t variant, int32_t mode, const int32_t *edge_begin,
                       const int32_t *edges, int32_t brlen_iters, const int32_t *retic_begin,
                       const int32_t *retics, int32_t prob_iters, double brprob_min, double brprob_max,
                       double prob_tolerance, int32_t n_threads, double *lnl, int32_t *status) {
  return guarded([&]() -> int {
    if (!batch || !edge_begin || !retic_begin || !lnl) throw std::invalid_argument("null argument");
    Batch &b = *reinterpret_cast<Batch *>(batch);
    const int n = static_cast<int>(b.engines.size());
    for (int i = 0; i < n; ++i) {
      if (edge_begin[i + 1] < edge_begin[i] || retic_begin[i + 1] < retic_begin[i])
        throw std::invalid_argument("candidate lists must be given by non-decreasing offsets");
      if ((edge_begin[i + 1] > edge_begin[i] && !edges) || (retic_begin[i + 1] > retic_begin[i] && !retics))
        throw std::invalid_argument("null candidate list");
    }
    if (brlen_iters < 0 || prob_iters < 0) throw std::invalid_argument("iteration counts must be >= 0");
    // every engine runs on its own stream: the workers' host waits overlap
    // and the device interleaves the candidates' kernels
    const int workers = std::max(1, std::min(n_threads > 0 ? n_threads : 16, n));
    std::atomic<int> next{0};
    std::vector<int> rcs(n, NRX_OK);
    std::vector<std::string> msgs(n);
    auto work = [&]() {
      (void)hipSetDevice(b.engines[0]->opt.device);
      for (int i = next++; i < n; i = next++) {
        Engine &e = *b.engines[i];
        rcs[i] = guarded([&]() -> int {
          double l = 0.0;
          int rc = NRX_OK;
          const int ne = edge_begin[i + 1] - edge_begin[i];
          if (ne > 0 && brlen_iters > 0)
            rc = optimize_branches_impl(e, variant, mode, edges + edge_begin[i], ne, brlen_iters, &l);
          const int nr = retic_begin[i + 1] - retic_begin[i];
          if (rc == NRX_OK && nr > 0 && prob_iters > 0)
            rc = optimize_reticulations_impl(e, variant, prob_iters, retics + retic_begin[i], nr, brprob_min,
                                             brprob_max, prob_tolerance, &l);
          return rc;
        });
        if (rcs[i] != NRX_OK) msgs[i] = g_last_error;
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < workers; ++t) pool.emplace_back(work);
    work();
    for (std::thread &t : pool) t.join();
    int first_bad = -1;
    for (int i = 0; i < n; ++i) {
      if (status) status[i] = rcs[i];
      if (rcs[i] != NRX_OK && first_bad < 0) first_bad = i;
    }
    // the optimised candidates scored together in one batched launch
    const int rc = nrx_batch_evaluate(batch, variant, lnl, nullptr, nullptr, nullptr);
    if (first_bad >= 0 && !status)
      return fail(rcs[first_bad], "candidate " + std::to_string(first_bad) + ": " + msgs[first_bad]);
    return rc;
  });
}

}  // extern "C"

namespace nrx {

// Shapes of the FP32 form's persistent kernels: 4-state data takes G ops
// per group, TS sites per workgroup (TS C threads, a multiple of 64) and H
// lanes per (site, category) row in the reverse sweep; 20-state data has 32
// sites per workgroup and tunes G only.
struct Shape {
  int g, ts, h;
  bool operator==(const Shape &o) const { return g == o.g && ts == o.ts && h == o.h; }
};

static std::vector<Shape> shape_candidates(const Engine &e) {
  std::vector<Shape> out;
  if (e.S == AA_STATES) {
    for (int g : {1, 2, 4}) out.push_back({g, AA_TILE, 1});
  } else {
    for (int h : {1, 2})
      for (int ts : {64, 32, 16})
        if (ts * e.C % 64 == 0)
          for (int g : {4, 2, 3, 8}) out.push_back({g, ts, h});
  }
  return out;
}

static void apply_shape(Engine &e, const Shape &s) {
  if (e.tune_group == s.g && e.tune_tile == s.ts && e.tune_lanes == s.h) return;
  e.tune_group = s.g;
  e.tune_tile = s.ts;
  e.tune_lanes = s.h;
  build_schedule(e);
  std::fill(e.clv_dirty.begin(), e.clv_dirty.end(), 1);
  e.cache_valid = false;
}

}  // namespace nrx

extern "C" {

int nrx_set_tuning(nrx_engine *eng, int32_t group_width, int32_t tile_sites, int32_t row_lanes) {
  return guarded([&]() -> int {
    if (!eng) throw std::invalid_argument("null argument");
    Engine &e = *static_cast<Engine *>(eng);
    if (e.opt.precision != NRX_FP32)
      return fail(NRX_ERR_UNSUPPORTED, "kernel shapes are tuned for the FP32 form only");
    const auto cands = shape_candidates(e);
    const Shape s{group_width, tile_sites, row_lanes};
    if (std::find(cands.begin(), cands.end(), s) == cands.end())
      throw std::invalid_argument("unsupported kernel shape for this engine");
    apply_shape(e, s);
    return NRX_OK;
  });
}

int nrx_get_tuning(nrx_engine *eng, nrx_tuning *out) {
  return guarded([&]() -> int {
    if (!eng || !out) throw std::invalid_argument("null argument");
    const Engine &e = *static_cast<Engine *>(eng);
    out->group_width = e.tune_group;
    out->tile_sites = e.S == AA_STATES ? AA_TILE : e.tune_tile;
    out->row_lanes = e.tune_lanes;
    out->ms = 0.0;
    return NRX_OK;
  });
}

int nrx_autotune(nrx_engine *eng, int32_t variant, int32_t reps, nrx_tuning *best,
                 nrx_tuning *results, int32_t max_results, int32_t *n_results) {
  return guarded([&]() -> int {
    if (!eng || !best || reps < 1) throw std::invalid_argument("bad autotune arguments");
    Engine &e = *static_cast<Engine *>(eng);
    if (e.opt.precision != NRX_FP32)
      return fail(NRX_ERR_UNSUPPORTED, "kernel shapes are tuned for the FP32 form only");
    const auto cands = shape_candidates(e);
    const int32_t saved_timing = e.opt.timing;
    e.opt.timing = 1;
    std::vector<double> ms(cands.size(), INFINITY);
    int rc = NRX_OK;
    try {
      for (size_t i = 0; i < cands.size() && rc == NRX_OK; ++i) {
        apply_shape(e, cands[i]);
        rc = evaluate(e, variant, false, true, true);  // warm-up (and graph capture)
        for (int r = 0; r < reps && rc == NRX_OK; ++r) {
          rc = evaluate(e, variant, false, true, true);
          ms[i] = std::min(ms[i], static_cast<double>(e.last_timing.total_ms));
        }
      }
    } catch (...) {
      e.opt.timing = saved_timing;
      throw;
    }
    e.opt.timing = saved_timing;
    if (rc != NRX_OK) return rc;
    const size_t bi = static_cast<size_t>(std::min_element(ms.begin(), ms.end()) - ms.begin());
    apply_shape(e, cands[bi]);
    *best = nrx_tuning{cands[bi].g, cands[bi].ts, cands[bi].h, ms[bi]};
    for (size_t i = 0; results && i < cands.size() && static_cast<int32_t>(i) < max_results; ++i)
      results[i] = nrx_tuning{cands[i].g, cands[i].ts, cands[i].h, ms[i]};
    if (n_results) *n_results = static_cast<int32_t>(cands.size());
    return NRX_OK;
  });
}

}  // extern "C"
