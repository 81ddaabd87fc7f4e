// Host-side check of the FP32 op schedule (netrax_amd/csrc/schedule.cpp):
// builds the grouped order and the reverse sweep's flags for one network
// and replays the sweep's memory semantics, reporting every violation.
//
// Input (stdin): n_nodes n_tips root n_edges n_retic, then n_edges edge
// parents, n_edges edge children, n_retic first and n_retic second parent
// edges, then pairs "group_width pass_slots" until end of input.
// Output: one line per pair, "ok <counts>" or "error <what>" lines.
#include <algorithm>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "../netrax_amd/csrc/schedule.hpp"

using namespace nrx;

static int check(const Dag &dag, int group, int n_slots) {
  const int nv = static_cast<int>(dag.var.size());
  std::vector<int> roots;
  for (const auto &c : dag.classes)
    if (std::find(roots.begin(), roots.end(), c.dtroot_slot) == roots.end()) roots.push_back(c.dtroot_slot);
  std::vector<int> root_of(nv, -1);
  for (size_t r = 0; r < roots.size(); ++r) root_of[roots[r]] = static_cast<int>(r);
  std::vector<DevOp> proto(nv);
  for (int v = 0; v < nv; ++v) {
    DevOp op{};
    op.out = v;
    op.nk = dag.var[v].nkids;
    for (int k = 0; k < 2; ++k) {
      op.kid[k] = k < op.nk ? dag.var[v].kid[k] : NO_CHILD;
      op.edge[k] = k < op.nk ? dag.var[v].edge[k] : 0;
    }
    op.root = root_of[v];
    proto[v] = op;
  }
  std::vector<DevOp> ops;
  std::vector<int32_t> gb;
  build_batched_order(dag, roots, proto, group, ops, gb);
  set_pass_slots(ops, gb, nv, n_slots);
  const int ng = static_cast<int>(gb.size()) - 1;
  int errors = 0;
  auto fail = [&](const std::string &what) {
    if (errors++ < 20) std::printf("error %s\n", what.c_str());
  };

  // order: each needed variant once, children in earlier groups, groups of
  // at most `group` ops, OP_GROUP_INDEP only where the group before is not consumed
  std::vector<int> group_of(nv, -1), consumers(nv, 0);
  for (int g = 0; g < ng; ++g) {
    if (gb[g + 1] - gb[g] > group || gb[g + 1] <= gb[g]) fail("group size");
    for (int i = gb[g]; i < gb[g + 1]; ++i) {
      if (group_of[ops[i].out] >= 0) fail("variant twice");
      group_of[ops[i].out] = g;
    }
  }
  for (int g = 0; g < ng; ++g)
    for (int i = gb[g]; i < gb[g + 1]; ++i)
      for (int k = 0; k < ops[i].nk; ++k) {
        const int kid = ops[i].kid[k];
        if (kid < 0) continue;
        consumers[kid]++;
        if (group_of[kid] < 0 || group_of[kid] >= g) fail("child not in an earlier group");
        if ((ops[gb[g]].flags & OP_GROUP_INDEP) && group_of[kid] == g - 1) fail("OP_GROUP_INDEP consumes the group before");
      }
  for (int r : roots)
    if (group_of[r] < 0) fail("root variant missing");

  // reverse sweep, groups backwards: first deliveries, early read-backs and
  // LDS slots (a slot read must see its own child with every delivery made;
  // a first slot write must find the slot free)
  std::vector<int> delivered(nv, 0), last_group(nv, -1), owner(PASS_SLOTS, -1);
  int n_pass = 0, n_reads = 0;
  for (int g = ng - 1, t = 0; g >= 0; --g, ++t) {
    for (int i = gb[g]; i < gb[g + 1]; ++i) {
      const DevOp &op = ops[i];
      if (((op.flags & OP_HAS_CONSUMERS) != 0) != (consumers[op.out] > 0)) fail("OP_HAS_CONSUMERS");
      if (op.flags & OP_A_PASS) {
        const int q = pass_slot(op.flags, 2);
        n_reads++;
        if (q >= n_slots || owner[q] != op.out || delivered[op.out] != consumers[op.out]) fail("slot read");
        if (q < PASS_SLOTS) owner[q] = -1;
      } else if (consumers[op.out] > 0 && delivered[op.out] != consumers[op.out]) {
        fail("own adjoint read before every delivery");
      }
    }
    std::vector<int> mine;
    for (int i = gb[g]; i < gb[g + 1]; ++i) {
      const DevOp &op = ops[i];
      for (int k = 0; k < op.nk; ++k) {
        const int kid = op.kid[k];
        if (kid < 0) continue;
        const bool first = (op.flags & (k == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1)) != 0;
        if (first != (delivered[kid] == 0)) fail("OP_FIRST_KID");
        const bool early = (op.flags & (k == 0 ? OP_ACC_EARLY0 : OP_ACC_EARLY1)) != 0;
        const bool in_group = std::find(mine.begin(), mine.end(), kid) != mine.end();
        if (early && (first || in_group || last_group[kid] < 0 || t - last_group[kid] < 1)) fail("OP_ACC_EARLY");
        if (op.flags & (k == 0 ? OP_PASS_KID0 : OP_PASS_KID1)) {
          const int q = pass_slot(op.flags, k);
          n_pass++;
          if (q >= n_slots) fail("slot index");
          else if (first && owner[q] != -1) fail("first slot write over a live slot");
          else if (!first && owner[q] != kid) fail("slot accumulate into another child's slot");
          if (q < PASS_SLOTS) owner[q] = kid;
          if (early) fail("early read-back of a slot delivery");
        }
        delivered[kid]++;
        mine.push_back(kid);
      }
    }
    for (int kid : mine) last_group[kid] = t;
  }
  for (int q = 0; q < PASS_SLOTS; ++q)
    if (owner[q] != -1) fail("slot never read");
  if (n_slots == 0 && n_pass > 0) fail("slot deliveries without slots");
  if (errors == 0)
    std::printf("ok ops %zu groups %d slot_deliveries %d slot_reads %d\n", ops.size(), ng, n_pass, n_reads);
  return errors;
}

// Plain form: level chunks split into runs of ops without a common child,
// replayed as concurrent deliveries: every run's ops find their own adjoint
// complete (all consumers delivered), a run never delivers twice to one
// child, and the first delivery to a child (and only it) writes.
static int check_plain(const Dag &dag) {
  const int nv = static_cast<int>(dag.var.size());
  int max_level = 0;
  for (const auto &v : dag.var) max_level = std::max(max_level, v.level);
  std::vector<DevOp> ops;
  std::vector<LevelChunk> chunks;
  for (int l = 1; l <= max_level; ++l) {
    LevelChunk ch{};
    ch.level = l;
    ch.op_begin = static_cast<int>(ops.size());
    for (int s = 0; s < nv; ++s)
      if (dag.var[s].level == l && dag.var[s].nkids > 0) {
        DevOp op{};
        op.out = s;
        op.nk = dag.var[s].nkids;
        for (int k = 0; k < 2; ++k) op.kid[k] = k < op.nk ? dag.var[s].kid[k] : NO_CHILD;
        ops.push_back(op);
      }
    ch.op_end = static_cast<int>(ops.size());
    if (ch.op_end > ch.op_begin) chunks.push_back(ch);
  }
  const std::vector<DevOp> before = ops;
  std::vector<LevelChunk> runs;
  std::vector<int> run_begin;
  split_level_runs(ops, chunks, runs, run_begin);
  int errors = 0;
  auto fail = [&](const std::string &what) {
    std::printf("error plain %s\n", what.c_str());
    ++errors;
  };
  if (run_begin.size() != chunks.size() + 1) fail("run table size");
  for (size_t c = 0; c < chunks.size() && errors == 0; ++c) {
    std::vector<int> a, b;
    for (int i = chunks[c].op_begin; i < chunks[c].op_end; ++i) {
      a.push_back(before[i].out);
      b.push_back(ops[i].out);
    }
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    if (a != b) fail("chunk ops changed");
    int at = chunks[c].op_begin;
    for (int r = run_begin[c]; r < run_begin[c + 1]; ++r) {
      if (runs[r].op_begin != at || runs[r].op_end <= runs[r].op_begin) fail("runs not contiguous");
      at = runs[r].op_end;
    }
    if (at != chunks[c].op_end) fail("runs do not cover the chunk");
  }
  std::vector<int> order;
  for (int c = static_cast<int>(chunks.size()) - 1; c >= 0; --c)
    for (int i = chunks[c].op_begin; i < chunks[c].op_end; ++i) order.push_back(i);
  set_adjoint_flags(ops, order, nv);
  std::vector<int> consumers(nv, 0), delivered(nv, 0);
  for (const DevOp &op : ops)
    for (int k = 0; k < op.nk; ++k)
      if (op.kid[k] >= 0) consumers[op.kid[k]]++;
  for (int c = static_cast<int>(chunks.size()) - 1; c >= 0 && errors == 0; --c)
    for (int r = run_begin[c]; r < run_begin[c + 1]; ++r) {
      std::vector<int> mine;
      for (int i = runs[r].op_begin; i < runs[r].op_end; ++i) {
        const DevOp &op = ops[i];
        if (delivered[op.out] != consumers[op.out]) fail("adjoint read before all deliveries");
        for (int k = 0; k < op.nk; ++k) {
          const int kid = op.kid[k];
          if (kid < 0) continue;
          if (std::find(mine.begin(), mine.end(), kid) != mine.end()) fail("two deliveries to one child in a run");
          mine.push_back(kid);
          const bool first = op.flags & (k == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1);
          if (first != (delivered[kid] == 0)) fail("first-delivery flag");
        }
      }
      for (int kid : mine) delivered[kid]++;
    }
  int widest = 0;
  for (const LevelChunk &ch : chunks) widest = std::max(widest, ch.op_end - ch.op_begin);
  if (errors == 0) std::printf("ok plain_runs %zu plain_chunks %zu plain_widest %d\n", runs.size(), chunks.size(), widest);
  return errors;
}

int main() {
  nrx_network_desc d{};
  std::cin >> d.n_nodes >> d.n_tips >> d.root >> d.n_edges >> d.n_retic;
  std::vector<int32_t> ep(d.n_edges), ec(d.n_edges), rf(std::max(1, d.n_retic)), rs(std::max(1, d.n_retic));
  for (auto &x : ep) std::cin >> x;
  for (auto &x : ec) std::cin >> x;
  for (int i = 0; i < d.n_retic; ++i) std::cin >> rf[i];
  for (int i = 0; i < d.n_retic; ++i) std::cin >> rs[i];
  d.edge_parent = ep.data();
  d.edge_child = ec.data();
  d.retic_first_edge = rf.data();
  d.retic_second_edge = rs.data();
  const Dag dag = compile_dag(d);
  int errors = check_plain(dag), group = 0, slots = 0;
  while (std::cin >> group >> slots) errors += check(dag, group, slots);
  return errors ? 1 : 0;
}
