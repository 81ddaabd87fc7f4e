// Schedules of the op DAG for the persistent FP32 kernels (see schedule.hpp).
#include "schedule.hpp"

#include <algorithm>
#include <stdexcept>
#include <utility>

namespace nrx {

// Sets OP_HAS_CONSUMERS / OP_FIRST_KID* for a list of ops walked by the
// adjoint in the order given by `adj_order` (indices into ops).
void set_adjoint_flags(std::vector<DevOp> &ops, const std::vector<int> &adj_order, int nv) {
  std::vector<char> has_cons(nv, 0), delivered(nv, 0);
  for (const DevOp &op : ops)
    for (int k = 0; k < op.nk; ++k)
      if (op.kid[k] >= 0) has_cons[op.kid[k]] = 1;
  for (int i : adj_order) {
    DevOp &op = ops[i];
    op.flags = has_cons[op.out] ? OP_HAS_CONSUMERS : 0;
    for (int k = 0; k < op.nk; ++k) {
      const int32_t kid = op.kid[k];
      if (kid >= 0 && !delivered[kid]) {
        op.flags |= k == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1;
        delivered[kid] = 1;
      }
    }
  }
}

void split_level_runs(std::vector<DevOp> &ops, const std::vector<LevelChunk> &chunks, std::vector<LevelChunk> &runs,
                      std::vector<int> &run_begin) {
  runs.clear();
  run_begin.assign(1, 0);
  std::vector<DevOp> left, next, ordered;
  std::vector<int32_t> taken;
  for (const LevelChunk &ch : chunks) {
    left.assign(ops.begin() + ch.op_begin, ops.begin() + ch.op_end);
    ordered.clear();
    while (!left.empty()) {
      LevelChunk run = ch;
      run.op_begin = ch.op_begin + static_cast<int>(ordered.size());
      taken.clear();
      next.clear();
      for (const DevOp &op : left) {
        bool clash = false;
        for (int k = 0; k < op.nk; ++k)
          if (op.kid[k] >= 0 && std::find(taken.begin(), taken.end(), op.kid[k]) != taken.end()) clash = true;
        if (clash) {
          next.push_back(op);
          continue;
        }
        for (int k = 0; k < op.nk; ++k)
          if (op.kid[k] >= 0) taken.push_back(op.kid[k]);
        ordered.push_back(op);
      }
      run.op_end = ch.op_begin + static_cast<int>(ordered.size());
      runs.push_back(run);
      left.swap(next);
    }
    std::copy(ordered.begin(), ordered.end(), ops.begin() + ch.op_begin);
    run_begin.push_back(static_cast<int>(runs.size()));
  }
}

// OP_ACC_EARLY*: for the FP32 form's reverse sweep (groups backwards, ops in
// index order), a delivery that accumulates into a child adjoint last
// written in an earlier group (not earlier in its own group), so the
// read-back is issued with the group's other loads.  Those loads follow the
// earlier group's stores in the wave's program order, and a wave's loads
// see its own earlier stores to the same address.
void set_acc_early(std::vector<DevOp> &ops, const std::vector<int32_t> &group_begin, int nv) {
  std::vector<int> last(nv, -1);  // sweep position of the group that last delivered to a slot
  const int ng = static_cast<int>(group_begin.size()) - 1;
  std::vector<int32_t> mine;
  for (int g = ng - 1, t = 0; g >= 0; --g, ++t) {
    mine.clear();
    for (int i = group_begin[g]; i < group_begin[g + 1]; ++i) {
      DevOp &op = ops[i];
      op.flags &= ~(OP_ACC_EARLY0 | OP_ACC_EARLY1);
      for (int k = 0; k < op.nk; ++k) {
        const int32_t kid = op.kid[k];
        if (kid < 0) continue;
        const bool first = (op.flags & (k == 0 ? OP_FIRST_KID0 : OP_FIRST_KID1)) != 0;
        const bool in_group = std::find(mine.begin(), mine.end(), kid) != mine.end();
        if (!first && !in_group && last[kid] >= 0 && t - last[kid] >= 1)
          op.flags |= k == 0 ? OP_ACC_EARLY0 : OP_ACC_EARLY1;
        mine.push_back(kid);
      }
    }
    for (int32_t kid : mine) last[kid] = t;
  }
}

// OP_PASS_* for the FP32 4-state reverse sweep (groups backwards): every
// delivery to a child held in one of n_slots (at most PASS_SLOTS) per-thread
// LDS slots goes there instead of HBM.  The child's first delivery writes
// the slot, later ones add to it, and the child's op reads it in its group's
// load phase.  A slot is held from the group of the child's first delivery
// to the group of the child's op; that op's load precedes the group's
// deliveries, so the slot can take a new child in that same group.
// Children are taken by earliest release (interval scheduling, which keeps
// the most children for a number of slots), each in the free slot released
// last.
void set_pass_slots(std::vector<DevOp> &ops, const std::vector<int32_t> &group_begin, int nv,
                           int n_slots) {
  const int ng = static_cast<int>(group_begin.size()) - 1;
  std::vector<int> op_of(nv, -1), release(nv, -1), start(nv, -1), slot_of(nv, -1);
  for (int g = ng - 1, t = 0; g >= 0; --g, ++t)
    for (int i = group_begin[g]; i < group_begin[g + 1]; ++i) {
      op_of[ops[i].out] = i;
      release[ops[i].out] = t;  // sweep position of the child's own op
      for (int k = 0; k < ops[i].nk; ++k)
        if (ops[i].kid[k] >= 0 && start[ops[i].kid[k]] < 0) start[ops[i].kid[k]] = t;
    }
  for (DevOp &op : ops) op.flags &= ~(OP_PASS_KID0 | OP_PASS_KID1 | OP_A_PASS | PASS_BITS);
  std::vector<int> cand;
  for (int v = 0; v < nv; ++v)
    if (start[v] >= 0 && op_of[v] >= 0 && release[v] > start[v]) cand.push_back(v);
  std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) { return release[x] < release[y]; });
  std::vector<int> busy_until(n_slots, -1);  // sweep position of the group that reads the slot
  for (int v : cand) {
    int best = -1;
    for (int q = 0; q < n_slots; ++q)
      if (busy_until[q] <= start[v] && (best < 0 || busy_until[q] > busy_until[best])) best = q;
    if (best < 0) continue;
    busy_until[best] = release[v];
    slot_of[v] = best;
  }
  for (DevOp &op : ops) {
    for (int k = 0; k < op.nk; ++k) {
      const int32_t kid = op.kid[k];
      if (kid < 0 || slot_of[kid] < 0) continue;
      op.flags |= (k == 0 ? OP_PASS_KID0 : OP_PASS_KID1) | (slot_of[kid] << (PASS_SHIFT + 3 * k));
      op.flags &= ~(k == 0 ? OP_ACC_EARLY0 : OP_ACC_EARLY1);  // no HBM read-back
    }
    if (slot_of[op.out] >= 0) op.flags |= OP_A_PASS | (slot_of[op.out] << (PASS_SHIFT + 6));
  }
}

void build_batched_order(const Dag &dag, const std::vector<int> &root_slots, const std::vector<DevOp> &proto,
                         int group, std::vector<DevOp> &ops, std::vector<int32_t> &group_begin) {
  const int nv = static_cast<int>(dag.var.size());
  std::vector<int> post;
  {
    std::vector<char> seen(nv, 0);
    std::vector<std::pair<int, int>> stack;
    for (int rs : root_slots) {
      if (seen[rs]) continue;
      stack.push_back({rs, 0});
      seen[rs] = 1;
      while (!stack.empty()) {
        auto &top = stack.back();
        const Variant &v = dag.var[top.first];
        if (top.second < v.nkids) {
          const int32_t kid = v.kid[top.second++];
          if (kid >= 0 && !seen[kid]) {
            seen[kid] = 1;
            stack.push_back({kid, 0});
          }
        } else {
          post.push_back(top.first);
          stack.pop_back();
        }
      }
    }
  }
  {
    std::vector<int> pos(nv, -1), deps(nv, 0);
    std::vector<std::vector<int>> parents(nv);
    for (size_t i = 0; i < post.size(); ++i) pos[post[i]] = static_cast<int>(i);
    for (int s : post)
      for (int k = 0; k < dag.var[s].nkids; ++k)
        if (dag.var[s].kid[k] >= 0) {
          deps[s]++;
          parents[dag.var[s].kid[k]].push_back(s);
        }
    // `ready` holds ops whose inputs all come from groups before the last
    // one formed, `fresh` those the last group made ready.  A group is taken
    // from `ready` when possible, so that it does not consume the group
    // before it (the kernels then overlap the two groups' memory traffic);
    // only when `ready` is empty does a group take `fresh` ops.
    auto by_pos = [&](int a, int b) { return pos[a] < pos[b]; };
    std::vector<int> ready, fresh;  // both kept sorted by post-order index
    for (int s : post)
      if (deps[s] == 0) ready.push_back(s);
    std::sort(ready.begin(), ready.end(), by_pos);
    std::vector<char> group_indep;
    ops.clear();
    group_begin.assign(1, 0);
    while (!ready.empty() || !fresh.empty()) {
      const bool indep = !ready.empty();
      std::vector<int> &src = indep ? ready : fresh;
      // up to `group` ops in post-order, skipping an op that shares a child
      // CLV with one already taken: two deliveries to one child inside a
      // group would make the second read back the first's store at once
      std::vector<int> grp;
      std::vector<int32_t> kids_taken;
      for (size_t i = 0; i < src.size() && static_cast<int>(grp.size()) < group;) {
        const Variant &v = dag.var[src[i]];
        bool clash = false;
        for (int k = 0; k < v.nkids; ++k)
          if (v.kid[k] >= 0 && std::find(kids_taken.begin(), kids_taken.end(), v.kid[k]) != kids_taken.end())
            clash = true;
        if (clash && !grp.empty()) {
          ++i;
          continue;
        }
        for (int k = 0; k < v.nkids; ++k)
          if (v.kid[k] >= 0) kids_taken.push_back(v.kid[k]);
        grp.push_back(src[i]);
        src.erase(src.begin() + static_cast<std::ptrdiff_t>(i));
      }
      // what the previous group made ready no longer depends on the group
      // after this one
      for (int s : fresh) ready.insert(std::upper_bound(ready.begin(), ready.end(), s, by_pos), s);
      fresh.clear();
      for (int s : grp) {
        ops.push_back(proto[s]);
        for (int par : parents[s])
          if (--deps[par] == 0) fresh.push_back(par);
      }
      std::sort(fresh.begin(), fresh.end(), by_pos);
      group_indep.push_back(indep && !group_indep.empty());
      group_begin.push_back(static_cast<int32_t>(ops.size()));
    }
    if (ops.size() != post.size()) throw std::logic_error("op grouping lost ops");
    std::vector<int> order;  // adjoint: groups backwards, ops in index order
    for (int g = static_cast<int>(group_begin.size()) - 2; g >= 0; --g)
      for (int i = group_begin[g]; i < group_begin[g + 1]; ++i) order.push_back(i);
    set_adjoint_flags(ops, order, nv);
    for (size_t g = 0; g < group_indep.size(); ++g)
      if (group_indep[g]) ops[group_begin[g]].flags |= OP_GROUP_INDEP;
    set_acc_early(ops, group_begin, nv);
  }
}

}  // namespace nrx
