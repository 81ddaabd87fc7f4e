// Schedules of the op DAG for the persistent FP32 kernels: the grouped op
// order and the per-op flags the reverse sweep reads (first deliveries,
// early read-backs, LDS pass slots).  Host code only, no HIP calls, so the
// schedules can be built and checked without a GPU.
#pragma once

#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace nrx {

// OP_HAS_CONSUMERS / OP_FIRST_KID* for ops walked by the adjoint in the
// order `adj_order` (indices into ops).
void set_adjoint_flags(std::vector<DevOp> &ops, const std::vector<int> &adj_order, int nv);

// Plain form: splits every level chunk into runs of ops that share no child
// (first fit, in op order) and reorders the chunk's ops so that each run is
// contiguous.  The ops of a run can deliver their adjoints concurrently;
// runs are returned as op ranges, per chunk in order, with run_begin[c]
// the first run of chunk c (size chunks + 1).
void split_level_runs(std::vector<DevOp> &ops, const std::vector<LevelChunk> &chunks, std::vector<LevelChunk> &runs,
                      std::vector<int> &run_begin);

// OP_ACC_EARLY* for the FP32 reverse sweep (groups backwards).
void set_acc_early(std::vector<DevOp> &ops, const std::vector<int32_t> &group_begin, int nv);

// OP_PASS_* / OP_A_PASS for the FP32 4-state reverse sweep with n_slots
// per-thread LDS slots (at most PASS_SLOTS).
void set_pass_slots(std::vector<DevOp> &ops, const std::vector<int32_t> &group_begin, int nv, int n_slots);

// The FP32 op order: depth-first post-order from the root CLVs, then groups
// of up to `group` mutually independent ops (greedy list schedule, lowest
// post-order index first, preferring ops that do not consume the group
// before).  proto[v] is the op computing variant v.  Sets the adjoint
// flags, OP_GROUP_INDEP and OP_ACC_EARLY*; pass slots are left to
// set_pass_slots.
void build_batched_order(const Dag &dag, const std::vector<int> &root_slots, const std::vector<DevOp> &proto,
                         int group, std::vector<DevOp> &ops, std::vector<int32_t> &group_begin);

}  // namespace nrx
