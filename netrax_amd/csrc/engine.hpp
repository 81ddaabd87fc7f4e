// Internal structures of the MI355X network-likelihood engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/netrax_gpu.h"

namespace nrx {

// Sites of every partition are padded to a multiple of this tile; one tile
// is the unit of work of every site-parallel kernel.
constexpr int TILE = 64;
constexpr int MAX_CATS = 8;

// FP32 form for 20-state data (matrix cores): a workgroup handles a 32-site
// tile, CLVs and adjoints are stored per tile as [C][20][32], and each
// (partition, edge, category) has three operand tables (P, P^T, P') of
// 64 lanes x 12 floats in the lane order of the matrix instruction.
constexpr int AA_STATES = 20;
constexpr int AA_TILE = 32;
constexpr int AA_TABLE = 64 * 12;
constexpr int AA_KINDS = 2;  // operand tables per (partition, edge, category): P, P^T

// ---------------------------------------------------------------- model
struct EigenModel {
  int S = 0;
  std::vector<double> eval;   // [S]
  std::vector<double> V;      // [S][S]  P = V diag(exp(eval r t)) Vinv
  std::vector<double> Vinv;   // [S][S]
  std::vector<double> freqs;  // [S]
};

std::vector<double> gamma_mean_rates(double alpha, int ncat);
void build_model(int S, const double *subst, const double *freqs, EigenModel &m);

// ---------------------------------------------------------------- op DAG
// Child code: >= 0 is a variant slot, < 0 is tip (-code - 1).
constexpr int32_t NO_CHILD = INT32_MIN;
inline int32_t tip_code(int tip) { return -tip - 1; }
inline bool is_tip_code(int32_t c) { return c < 0 && c != NO_CHILD; }
inline int tip_of(int32_t c) { return -c - 1; }

struct Variant {
  int node = -1;
  int nkids = 0;
  int32_t kid[2] = {NO_CHILD, NO_CHILD};
  int32_t edge[2] = {-1, -1};
  int level = 0;
};

struct RootClass {
  std::vector<uint32_t> configs;  // member configurations
  int dtroot_slot = -1;           // variant at the displayed-tree root
  std::vector<uint64_t> edges;    // alive edge set (bitset over edges)
};

struct Dag {
  int n_retic = 0;
  std::vector<Variant> var;
  std::vector<RootClass> classes;
};

// Enumerates the 2^r displayed-tree configurations and deduplicates CLVs
// by (node, alive subtree).  Throws std::invalid_argument on bad input.
Dag compile_dag(const nrx_network_desc &net);

// ---------------------------------------------------------------- device
// One CLV op.  The forward pass computes CLV `out` from its children; the
// adjoint pass (push form) reads the adjoint of `out` and delivers the
// back-propagated vectors to the children.
struct DevOp {
  int32_t out;
  int32_t nk;
  int32_t kid[2];    // child codes
  int32_t edge[2];   // edge to each child
  int32_t root;      // index of the root-CLV site sum this op feeds, or -1
  int32_t flags;     // OP_HAS_CONSUMERS | OP_FIRST_KID0 | OP_FIRST_KID1
  int32_t gslot[2];  // gradient-partial row of (this op, child k)
};
constexpr int32_t OP_HAS_CONSUMERS = 1;  // adjoint of `out` was accumulated by consumers
constexpr int32_t OP_FIRST_KID0 = 2;     // first consumer of kid[0] in adjoint order
constexpr int32_t OP_FIRST_KID1 = 4;
// FP32 form, first op of a group: no op of the group consumes a CLV of the
// group before it, so a kernel may load this group's inputs before the
// previous group's stores (and, walking backwards, load the previous
// group's inputs before this group delivers)
constexpr int32_t OP_GROUP_INDEP = 8;
// FP32 form: child k's adjoint already holds deliveries made in earlier
// groups of the reverse sweep and none in this group, so the accumulated
// value may be loaded with the group's other inputs (the kernels issue a
// group's loads after the previous group's stores)
constexpr int32_t OP_ACC_EARLY0 = 16;
constexpr int32_t OP_ACC_EARLY1 = 32;
// FP32 form, 4 states: a child whose adjoint lives in a per-thread LDS slot
// instead of HBM.  OP_PASS_KIDk: deliver child k's adjoint to slot
// pass_slot(flags, k) (written by the first delivery, added to by later
// ones); OP_A_PASS: this op's own adjoint is in slot pass_slot(flags, 2)
constexpr int32_t OP_PASS_KID0 = 64;
constexpr int32_t OP_PASS_KID1 = 128;
constexpr int32_t OP_A_PASS = 256;
constexpr int PASS_SLOTS = 5;  // at most; fewer where LDS would cost residency
constexpr int PASS_SHIFT = 12;  // slot of use u (0, 1: children, 2: own) in bits 12 + 3u
constexpr int32_t PASS_BITS = 0x1FF << PASS_SHIFT;
inline __host__ __device__ int pass_slot(int32_t flags, int u) { return (flags >> (PASS_SHIFT + 3 * u)) & 7; }

// A launch of the plain form: the ops [op_begin, op_end) of one pruning
// level.  The forward runs one workgroup per (tile, PLAIN_G consecutive
// ops); the adjoint one workgroup per tile walking all ops in order.
struct LevelChunk {
  int level, op_begin, op_end;
};
constexpr int PLAIN_G = 4;  // forward ops per workgroup
// adjoint: ops of a level whose P matrices are staged in LDS (48 KiB)
constexpr int plain_staged_ops(int C) { return 49152 / (2 * C * 16 * 8); }

// One step of a root-to-branch path: the op (its output slot), the child
// the path continues into and the other child (NO_CHILD for a one-child
// op), with their edges; children are slots or tip codes as in DevOp.
struct PathStep {
  int32_t out, child, sib;
  int32_t edge_child, edge_sib;
  int32_t nk;
};
// P blocks of a path's first PATH_STAGED_STEPS steps are staged in LDS by
// the path kernels (4 states)
constexpr int PATH_STAGED_STEPS = 24;
inline size_t path_lds_bytes(int S, int C, int max_steps) {
  return S == 4 ? sizeof(double) * 2 * C * S * S * static_cast<size_t>(std::min(max_steps, PATH_STAGED_STEPS)) : 0;
}

// Device-side Newton-Raphson of one branch length (nrx_optimize_branches):
// the safeguarded iteration's state, and per displayed-tree class what the
// derivative mixture needs.
struct NewtonState {
  double x, a, b;      // current length and the bracket of the maximiser
  double x0, lo, hi, tol;
  double d1, d2;       // last derivatives
  double lnl_old, lnl_new;
  int32_t n_paths, done, iters, accepted;
  int32_t max_iters, pad;
};
struct DerivClassDev {
  double logprob;
  int32_t use;   // passes min_tree_logprob
  int32_t has;   // its alive edges contain the branch
  int32_t path;  // sumtable row, or -1
  int32_t pad;
};

struct DevClass {
  double logprob;
  int32_t root_index;
  int32_t use;
};

// Everything a kernel needs about the site axis and the models.
struct DevLayout {
  int64_t n_sites_pad;   // sites over all partitions, tile padded
  int32_t n_tiles;
  int32_t n_tiles_grid;  // n_tiles rounded up to a multiple of 8 (XCDs)
  int32_t n_parts;
  int32_t n_edges;
  int32_t S;
  int32_t C;
};

// Device view of one FP32-form engine, read by the persistent kernels; a
// launch walks a table of (network, tile) work items, so one launch can
// score one network or a batch of them.
struct NetDesc {
  DevLayout L;
  const DevOp *ops;
  const int32_t *group_begin;
  int32_t n_groups;
  int32_t n_root;
  int32_t n_uses;  // gradient rows
  int32_t n_pass;  // LDS pass slots of the 4-state reverse sweep (layout of its dynamic LDS)
  float *clv;
  int32_t *scl;
  float *adj;
  int32_t *adj_exp;
  const uint8_t *tips;
  const float *Pf;
  const float *Qf;
  const float *Pa;  // 20-state data: operand tables [part][edge][C][kind][AA_TABLE]
  const int32_t *tile_part;
  const double *freqs;
  const double *weights;
  const double *beta;
  double *root_partials;
  double *grad_partials;
};

struct WorkItem {
  int32_t net;
  int32_t tile;
};

// Per-network parameters of the small kernels when a batch is scored.
struct SmallDesc {
  DevLayout L;
  // P matrices
  const double *V, *Vinv, *ev, *rates, *brlen;
  double *P, *dP, *d2P;
  float *Pf;
  // reductions and mixture
  int32_t n_root, nc, R, n_uses;
  const int32_t *part_tile_begin;
  const DevClass *classes;
  const double *class_dlogprob;
  double *root_partials, *root_sums, *grad_partials, *use_sums, *beta;
  const int32_t *edge_use_begin, *edge_use_list;
  double *res;
  int32_t rl_total, rl_per_part, rl_root_logl, rl_tree_logl, rl_alpha, rl_dbrlen,
      rl_part_dbrlen, rl_dprob;
  int32_t pw;           // root partials per 64-site tile
  int32_t gpw, gbpt, pad;  // gradient partials: waves per site block, blocks per tile
};

struct Engine;

// Batch launchers (kernels.hip): one launch per kernel for n networks.
void launch_pmatrices_multi(hipStream_t s, const SmallDesc *d, int n, int max_elems);
void launch_root_reduce_multi(hipStream_t s, const SmallDesc *d, int n, int max_rows);
void launch_mixture_multi(hipStream_t s, const SmallDesc *d, int n, int variant);
void launch_grad_reduce_multi(hipStream_t s, const SmallDesc *d, int n, int max_rows, int max_edges);
void launch_gather_results(hipStream_t s, const SmallDesc *d, int n, const int64_t *offsets,
                           double *out, int max_size);
// G = ops per group, TS = sites per workgroup (the autotuned shape)
void launch_forward_persistent_multi(hipStream_t stream, int C, int G, int TS, const NetDesc *nets,
                                     const WorkItem *work, int n_work, size_t lds);
// H = lanes per (site, category) row of the 4-state reverse sweep (1 or 2)
void launch_adjoint_persistent_multi(hipStream_t stream, int C, int G, int TS, int H, const NetDesc *nets,
                                     const WorkItem *work, int n_work, size_t lds);
// dynamic LDS of the FP32 4-state adjoint (P/Q tables and pass slots)
size_t adjoint_persistent_lds(int E, int C, int TS, int n_pass);
int adjoint_pass_slots(int C, int G, int TS, int H, int E, int64_t n_work);

// Kernel launchers (kernels.hip)
void launch_pmatrices(Engine &e);
void launch_forward_level(Engine &e, const DevOp *ops, int n_ops, const LevelChunk &ch);
// 4 states: the levels of `walk` (device array of op ranges into `ops`,
// lowest level first) in one launch
void launch_forward_walk(Engine &e, const DevOp *ops, const LevelChunk *walk, int n_walk);
void launch_root_partials(Engine &e, bool seeds);
// root partials of n roots: physical rows `rows`, root indices `index`
void launch_root_partials(Engine &e, bool seeds, const int32_t *rows, const int32_t *index, int n);
void launch_mixture(Engine &e, int variant, int partials_per_tile);
// 20 states: one run of ops, taken from `ops` (same positions as in the
// op list)
void launch_adjoint_level(Engine &e, const LevelChunk &ch, const DevOp *ops);
// 4 states: level chunks c_hi down to c_lo in one launch
void launch_adjoint_levels(Engine &e, const DevOp *ops, int c_hi, int c_lo);
void launch_grad_reduce(Engine &e);
// nets: a descriptor other than the engine's own (its invalid ops only)
void launch_forward_persistent(Engine &e, const NetDesc *nets = nullptr);
void launch_adjoint_persistent(Engine &e);
void launch_mfma_tables(Engine &e);
void launch_forward_mfma(Engine &e, const NetDesc *nets = nullptr);
void launch_adjoint_mfma(Engine &e);
void launch_sumtables(Engine &e, int n_paths, int max_steps);
// FP64 4-state: per use, chunk and category the summed z x^T (dlnL/dP);
// per root CLV and chunk the frequencies' direct share (a chunk is one of
// n_split slices of a partition's tiles, chunk = partition * n_split + slice)
void launch_model_grad(Engine &e, int n_split, double *g_out, double *r_out);
// Newton-Raphson of a branch length on the device: `rounds` rounds of
// (evaluate the sumtables at the state's length, reduce, Newton step; no-ops
// once the state is done or has taken its max_iters steps), then the
// evaluation at the current length and the acceptance check.  The number
// of paths and every length come from the device state, so the sequence is
// captured once and replayed per branch (and again while not done).
void enqueue_newton(Engine &e, int variant, int mode, int rounds, int max_paths, NewtonState *st,
                    const DerivClassDev *cls, const double *base_logl);
// allocation-free branch derivatives: a fixed device array of this many
// bytes per GPU holds every scratch buffer of nrx_brlen_derivatives_noalloc
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(16) << 20;
void *noalloc_scratch();  // its device address on the current device
// max_steps: the longest path's step count (sizes the staged P in LDS)
void launch_deriv_fused(Engine &e, const PathStep *steps, const int32_t *path_begin, int n_paths, int max_steps,
                        double brlen, double *partials, double *sums);
void launch_sumtable_eval(Engine &e, int n_paths, double brlen);

struct HipCheck {
  static void check(hipError_t err, const char *what);
};
#define NRX_HIP(x) ::nrx::HipCheck::check((x), #x)

}  // namespace nrx
