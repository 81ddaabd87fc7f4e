/*
 * netrax_gpu.h -- C ABI of the MI355X network-likelihood engine.
 *
 * Drop-in boundary for NetRAX's likelihood layer (src/likelihood/).  Every
 * entry point is extern "C", takes plain pointers and sizes, returns an
 * integer status (NRX_OK == 0) and never throws; the message of the last
 * failure on the calling thread is available from nrx_last_error().
 *
 * Reference interfaces replaced (paths relative to the NetRAX source tree):
 *   nrx_create / nrx_destroy
 *       AnnotatedNetwork construction + createNetworkPllTreeinfo
 *       (src/graph/AnnotatedNetwork.cpp:80-185, src/RaxmlWrapper.cpp:539-669):
 *       the engine owns CLVs, scalers and P matrices in HBM.
 *   nrx_set_model
 *       pll model setup (subst rates, frequencies, Gamma alpha) done by
 *       raxml-ng's Model / pllmod_treeinfo (RaxmlWrapper.cpp:189-214).
 *   nrx_set_branch_lengths
 *       fake_treeinfo->linked_branch_lengths + invalidatePmatrixIndex
 *       (RaxmlWrapper.cpp:652-663, src/helper/InvalidationHelper.cpp:157-172).
 *   nrx_set_reticulation_probs
 *       setReticulationProb (src/optimization/ReticulationOptimization.cpp:25-38).
 *   nrx_loglikelihood
 *       double computeLoglikelihood(AnnotatedNetwork&, int incremental,
 *       int update_pmatrices)  (src/likelihood/LikelihoodComputation.hpp:17-18)
 *       and the pll-modules hook network_logl_wrapper (RaxmlWrapper.cpp:21-26).
 *   nrx_gradient
 *       no single reference function: replaces the E calls of
 *       computePartitionSumtables + computeLoglikelihoodDerivatives that a
 *       full gradient costs in the reference, and adds d lnL / d p_r
 *       (the reference optimises p_r by Brent, ReticulationOptimization.cpp:68-101).
 *   nrx_brlen_derivatives
 *       computePartitionSumtables(ann_network, pmatrix_index) followed by
 *       computeLoglikelihoodDerivatives(ann_network, sumtables, pmatrix_index)
 *       (src/likelihood/LikelihoodDerivatives.hpp:83-88), evaluated at a given
 *       branch length as pll's Newton-Raphson callback does
 *       (src/optimization/BranchLengthOptimization.cpp:171-200).
 */
#ifndef NETRAX_GPU_H
#define NETRAX_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRX_OK 0
#define NRX_ERR_ARG 1        /* invalid argument / shape */
#define NRX_ERR_NEGINF 2     /* network likelihood is -infinity */
#define NRX_ERR_DEVICE 3     /* HIP runtime failure */
#define NRX_ERR_UNSUPPORTED 4
#define NRX_ERR_STATE 5      /* call out of order (e.g. no model set) */

/* LikelihoodVariant (src/likelihood/LikelihoodVariant.hpp:3-7) */
#define NRX_AVERAGE_DISPLAYED_TREES 0
#define NRX_BEST_DISPLAYED_TREE 1

/* arithmetic of the CLV products */
#define NRX_FP64 0 /* plain form: FP64 CLVs; one forward and one reverse-sweep
                      launch walking the pruning levels per site tile */
#define NRX_FP32 1 /* batched form: FP32 CLVs + per-site power-of-two
                      exponents, persistent site-tile kernels, FP64 sums */

/* branch-derivative normalisation for nrx_brlen_derivatives */
#define NRX_DERIV_REFERENCE 0 /* sums over trees containing the branch only
                                 (LikelihoodDerivatives.cpp:153-183, 325) */
#define NRX_DERIV_TRUE 1      /* true derivatives of the network lnL */

typedef struct nrx_engine nrx_engine;

/* Network topology with NetRAX's index layout (src/io/NetworkIO.cpp:114-185):
 * edge e goes edge_parent[e] -> edge_child[e]; reticulation r has parent
 * edges retic_first_edge[r] and retic_second_edge[r]. */
typedef struct {
  int32_t n_nodes;
  int32_t n_tips; /* tips are nodes 0 .. n_tips-1 */
  int32_t root;
  int32_t n_edges;
  int32_t n_retic;
  const int32_t *edge_parent;       /* [n_edges] */
  const int32_t *edge_child;        /* [n_edges] */
  const int32_t *retic_first_edge;  /* [n_retic] */
  const int32_t *retic_second_edge; /* [n_retic] */
} nrx_network_desc;

/* One partition of compressed site patterns. */
typedef struct {
  int32_t n_sites;           /* number of patterns */
  int32_t states;            /* 4 (DNA) or 20 (protein) */
  int32_t rate_cats;         /* discrete Gamma categories (1..8) */
  const uint32_t *tip_masks; /* [n_tips][n_sites]; bit i set = state i allowed */
  const uint32_t *weights;   /* [n_sites] pattern weights */
} nrx_partition_desc;

typedef struct {
  int32_t precision;       /* NRX_FP64 or NRX_FP32 */
  int32_t device;          /* HIP device ordinal */
  double min_tree_logprob; /* NetraxOptions::min_interesting_tree_logprob, log(1e-6) */
  double brlen_min;        /* branch lengths are clamped to [brlen_min, brlen_max] */
  double brlen_max;
  int32_t use_graph;       /* FP32 form: each (variant, gradient) evaluation is captured
                              once in a hipGraph and replayed as one graph launch (re-
                              captured after a schedule change); per-kernel timings are
                              then not recorded, only the total.  A batch follows the
                              option of its first engine. */
  int32_t timing;          /* record HIP events around kernels (nrx_last_timing) */
} nrx_options;

typedef struct {
  int32_t n_variants;      /* distinct (node, displayed subtree) CLVs */
  int32_t n_ops_needed;    /* CLV ops evaluated per full forward pass */
  int32_t n_levels;        /* pruning levels */
  int32_t n_classes;       /* root variants (distinct displayed trees) */
  int32_t n_root_clvs;     /* distinct CLVs the root kernel reads */
  int32_t n_uses;          /* (parent op, child) pairs walked by the adjoint */
  int32_t n_inner_uses;    /* those pairs whose child is a CLV (not a tip) */
  int32_t n_lds_handoffs;  /* FP32 4-state adjoint: deliveries into a child's LDS
                              pass slot instead of HBM */
  int64_t n_sites_padded;  /* sites over all partitions after tile padding */
  /* algorithmic HBM bytes over the real (unpadded) sites: each CLV, scaler
   * and tip row counted once per read or write an op makes */
  double fwd_op_bytes;     /* pruning ops of one full forward pass */
  double root_bytes;       /* displayed-tree root site sums */
  double adj_bytes;        /* adjoint ops of one reverse pass */
  double fwd_bytes;        /* P matrices + pruning + root */
  double grad_bytes;       /* fwd_bytes + adj_bytes */
  double clv_bytes;        /* device bytes held by CLVs, scalers and adjoints */
} nrx_stats;

/* Device time of the last nrx_loglikelihood / nrx_gradient call (HIP
 * events on the engine's stream; requires nrx_options.timing). */
typedef struct {
  double total_ms;
  double fwd_ms;           /* pruning kernel(s), summed over launches */
  int32_t fwd_launches;
  int32_t adj_launches;
  double adj_ms;           /* adjoint kernel(s), summed over launches */
} nrx_timing;

void nrx_default_options(nrx_options *opt);
const char *nrx_last_error(void);

int nrx_create(const nrx_network_desc *net, const nrx_partition_desc *parts,
               int32_t n_parts, const nrx_options *opt, nrx_engine **out);
int nrx_destroy(nrx_engine *eng);

/* subst_rates: states*(states-1)/2 exchangeabilities (upper triangle,
 * row-major: AC AG AT CG CT GT for DNA); freqs: states values > 0;
 * alpha: Gamma shape (ignored for one category). */
int nrx_set_model(nrx_engine *eng, int32_t partition, const double *subst_rates,
                  const double *freqs, double alpha);
/* explicit category rates (mean 1 expected) instead of Gamma(alpha) */
int nrx_set_category_rates(nrx_engine *eng, int32_t partition, const double *rates);
int nrx_set_branch_lengths(nrx_engine *eng, const double *brlen /* [n_edges] */);
int nrx_set_reticulation_probs(nrx_engine *eng,
                               const double *first_parent_prob /* [n_retic] */);

/* Network lnL.  incremental = 0 recomputes every CLV; 1 recomputes only
 * CLVs depending on branch lengths changed since the last call.
 * update_pmatrices = 0 keeps the current P matrices.  per_partition may be
 * NULL.  Returns NRX_ERR_NEGINF when the result is -infinity
 * (ImprovedLoglikelihood.cpp:637-640). */
int nrx_loglikelihood(nrx_engine *eng, int32_t variant, int32_t incremental,
                      int32_t update_pmatrices, double *total, double *per_partition);

/* lnL plus dlnL/d(branch length) for every edge and dlnL/d(first-parent
 * probability) for every reticulation, from one forward and one adjoint
 * pass.  per_partition_dbrlen ([n_parts][n_edges]) may be NULL. */
int nrx_gradient(nrx_engine *eng, int32_t variant, double *total, double *dbrlen,
                 double *dprob, double *per_partition_dbrlen);

/* First and second derivative of lnL w.r.t. the length of one branch,
 * evaluated at branch length `brlen` (clamped to [brlen_min, brlen_max] as
 * nrx_set_branch_lengths clamps) with every other parameter fixed.
 * mode: NRX_DERIV_REFERENCE or NRX_DERIV_TRUE.  part_d1/part_d2 may be NULL. */
int nrx_brlen_derivatives(nrx_engine *eng, int32_t variant, int32_t edge,
                          int32_t mode, double brlen, double *d1, double *d2,
                          double *part_d1, double *part_d2);

/* nrx_brlen_derivatives without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving or optimisation loops, harnesses that
 * forbid hipMalloc).  Same arguments, results and errors as
 * nrx_brlen_derivatives, evaluated on the engine's state as left by the
 * last nrx_loglikelihood / nrx_gradient of the same variant: NRX_ERR_STATE
 * when parameters changed since (refreshing the CLVs could allocate).  Every
 * scratch buffer (path tables, per-wave partial sums) lives in one fixed
 * device array of 16 MiB per GPU, held until the process ends; an engine
 * whose derivative scratch would exceed it is refused with NRX_ERR_ARG
 * (about 32 displayed trees x 350k sites at 4 rate categories fit).  The
 * paths are walked again on every call (no sumtables are kept).  Calls that
 * overlap on one GPU (from several threads or streams) share that array and
 * race: run them one at a time per device. */
int nrx_brlen_derivatives_noalloc(nrx_engine *eng, int32_t variant, int32_t edge,
                                  int32_t mode, double brlen, double *d1, double *d2,
                                  double *part_d1, double *part_d2);

/* lnL and its gradient with respect to the substitution model of every
 * partition: d_subst [n_parts][S(S-1)/2] (exchangeabilities, upper
 * triangle row-major as nrx_set_model takes them), d_freqs [n_parts][S]
 * (each frequency as an independent argument of nrx_set_model, through the
 * rate matrix and the root), d_alpha [n_parts] (Gamma shape; NaN after
 * nrx_set_category_rates) and d_rates [n_parts][C] (the category rates);
 * d_alpha and d_rates may be NULL.  No single reference function: replaces
 * the numerical model optimisation raxml-ng runs through the pll-modules
 * hook (src/optimization/ModelOptimization.cpp:25-97, likelihood target
 * network_logl_wrapper, src/RaxmlWrapper.cpp:21-26, 647-648) with analytic
 * derivatives.  One forward and one adjoint pass, then a pass that forms
 * dlnL/dP per (edge, category) from the adjoints, chained through the
 * eigendecomposition of the rate matrix on the host.  FP64 form, 4-state
 * data (NRX_ERR_UNSUPPORTED otherwise). */
int nrx_model_gradient(nrx_engine *eng, int32_t variant, double *total, double *d_subst,
                       double *d_freqs, double *d_alpha, double *d_rates);

/* Newton-Raphson branch-length sweep with the iteration on the device.
 * Replaces optimize_branches / optimize_branch /
 * optimize_branch_newton_raphson (src/optimization/BranchLengthOptimization.cpp:
 * 202-241, 317-330, 345-420, 422-477): every listed edge (duplicates
 * dropped, taken in index order) is optimised once, within [brlen_min,
 * brlen_max] to a tolerance of brlen_min / 10, from its sumtables' first
 * and second derivatives (mode: NRX_DERIV_REFERENCE or NRX_DERIV_TRUE); a
 * length that does not improve lnL is given up.  Per edge: the forward
 * pass and sumtables on the host's order, then a captured graph runs six
 * Newton rounds (FP64 derivative mixture, safeguarded step and bracket on
 * the device; rounds after convergence do nothing) and the acceptance
 * check, and the host waits once; a branch not converged after six rounds
 * takes another launch, up to max_iters rounds in all.
 * *logl (may be NULL) receives the final network lnL; fails with
 * NRX_ERR_STATE ("Overall loglikelihood got worse") when the sweep lost
 * more than 1e-3.  The new lengths are read with nrx_get_branch_lengths. */
int nrx_optimize_branches(nrx_engine *eng, int32_t variant, int32_t mode, const int32_t *edges,
                          int32_t n_edges, int32_t max_iters, double *logl);
/* The engine's branch lengths [n_edges], as clamped. */
int nrx_get_branch_lengths(nrx_engine *eng, double *brlen);

/* lnL and its first and second derivatives in every reticulation
 * probability (d1, d2 [n_reticulations]; either may be NULL), from one
 * evaluation that reuses every valid CLV: a probability change only
 * re-weights the displayed trees.  Derivative counterpart of the Brent
 * target brent_target_networks_prob
 * (src/optimization/ReticulationOptimization.cpp:42-47). */
int nrx_prob_derivatives(nrx_engine *eng, int32_t variant, double *total, double *d1, double *d2);

/* Replaces optimize_reticulations / optimize_reticulation
 * (src/optimization/ReticulationOptimization.cpp:68-119): rounds over the
 * reticulations, at most max_iters, ending early when a round leaves lnL
 * unchanged; each probability is moved inside [brprob_min, brprob_max]
 * (NetraxOptions.hpp:102-103) by a safeguarded Newton iteration on the
 * exact derivative in place of Brent's method, to an interval of width
 * `tolerance`, keeping the best lnL seen.  logl (may be NULL) receives the
 * final lnL; nrx_get_reticulation_probs reads the probabilities. */
int nrx_optimize_reticulations(nrx_engine *eng, int32_t variant, int32_t max_iters, double brprob_min,
                               double brprob_max, double tolerance, double *logl);
/* The engine's reticulation probabilities (first-parent) [n_reticulations]. */
int nrx_get_reticulation_probs(nrx_engine *eng, double *probs);

/* Per-(partition, displayed tree) log-likelihoods of the last forward pass:
 * tree_logl [n_parts][n_classes], class_logprob [n_classes]. */
int nrx_tree_logls(nrx_engine *eng, double *tree_logl, double *class_logprob);

int nrx_get_stats(nrx_engine *eng, nrx_stats *out);

/* Batches of networks (FP32 form).  Replaces the sequential candidate loop
 * of a search step (src/search/Filtering.cpp:236-270 scores candidates one
 * at a time): every engine of the batch is scored by one launch per kernel.
 * All engines must use NRX_FP32, 4-state data and the same number of rate
 * categories on the same device, must outlive the batch and may appear
 * only once in it (a repeated engine is refused: NRX_ERR_ARG).  With
 * use_graph set on the first engine, the whole batch evaluation (P matrices,
 * persistent forward, reductions, mixture, persistent adjoint, gradient
 * totals, result gather) is one hipGraph launch. */
typedef struct nrx_batch nrx_batch;

int nrx_batch_create(nrx_engine *const *engines, int32_t n_engines, nrx_batch **out);
int nrx_batch_destroy(nrx_batch *batch);
/* lnl[i] for every engine; when dbrlen (and dprob) are non-NULL also the
 * full gradient: dbrlen is the concatenation of each engine's n_edges
 * values, dprob of its n_retic values, in engine order.  device_ms (may be
 * NULL) receives the device time of the whole batch. */
int nrx_batch_evaluate(nrx_batch *batch, int32_t variant, double *lnl, double *dbrlen,
                       double *dprob, double *device_ms);

/* Optimises every candidate of the batch, then scores them all in one
 * batched evaluation (lnl [n]): optimizeAfterMove's per-candidate steps
 * (src/search/Filtering.cpp:158-208) ahead of the scoring loop of
 * filterCandidates (:236-270).  Candidate i's branches are
 * edges[edge_begin[i] .. edge_begin[i+1]), swept by the device Newton
 * iteration of nrx_optimize_branches with brlen_iters iterations (1: the
 * prefilter's single step, 32: a full sweep), then its reticulations
 * retics[retic_begin[i] .. retic_begin[i+1]) by nrx_optimize_reticulations'
 * search with prob_iters rounds.  Each engine runs on its own stream;
 * n_threads host workers (<= 0: up to 16) drive the candidates so their
 * host waits overlap.  status [n] (may be NULL) receives each candidate's
 * return code; without it the first failure is returned after scoring. */
int nrx_batch_optimize(nrx_batch *batch, int32_t variant, int32_t mode, const int32_t *edge_begin,
                       const int32_t *edges, int32_t brlen_iters, const int32_t *retic_begin,
                       const int32_t *retics, int32_t prob_iters, double brprob_min, double brprob_max,
                       double prob_tolerance, int32_t n_threads, double *lnl, int32_t *status);

/* Memory budget of the FP64 form: bytes for the CLV, adjoint and scaler
 * rows (0: every CLV resident, the default).  Under a budget smaller than
 * the resident rows the level schedule is cut into segments; CLVs read by
 * a later segment keep their rows and each segment's other CLVs share a
 * pool, so the forward pass keeps only those and the reverse sweep
 * recomputes each segment's pooled CLVs before its adjoint (at most one
 * more forward pass).  lnL and gradient are bit-identical to the resident
 * evaluation.  Branch derivatives, sumtables and model gradients need
 * every CLV and return NRX_ERR_UNSUPPORTED while a budget is active.
 * NRX_ERR_ARG when even the finest segmentation does not fit (the previous
 * setting stays).  No single reference counterpart: the reference keeps
 * per-displayed-tree CLVs and recomputes along re-rooted paths instead
 * (src/likelihood/VirtualRerooting.cpp:192-252). */
int nrx_set_memory_budget(nrx_engine *eng, uint64_t bytes);
typedef struct {
  int32_t segments;      /* 1: every CLV resident */
  int32_t rows;          /* CLV rows allocated */
  int32_t boundary_rows; /* rows kept across segments */
  int32_t resident_rows; /* rows without a budget */
  double row_bytes;      /* bytes of one CLV + adjoint + scaler row */
} nrx_memory_plan_info;
int nrx_memory_plan(nrx_engine *eng, nrx_memory_plan_info *out);

/* Device timing of the last evaluation (zeros unless timing is enabled). */
int nrx_last_timing(nrx_engine *eng, nrx_timing *out);

/* Shape of the FP32 form's persistent kernels.  No counterpart in the
 * reference (libpll's kernels have no launch shape); the autotuner picks
 * it per engine.  4-state data: group_width (2, 3, 4 or 8) independent ops
 * per group, whose CLV loads are issued together, tile_sites (16, 32
 * or 64, with tile_sites * rate_cats a multiple of 64) sites per
 * workgroup, i.e. tile_sites * rate_cats / 64 wavefronts per site block,
 * and row_lanes (1 or 2) lanes per (site, category) row of 4 states in the
 * reverse sweep (2: each lane holds two states, twice the wavefronts).
 * 20-state data: group_width 1, 2 or 4, 32 sites per workgroup (one
 * matrix-instruction column block) and row_lanes 1. */
typedef struct {
  int32_t group_width;
  int32_t tile_sites;
  int32_t row_lanes;
  double ms; /* device time of one lnL + gradient evaluation (autotune) */
} nrx_tuning;

/* Times one lnL + gradient evaluation (best of `reps`) for every candidate
 * shape, keeps the fastest for the engine and returns it in *best;
 * results[0 .. min(max_results, *n_results)) (may be NULL) receive every
 * candidate's time.  Leaves the engine's CLVs recomputed. */
int nrx_autotune(nrx_engine *eng, int32_t variant, int32_t reps, nrx_tuning *best,
                 nrx_tuning *results, int32_t max_results, int32_t *n_results);
/* Sets a shape (e.g. one found by an earlier nrx_autotune); NRX_ERR_ARG if
 * the engine's data cannot take it.  Engines of one batch must share it. */
int nrx_set_tuning(nrx_engine *eng, int32_t group_width, int32_t tile_sites, int32_t row_lanes);
int nrx_get_tuning(nrx_engine *eng, nrx_tuning *out);

#ifdef __cplusplus
}
#endif

#endif /* NETRAX_GPU_H */
