// netrax_likelihood.hpp -- header-only C++ mirror of NetRAX's likelihood
// interface over the engine's C ABI (netrax_gpu.h).
//
// Names, argument meaning and error behaviour follow the reference:
//   computeLoglikelihood(ann, incremental, update_pmatrices)
//       src/likelihood/LikelihoodComputation.hpp:17-18; throws
//       std::runtime_error("Invalid network likelihood: negative infinity \n")
//       as ImprovedLoglikelihood.cpp:637-640 does.
//   computePartitionSumtables(ann, pmatrix_index) /
//   computeLoglikelihoodDerivatives(ann, sumtables, pmatrix_index)
//       src/likelihood/LikelihoodDerivatives.hpp:83-88; LoglDerivatives as
//       LikelihoodDerivatives.hpp:76-81.  The sumtables stay on the device;
//       the returned object only names the branch they belong to.
//   LikelihoodVariant  src/likelihood/LikelihoodVariant.hpp:3-7.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "netrax_gpu.h"

namespace netrax {

enum class LikelihoodVariant {
  AVERAGE_DISPLAYED_TREES = NRX_AVERAGE_DISPLAYED_TREES,
  BEST_DISPLAYED_TREE = NRX_BEST_DISPLAYED_TREE,
  SARAH_PSEUDO = 2  // not provided by the engine
};

struct LoglDerivatives {
  double logl_prime = 0.0;
  double logl_prime_prime = 0.0;
  std::vector<double> partition_logl_prime;
  std::vector<double> partition_logl_prime_prime;
};

inline void nrx_check(int rc) {
  if (rc != NRX_OK) throw std::runtime_error(nrx_last_error());
}

// The part of AnnotatedNetwork (src/graph/AnnotatedNetwork.hpp:42-89) the
// likelihood path reads: the engine handle (which owns CLVs, scalers and P
// matrices on the device), the variant, linked branch lengths and the
// per-partition results.
struct GpuAnnotatedNetwork {
  nrx_engine *engine = nullptr;
  LikelihoodVariant likelihood_variant = LikelihoodVariant::AVERAGE_DISPLAYED_TREES;
  std::vector<double> linked_branch_lengths;  // [n_edges], clamped as the engine clamps
  std::vector<double> partition_loglh;        // [n_parts]
  double cached_logl = 0.0;
  bool cached_logl_valid = false;
  double brlen_min = 1e-6, brlen_max = 100.0;

  GpuAnnotatedNetwork() = default;
  GpuAnnotatedNetwork(const nrx_network_desc &net, const std::vector<nrx_partition_desc> &parts,
                      const nrx_options *opt = nullptr) {
    nrx_options o;
    nrx_default_options(&o);
    if (opt) o = *opt;
    brlen_min = o.brlen_min;
    brlen_max = o.brlen_max;
    nrx_check(nrx_create(&net, parts.data(), static_cast<int32_t>(parts.size()), &o, &engine));
    partition_loglh.assign(parts.size(), 0.0);
    linked_branch_lengths.assign(net.n_edges, 0.1);
  }
  GpuAnnotatedNetwork(const GpuAnnotatedNetwork &) = delete;
  GpuAnnotatedNetwork &operator=(const GpuAnnotatedNetwork &) = delete;
  ~GpuAnnotatedNetwork() {
    if (engine) nrx_destroy(engine);
  }

  // createNetworkPllTreeinfo clamps to [brlen_min, brlen_max]
  // (RaxmlWrapper.cpp:652-663); the stored lengths are the clamped ones
  void setBranchLengths(const std::vector<double> &b) {
    nrx_check(nrx_set_branch_lengths(engine, b.data()));
    linked_branch_lengths = b;
    for (double &x : linked_branch_lengths) x = x < brlen_min ? brlen_min : (x > brlen_max ? brlen_max : x);
    cached_logl_valid = false;
  }
  void setReticulationProbs(const std::vector<double> &first_parent_prob) {
    nrx_check(nrx_set_reticulation_probs(engine, first_parent_prob.data()));
    cached_logl_valid = false;
  }
};

inline double computeLoglikelihood(GpuAnnotatedNetwork &ann_network, int incremental = 1,
                                   int update_pmatrices = 1) {
  if (ann_network.likelihood_variant == LikelihoodVariant::SARAH_PSEUDO)
    throw std::runtime_error("SARAH_PSEUDO is not provided by the engine");
  double total = 0.0;
  const int rc = nrx_loglikelihood(ann_network.engine,
                                   static_cast<int32_t>(ann_network.likelihood_variant),
                                   incremental, update_pmatrices, &total,
                                   ann_network.partition_loglh.data());
  if (rc == NRX_ERR_NEGINF)
    throw std::runtime_error("Invalid network likelihood: negative infinity \n");
  nrx_check(rc);
  ann_network.cached_logl = total;
  ann_network.cached_logl_valid = true;
  return total;
}

struct SumtableHandle {
  unsigned int pmatrix_index;
};

inline std::vector<std::vector<SumtableHandle>> computePartitionSumtables(
    GpuAnnotatedNetwork &ann_network, unsigned int pmatrix_index) {
  if (pmatrix_index >= ann_network.linked_branch_lengths.size())
    throw std::runtime_error("pmatrix_index out of range");
  return std::vector<std::vector<SumtableHandle>>(ann_network.partition_loglh.size(),
                                                  {SumtableHandle{pmatrix_index}});
}

// Derivatives at the branch's current length in ann_network (pll's
// Newton-Raphson callback sets the length before each call).
inline LoglDerivatives computeLoglikelihoodDerivatives(
    GpuAnnotatedNetwork &ann_network, const std::vector<std::vector<SumtableHandle>> &sumtables,
    unsigned int pmatrix_index, int mode = NRX_DERIV_REFERENCE) {
  for (const auto &p : sumtables)
    for (const auto &s : p)
      if (s.pmatrix_index != pmatrix_index)
        throw std::runtime_error("sumtables belong to another branch");
  LoglDerivatives res;
  const size_t P = ann_network.partition_loglh.size();
  res.partition_logl_prime.assign(P, 0.0);
  res.partition_logl_prime_prime.assign(P, 0.0);
  nrx_check(nrx_brlen_derivatives(ann_network.engine,
                                  static_cast<int32_t>(ann_network.likelihood_variant),
                                  static_cast<int32_t>(pmatrix_index), mode,
                                  ann_network.linked_branch_lengths[pmatrix_index],
                                  &res.logl_prime, &res.logl_prime_prime,
                                  res.partition_logl_prime.data(),
                                  res.partition_logl_prime_prime.data()));
  return res;
}

}  // namespace netrax
