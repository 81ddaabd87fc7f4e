// Scores tiny.nw (NetRAX test/sample_networks) with the C++ mirror of the
// reference interface: netrax::computeLoglikelihood and the branch
// derivatives, printing one line "lnl d1 d2" per variant.  Exit code 2 and
// the engine's message on failure (e.g. no GPU).
//
// ((A:2,(B:1)X#H1:0::0.3)Q:2,(C:2,X#H1:0::0.7)R:2);  in NetRAX's layout:
// tips A,B,C = 0,1,2; inner R=3, Q=4, root=5; reticulation X=6 with parent
// edges 5 (first, from Q) and 6 (from R).
#include <cstdio>
#include <vector>

#include "../../include/netrax_likelihood.hpp"

int main() {
  const int32_t edge_parent[] = {4, 6, 3, 5, 5, 4, 3};
  const int32_t edge_child[] = {0, 1, 2, 3, 4, 6, 6};
  const int32_t first[] = {5}, second[] = {6};
  nrx_network_desc net{7, 3, 5, 7, 1, edge_parent, edge_child, first, second};
  // alignment of tiny_fake_alignment.txt as state masks (A=1 C=2 G=4 T=8, gap=15)
  const char *seqs[3] = {"AAGT-CG----GTAGCA", "AACCGGTAGATCA-GCA", "AACCGGTAGCA-AGCAT"};
  const int n = 17;
  std::vector<uint32_t> masks(3 * n), weights(n, 1);
  for (int t = 0; t < 3; ++t)
    for (int s = 0; s < n; ++s) {
      const char ch = seqs[t][s];
      masks[t * n + s] = ch == 'A' ? 1 : ch == 'C' ? 2 : ch == 'G' ? 4 : ch == 'T' ? 8 : 15;
    }
  std::vector<nrx_partition_desc> parts{{n, 4, 4, masks.data(), weights.data()}};
  const double rates[6] = {1.0, 2.0, 1.0, 1.0, 2.0, 1.0};
  const double freqs[4] = {0.1, 0.2, 0.3, 0.4};
  try {
    netrax::GpuAnnotatedNetwork ann(net, parts);
    netrax::nrx_check(nrx_set_model(ann.engine, 0, rates, freqs, 0.7));
    ann.setBranchLengths({2.0, 1.0, 2.0, 2.0, 2.0, 0.1, 0.2});
    ann.setReticulationProbs({0.3});
    for (auto v : {netrax::LikelihoodVariant::AVERAGE_DISPLAYED_TREES,
                   netrax::LikelihoodVariant::BEST_DISPLAYED_TREE}) {
      ann.likelihood_variant = v;
      const double lnl = netrax::computeLoglikelihood(ann, 0, 1);
      auto st = netrax::computePartitionSumtables(ann, 1);
      netrax::LoglDerivatives d = netrax::computeLoglikelihoodDerivatives(ann, st, 1);
      std::printf("%.17g %.17g %.17g\n", lnl, d.logl_prime, d.logl_prime_prime);
    }
  } catch (const std::exception &ex) {
    std::fprintf(stderr, "error: %s\n", ex.what());
    return 2;
  }
  return 0;
}
