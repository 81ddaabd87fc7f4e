"""MI355X-native network-likelihood engine for NetRAX's likelihood path.

The compute path is ``libnetrax_gpu.so`` (HIP kernels for gfx950 behind the
C ABI declared in include/netrax_gpu.h); this package holds the ctypes
binding, the host-side mirror of the reference interface, the network
model and the seeded input generators.
"""
from .likelihood import (AnnotatedNetwork, LikelihoodVariant, LoglDerivatives, NetraxOptions,
                         NetworkBatch, Precision, computeLoglikelihood, computeLoglikelihoodDerivatives,
                         computePartitionSumtables)
from .network import Network, parse_newick, random_network
from .simulate import Model, Partition, Workload

__all__ = [
    "AnnotatedNetwork", "NetworkBatch", "LikelihoodVariant", "LoglDerivatives", "NetraxOptions", "Precision",
    "computeLoglikelihood", "computeLoglikelihoodDerivatives", "computePartitionSumtables",
    "Network", "parse_newick", "random_network", "Model", "Partition", "Workload",
]
