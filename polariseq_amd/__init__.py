"""polariseq_amd: MI355X-native neighbour-graph step for Polariseq.

NN-Descent kNN on a PCA embedding, and an exact tiled kNN, as hand-written HIP
kernels for gfx950 behind a C-ABI (include/psq_knn.h, lib/libpsq_knn.so).
The Python functions mirror the reference binding's ``nndescent_dense`` and
``neighbors_dense`` (crates/polariseq-python/src/lib.rs:471-566), and its
``fuzzy_connectivities`` / ``fuzzy_connectivities_into`` (lib.rs:568-642), the
graph step's consumer inside ``pp.neighbors``; ``neighbors_graph`` is the graph
computation of ``pp.neighbors`` itself (python/polariseq/__init__.py:2514-2600).
``pca_sparse`` / ``pca_sparse_scaled`` are the binding's sparse PCA calls
(lib.rs:395-467), the step that produces the embedding the graph is built on.
``RocmBackend`` implements the reference's ``ComputeBackend`` trait (pca, neighbors).
"""

from .neighbors import (  # noqa: F401
    batched_report,
    effective_k,
    knn_exact_batched,
    knn_exact_batched_device,
    knn_exact,
    knn_exact_device,
    last_stats,
    nndescent_dense,
    nndescent_device,
    neighbors_dense,
    neighbors_dense_device,
    release_workspace,
)

from .fuzzy import (  # noqa: F401,E402
    fuzzy_connectivities,
    fuzzy_connectivities_device,
    fuzzy_connectivities_into,
)
from .encoder import Encoder  # noqa: F401,E402
from .graph import (  # noqa: F401,E402
    cosine_distances_device,
    neighbors_graph,
    unit_rows_device,
)
from .pca import (  # noqa: F401,E402
    pca_csr_device,
    pca_dense,
    pca_last_report,
    pca_sparse,
    pca_sparse_scaled,
)
from .backend import RocmBackend  # noqa: F401,E402

__version__ = "0.1.0"
