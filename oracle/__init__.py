"""CPU oracle for the residual-correlation cover-tree path (TEST INFRASTRUCTURE).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker or as the timed CPU
baseline.  The product package ``covertreex_amd`` never imports it.

Contents
--------
- ``liboracle.so`` (built from ``oracle/*.c`` by ``oracle/Makefile``): a plain-C
  restatement of the reference's rust-hilbert build and query in the
  reference's f32 evaluation order (see ``oracle/oracle.h``).
- :mod:`oracle.fp64`: FP64 NumPy Nystrom factor and FP64 brute-force residual
  kNN evaluated on the same stored FP32 inputs.
"""

from .lib import (  # noqa: F401
    EXP_CR,
    EXP_LIBM,
    OracleTree,
    build_residual_tree,
    build_residual_tree_with_insert,
    build_euclidean_tree,
    build_euclidean_tree_with_insert,
    delete_rows,
    dist_batch_pairs,
    dist_sq_pairs,
    greedy_mis,
    grid_adjacency,
    hilbert_codes,
    hilbert_order,
    knn_euclidean,
    knn_residual,
    luby_mis,
    morton_order,
    si_cache,
    subtree_bounds,
    tree_from_arrays,
)
