"""covertreex_amd: MI355X engine for covertreex's residual-correlation cover tree.

Drop-in for the reference's residual path behind ``cover_tree()`` /
``TreeEngine`` with ``engine="hip"``.  All compute runs in
``libcovertreex_hip.so`` (hand-written gfx950 kernels behind a C-ABI); the
package fails loudly when that library is missing.
"""

from .engine import (  # noqa: F401
    DEFAULT_ENGINE,
    SUPPORTED_ENGINES,
    CoverTree,
    HipEngine,
    RuntimeConfig,
    TreeEngine,
    build_tree,
    get_engine,
    register_engine,
)
from .factory import BuiltCoverTree, cover_tree  # noqa: F401
from .kernels import RBF, Kernel, Matern52  # noqa: F401
from .native import HipTree, compute_subtree_bounds, hilbert_order  # noqa: F401
from .residual import ResidualBackend, build_residual_backend  # noqa: F401

__version__ = "0.1.0"
