"""The rocm device package, the counterpart of sqaod.cuda
(sqaodpy/sqaod/cuda/__init__.py)."""
from . import _cext  # noqa: F401  (fails loudly when libsqaod_rocm.so is missing)
from .dense_graph_annealer import dense_graph_annealer, DenseGraphAnnealer  # noqa: F401
from .device import create_device, Device  # noqa: F401
from .dense_graph_bf_searcher import dense_graph_bf_searcher, DenseGraphBFSearcher  # noqa: F401
from .bipartite_graph_annealer import bipartite_graph_annealer, BipartiteGraphAnnealer  # noqa: F401
from .bipartite_graph_bf_searcher import bipartite_graph_bf_searcher, BipartiteGraphBFSearcher  # noqa: F401
from . import formulas  # noqa: F401
from .dense_batch import dense_batch, DenseBatch  # noqa: F401
from .serve import server, Server  # noqa: F401
