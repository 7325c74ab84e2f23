"""gvc -- MI355X (gfx950) host library for the differentiable Gaussian-mixture
evaluator, its mesh sampler and its 3-NN initialiser.

``gvc._lib`` binds the C ABI of libgvc.so (include/gvc.h); ``gvc.mixture`` wraps
the mixture entry points for torch tensors; the drop-in packages
``bvh_diff_gaussian_rasterization``, ``gpu_mesh_sampling`` and ``simple_knn`` sit
beside this package and import it.
"""
from ._lib import GvcError, lib  # noqa: F401
