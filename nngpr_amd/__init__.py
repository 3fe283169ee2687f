"""nngpr_amd: Nearest-Neighbour GP regression (Vecchia) on MI355X (gfx950) HIP kernels.

Drop-in for the hot path of the `nngpr` package: TorchNngpr / NumpyNngpr with
fit / predict / log_marginal_likelihood over White + Constant * RBF kernels, backed
by libnngp_hip.so (C ABI in include/nngp.h).
"""
from . import batched_kernels
from .base_nngpr import HipWorker, Nngpr
from .numpy_nngpr import NumpyNngpr
from .torch_nngpr import TorchNngpr

__all__ = ["batched_kernels", "HipWorker", "Nngpr", "NumpyNngpr", "TorchNngpr"]
__version__ = "0.1.0"
