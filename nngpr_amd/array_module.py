"""numpy-like array modules used by the batched kernels and the worker surface.

Mirrors the reference's ArrayModule contract (base_nngpr.py:1467-1555):
to_numpy / from_numpy / to_float32 / cdist plus the numpy functions the batched
kernels use (moveaxis, swapaxes, log, sqrt, exp, arange, where, sum, tile, eye,
full, zeros, ones, empty, asarray, concatenate).
"""
import numpy as np
import torch
from scipy.spatial.distance import cdist as _sp_cdist

_NP2TORCH = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.float16): torch.float16,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.bool_): torch.bool,
}


class NumpyModule:
    """numpy itself plus the ArrayModule extras."""

    def __getattr__(self, name):
        return getattr(np, name)

    def to_numpy(self, x):
        return np.asarray(x)

    def from_numpy(self, x):
        return np.asarray(x)

    def to_float32(self, x):
        return x.astype(np.float32, copy=False)

    def cdist(self, x, y):
        out = np.empty((*x.shape[:-2], x.shape[-2], y.shape[-2]), dtype=x.dtype)
        for index in np.ndindex(x.shape[:-2]):
            out[index] = _sp_cdist(x[index], y[index])
        return out


class TorchModule:
    """torch on one device, with numpy-style call signatures."""

    def __init__(self, device):
        self.torch_device = torch.device(device)

    def __getstate__(self):
        return {"torch_device": str(self.torch_device)}

    def __setstate__(self, state):
        self.torch_device = torch.device(state["torch_device"])

    @staticmethod
    def _dt(dtype):
        if dtype is None:
            return None
        if isinstance(dtype, torch.dtype):
            return dtype
        return _NP2TORCH[np.dtype(dtype)]

    def to_numpy(self, x):
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
        return np.asarray(x)

    def from_numpy(self, x):
        if isinstance(x, torch.Tensor):
            return x.to(self.torch_device)
        return torch.from_numpy(np.ascontiguousarray(x)).to(self.torch_device)

    def to_float32(self, x):
        return x.to(torch.float32)

    def cdist(self, x, y):
        return torch.cdist(x, y)

    def empty(self, shape, dtype=np.float64):
        return torch.empty(shape, dtype=self._dt(dtype), device=self.torch_device)

    def zeros(self, shape, dtype=np.float64):
        return torch.zeros(shape, dtype=self._dt(dtype), device=self.torch_device)

    def ones(self, shape, dtype=np.float64):
        return torch.ones(shape, dtype=self._dt(dtype), device=self.torch_device)

    def full(self, shape, value, dtype=np.float64):
        return torch.full(tuple(shape), value, dtype=self._dt(dtype), device=self.torch_device)

    def ones_like(self, x):
        return torch.ones_like(x)

    def zeros_like(self, x):
        return torch.zeros_like(x)

    def eye(self, n, dtype=np.float64):
        return torch.eye(n, dtype=self._dt(dtype), device=self.torch_device)

    def arange(self, *args, dtype=None):
        return torch.arange(*args, dtype=self._dt(dtype), device=self.torch_device)

    def asarray(self, x, dtype=None, copy=None):
        dt = self._dt(dtype)
        if isinstance(x, torch.Tensor):
            y = x.to(device=self.torch_device, dtype=dt if dt is not None else x.dtype)
            return y.clone() if copy else y
        return torch.as_tensor(np.asarray(x), dtype=dt, device=self.torch_device)

    def tile(self, x, reps):
        return x.repeat(*reps) if len(reps) >= x.ndim else x.repeat(*((1,) * (x.ndim - len(reps)) + tuple(reps)))

    def moveaxis(self, x, a, b):
        return torch.movedim(x, a, b)

    def swapaxes(self, x, a, b):
        return torch.transpose(x, a, b)

    def concatenate(self, xs, axis=0):
        return torch.cat(list(xs), dim=axis)

    def where(self, *args):
        return torch.where(*args)

    def sum(self, x, axis=None):
        if axis is None:
            return torch.sum(x)
        return torch.sum(x, dim=axis)

    def log(self, x):
        return torch.log(x)

    def sqrt(self, x):
        return torch.sqrt(x)

    def exp(self, x):
        return torch.exp(x)

    def sort(self, x, axis=-1):
        return torch.sort(x, dim=axis).values
