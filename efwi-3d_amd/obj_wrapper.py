"""scipy.optimize adapter for a torch module that returns a scalar misfit:
the drop-in for the reference's obj_wrapper.PyTorchObjective
(obj_wrapper.py:10-100), with the same constructor, attributes and methods
(x0, bounds, fun, jac, is_new, cache, unpack_parameters, pack_grads,
pack_bounds) so that 3D_inversion_BFGS.py runs unchanged.

Parameters are packed in named_parameters() order into one float64 vector;
fun(x) and jac(x) share one evaluation of the misfit and its gradient per new
x (a point within 1e-8 of the cached one reuses it); `f` holds the last
misfit, as the reference's save_prog callback reads it. Parameters that live
on a GPU are updated in place on their device.
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np
import torch


class PyTorchObjective:
    def __init__(self, obj, loss):
        self.obj = obj
        self.loss = loss
        params = OrderedDict(obj.named_parameters())
        self.param_shapes = {n: p.size() for n, p in params.items()}
        self.x0 = np.concatenate([p.detach().cpu().numpy().ravel() for p in params.values()]).astype(np.float64)
        bounds = getattr(obj, "Bounds", {})
        self.bounds = self.pack_bounds() if bounds else None

    def unpack_parameters(self, x):
        """Split the flat vector into {name: tensor} of the parameters' shapes."""
        out = OrderedDict()
        i = 0
        for n, shape in self.param_shapes.items():
            k = int(np.prod(shape)) if len(shape) else 1
            out[n] = torch.from_numpy(np.asarray(x[i:i + k]).reshape(tuple(shape)))
            i += k
        return out

    def pack_grads(self):
        return np.concatenate([p.grad.detach().cpu().numpy().ravel()
                               for p in self.obj.parameters()]).astype(np.float64)

    def pack_bounds(self):
        from scipy import optimize
        lo = [np.asarray(self.obj.Bounds[n][0]).ravel() for n in self.param_shapes]
        hi = [np.asarray(self.obj.Bounds[n][1]).ravel() for n in self.param_shapes]
        return optimize.Bounds(np.concatenate(lo).astype(np.float64), np.concatenate(hi).astype(np.float64))

    def is_new(self, x):
        if not hasattr(self, "cached_x"):
            return True
        return float(np.max(np.abs(np.asarray(x) - np.asarray(self.cached_x)))) > 1e-8

    def cache(self, x):
        """Load x into the module, evaluate the misfit and its gradient."""
        values = self.unpack_parameters(x)
        with torch.no_grad():
            for n, p in self.obj.named_parameters():
                p.copy_(values[n].to(device=p.device, dtype=p.dtype))
        self.cached_x = np.array(x, copy=True)
        self.obj.zero_grad()
        f = self.loss()
        self.f = f.item()
        f.backward()
        self.g = self.pack_grads()

    def fun(self, x):
        if self.is_new(x):
            self.cache(x)
        return self.f

    def jac(self, x):
        if self.is_new(x):
            self.cache(x)
        return self.g
