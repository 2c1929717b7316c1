"""obj_wrapper.PyTorchObjective (the scipy adapter of the reference's
driver) on a toy module: packing order, caching, bounds and a scipy L-BFGS-B
run to the known minimum."""
import numpy as np
import torch
from scipy import optimize

from obj_wrapper import PyTorchObjective


class _Toy(torch.nn.Module):
    def __init__(self, bounds=None):
        super().__init__()
        self.a = torch.nn.Parameter(torch.zeros(2, 3))
        self.b = torch.nn.Parameter(torch.zeros(4))
        self.register_buffer("target", torch.arange(10, dtype=torch.float32))
        self.Bounds = bounds or {}
        self.calls = 0

    def forward(self):
        self.calls += 1
        x = torch.cat([self.a.ravel(), self.b])
        return ((x - self.target) ** 2).sum()


def test_pack_unpack_and_cache():
    m = _Toy()
    obj = PyTorchObjective(m, lambda: m())
    assert obj.x0.dtype == np.float64 and obj.x0.shape == (10,) and obj.bounds is None
    x = np.arange(10, dtype=np.float64) * 0.5
    parts = obj.unpack_parameters(x)
    assert list(parts) == ["a", "b"] and tuple(parts["a"].shape) == (2, 3)
    f = obj.fun(x)
    g = obj.jac(x)
    assert m.calls == 1
    assert np.isclose(f, float(((x - np.arange(10)) ** 2).sum()))
    assert np.allclose(g, 2 * (x - np.arange(10)))
    obj.fun(x + 1e-9)
    assert m.calls == 1
    obj.fun(x + 1e-3)
    assert m.calls == 2


def test_scipy_lbfgsb_reaches_the_bounded_minimum():
    lo = {"a": np.zeros((2, 3)), "b": np.zeros(4)}
    hi = {"a": np.full((2, 3), 4.0), "b": np.full(4, 100.0)}
    m = _Toy(bounds={k: (lo[k], hi[k]) for k in lo})
    obj = PyTorchObjective(m, lambda: m())
    r = optimize.minimize(obj.fun, obj.x0, method="L-BFGS-B", jac=obj.jac, bounds=obj.bounds,
                          options={"maxiter": 50})
    want = np.minimum(np.arange(10.0), [4.0] * 6 + [100.0] * 4)
    assert np.allclose(r.x, want, atol=1e-5)
