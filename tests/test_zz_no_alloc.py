"""efwi_nn_wgrad_noalloc (partial sums in a static device array, no
allocation) against efwi_nn_wgrad (partial sums in a caller's workspace),
through the C ABI: the same weight and bias gradients (up to the order of the
atomic FP32 additions both entry points make) on two calls in a row, the second
on a smaller volume, and a clear refusal when the partial sums would not fit
the static array."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _p(t):
    return C.c_void_p(t.data_ptr())


def _s():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _bf16(rng, *shape):
    return torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(DEV).to(torch.bfloat16)


def _both(lib, dims, mc, kc, per, rng):
    nz, nx, ny = dims
    V = nz * nx * ny
    dpre = _bf16(rng, V, mc)
    src = _bf16(rng, V, kc)
    need = int(lib.efwi_nn_wgrad_scratch(mc, kc, nz, nx, ny, per))
    ws = torch.empty(need, device=DEV)
    ref = torch.full((27, mc, kc), np.nan, device=DEV)
    got = torch.full((27, mc, kc), np.nan, device=DEV)
    ref_b = torch.full((mc,), np.nan, device=DEV)
    got_b = torch.full((mc,), np.nan, device=DEV)
    print(f"[gpu] wgrad noalloc {dims} {mc}x{kc} per={per} partials={need}", flush=True)
    assert lib.efwi_nn_wgrad(_p(dpre), mc, mc, _p(src), kc, kc, _p(ws), need, _p(ref), _p(ref_b), nz, nx, ny, per,
                             _s()) == 0
    assert lib.efwi_nn_wgrad_noalloc(_p(dpre), mc, mc, _p(src), kc, kc, _p(got), _p(got_b), nz, nx, ny, per,
                                     _s()) == 0
    torch.cuda.synchronize()
    return (np.concatenate([ref.cpu().numpy().ravel(), ref_b.cpu().numpy()]),
            np.concatenate([got.cpu().numpy().ravel(), got_b.cpu().numpy()]))


def test_noalloc_wgrad_equals_workspace_wgrad():
    from efwi import _native as N
    lib = N.lib()
    rng = np.random.default_rng(21)
    # first call: a larger volume; second: a smaller one (the static array
    # keeps the first call's partials past the second's range)
    for dims, mc, kc, per in (((8, 12, 40), 32, 48, 8), ((6, 8, 20), 16, 32, 1)):
        assert lib.efwi_nn_wgrad_scratch(mc, kc, *dims, per) <= 4 << 20
        ref, got = _both(lib, dims, mc, kc, per, rng)
        assert np.isfinite(ref).all() and np.abs(ref).max() > 0
        # the same kernels; the per-group sums of the partials are added to dw
        # with atomics in either entry point, in an order that varies between
        # calls once there are more than two groups of 32 chunks: FP32
        # rounding of the order, far inside the 1e-5 bar
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"rel-L2 {err:.2e}, bit-identical {np.array_equal(got, ref)}")
        assert err <= 1e-6, dims


def test_noalloc_wgrad_refuses_partials_beyond_its_array():
    from efwi import _native as N
    lib = N.lib()
    max_floats = 4 << 20
    nz, nx, ny, mc, kc, per = 64, 64, 64, 64, 96, 1
    assert lib.efwi_nn_wgrad_scratch(mc, kc, nz, nx, ny, per) > max_floats
    dpre = torch.zeros(16, mc, dtype=torch.bfloat16, device=DEV)
    src = torch.zeros(16, kc, dtype=torch.bfloat16, device=DEV)
    dw = torch.zeros(27, mc, kc, device=DEV)
    # EFWI_NN_ERR_SCRATCH; nothing is launched (the sizes are checked first)
    assert lib.efwi_nn_wgrad_noalloc(_p(dpre), mc, mc, _p(src), kc, kc, _p(dw), None, nz, nx, ny, per, _s()) == 3
    # invalid arguments stay EFWI_NN_ERR_ARG
    assert lib.efwi_nn_wgrad_noalloc(_p(dpre), mc, mc, _p(src), kc, 8, _p(dw), None, 1, 1, 16, 1, _s()) == 1
