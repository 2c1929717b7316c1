"""The learned preconditioner's BF16 matrix-core kernels against the CPU
oracle (oracle/nn_oracle.py), and the whole U-Net's forward and weight
gradients against the oracle's network at the same rounding points.

Inputs are BF16 values; the oracle accumulates them in float64. Bars:
convolution outputs (rounded to BF16 on the GPU) within 1 BF16 ulp of the
rounded oracle value, plus a relative 1e-6 of the row's magnitude for the
FP32 accumulation order; weight gradients (FP32) within 1e-5 relative L2.
Inside the network every layer, fed the GPU's own input activations, gives
the correctly rounded oracle value except at near-ties of the FP32
accumulation (at most 1 element in 10^3, never beyond 1 ulp). End to end,
the network's output and per-layer weight gradients stay within 5e-3
relative L2 of the oracle's network: a single flipped BF16 rounding moves
every later layer, so this bar measures propagation, not kernel error
(measured on MI355X at 8x16x32 voxels: at most 2.8e-3; at 8x12x16 one flip in
the first layers moves the first layer's weight gradient by 8e-3)."""
import numpy as np
import pytest
import torch

import nn_oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _bf(rng, *shape, scale=1.0):
    return O.bf16_round((rng.standard_normal(shape) * scale).astype(np.float32))


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(DEV).to(torch.bfloat16)


def _host(t):
    return t.float().cpu().numpy()


def _ulps_ok(got, ref, rel=1e-6):
    ref_b = O.bf16_round(ref.astype(np.float32))
    ulp = np.abs(ref_b).astype(np.float64) * 2.0 ** -7 + 1e-30
    row = np.abs(ref).max() if ref.size else 1.0
    return np.abs(got - ref_b) <= ulp + rel * row + 1e-12


def _p(t, off=0):
    import ctypes as C
    return C.c_void_p(t.data_ptr() + off * t.element_size())


def _s():
    import ctypes as C
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize("dims,kc,nc,relu", [((6, 10, 37), 16, 16, 1), ((4, 8, 16), 32, 64, 0),
                                             ((5, 6, 33), 48, 16, 1), ((4, 4, 20), 96, 32, 1),
                                             # rows of 48 voxels or more: 64-voxel wavefront tiles, the last partial
                                             ((3, 4, 70), 16, 64, 1), ((2, 3, 130), 48, 32, 0),
                                             # 1024 bricks: enough for 2 to 4 output blocks per workgroup
                                             # (smaller volumes take one block per workgroup)
                                             ((32, 128, 128), 16, 64, 1), ((32, 128, 128), 32, 48, 0),
                                             ((32, 128, 128), 32, 32, 1)])
def test_conv3_forward_and_input_gradient(dims, kc, nc, relu):
    from efwi import _native as N
    from efwi.unet import nn_check
    rng = np.random.default_rng(kc + nc)
    nz, nx, ny = dims
    V = nz * nx * ny
    x = _bf(rng, nz, nx, ny, kc)
    w = _bf(rng, 27, nc, kc, scale=0.1)
    b = rng.standard_normal(nc).astype(np.float32)
    print(f"[gpu] conv3 {dims} {kc}->{nc}", flush=True)
    xd, wd = _dev(x.reshape(V, kc)), _dev(w)
    bd = torch.from_numpy(b).to(DEV)
    out = torch.zeros(V, nc, dtype=torch.bfloat16, device=DEV)
    nn_check(N.lib().efwi_nn_conv3(_p(xd), kc, kc, _p(wd), _p(bd), _p(out), nc, nc, None, 0, nz, nx, ny, 0, relu, _s()),
             "conv")
    ref = O.conv3(x, w, bias=b, relu=bool(relu)).reshape(V, nc)
    got = _host(out)
    assert _ulps_ok(got, ref).all(), np.abs(got - ref).max()
    # input gradient with a ReLU mask, written into a channel slice of a wider tensor
    g = _bf(rng, nz, nx, ny, nc)
    mask = _bf(rng, nz, nx, ny, kc)
    wide = torch.zeros(V, kc + 16, dtype=torch.bfloat16, device=DEV)
    md = _dev(mask.reshape(V, kc))
    gd = _dev(g.reshape(V, nc))
    nn_check(N.lib().efwi_nn_conv3(_p(gd), nc, nc, _p(wd), None, _p(wide, 16), kc + 16, kc,
                                   _p(md), kc, nz, nx, ny, 1, 0, _s()), "dgrad")
    ref = (O.conv3_dgrad(g, w) * (mask > 0)).reshape(V, kc)
    got = _host(wide)
    assert not got[:, :16].any()
    assert _ulps_ok(got[:, 16:], ref).all(), np.abs(got[:, 16:] - ref).max()
    # the same input gradient as a forward convolution with the flipped weights
    wt = torch.empty_like(wd)
    nn_check(N.lib().efwi_nn_flip_weights(_p(wd), kc, nc, _p(wt), _s()), "flip")
    assert np.array_equal(_host(wt).reshape(27, kc, nc), w[::-1].transpose(0, 2, 1))
    wide.zero_()
    nn_check(N.lib().efwi_nn_conv3(_p(gd), nc, nc, _p(wt), None, _p(wide, 16), kc + 16, kc,
                                   _p(md), kc, nz, nx, ny, 0, 0, _s()), "dgrad via flipped weights")
    got = _host(wide)
    assert not got[:, :16].any()
    assert _ulps_ok(got[:, 16:], ref).all(), np.abs(got[:, 16:] - ref).max()
    # a slice 4 channels in (8 bytes: outputs written channel by channel, not
    # as 16-byte granules): the same values
    wide.zero_()
    nn_check(N.lib().efwi_nn_conv3(_p(gd), nc, nc, _p(wt), None, _p(wide, 4), kc + 16, kc,
                                   _p(md), kc, nz, nx, ny, 0, 0, _s()), "dgrad into an 8-byte aligned slice")
    got2 = _host(wide)
    assert not got2[:, :4].any() and not got2[:, 4 + kc:].any()
    assert np.array_equal(got2[:, 4:4 + kc], got[:, 16:])


def test_flip_weights_of_several_layers_in_one_launch():
    import ctypes as C
    from efwi import _native as N
    from efwi.unet import nn_check
    rng = np.random.default_rng(3)
    shapes = [(16, 16), (48, 16), (32, 96)]
    ws = [rng.standard_normal((27, nc, kc)).astype(np.float32) for kc, nc in shapes]
    wd = [torch.from_numpy(w).to(DEV).to(torch.bfloat16) for w in ws]
    wt = [torch.full_like(w, 5.0) for w in wd]
    n = len(shapes)
    arr = lambda ts: (C.c_void_p * n)(*[t.data_ptr() for t in ts])
    kcs = (C.c_int * n)(*[kc for kc, _ in shapes])
    ncs = (C.c_int * n)(*[nc for _, nc in shapes])
    print("[gpu] flip_weights_n 3 layers", flush=True)
    nn_check(N.lib().efwi_nn_flip_weights_n(n, arr(wd), kcs, ncs, arr(wt), _s()), "flip n")
    for (kc, nc), w, t in zip(shapes, wd, wt):
        ref = _host(w).reshape(27, nc, kc)[::-1].transpose(0, 2, 1)
        assert np.array_equal(_host(t).reshape(27, kc, nc), ref)
    # counts outside 1..16 and aliasing are refused before any launch
    assert N.lib().efwi_nn_flip_weights_n(0, arr(wd), kcs, ncs, arr(wt), _s()) == 1
    assert N.lib().efwi_nn_flip_weights_n(17, arr(wd), kcs, ncs, arr(wt), _s()) == 1
    assert N.lib().efwi_nn_flip_weights_n(n, arr(wd), kcs, ncs, arr(wd), _s()) == 1


@pytest.mark.parametrize("dims,k1,k2,n1,n2,up", [((6, 10, 70), 32, 16, 16, 0, 0), ((5, 8, 40), 64, 32, 32, 0, 0),
                                                  ((4, 6, 70), 16, 32, 32, 16, 0), ((6, 8, 33), 16, 0, 32, 64, 0),
                                                  ((16, 128, 64), 32, 16, 32, 16, 0), ((6, 10, 70), 32, 16, 16, 0, 1),
                                                  ((8, 12, 130), 64, 32, 32, 0, 1), ((16, 128, 64), 32, 16, 16, 0, 1)])
def test_convolutions_on_split_channels(dims, k1, k2, n1, n2, up):
    """efwi_nn_conv3_cat and efwi_nn_wgrad_cat on an input kept as two
    tensors (the first, with up, a half-size volume read through its nearest
    upsampling), and an output written as two, give efwi_nn_conv3's and
    efwi_nn_wgrad's results on the joined tensors bit for bit, with and without
    a mask."""
    from efwi import _native as N
    from efwi.unet import nn_check
    lib = N.lib()
    rng = np.random.default_rng(k1 * 7 + k2 + n1 + n2 + up)
    nz, nx, ny = dims
    V = nz * nx * ny
    kc, nc = k1 + k2, n1 + n2
    if up:
        coarse = _bf(rng, nz // 2, nx // 2, ny // 2, k1)
        x1 = _dev(coarse.reshape(-1, k1))
        x2h = _bf(rng, nz, nx, ny, k2).reshape(V, k2)
        x = _dev(np.concatenate([O.up2(coarse).reshape(V, k1), x2h], axis=1))
        x2 = _dev(x2h)
    else:
        x = _dev(_bf(rng, nz, nx, ny, kc).reshape(V, kc))
        x1 = x[:, :k1].contiguous()
        x2 = x[:, k1:].contiguous() if k2 else None
    w = _dev(_bf(rng, 27, nc, kc, scale=0.1))
    b = torch.from_numpy(rng.standard_normal(nc).astype(np.float32)).to(DEV)
    m = _dev(_bf(rng, nz, nx, ny, nc).reshape(V, nc))
    print(f"[gpu] split-channel conv {dims} {k1}+{k2} -> {n1}+{n2}", flush=True)
    for mask in (None, m):
        ref = torch.zeros(V, nc, dtype=torch.bfloat16, device=DEV)
        mk = None if mask is None else _p(mask)
        nn_check(lib.efwi_nn_conv3(_p(x), kc, kc, _p(w), _p(b), _p(ref), nc, nc, mk, nc, nz, nx, ny, 0, 1, _s()),
                 "conv")
        o1 = torch.full((V, n1), 3.0, dtype=torch.bfloat16, device=DEV)
        o2 = torch.full((V, n2), 3.0, dtype=torch.bfloat16, device=DEV) if n2 else None
        nn_check(lib.efwi_nn_conv3_cat(_p(x1), k1, k1, up, None if x2 is None else _p(x2), k2, k2, _p(w), _p(b),
                                       _p(o1), n1, n1, None if o2 is None else _p(o2), n2, n2, mk, nc, nz, nx, ny,
                                       0, 1, _s()), "conv cat")
        assert torch.equal(o1, ref[:, :n1])
        if n2:
            assert torch.equal(o2, ref[:, n1:])
            # the second output as a slice 4 channels into a wider tensor (8-byte
            # aligned: both outputs then leave channel by channel)
            wide = torch.full((V, n2 + 16), 3.0, dtype=torch.bfloat16, device=DEV)
            o1.fill_(3.0)
            nn_check(lib.efwi_nn_conv3_cat(_p(x1), k1, k1, up, None if x2 is None else _p(x2), k2, k2, _p(w),
                                           _p(b), _p(o1), n1, n1, _p(wide, 4), n2 + 16, n2, mk, nc, nz, nx, ny, 0,
                                           1, _s()), "conv cat into a slice")
            assert torch.equal(o1, ref[:, :n1])
            assert torch.equal(wide[:, 4:4 + n2], ref[:, n1:])
            assert bool((wide[:, :4] == 3.0).all()) and bool((wide[:, 4 + n2:] == 3.0).all())
    if k2:
        g = _dev(_bf(rng, nz, nx, ny, nc).reshape(V, nc))
        per = 2
        need = int(lib.efwi_nn_wgrad_scratch(nc, kc, nz, nx, ny, per))
        scr = torch.empty(need, device=DEV)
        dw, db = torch.zeros(27, nc, kc, device=DEV), torch.zeros(nc, device=DEV)
        dw2, db2 = torch.full_like(dw, 5.0), torch.full_like(db, 5.0)
        nn_check(lib.efwi_nn_wgrad(_p(g), nc, nc, _p(x), kc, kc, _p(scr), need, _p(dw), _p(db), nz, nx, ny, per,
                                   _s()), "wgrad")
        nn_check(lib.efwi_nn_wgrad_cat(_p(g), nc, nc, _p(x1), k1, k1, up, _p(x2), k2, k2, _p(scr), need, _p(dw2),
                                       _p(db2), nz, nx, ny, per, _s()), "wgrad cat")
        torch.cuda.synchronize()
        # the atomic sums of the partial groups may add in another order
        assert torch.allclose(dw2, dw, rtol=1e-6, atol=1e-6 * float(dw.abs().max()))
        assert torch.allclose(db2, db, rtol=1e-6, atol=1e-6 * float(db.abs().max()))
        # a chunk may not straddle the two inputs: k1 not a multiple of 16 is refused
        assert lib.efwi_nn_wgrad_cat(_p(g), nc, nc, _p(x1), k1, 8, up, _p(x2), k2, k2, _p(scr), need, _p(dw2),
                                     None, nz, nx, ny, per, _s()) == 1


@pytest.mark.parametrize("dims,mc,kc", [((6, 10, 37), 16, 16), ((4, 8, 16), 64, 32), ((4, 4, 20), 32, 96)])
def test_conv3_weight_gradient(dims, mc, kc):
    from efwi import _native as N
    from efwi.unet import nn_check
    rng = np.random.default_rng(mc * kc)
    nz, nx, ny = dims
    V = nz * nx * ny
    dpre = _bf(rng, nz, nx, ny, mc)
    x = _bf(rng, nz, nx, ny, kc)
    per = 3
    need = int(N.lib().efwi_nn_wgrad_scratch(mc, kc, nz, nx, ny, per))
    scratch = torch.empty(need, device=DEV)
    dw = torch.zeros(27, mc, kc, device=DEV)
    print(f"[gpu] wgrad {dims} {mc}x{kc}", flush=True)
    # (every device operand is held in a variable: a temporary would be freed
    # before the asynchronous kernel reads it)
    dd, xd = _dev(dpre.reshape(V, mc)), _dev(x.reshape(V, kc))
    db = torch.full((mc,), 7.0, device=DEV)
    nn_check(N.lib().efwi_nn_wgrad(_p(dd), mc, mc, _p(xd), kc, kc, _p(scratch), need, _p(dw), _p(db), nz, nx, ny, per,
                                   _s()), "wgrad")
    ref = O.conv3_wgrad(dpre, x)
    got = dw.cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    # the bias gradient summed on the side: sum over the voxels of dpre
    ref_b = dpre.reshape(V, mc).astype(np.float64).sum(axis=0)
    assert np.linalg.norm(db.cpu().numpy() - ref_b) <= 1e-5 * np.linalg.norm(ref_b)
    # without a bias output the weight gradient is the same
    dw2 = torch.full_like(dw, 3.0)
    nn_check(N.lib().efwi_nn_wgrad(_p(dd), mc, mc, _p(xd), kc, kc, _p(scratch), need, _p(dw2), None, nz, nx, ny, per,
                                   _s()), "wgrad without bias")
    assert torch.equal(dw2, dw)
    # too small a scratch is refused
    assert N.lib().efwi_nn_wgrad(_p(dd), mc, mc, _p(xd), kc, kc, _p(scratch), need - 1, _p(dw), _p(db), nz, nx, ny,
                                 per, _s()) == 1


@pytest.mark.parametrize("c,ld,off,nvox", [(16, 48, 32, 12345), (64, 64, 0, 3000), (32, 96, 64, 777), (4, 16, 2, 999)])
def test_chan_sum(c, ld, off, nvox):
    """Bias gradients: the 16-byte-load kernel (c, ld multiples of 8, aligned
    slice) and the general one (c = 4, slice at an odd offset of 2 channels)."""
    from efwi import _native as N
    from efwi.unet import nn_check
    rng = np.random.default_rng(c + ld)
    x = _bf(rng, nvox, ld)
    xd = _dev(x)
    out = torch.full((c,), np.nan, device=DEV)
    nn_check(N.lib().efwi_nn_chan_sum(_p(xd, off), ld, c, nvox, _p(out), _s()), "chan_sum")
    ref = x[:, off:off + c].astype(np.float64).sum(0)
    got = out.cpu().numpy()
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-4 * np.abs(x).sum(0).max() / nvox * np.sqrt(nvox)), (got, ref)


@pytest.mark.parametrize("c,off", [(4, 2), (24, 8), (8, 0)])
def test_pool_up_add_granule_and_scalar_paths(c, off):
    """Pooling, upsampling, their gradients and the masked add on channel
    slices at an offset `off` of a wider tensor: 16-byte granules where c and
    the offset are multiples of 8 channels, the per-channel kernels otherwise;
    bit-identical to each other's arithmetic (checked against the oracle)."""
    from efwi import _native as N
    from efwi.unet import nn_check
    lib = N.lib()
    rng = np.random.default_rng(c + off)
    nz, nx, ny = 4, 6, 12
    V, ld = nz * nx * ny, c + off + 8
    x = _bf(rng, nz, nx, ny, c)
    wide = np.zeros((V, ld), np.float32)
    wide[:, off:off + c] = x.reshape(V, c)
    xd = _dev(wide)
    print(f"[gpu] granule paths c={c} off={off}", flush=True)
    p = torch.zeros(V // 8, ld, dtype=torch.bfloat16, device=DEV)
    nn_check(lib.efwi_nn_pool2(_p(xd, off), ld, _p(p, off), ld, c, nz, nx, ny, _s()), "pool")
    assert np.array_equal(_host(p)[:, off:off + c], O.bf16_round(O.pool2(x).reshape(-1, c).astype(np.float32)))
    assert not _host(p)[:, :off].any() and not _host(p)[:, off + c:].any()
    gc = _bf(rng, nz // 2, nx // 2, ny // 2, c)
    gwide = np.zeros((V // 8, ld), np.float32)
    gwide[:, off:off + c] = gc.reshape(-1, c)
    gcd = _dev(gwide)
    acc = _dev(wide)
    nn_check(lib.efwi_nn_pool2_bwd(_p(gcd, off), ld, _p(acc, off), ld, c, nz, nx, ny, 1, None, 0, _s()), "pool bwd")
    assert _ulps_ok(_host(acc)[:, off:off + c], (x + O.pool2_bwd(gc)).reshape(V, c)).all()
    # with the ReLU mask of x: the same rounded sums, zero where x <= 0
    accm = _dev(wide)
    nn_check(lib.efwi_nn_pool2_bwd(_p(gcd, off), ld, _p(accm, off), ld, c, nz, nx, ny, 1, _p(xd, off), ld, _s()),
             "pool bwd masked")
    hm = _host(accm)
    assert np.array_equal(hm[:, off:off + c], np.where(wide[:, off:off + c] > 0, _host(acc)[:, off:off + c], 0.0))
    assert np.array_equal(hm[:, :off], wide[:, :off]) and np.array_equal(hm[:, off + c:], wide[:, off + c:])
    up = torch.zeros(V, ld, dtype=torch.bfloat16, device=DEV)
    nn_check(lib.efwi_nn_up2(_p(gcd, off), ld, _p(up, off), ld, c, nz, nx, ny, _s()), "up")
    assert np.array_equal(_host(up)[:, off:off + c], O.up2(gc).reshape(V, c))
    dc = torch.zeros(V // 8, ld, dtype=torch.bfloat16, device=DEV)
    nn_check(lib.efwi_nn_up2_bwd(_p(xd, off), ld, _p(dc, off), ld, c, nz, nx, ny, _p(gcd, off), ld, _s()), "up bwd")
    assert _ulps_ok(_host(dc)[:, off:off + c], (O.up2_bwd(x) * (gc > 0)).reshape(-1, c)).all()
    a = _dev(wide)
    nn_check(lib.efwi_nn_add(_p(a, off), ld, _p(xd, off), ld, c, V, _p(xd, off), ld, _s()), "add")
    assert _ulps_ok(_host(a)[:, off:off + c], (2 * x * (x > 0)).reshape(V, c)).all()


def test_pool_up_add_head_loss_adam_pack():
    from efwi import _native as N
    from efwi.unet import nn_check
    lib = N.lib()
    rng = np.random.default_rng(5)
    nz, nx, ny, c = 4, 6, 10, 16
    V = nz * nx * ny
    x = _bf(rng, nz, nx, ny, c)
    xd = _dev(x.reshape(V, c))
    # pooling (into a channel slice) and its gradient (accumulating)
    p = torch.zeros(V // 8, 32, dtype=torch.bfloat16, device=DEV)
    nn_check(lib.efwi_nn_pool2(_p(xd), c, _p(p, 16), 32, c, nz, nx, ny, _s()), "pool")
    assert _ulps_ok(_host(p)[:, 16:], O.pool2(x).reshape(-1, c)).all()
    gc = _bf(rng, nz // 2, nx // 2, ny // 2, c)
    gcd = _dev(gc.reshape(-1, c))
    acc = _bf(rng, nz, nx, ny, c)
    accd = _dev(acc.reshape(V, c))
    acc0 = _dev(acc.reshape(V, c))
    nn_check(lib.efwi_nn_pool2_bwd(_p(gcd), c, _p(accd), c, c, nz, nx, ny, 1, None, 0, _s()), "pool bwd")
    assert _ulps_ok(_host(accd), (acc + O.pool2_bwd(gc)).reshape(V, c)).all()
    # upsampling and its gradient with a mask
    up = torch.zeros(V, c, dtype=torch.bfloat16, device=DEV)
    nn_check(lib.efwi_nn_up2(_p(gcd), c, _p(up), c, c, nz, nx, ny, _s()), "up")
    assert np.array_equal(_host(up), O.up2(gc).reshape(V, c))
    m = _bf(rng, nz // 2, nx // 2, ny // 2, c)
    mdv = _dev(m.reshape(-1, c))
    dc = torch.zeros(V // 8, c, dtype=torch.bfloat16, device=DEV)
    nn_check(lib.efwi_nn_up2_bwd(_p(xd), c, _p(dc), c, c, nz, nx, ny, _p(mdv), c, _s()), "up bwd")
    assert _ulps_ok(_host(dc), (O.up2_bwd(x) * (m > 0)).reshape(-1, c)).all()
    # add with mask, and mask alone
    a = _bf(rng, V, c)
    ad = _dev(a)
    nn_check(lib.efwi_nn_add(_p(ad), c, _p(xd), c, c, V, _p(acc0), c, _s()), "add")
    assert _ulps_ok(_host(ad), (a + x.reshape(V, c)) * (acc.reshape(V, c) > 0)).all()
    # mask alone, over the first V/8 voxels (the mask tensor's size)
    nn_check(lib.efwi_nn_add(_p(ad), c, None, 0, c, V // 8, _p(mdv), c, _s()), "mask")
    assert _ulps_ok(_host(ad)[:V // 8], (O.bf16_round(((a + x.reshape(V, c)) * (acc.reshape(V, c) > 0))
                                                      .astype(np.float32))[:V // 8] * (m.reshape(-1, c) > 0))).all()
    # head and its gradient
    act = np.maximum(x, 0).reshape(V, c)
    w = rng.standard_normal(16).astype(np.float32)
    wd, bd = torch.from_numpy(w).to(DEV), torch.tensor([0.25], device=DEV)
    out = torch.zeros(V, device=DEV)
    actd = _dev(act)
    nn_check(lib.efwi_nn_head(_p(actd), c, c, _p(wd), _p(bd), _p(out), V, _s()), "head")
    assert np.allclose(out.cpu().numpy(), O.head(act, w, 0.25), rtol=1e-5, atol=1e-5)
    t = rng.standard_normal(V).astype(np.float32)
    mk = (rng.uniform(size=V) > 0.4).astype(np.float32)
    td, mkd = torch.from_numpy(t).to(DEV), torch.from_numpy(mk).to(DEV)
    dout = torch.zeros(V, device=DEV)
    loss = torch.zeros(1, dtype=torch.float64, device=DEV)
    nn_check(lib.efwi_nn_mse(_p(out), _p(td), _p(mkd), V, float(mk.sum()), _p(dout), _p(loss), _s()), "mse")
    lref, dref = O.mse(out.cpu().numpy(), t, mk, mk.sum())
    assert np.isclose(float(loss), lref, rtol=1e-6)
    assert np.allclose(dout.cpu().numpy(), dref, rtol=1e-5, atol=1e-9)
    dact = torch.zeros(V, c, dtype=torch.bfloat16, device=DEV)
    dw, db = torch.zeros(16, device=DEV), torch.zeros(1, device=DEV)
    nn_check(lib.efwi_nn_head_bwd(_p(actd), c, c, _p(wd), _p(dout), _p(dact), c, _p(dw), _p(db), V, _s()), "head bwd")
    rda, rdw, rdb = O.head_bwd(act, w, dout.cpu().numpy())
    # multiplicative output and its gradient
    ug = torch.zeros(V, device=DEV)
    nn_check(lib.efwi_nn_gate(_p(xd), c, _p(out), _p(ug), V, _s()), "gate")
    assert np.allclose(ug.cpu().numpy(), O.gate(x.reshape(V, c), out.cpu().numpy()), rtol=1e-5, atol=1e-6)
    dgo = torch.zeros(V, device=DEV)
    nn_check(lib.efwi_nn_gate_bwd(_p(out), _p(ug), _p(dout), _p(dgo), V, _s()), "gate bwd")
    assert np.allclose(dgo.cpu().numpy(), O.gate_bwd(out.cpu().numpy(), ug.cpu().numpy(), dout.cpu().numpy()),
                       rtol=1e-6, atol=1e-12)
    assert _ulps_ok(_host(dact), rda).all()
    assert np.allclose(dw.cpu().numpy(), rdw, rtol=1e-4, atol=1e-7)
    assert np.isclose(float(db), rdb, rtol=1e-4, atol=1e-7)
    # Adam, two steps, and the bf16 copy
    n = 1000
    w0 = rng.standard_normal(n).astype(np.float32)
    gs = [rng.standard_normal(n).astype(np.float32) for _ in range(2)]
    wdv = torch.from_numpy(w0.copy()).to(DEV)
    m1, m2 = torch.zeros(n, device=DEV), torch.zeros(n, device=DEV)
    w16 = torch.zeros(n, dtype=torch.bfloat16, device=DEV)
    rw, r1, r2 = w0.astype(np.float64), np.zeros(n), np.zeros(n)
    for k, g in enumerate(gs, 1):
        gdv = torch.from_numpy(g).to(DEV)
        nn_check(lib.efwi_nn_adam(_p(wdv), _p(gdv), _p(m1), _p(m2), n, 1e-2, 0.9, 0.999, 1e-8, k, _p(w16), _s()),
                 "adam")
        rw, r1, r2 = O.adam(rw, g, r1, r2, k, 1e-2)
    assert np.allclose(wdv.cpu().numpy(), rw, rtol=1e-5, atol=1e-6)
    assert np.array_equal(_host(w16), O.bf16_round(wdv.cpu().numpy()))
    # pack and absmax
    g3 = rng.standard_normal((12, 13, 14)).astype(np.float32)
    il = rng.uniform(0.01, 3.0, size=(12, 13, 14)).astype(np.float32)
    g3d, ild = torch.from_numpy(g3).to(DEV), torch.from_numpy(il).to(DEV)
    mx = torch.zeros(1, device=DEV)
    nn_check(lib.efwi_nn_absmax(_p(g3d), g3.size, _p(mx), _s()), "absmax")
    assert float(mx) == np.abs(g3).max()
    feat = torch.zeros(8 * 9 * 10, 16, dtype=torch.bfloat16, device=DEV)
    nn_check(lib.efwi_nn_pack(_p(g3d), _p(ild), 12, 13, 14, 2, 0.5, 1 / 3.0, _p(feat), _s()), "pack")
    assert _ulps_ok(_host(feat), O.pack(g3, il, 2, 0.5, 1 / 3.0).reshape(-1, 16), rel=1e-5).all()


def _oracle_network(net, feat, dims, target, mask, count):
    from unet_reference import oracle_network
    return oracle_network({k: v.cpu().numpy() for k, v in net.w.items()},
                          {k: v.cpu().numpy() for k, v in net.b.items()}, net.wh.cpu().numpy(),
                          float(net.bh.cpu()[0]), _host(feat), dims, target, mask, count)


def _layer_inputs(net, feat, dims):
    """(input tensors, which are joined along channels, output tensor,
    volume) of every convolution of the network's last forward."""
    B = net._bufs[dims]
    d1, d2, d3 = dims, tuple(x // 2 for x in dims), tuple(x // 4 for x in dims)
    return {"enc1a": ((feat,), B["a1"], d1), "enc1b": ((B["a1"],), B["e1"], d1),
            "enc2a": ((B["p1"],), B["a2"], d2), "enc2b": ((B["a2"],), B["e2"], d2),
            "bot_a": ((B["p2"],), B["a3"], d3), "bot_b": ((B["a3"],), B["b3"], d3),
            "dec2a": ((("up", B["b3"], d3), B["e2"]), B["d2a"], d2), "dec2b": ((B["d2a"],), B["d2b"], d2),
            "dec1a": ((("up", B["d2b"], d2), B["e1"]), B["d1a"], d1), "dec1b": ((B["d1a"],), B["d1b"], d1)}


def _joined(srcs):
    """Host array of the channels of srcs side by side; ("up", t, dims) is the
    nearest upsampling of t, a (dims) volume."""
    parts = []
    for t in srcs:
        if isinstance(t, tuple):
            _, c, d = t
            parts.append(O.up2(_host(c).reshape(*d, -1)).reshape(-1, c.shape[1]))
        else:
            parts.append(_host(t))
    return np.concatenate(parts, axis=1)


def _network_case(dims, seed=9):
    from efwi.unet import UNet3D
    rng = np.random.default_rng(seed)
    V = dims[0] * dims[1] * dims[2]
    net = UNet3D(0, seed=1)
    # nonzero biases so that every ReLU mask has both signs, and a head of
    # trained size so that the gate's scaling varies
    for k in net.b:
        net.b[k].copy_(torch.from_numpy(rng.uniform(-0.05, 0.1, net.b[k].numel()).astype(np.float32)))
    net.wh.copy_(torch.from_numpy(rng.standard_normal(16).astype(np.float32) * 0.25))
    f0 = np.zeros((V, 16), np.float32)
    f0[:, :2] = rng.standard_normal((V, 2))
    target = rng.standard_normal(V).astype(np.float32)
    return net, _dev(f0), target, np.ones(V, np.float32), V


@pytest.mark.parametrize("dims", [(8, 12, 16), (8, 16, 32)])
def test_unet_layers_on_network_activations_match_oracle(dims):
    """Every convolution of a forward pass against the oracle convolution of
    the same (GPU-computed) input activations: correctly rounded except at
    near-ties of the FP32 accumulation."""
    from efwi.unet import LAYERS
    net, feat, _, _, _ = _network_case(dims)
    print(f"[gpu] U-Net layers at {dims}", flush=True)
    net.forward(feat, dims)
    torch.cuda.synchronize()
    spec = _layer_inputs(net, feat, dims)
    for name, cin, cout, _ in LAYERS:
        srcs, dst, d = spec[name]
        x = _joined(srcs).reshape(*d, cin)
        w = _host(net.w16[name]).reshape(27, cout, cin)
        ref = O.conv3(x, w, bias=net.b[name].cpu().numpy(), relu=True).reshape(-1, cout)
        got = _host(dst)
        assert _ulps_ok(got, ref).all(), name
        assert np.mean(got != O.bf16_round(ref.astype(np.float32))) <= 1e-3, name


def test_unet_forward_and_weight_gradients_match_oracle():
    dims = (8, 16, 32)
    net, feat, target, mask, V = _network_case(dims)
    print(f"[gpu] U-Net {dims}, {net.nparams()} parameters", flush=True)
    out = net.forward(feat, dims)
    loss, dout = net.loss(out, torch.from_numpy(target).to(DEV), torch.from_numpy(mask).to(DEV), V)
    net.backward(dout)
    torch.cuda.synchronize()
    ref_out, G = _oracle_network(net, feat, dims, target, mask, V)
    got_out = out.cpu().numpy()
    assert np.linalg.norm(got_out - ref_out.ravel()) <= 5e-3 * np.linalg.norm(ref_out)
    worst = {}
    for k, ref in G.items():
        got = net.grads[k].cpu().numpy().reshape(np.shape(ref))
        worst[k] = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-30))
    print("[gpu] relative L2 of the weight gradients:", {k: round(v, 5) for k, v in worst.items()}, flush=True)
    for k, e in worst.items():
        assert e < 5e-3, (k, e)


@pytest.mark.parametrize("log_scaling", [False, True])
def test_fused_loss_and_head_gradient_equal_the_separate_calls(log_scaling):
    """train_step's single loss-and-head-gradient kernel against loss() then
    backward(): every convolution's gradients bit for bit (the head's input
    gradient is the same per voxel), the head's own gradients and the loss up
    to the order of their sums."""
    dims = (8, 16, 32)
    net, feat, target, mask, V = _network_case(dims)
    tg, mk = torch.from_numpy(target).to(DEV), torch.from_numpy(mask * 0.5).to(DEV)
    print(f"[gpu] fused loss tail, log_scaling={log_scaling}", flush=True)
    out = net.forward(feat, dims)
    o = net._bufs[dims]["o"]
    loss, dout = net.loss(o if log_scaling else out, tg, mk, V)
    net.backward(dout, through_gate=not log_scaling)
    torch.cuda.synchronize()
    ref = {k: v.clone() for k, v in net.grads.items()}
    ref_loss = float(loss)
    for v in net.grads.values():
        v.fill_(7.0)
    # train_step = forward, the fused tail, the body, Adam: skip Adam by lr 0
    steps = net.steps
    got_loss = float(net.train_step(feat, tg, mk, V, dims, lr=0.0, log_scaling=log_scaling))
    torch.cuda.synchronize()
    assert net.steps == steps + 1
    assert abs(got_loss - ref_loss) <= 1e-12 * abs(ref_loss)
    for k, v in ref.items():
        if k.startswith("head."):
            assert torch.allclose(net.grads[k], v, rtol=1e-5, atol=1e-6 * float(v.abs().max())), k
        else:
            assert torch.equal(net.grads[k], v), k


def test_unet_training_lowers_the_loss():
    """A short training run on the GPU's own kernels: learn the pseudo-Hessian
    scaling g / (sqrt(illumination) + eps) of the two input channels on random
    smooth volumes; the loss must fall well below its start."""
    from scipy.ndimage import gaussian_filter
    from efwi.unet import UNet3D
    rng = np.random.default_rng(2)
    dims = (16, 16, 16)
    V = dims[0] * dims[1] * dims[2]
    net = UNet3D(0, seed=3)
    data = []
    for _ in range(4):
        g = gaussian_filter(rng.standard_normal(dims), 1.5)
        il = np.exp(gaussian_filter(rng.standard_normal(dims), 3.0) * 2)
        g /= np.abs(g).max()
        il /= il.max()
        u = g / (np.sqrt(il) + 0.05)
        u /= np.abs(u).max()
        f = np.zeros((V, 16), np.float32)
        f[:, 0] = g.ravel()
        f[:, 1] = np.log10(il.ravel() + 1e-6) / 6 + 1
        data.append((_dev(f), torch.from_numpy(u.ravel().astype(np.float32)).to(DEV)))
    mask = torch.ones(V, device=DEV)
    losses = []
    for it in range(150):
        feat, t = data[it % 4]
        losses.append(float(net.train_step(feat, t, mask, V, dims, lr=3e-3)))
    print(f"[gpu] U-Net training loss {losses[0]:.3e} -> {np.mean(losses[-8:]):.3e}", flush=True)
    assert np.isfinite(losses).all()
    assert np.mean(losses[-8:]) < 0.3 * np.mean(losses[:4])
