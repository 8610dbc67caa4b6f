"""bfloat16 MFMA convolution kernels of the learned preconditioner against
torch's fp32 convolutions on the same bfloat16-rounded operands, and the
whole U-Net's inference against the torch network."""
import copy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _conv(lib, x, wt, b, h, w, cin, cout, ks, transposed=False, gelu=False, f32=True,
          out=None, stride=None, off=0):
    import torch
    from psvwave_amd import _lib
    if out is None:
        oh, ow = (2 * h, 2 * w) if transposed else (h, w)
        out = torch.full((oh, ow, cout), float("nan"), device="cuda",
                         dtype=torch.float32 if f32 else torch.bfloat16)
        stride = cout
    _lib.check(lib.psv_conv2d_bf16(x.data_ptr(), wt.data_ptr(), b.data_ptr(), out.data_ptr(),
                                   int(f32), h, w, cin, cout, ks, int(transposed), int(gelu),
                                   stride, off, torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("h,w,cin,cout,ks,gelu", [(13, 7, 8, 5, 3, False),
                                                  (32, 32, 16, 32, 3, True),
                                                  (20, 45, 64, 40, 3, False),
                                                  (9, 11, 24, 70, 1, True),
                                                  (40, 24, 256, 128, 3, True)])
def test_conv_matches_torch(h, w, cin, cout, ks, gelu):
    import torch
    from psvwave_amd import _lib
    lib = _lib.load()
    g = torch.Generator().manual_seed(h * 1000 + cout)
    x = torch.randn(1, cin, h, w, generator=g).bfloat16().float()
    wt = (torch.randn(cout, cin, ks, ks, generator=g) / (cin * ks * ks) ** 0.5).bfloat16().float()
    b = torch.randn(cout, generator=g)
    ref = torch.nn.functional.conv2d(x.double(), wt.double(), b.double(), padding=ks // 2)
    if gelu:
        ref = torch.nn.functional.gelu(ref)
    ref = ref[0].permute(1, 2, 0).float()
    xd = x[0].permute(1, 2, 0).contiguous().to("cuda", torch.bfloat16)
    wd = wt.permute(0, 2, 3, 1).reshape(cout, -1).contiguous().to("cuda", torch.bfloat16)
    out = _conv(lib, xd, wd, b.cuda(), h, w, cin, cout, ks, gelu=gelu).cpu()
    assert torch.allclose(out, ref, rtol=1e-5, atol=5e-5), (out - ref).abs().max()
    out16 = _conv(lib, xd, wd, b.cuda(), h, w, cin, cout, ks, gelu=gelu, f32=False).cpu()
    assert torch.equal(out16, out.bfloat16())  # same accumulator, rounded to nearest even


def test_transposed_conv_into_channel_slice():
    import torch
    from psvwave_amd import _lib
    lib = _lib.load()
    h, w, cin, cout, extra = 10, 6, 64, 32, 16
    g = torch.Generator().manual_seed(5)
    x = torch.randn(1, cin, h, w, generator=g).bfloat16().float()
    wt = (torch.randn(cin, cout, 2, 2, generator=g) / cin ** 0.5).bfloat16().float()
    b = torch.randn(cout, generator=g)
    ref = torch.nn.functional.conv_transpose2d(x.double(), wt.double(), b.double(), stride=2)
    ref = ref[0].permute(1, 2, 0).float()
    xd = x[0].permute(1, 2, 0).contiguous().to("cuda", torch.bfloat16)
    wd = wt.permute(2, 3, 1, 0).reshape(4 * cout, cin).contiguous().to("cuda", torch.bfloat16)
    out = torch.full((2 * h, 2 * w, cout + extra), 7.0, device="cuda")
    _conv(lib, xd, wd, b.cuda(), h, w, cin, cout, 1, transposed=True, out=out,
          stride=cout + extra, off=extra)
    out = out.cpu()
    assert torch.allclose(out[..., extra:], ref, rtol=1e-5, atol=1e-5)
    assert torch.all(out[..., :extra] == 7.0)


def test_maxpool_and_validation():
    import torch
    from psvwave_amd import _lib
    lib = _lib.load()
    x = torch.randn(12, 18, 40, generator=torch.Generator().manual_seed(2)).bfloat16()
    xd = x.cuda()
    out = torch.empty(6, 9, 40, device="cuda", dtype=torch.bfloat16)
    _lib.check(lib.psv_maxpool2_bf16(xd.data_ptr(), out.data_ptr(), 12, 18, 40,
                                     torch.cuda.current_stream().cuda_stream))
    ref = torch.nn.functional.max_pool2d(x.permute(2, 0, 1).float(), 2).permute(1, 2, 0)
    assert torch.equal(out.cpu().float(), ref)
    with pytest.raises(ValueError, match="multiple of 8"):
        _lib.check(lib.psv_conv2d_bf16(xd.data_ptr(), xd.data_ptr(), xd.data_ptr(),
                                       out.data_ptr(), 0, 4, 4, 12, 8, 3, 0, 0, 8, 0, None))
    with pytest.raises(ValueError, match="out_stride"):
        _lib.check(lib.psv_conv2d_bf16(xd.data_ptr(), xd.data_ptr(), xd.data_ptr(),
                                       out.data_ptr(), 0, 4, 4, 8, 8, 3, 0, 0, 12, 8, None))
    with pytest.raises(ValueError, match="even"):
        _lib.check(lib.psv_maxpool2_bf16(xd.data_ptr(), out.data_ptr(), 5, 4, 8, None))


def test_unet_inference_matches_torch():
    import torch
    from psvwave_amd import precond
    torch.manual_seed(0)
    net = precond.make_unet().cuda().eval()
    hip = precond.HipUNet(net)
    rng = np.random.default_rng(3)
    g = rng.standard_normal((3, 150, 77)).astype(np.float32)
    x = torch.tensor(precond.normalise_gradient(g))[None].cuda()
    with torch.no_grad():
        ref = copy.deepcopy(net).cpu().double()(x.cpu().double())[0].float().cuda()
        with torch.autocast(device_type="cuda", dtype=torch.bfloat16):
            ref16 = net(x)[0].float()
    y = hip(precond.normalise_gradient(g))
    assert y.shape == (3, 150, 77) and torch.isfinite(y).all()
    rel = lambda a, b: float((a - b).norm() / b.norm())
    # bfloat16 activations between layers: a few 1e-3 relative on the output,
    # the same size as torch's own bfloat16 autocast error
    assert rel(y, ref) < 2e-2, (rel(y, ref), rel(ref16, ref))
    assert rel(y, ref) < 3 * rel(ref16, ref) + 1e-3


def _nhwc(t):
    return t.permute(0, 2, 3, 1).contiguous()


def _bf(t):
    import torch
    return t.to(torch.bfloat16).float()


@pytest.mark.parametrize("batch,h,w,cin,cout,ks", [(2, 12, 10, 16, 24, 3), (3, 8, 6, 8, 3, 1),
                                                   (1, 20, 14, 64, 40, 3)])
def test_training_convolutions_match_autograd(batch, h, w, cin, cout, ks):
    """Forward with kept pre-activation, input gradient through the flipped
    weights with the GELU derivative fused, weight and bias gradients."""
    import torch
    from psvwave_amd import _lib
    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    g = torch.Generator().manual_seed(batch * 100 + cout)
    x = _bf(torch.randn(batch, cin, h, w, generator=g)).double().requires_grad_(True)
    wt = _bf(torch.randn(cout, cin, ks, ks, generator=g) / (cin * ks * ks) ** 0.5).double()
    wt.requires_grad_(True)
    b = torch.randn(cout, generator=g).double().requires_grad_(True)
    dy = _bf(torch.randn(batch, cout, h, w, generator=g)).double()
    z = torch.nn.functional.conv2d(x, wt, b, padding=ks // 2)
    zb = _bf(z.detach())  # the kernel keeps the bfloat16 pre-activation
    a = torch.nn.functional.gelu(zb)
    (z * dy).sum().backward()  # gradient at z is dy (the GELU handled below)
    bf = torch.bfloat16
    xd = _nhwc(x.detach()).to("cuda", bf)
    wf = wt.detach().permute(0, 2, 3, 1).reshape(cout, -1).to("cuda", bf).contiguous()
    pre = torch.empty(batch, h, w, cout, device="cuda", dtype=bf)
    act = torch.empty(batch, h, w, cout, device="cuda", dtype=bf)
    bd = b.detach().float().cuda()  # every operand stays referenced until the kernel has run
    _lib.check(lib.psv_conv_bf16(xd.data_ptr(), cin, wf.data_ptr(), bd.data_ptr(),
                                 act.data_ptr(), 0, cout, 0, pre.data_ptr(), None,
                                 0, batch, h, w, cin, cout, ks, 0, 1, stream))
    assert torch.allclose(pre.cpu().float(), _nhwc(zb).float(), rtol=2 ** -7, atol=1e-6)
    assert torch.allclose(act.cpu().float(), _nhwc(a).float(), rtol=2 ** -7, atol=1e-5)
    # input gradient through the layer, times gelu'(z') of a made-up
    # previous pre-activation zp
    zp = _bf(torch.randn(batch, cin, h, w, generator=g)).double()
    gp = (torch.distributions.Normal(0., 1.).cdf(zp.float()).double() +
          zp * torch.exp(-0.5 * zp * zp) / (2 * torch.pi) ** 0.5)
    ref_dx = _nhwc(x.grad * gp).float()
    o8 = -(-cout // 8) * 8
    wd = torch.zeros(cin, ks, ks, o8, dtype=torch.float64)
    wd[..., :cout] = wt.detach().flip(2, 3).permute(1, 2, 3, 0)
    wd = wd.reshape(cin, -1).to("cuda", bf).contiguous()
    dyd = torch.zeros(batch, h, w, o8, device="cuda", dtype=bf)
    dyd[..., :cout] = _nhwc(dy).to("cuda", bf)
    dx = torch.empty(batch, h, w, cin, device="cuda")
    zpd = _nhwc(zp).to("cuda", bf)
    _lib.check(lib.psv_conv_bf16(dyd.data_ptr(), o8, wd.data_ptr(), None, dx.data_ptr(), 1, cin,
                                 0, None, zpd.data_ptr(), cin, batch, h, w,
                                 o8, cin, ks, 0, 0, stream))
    assert torch.allclose(dx.cpu(), ref_dx, rtol=1e-4, atol=1e-4), (dx.cpu() - ref_dx).abs().max()
    ws = torch.empty(1 << 20, device="cuda")
    for workspace in (ws, None):
        dw = torch.full((cout, ks * ks * cin), float("nan"), device="cuda")
        dbw = torch.full((cout,), float("nan"), device="cuda")
        _lib.check(lib.psv_conv_wgrad_bf16(dyd.data_ptr(), o8, 0, cout, xd.data_ptr(), cin, cin,
                                           ks, batch, h, w, dw.data_ptr(), dbw.data_ptr(),
                                           None if workspace is None else workspace.data_ptr(),
                                           0 if workspace is None else workspace.numel(),
                                           stream))
        got = dw.cpu().view(cout, ks, ks, cin).permute(0, 3, 1, 2)
        assert torch.allclose(got.double(), wt.grad, rtol=1e-4, atol=1e-4), \
            (got.double() - wt.grad).abs().max()
        assert torch.allclose(dbw.cpu().double(), b.grad, rtol=1e-5, atol=1e-4)
    db = torch.empty(cout, device="cuda")
    _lib.check(lib.psv_channel_sum_bf16(dyd.data_ptr(), o8, 0, cout, batch * h * w,
                                        db.data_ptr(), ws.data_ptr(), ws.numel(), stream))
    assert torch.allclose(db.cpu().double(), b.grad, rtol=1e-5, atol=1e-4)


def test_transposed_convolution_gradients():
    import torch
    from psvwave_amd import _lib
    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream
    bf = torch.bfloat16
    batch, h, w, cin, cout, extra = 2, 6, 5, 32, 16, 8
    g = torch.Generator().manual_seed(11)
    x = _bf(torch.randn(batch, cin, h, w, generator=g)).double().requires_grad_(True)
    wt = _bf(torch.randn(cin, cout, 2, 2, generator=g) / cin ** 0.5).double().requires_grad_(True)
    dy = _bf(torch.randn(batch, cout, 2 * h, 2 * w, generator=g)).double()
    (torch.nn.functional.conv_transpose2d(x, wt, stride=2) * dy).sum().backward()
    # dy sits in the first cout channels of a wider buffer (a concatenation)
    dyd = torch.zeros(batch, 2 * h, 2 * w, cout + extra, device="cuda", dtype=bf)
    dyd[..., :cout] = _nhwc(dy).to("cuda", bf)
    wd = wt.detach().permute(0, 2, 3, 1).reshape(cin, 4 * cout).to("cuda", bf).contiguous()
    zp = torch.zeros(batch, h, w, cin, device="cuda", dtype=bf)  # gelu'(0) = 1/2
    dx = torch.empty(batch, h, w, cin, device="cuda")
    _lib.check(lib.psv_conv_bf16(dyd.data_ptr(), cout + extra, wd.data_ptr(), None, dx.data_ptr(),
                                 1, cin, 0, None, zp.data_ptr(), cin, batch, h, w, cout, cin, 1,
                                 2, 0, stream))
    ref = _nhwc(x.grad).float() * 0.5
    assert torch.allclose(dx.cpu(), ref, rtol=1e-4, atol=1e-4)
    xd = _nhwc(x.detach()).to("cuda", bf)
    dw = torch.empty(4 * cout, cin, device="cuda")
    db = torch.empty(4 * cout, device="cuda")
    ws = torch.empty(1 << 20, device="cuda")
    _lib.check(lib.psv_conv_wgrad_bf16(dyd.data_ptr(), cout + extra, 1, cout, xd.data_ptr(), cin,
                                       cin, 1, batch, h, w, dw.data_ptr(), db.data_ptr(),
                                       ws.data_ptr(), ws.numel(), stream))
    got = dw.cpu().view(2, 2, cout, cin).permute(3, 2, 0, 1).double()
    assert torch.allclose(got, wt.grad, rtol=1e-4, atol=1e-4)
    ref_db = dy.sum(dim=(0, 2, 3))
    assert torch.allclose(db.cpu().double().view(4, cout).sum(0), ref_db, rtol=1e-5, atol=1e-4)


def test_pool_backward_routes_like_torch():
    import torch
    from psvwave_amd import _lib
    lib = _lib.load()
    bf = torch.bfloat16
    batch, h, w, c, cs = 2, 8, 6, 16, 24
    g = torch.Generator().manual_seed(4)
    act = torch.randn(batch, c, h, w, generator=g).round()  # many ties within windows
    act = act.to(bf).float().requires_grad_(True)
    dpool = torch.randn(batch, c, h // 2, w // 2, generator=g)
    torch.nn.functional.max_pool2d(act, 2).backward(dpool)
    skip = torch.randn(batch, h, w, cs, generator=g).to(bf)
    z = torch.randn(batch, h, w, c, generator=g).to(bf)
    zf = z.float()
    gp = torch.distributions.Normal(0., 1.).cdf(zf) + zf * torch.exp(-0.5 * zf * zf) / (2 * torch.pi) ** 0.5
    ref = ((_nhwc(act.grad) + skip[..., cs - c:].float()) * gp).to(bf)
    out = torch.empty(batch, h, w, c, device="cuda", dtype=bf)
    ad, zd = _nhwc(act.detach()).to("cuda", bf), z.cuda()
    dpd, sd = _nhwc(dpool).cuda(), skip.cuda()
    _lib.check(lib.psv_maxpool2_backward_bf16(
        ad.data_ptr(), zd.data_ptr(), dpd.data_ptr(), sd.data_ptr(), cs, cs - c, out.data_ptr(),
        batch, h, w, c, torch.cuda.current_stream().cuda_stream))
    d = (out.cpu().float() - ref.float()).abs()
    assert float(d.max()) <= 2 ** -7 * float(ref.float().abs().max())  # gelu' rounding only


def test_unet_training_gradients_match_torch():
    """One training step's gradients from the kernels against torch autograd
    of the same network in fp64 (bfloat16 operands: compared with torch's own
    bfloat16-autocast error)."""
    import torch
    from psvwave_amd import precond
    torch.manual_seed(1)
    net = precond.make_unet().cuda()
    tr = precond.HipUNetTrainer(net)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((2, 3, 40, 30)).astype(np.float32)
    y = torch.tensor(0.05 * rng.standard_normal((2, 3, 40, 30)), dtype=torch.float32).cuda()
    x = torch.tensor(precond.normalise_gradient(g)).cuda()
    mask = torch.zeros(40, 30, device="cuda")
    mask[5:35, 4:26] = 1.0  # an interior region, as the kernel mask of the full grid
    loss, grads = tr.loss_and_grads(x, y, mask)
    ref_net = copy.deepcopy(net).cpu().double()
    pred = ref_net(x.cpu().double())
    ref_loss = (((pred - y.cpu().double()) * mask.cpu().double()) ** 2).mean() / (
        mask.cpu().double().mean() * 1e-3)
    ref_loss.backward()
    net.zero_grad()
    with torch.autocast(device_type="cuda", dtype=torch.bfloat16):
        p16 = net(x)
    l16 = (((p16.float() - y) * mask) ** 2).mean() / (mask.mean() * 1e-3)
    l16.backward()
    names = {m: n for n, m in net.named_modules()}
    ref_mods = dict(ref_net.named_modules())
    assert len(grads) == 18  # 6 encoder, 2 middle, 3 up, 6 decoder, 1 output
    tot = tot16 = totref = 0.0
    for mod, gw, gb in grads:
        rm = ref_mods[names[mod]]
        for got, ref, g16 in ((gw, rm.weight.grad, mod.weight.grad), (gb, rm.bias.grad, mod.bias.grad)):
            ref = ref.float().cuda()
            assert got.shape == ref.shape
            tot += float(((got - ref) ** 2).sum())
            tot16 += float(((g16 - ref) ** 2).sum())
            totref += float((ref ** 2).sum())
            rel = float((got - ref).norm() / ref.norm())
            assert rel < 0.1, (names[mod], rel)
    rel, rel16 = (tot / totref) ** 0.5, (tot16 / totref) ** 0.5
    assert abs(float(loss.detach()) - float(ref_loss.detach())) < 0.02 * float(ref_loss.detach())
    assert rel < 3 * rel16 + 0.01, (rel, rel16)


def test_mirror_averaged_direction():
    """learned_direction_hip(mirror=True): mean of the prediction and the
    mirrored prediction of the mirrored gradient, times the fields."""
    import torch
    from psvwave_amd import precond
    torch.manual_seed(2)
    hip = precond.HipUNet(precond.make_unet(base=8).cuda().eval())
    rng = np.random.default_rng(7)
    g = [rng.standard_normal((40, 30)) for _ in range(3)]
    fields = [1000 + rng.random((40, 30)) for _ in range(3)]
    x = precond.normalise_gradient(np.stack(g))
    y = hip(x).cpu().numpy().astype(np.float64)
    ym = hip(np.ascontiguousarray(x[:, ::-1])).cpu().numpy().astype(np.float64)[:, ::-1]
    d = precond.learned_direction_hip(hip, g, fields, mirror=True)
    for k in range(3):
        assert np.array_equal(d[k], 0.5 * (y[k] + ym[k]) * fields[k])
    d0 = precond.learned_direction_hip(hip, g, fields)
    assert all(np.array_equal(a, b * f) for a, b, f in zip(d0, y, fields))


def test_ensemble_direction_is_the_mean():
    import torch
    from psvwave_amd import precond
    unets = []
    for seed in (3, 4):
        torch.manual_seed(seed)
        unets.append(precond.HipUNet(precond.make_unet(base=8).cuda().eval()))
    rng = np.random.default_rng(8)
    g = [rng.standard_normal((40, 30)) for _ in range(3)]
    fields = [1000 + rng.random((40, 30)) for _ in range(3)]
    d = precond.learned_direction_ensemble(unets, g, fields, mirror=True)
    d0, d1 = (precond.learned_direction_hip(u, g, fields, mirror=True) for u in unets)
    for k in range(3):
        assert np.array_equal(d[k], (d0[k] + d1[k]) / 2)
