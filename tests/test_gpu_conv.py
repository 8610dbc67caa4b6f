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
