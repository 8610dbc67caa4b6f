"""Host-side pieces of the learned preconditioner: workload geometry,
normalisation and the network's shape contract."""
import numpy as np
import pytest

from psvwave_amd import precond


def test_models_and_mask():
    bg, true = precond.c5_models(np.random.default_rng(0))
    assert bg[0].shape == (316, 188)
    mask = precond.kernel_mask(bg[0].shape)
    for b, t in zip(bg, true):
        rel = (t - b) / b
        assert np.all(rel[mask == 0] == 0)
        assert np.abs(rel).max() <= 0.05 + 1e-12
    cfg = precond.c5_config()
    vmax = max(f.max() for f in (bg[0], true[0]))
    assert vmax * cfg["domain"]["dt"] / cfg["domain"]["dx"] < 0.6  # CFL


def test_normalisation_is_per_sample_and_parameter():
    rng = np.random.default_rng(1)
    g = rng.standard_normal((2, 3, 20, 10)) * np.array([1e3, 1e-5, 7.0])[None, :, None, None]
    n = precond.normalise_gradient(g)
    rms = np.sqrt((n.astype(np.float64) ** 2).mean(axis=(2, 3)))
    assert np.allclose(rms, 1.0, rtol=1e-5)
    assert np.array_equal(precond.normalise_gradient(g[0]), n[0])


def test_unet_shapes():
    torch = pytest.importorskip("torch")
    net = precond.make_unet(base=4)
    y = net(torch.zeros(1, 3, 316, 188))
    assert y.shape == (1, 3, 316, 188)


def _host_trainer():
    """HipUNetTrainer's layout helpers without a GPU (no library calls)."""
    tr = precond.HipUNetTrainer.__new__(precond.HipUNetTrainer)
    tr.device = "cpu"
    return tr


def _bf(t):
    import torch
    return t.to(torch.bfloat16).double()


@pytest.mark.parametrize("ci,co,k", [(3, 32, 3), (64, 40, 3), (32, 3, 1)])
def test_conv_layouts_match_torch(ci, co, k):
    """Forward weights [co][(ky k + kx) c8 + ci] and flipped input-gradient
    weights [ci][(ky k + kx) o8 + co], checked as torch convolutions."""
    torch = pytest.importorskip("torch")
    g = torch.Generator().manual_seed(ci * 7 + co)
    conv = torch.nn.Conv2d(ci, co, k, padding=k // 2)
    with torch.no_grad():
        conv.weight.copy_(torch.randn(conv.weight.shape, generator=g))
    L = _host_trainer()._conv_layout(conv)
    c8, o8 = L["c8"], L["o8"]
    assert c8 % 8 == 0 and o8 % 8 == 0 and c8 >= ci and o8 >= co
    w = _bf(conv.weight.detach())
    wf = L["wf"].double().view(co, k, k, c8)
    assert torch.equal(wf[..., :ci].permute(0, 3, 1, 2), w)
    assert torch.all(wf[..., ci:] == 0)
    wd = L["wd"].double().view(ci, k, k, o8)
    assert torch.all(wd[..., co:] == 0)
    # the input gradient of the convolution is a convolution of the output
    # gradient with wd (as a (ci, co, k, k) weight), "same" padding
    x = torch.randn(1, ci, 9, 7, generator=g, dtype=torch.float64, requires_grad=True)
    dy = torch.randn(1, co, 9, 7, generator=g, dtype=torch.float64)
    (torch.nn.functional.conv2d(x, w, padding=k // 2) * dy).sum().backward()
    dx = torch.nn.functional.conv2d(dy, wd[..., :co].permute(0, 3, 1, 2), padding=k // 2)
    assert torch.allclose(dx, x.grad, rtol=1e-12, atol=1e-12)


def test_transposed_layouts_and_gradient_mapping():
    """2x2 stride-2 transposed convolution: forward [(2a + b) co + o][ci],
    input gradient [ci][(2a + b) co + o]; kernel-layout gradients map back to
    torch's weight shapes."""
    torch = pytest.importorskip("torch")
    g = torch.Generator().manual_seed(3)
    ct = torch.nn.ConvTranspose2d(16, 8, 2, stride=2)
    with torch.no_grad():
        ct.weight.copy_(torch.randn(ct.weight.shape, generator=g))
    tr = _host_trainer()
    L = tr._convt_layout(ct)
    w = _bf(ct.weight.detach())  # (ci, co, 2, 2)
    x = torch.randn(1, 16, 5, 4, generator=g, dtype=torch.float64)
    y = torch.nn.functional.conv_transpose2d(x, w, stride=2)  # (1, 8, 10, 8)
    wf = L["wf"].double()  # [(2a + b) 8 + o][ci]
    for a in range(2):
        for b in range(2):
            blk = wf[(2 * a + b) * 8:(2 * a + b + 1) * 8]  # (co, ci)
            got = torch.einsum("oc,chw->ohw", blk, x[0])
            assert torch.allclose(got, y[0, :, a::2, b::2], rtol=1e-12, atol=1e-12)
    assert torch.equal(L["wd"].double().view(16, 2, 2, 8).permute(0, 3, 1, 2), w)
    # gradients in kernel layout -> weight shapes (the inverse of the layouts)
    conv = torch.nn.Conv2d(8, 24, 3, padding=1)
    Lc = tr._conv_layout(conv)
    layouts = {"up": [L]}
    dw_conv = Lc["wf"].float().clone()
    dw_ct = L["wf"].float().clone()
    out = dict((id(m), (gw, gb)) for m, gw, gb in precond.HipUNetTrainer._to_param_grads(
        layouts, {conv: (dw_conv, torch.zeros(24)), ct: (dw_ct, torch.zeros(8))}))
    assert torch.equal(out[id(conv)][0].double(), _bf(conv.weight.detach()))
    assert torch.equal(out[id(ct)][0].double(), w)
