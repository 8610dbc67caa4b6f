"""The learned-preconditioner pipeline at a small scale: training pairs from
the device solver, a few optimisation steps of the U-Net on the library's
kernels, and FWI without it and with it (inference on the library's kernels
and, for comparison, torch's): each accepted step lowers the misfit, and
the plain direction is the fallback."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch_direction(net, g, fields):
    """The same update direction with the network run by torch's bfloat16
    autocast (comparison only; the library's path is HipUNet)."""
    import torch
    from psvwave_amd import precond
    x = torch.tensor(precond.normalise_gradient(np.stack(g))[None], device="cuda")
    with torch.no_grad(), torch.autocast(device_type="cuda", dtype=torch.bfloat16):
        y = net.to("cuda")(x)
    y = y.float()[0].cpu().numpy().astype(np.float64)
    return [yy * f for yy, f in zip(y, fields)]


def test_precond_pipeline_small():
    torch = pytest.importorskip("torch")
    import psvwave_amd
    from psvwave_amd import precond
    cfg = precond.c5_config(nt=600)
    g, p = precond.gradient_pairs(6, seed=1, batch_size=4, cfg=cfg)
    assert g.shape == (6, 3, 316, 188) and p.shape == g.shape
    # pair 5 (second batch) is the gradient one model's run_model gives
    rng = np.random.default_rng([1, 5])
    bg, true = precond.c5_models(rng)
    one = psvwave_amd.fdModel(cfg)
    one.set_parameter_fields(*true)
    for s in range(one.n_shots):
        one.forward_simulate(s, store_fields=False)
    one.set_observed_data(*one.get_synthetic_data())
    one.set_parameter_fields(*bg)
    one.run_model(False, True)
    assert np.array_equal(g[5], np.stack(one.get_kernels()).astype(np.float32))
    del one
    assert np.all(np.isfinite(g)) and np.abs(g).max() > 0
    mask = precond.kernel_mask((316, 188))
    assert np.all(p[:, :, mask == 0] == 0)
    # the gradient vanishes outside the kernel region, as the reference's
    assert np.all(g[:, :, mask == 0] == 0)
    # pairs part of the way to the true model: same models, smaller remaining
    # perturbations, different gradients
    g2, p2 = precond.gradient_pairs(2, seed=1, cfg=cfg, along_path=1.0)
    for k in range(2):
        assert np.abs(p2[k]).max() < np.abs(p[k]).max()
        assert np.all(np.sign(p2[k]) == np.sign(p[k]))
        assert not np.array_equal(g2[k], g[k])
    torch.manual_seed(0)
    net = precond.make_unet(base=8)
    hist = precond.train_hip(net, g, p, epochs=3, batch=3)  # the library's kernels
    assert np.isfinite(hist).all()
    net.eval()
    unet = precond.HipUNet(net)
    bg, true = precond.c5_models(np.random.default_rng(3))
    m = psvwave_amd.fdModel(cfg)
    m.set_parameter_fields(*true)
    for s in range(m.n_shots):
        m.forward_simulate(s, store_fields=False)
    m.set_observed_data(*m.get_synthetic_data())
    for direction in (None, lambda gg, f: precond.learned_direction_hip(unet, gg, f),
                      lambda gg, f: _torch_direction(net, gg, f)):
        h, fields = precond.fwi(m, bg, iterations=3, direction=direction)
        misfits = [x[1] for x in h]
        assert all(b <= a for a, b in zip(misfits, misfits[1:]))
        assert misfits[-1] < misfits[0]
        assert precond.time_to_misfit(h, misfits[-1]) is not None
