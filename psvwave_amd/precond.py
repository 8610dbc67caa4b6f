"""Learned gradient preconditioner and the FWI driver that compares it with
plain steepest descent (time-to-misfit, learned vs plain).

Workload: a 256 x 128 inner grid plus a 30-point taper (316 x 188),
dx = dz = 10 m, dt = 1 ms, nt = 2500, 8 Hz, 6 shots and 48 receivers along
the surface.  Models are a smooth Gaussian-random-field background plus a
+-5 % random perturbation inside the kernel region.

* ``gradient_pairs``: (gradient at the background, true perturbation) pairs,
  computed for many models at once by the batched solver (``Batch``: one
  launch per time step over every member, forward and adjoint with kernel
  correlation).
* ``UNet``: a 4-level U-Net (32 base channels) mapping the normalised
  velocity kernels to a relative model update.  ``train_hip`` /
  ``HipUNetTrainer`` train it and ``HipUNet`` runs it on this library's
  bfloat16 MFMA kernels (psv_conv.hip: convolutions, input and weight
  gradients, pooling) with fp32 accumulation and fp32 master weights.
* ``fwi``: steepest-descent FWI with a backtracking line search on the
  plain-forward misfit; with a preconditioner, its update direction is
  tried first and the plain gradient is the fallback when the line search
  finds no decrease along it.

The gradient is the reference's (``get_kernels`` after the adjoint,
src/fdModel.cpp:789-969, 1164-1183); the FWI loop follows the descent loop of
tests/test_descent.py:106-129 with a line search added.
"""
import time

import numpy as np

from .synthetic import gaussian_random_field

NP_BOUNDARY = 30
NX_INNER, NZ_INNER = 256, 128
NZ_INNER_BOUNDARY = 8  # sources and receivers sit above the kernel region


def c5_config(nt=2500, n_shots=6, nr=48):
    """Configuration dict of the preconditioner workload (316 x 188 grid)."""
    nxi, nzi = NX_INNER, NZ_INNER
    return {
        "domain": dict(nt=nt, nx_inner=nxi, nz_inner=nzi, nx_inner_boundary=0,
                       nz_inner_boundary=NZ_INNER_BOUNDARY, dx=10.0, dz=10.0, dt=1e-3),
        "boundary": dict(np_boundary=NP_BOUNDARY, np_factor=0.015),
        "medium": dict(scalar_rho=1900.0, scalar_vp=2300.0, scalar_vs=1150.0),
        "sources": dict(peak_frequency=8.0, n_sources=n_shots, n_shots=n_shots,
                        source_timeshift=0.0, delay_cycles_per_shot=0,
                        moment_angles=[30.0 + 20.0 * k for k in range(n_shots)],
                        ix_sources=[int((k + 0.5) * nxi / n_shots) for k in range(n_shots)],
                        iz_sources=[4] * n_shots,
                        which_source_to_fire_in_which_shot=[[k] for k in range(n_shots)]),
        "receivers": dict(nr=nr, ix_receivers=[int((k + 0.5) * nxi / nr) for k in range(nr)],
                          iz_receivers=[2] * nr),
        "inversion": dict(snapshot_interval=10),
        "basis": dict(npx=1, npz=1),
        "output": dict(observed_data_folder=".", stf_folder="."),
    }


def kernel_mask(shape):
    """1 on the kernel region (where the reference accumulates kernels)."""
    m = np.zeros(shape)
    m[NP_BOUNDARY:NP_BOUNDARY + NX_INNER,
      NP_BOUNDARY + NZ_INNER_BOUNDARY:NP_BOUNDARY + NZ_INNER - NZ_INNER_BOUNDARY] = 1.0
    return m


def c5_models(rng):
    """(background, true) models: smooth background (vp 2000-2600 m/s,
    vs/vp 0.48-0.52, rho 1800-2050 kg/m3, correlation length 40 points)
    and the true model = background * (1 + perturbation), the perturbation
    a +-5 % field (correlation length 8-16 points) on the kernel region."""
    shape = (NX_INNER + 2 * NP_BOUNDARY, NZ_INNER + 2 * NP_BOUNDARY)
    vp = 2000.0 + 600.0 * gaussian_random_field(shape, 40, rng)
    vs = vp * (0.48 + 0.04 * gaussian_random_field(shape, 40, rng))
    rho = 1800.0 + 250.0 * gaussian_random_field(shape, 40, rng)
    mask = kernel_mask(shape)
    corr = float(rng.uniform(8, 16))
    p = 0.1 * (gaussian_random_field(shape, corr, rng) - 0.5) * mask
    q = 0.1 * (gaussian_random_field(shape, corr, rng) - 0.5) * mask
    true = (vp * (1 + p), vs * (1 + q), rho * (1 + 0.5 * p))
    return (vp, vs, rho), true


def _simulate(model, fields, store):
    model.set_parameter_fields(*fields)
    for s in range(model.n_shots):
        model.forward_simulate(s, store_fields=store)


def gradient(model, two_lanes=True):
    """Velocity kernels at the model's current parameters against its
    observed data (the misfit is left in model.misfit).  two_lanes: through
    run_model (src/fdModel.cpp:1291-1308), which overlaps the forward of the
    next shot with the adjoint of the current one on two HIP streams; the
    per-shot calls otherwise (bitwise the same result)."""
    if two_lanes:
        model.run_model(False, True)
        return model.get_kernels()
    for s in range(model.n_shots):
        model.forward_simulate(s, store_fields=True)
    model.calculate_l2_misfit()
    model.calculate_l2_adjoint_sources()
    model.reset_kernels()
    for s in range(model.n_shots):
        model.adjoint_simulate(s)
    model.map_kernels_to_velocity()
    return model.get_kernels()


def misfit_at(model, fields):
    """Misfit at `fields`: forward of every shot (two HIP streams, through
    run_model without the adjoint) and the L2 misfit."""
    model.set_parameter_fields(*fields)
    model.run_model(False, False)
    return model.misfit


def gradient_pairs(n, seed=7, batch_size=256, cfg=None, log=None, along_path=0.0):
    """n (gradient, perturbation) pairs as float32 arrays of shape
    (n, 3, nx, nz): the velocity kernels at the background model against
    data of the true model, and (true - background) / background.  With
    along_path = f > 0, a fraction f of the pairs is taken instead at a
    model part of the way to the true one, m = bg + t (true - bg) with
    t ~ U(0.2, 0.8): the kernels there and the remaining (true - m) / m --
    what the network sees in later FWI iterations.

    The device solver runs `batch_size` pairs at once as one Batch: the
    true models' forward runs give the observed data, then Batch.gradients
    computes every member's gradient (bitwise each model's run_model) with
    one launch per time step over all members.  The next batch's models are
    drawn on a host thread meanwhile.  Pair k depends only on (seed, k)."""
    from concurrent.futures import ThreadPoolExecutor
    from .batch import Batch
    cfg = cfg or c5_config()

    def draw(k):
        rng = np.random.default_rng([seed, k])
        bg, true = c5_models(rng)
        if rng.random() < along_path:
            t = rng.uniform(0.2, 0.8)
            bg = tuple(b + t * (tr - b) for b, tr in zip(bg, true))
        return bg, true

    def draw_all(ks):
        return [draw(k) for k in ks]

    size = max(1, min(batch_size, n))
    batch = Batch([cfg] * size)
    shape = (batch.infos[0]["nx"], batch.infos[0]["nz"])
    out_g = np.empty((n, 3) + shape, dtype=np.float32)
    out_p = np.empty((n, 3) + shape, dtype=np.float32)
    starts = list(range(0, n, size))
    with ThreadPoolExecutor(1) as pool:
        nxt = pool.submit(draw_all, range(0, min(n, size)))
        for i, k0 in enumerate(starts):
            models = nxt.result()
            if i + 1 < len(starts):
                k1 = starts[i + 1]
                nxt = pool.submit(draw_all, range(k1, min(n, k1 + size)))
            for j, (_, true) in enumerate(models):
                batch.set_parameter_fields(j, *true)
            batch.forward_all()
            for j, (bg, _) in enumerate(models):
                batch.set_observed_data(j, *batch.get_synthetic_data(j))
                batch.set_parameter_fields(j, *bg)
            batch.gradients()
            for j, (bg, true) in enumerate(models):
                out_g[k0 + j] = np.stack(batch.get_kernels(j))
                out_p[k0 + j] = np.stack([(t - b) / b for t, b in zip(true, bg)])
            if log:
                log(f"training pairs {k0 + len(models)}/{n}")
    return out_g, out_p


def normalise_gradient(g):
    """Per-sample, per-parameter scaling of velocity kernels to unit RMS on
    the kernel region (the network sees shapes, not magnitudes)."""
    g = np.asarray(g, dtype=np.float32)
    single = g.ndim == 3
    if single:
        g = g[None]
    rms = np.sqrt(np.mean(g.astype(np.float64) ** 2, axis=(2, 3), keepdims=True))
    out = (g / np.where(rms > 0, rms, 1.0)).astype(np.float32)
    return out[0] if single else out


def _unet_class():
    import torch
    from torch import nn

    def block(ci, co):
        return nn.Sequential(nn.Conv2d(ci, co, 3, padding=1), nn.GELU(),
                             nn.Conv2d(co, co, 3, padding=1), nn.GELU())

    class UNet(nn.Module):
        """4-level U-Net (two 3 x 3 convolutions per level, channel widths
        `widths`, by default 32, 48, 64, 128: 0.58 M parameters); input and
        output 3 channels (vp, vs, rho) on the full (nx, nz) grid,
        zero-padded to a multiple of 16."""

        def __init__(self, widths=(32, 48, 64, 128)):
            super().__init__()
            c = list(widths)
            self.enc = nn.ModuleList([block(3, c[0]), block(c[0], c[1]), block(c[1], c[2])])
            self.mid = block(c[2], c[3])
            self.up = nn.ModuleList([nn.ConvTranspose2d(c[3], c[2], 2, stride=2),
                                     nn.ConvTranspose2d(c[2], c[1], 2, stride=2),
                                     nn.ConvTranspose2d(c[1], c[0], 2, stride=2)])
            self.dec = nn.ModuleList([block(2 * c[2], c[2]), block(2 * c[1], c[1]),
                                      block(2 * c[0], c[0])])
            self.out = nn.Conv2d(c[0], 3, 1)

        def forward(self, x):
            nx, nz = x.shape[-2:]
            px, pz = (-nx) % 16, (-nz) % 16
            x = nn.functional.pad(x, (0, pz, 0, px))
            skips = []
            for e in self.enc:
                x = e(x)
                skips.append(x)
                x = nn.functional.max_pool2d(x, 2)
            x = self.mid(x)
            for u, d, s in zip(self.up, self.dec, reversed(skips)):
                x = d(torch.cat([u(x), s], dim=1))
            return self.out(x)[..., :nx, :nz]

    return UNet


def make_unet(widths=(32, 48, 64, 128), base=None):
    """The preconditioner network with channel widths `widths` (four
    levels), or (base, 2 base, 4 base, 8 base) when `base` is given --
    base = 32 is the wider 1.9 M-parameter variant.  The library's
    convolution kernels need widths that are multiples of 8."""
    if base is not None:
        widths = (base, 2 * base, 4 * base, 8 * base)
    if len(widths) != 4:
        raise ValueError("four channel widths")
    return _unet_class()(tuple(int(c) for c in widths))


def fwi(model, start, iterations=30, direction=None, step0=0.01, max_halvings=6,
        target=None, log=None, lcap=1.0, expand=False):
    """Steepest-descent FWI from `start` against model's observed data.

    Each iteration computes the gradient and tries, in order, the
    preconditioned update `direction(g, fields)` (when given; first at its
    own scale, later at 1.5 x the last accepted fraction of it) and the
    plain direction (-g, each parameter block scaled so a step
    changes it by at most step0 times its largest value, times an adaptive
    step factor), halving each until the plain-forward misfit decreases (the
    plain direction is the fallback).  The learned scale is capped at lcap
    times the network's own; with expand, an accepted first trial is
    followed by one at twice the step, kept when it lowers the misfit
    further.  Returns the history: list of (elapsed seconds, misfit, used
    fallback)."""
    fields = [np.array(f, dtype=np.float64) for f in start]
    model.set_parameter_fields(*fields)
    t0 = time.perf_counter()
    hist = []
    factor = 1.0   # plain step factor (adaptive)
    lfactor = 1.0  # learned step scale (adaptive, at most the network's own)
    mask = kernel_mask(fields[0].shape)
    for it in range(iterations):
        g = gradient(model)
        f0 = model.misfit
        if it == 0:
            hist.append((time.perf_counter() - t0, f0, False))
        cands = []
        if direction is not None:
            cands.append(("learned", direction(g, fields)))
        # plain direction: -g with each parameter block scaled to its own
        # largest value, so one step changes vp, vs and rho by the same
        # fraction (raw kernels differ by orders of magnitude between them)
        cands.append(("plain", [-gg / max(np.abs(gg).max(), 1e-300) * f.max()
                                for gg, f in zip(g, fields)]))
        accepted = False
        for name, p in cands:
            p = [pp * mask for pp in p]
            if name == "learned":
                # the network predicts the update itself: start from its own
                # scale, or from 1.5 x the last accepted fraction of it
                scale, a = 1.0, lfactor
            else:
                scale = step0
                a = scale * factor
            for k in range(max_halvings):
                trial = [f + a * pp for f, pp in zip(fields, p)]
                ft = misfit_at(model, trial)
                if ft < f0:
                    if expand and k == 0:
                        trial2 = [f + 2 * a * pp for f, pp in zip(fields, p)]
                        ft2 = misfit_at(model, trial2)
                        if ft2 < ft:
                            trial, ft, a = trial2, ft2, 2 * a
                    fields = trial
                    if name == "plain":
                        factor = min(2.0 * a / scale, 4.0)
                    else:
                        lfactor = min(1.5 * a, lcap)
                    accepted = True
                    break
                a *= 0.5
            if not accepted and name == "learned":
                lfactor = max(a, 1.0 / 64)
            if accepted:
                hist.append((time.perf_counter() - t0, ft, name == "plain" and direction is not None))
                break
        if not accepted:
            hist.append((time.perf_counter() - t0, f0, True))
            factor *= 0.25
        model.set_parameter_fields(*fields)
        if log:
            log(f"iteration {it + 1}: misfit {hist[-1][1]:.6e} ({hist[-1][0]:.2f} s)")
        if target is not None and hist[-1][1] <= target:
            break
    return hist, fields


def time_to_misfit(hist, target):
    """Seconds until the misfit first reaches `target` (None if never)."""
    for t, f, _ in hist:
        if f <= target:
            return t
    return None


class HipUNet:
    """Inference of a trained ``UNet`` with this library's bfloat16 MFMA
    kernels (``psv_conv2d_bf16``, ``psv_maxpool2_bf16``): NHWC bfloat16
    activations, fp32 accumulation, bias and erf GELU fused; torch only
    allocates the buffers and does the concatenations' copies."""

    def __init__(self, net, device="cuda"):
        import torch
        from . import _lib
        self._lib = _lib.load()
        self._check = _lib.check
        self.device = device
        bf = torch.bfloat16

        def conv_w(conv):
            w = conv.weight.detach().float().cpu()  # (co, ci, k, k)
            co, ci, kh, kw = w.shape
            cp = -(-ci // 8) * 8
            wp = torch.zeros(co, kh, kw, cp)
            wp[..., :ci] = w.permute(0, 2, 3, 1)
            return (wp.reshape(co, -1).to(device, bf).contiguous(),
                    conv.bias.detach().float().to(device).contiguous(), cp, co, kh)

        def convt_w(ct):
            w = ct.weight.detach().float().cpu()  # (ci, co, 2, 2)
            ci, co = w.shape[:2]
            wp = w.permute(2, 3, 1, 0).reshape(4 * co, ci)  # n = (2a + b) co + o
            return (wp.to(device, bf).contiguous(),
                    ct.bias.detach().float().to(device).contiguous(), ci, co, 1)

        self.enc = [(conv_w(b[0]), conv_w(b[2])) for b in net.enc]
        self.mid = (conv_w(net.mid[0]), conv_w(net.mid[2]))
        self.up = [convt_w(u) for u in net.up]
        self.dec = [(conv_w(b[0]), conv_w(b[2])) for b in net.dec]
        self.out = conv_w(net.out)

    def _conv(self, x, layer, h, w, gelu=True, out=None, stride=None, off=0,
              f32=False, transposed=False):
        import torch
        wt, b, cin, cout, k = layer
        if out is None:
            oh, ow = (2 * h, 2 * w) if transposed else (h, w)
            out = torch.empty(oh, ow, cout, device=self.device,
                              dtype=torch.float32 if f32 else torch.bfloat16)
            stride = cout
        self._check(self._lib.psv_conv2d_bf16(
            x.data_ptr(), wt.data_ptr(), b.data_ptr(), out.data_ptr(), int(f32), h, w, cin,
            cout, 1 if transposed else k, int(transposed), int(gelu), stride, off,
            torch.cuda.current_stream().cuda_stream))
        return out

    def _pool(self, x, h, w, c):
        import torch
        out = torch.empty(h // 2, w // 2, c, device=self.device, dtype=torch.bfloat16)
        self._check(self._lib.psv_maxpool2_bf16(x.data_ptr(), out.data_ptr(), h, w, c,
                                                torch.cuda.current_stream().cuda_stream))
        return out

    def __call__(self, x):
        """x: (3, nx, nz) array or tensor (normalised gradient); returns the
        (3, nx, nz) float32 tensor the network predicts."""
        import torch
        x = torch.as_tensor(np.asarray(x, dtype=np.float32) if not torch.is_tensor(x) else x)
        nx, nz = x.shape[-2:]
        h, w = nx + (-nx) % 16, nz + (-nz) % 16
        cp = self.enc[0][0][2]
        a = torch.zeros(h, w, cp, device=self.device, dtype=torch.bfloat16)
        a[:nx, :nz, :3] = x.to(self.device).permute(1, 2, 0).to(torch.bfloat16)
        skips = []
        for c1, c2 in self.enc:
            a = self._conv(self._conv(a, c1, h, w), c2, h, w)
            skips.append((a, h, w))
            a = self._pool(a, h, w, c2[3])
            h, w = h // 2, w // 2
        a = self._conv(self._conv(a, self.mid[0], h, w), self.mid[1], h, w)
        for up, (c1, c2), (s, sh, sw) in zip(self.up, self.dec, reversed(skips)):
            cs = s.shape[-1]
            cat = torch.empty(sh, sw, up[3] + cs, device=self.device, dtype=torch.bfloat16)
            self._conv(a, up, h, w, gelu=False, out=cat, stride=up[3] + cs, off=0,
                       transposed=True)
            cat[:, :, up[3]:] = s
            h, w = sh, sw
            a = self._conv(self._conv(cat, c1, h, w), c2, h, w)
        y = self._conv(a, self.out, h, w, gelu=False, f32=True)
        return y[:nx, :nz].permute(2, 0, 1).contiguous()


def learned_direction_hip(unet, g, fields, mirror=False):
    """Model update direction from the network run by HipUNet: the
    predicted relative perturbation times the current parameters (a descent
    direction up to the line search's scale).  mirror: the
    average of the prediction and the mirrored prediction for the mirrored
    gradient (the network is trained with mirror augmentation in x)."""
    x = normalise_gradient(np.stack(g))
    y = unet(x).cpu().numpy().astype(np.float64)
    if mirror:
        ym = unet(np.ascontiguousarray(x[:, ::-1])).cpu().numpy().astype(np.float64)
        y = 0.5 * (y + ym[:, ::-1])
    return [yy * f for yy, f in zip(y, fields)]


def learned_direction_ensemble(unets, g, fields, mirror=True):
    """Mean of learned_direction_hip over several trained networks (the
    outcome varies widely between trainings; averaging does not always
    help)."""
    ds = [learned_direction_hip(u, g, fields, mirror=mirror) for u in unets]
    return [sum(d[k] for d in ds) / len(ds) for k in range(len(fields))]


class HipUNetTrainer:
    """Training of a ``UNet`` with this library's bfloat16 MFMA kernels:
    forward convolutions keep their bfloat16 pre-activations, the backward
    pass runs input gradients as convolutions with flipped weights (GELU
    derivative fused), weight gradients as pixel-reduction MFMA GEMMs and
    the pooling/skip gradients in one fused kernel.  The loss and AdamW run
    in torch on fp32 tensors (elementwise work on the small output and the
    1.9 M parameters)."""

    WORKSPACE_FLOATS = 1 << 25  # 128 MiB of split partial sums

    def __init__(self, net, lr=2e-3, weight_decay=1e-4, epochs=20, device="cuda"):
        import torch
        from . import _lib
        self.lib = _lib.load()
        self.check = _lib.check
        self.device = device
        self.net = net.to(device)
        self.opt = torch.optim.AdamW(net.parameters(), lr=lr, weight_decay=weight_decay)
        self.sched = torch.optim.lr_scheduler.CosineAnnealingLR(self.opt, T_max=epochs)
        self.ws = torch.empty(self.WORKSPACE_FLOATS, device=device)

    # ---- weight layouts -------------------------------------------------
    @staticmethod
    def _pad8(c):
        return -(-c // 8) * 8

    def _conv_layout(self, conv):
        """Forward [co][(ky ks + kx) cin8 + ci] and input-gradient
        [ci][(ky ks + kx) co8 + co] (flipped taps) weights, bfloat16."""
        import torch
        w = conv.weight.detach()
        co, ci, k, _ = w.shape
        c8, o8 = self._pad8(ci), self._pad8(co)
        f = torch.zeros(co, k, k, c8, device=self.device)
        f[..., :ci] = w.permute(0, 2, 3, 1)
        d = torch.zeros(ci, k, k, o8, device=self.device)
        d[..., :co] = w.flip(2, 3).permute(1, 2, 3, 0)
        return dict(kind="conv", mod=conv, ci=ci, c8=c8, co=co, o8=o8, k=k,
                    wf=f.reshape(co, -1).to(torch.bfloat16).contiguous(),
                    wd=d.reshape(ci, -1).to(torch.bfloat16).contiguous(),
                    b=conv.bias.detach().float().contiguous())

    def _convt_layout(self, ct):
        import torch
        w = ct.weight.detach()  # (ci, co, 2, 2)
        ci, co = w.shape[:2]
        return dict(kind="convt", mod=ct, ci=ci, co=co,
                    wf=w.permute(2, 3, 1, 0).reshape(4 * co, ci).to(torch.bfloat16).contiguous(),
                    wd=w.permute(0, 2, 3, 1).reshape(ci, 4 * co).to(torch.bfloat16).contiguous(),
                    b=ct.bias.detach().float().contiguous())

    def _layouts(self):
        n = self.net
        return dict(enc=[(self._conv_layout(b[0]), self._conv_layout(b[2])) for b in n.enc],
                    mid=(self._conv_layout(n.mid[0]), self._conv_layout(n.mid[2])),
                    up=[self._convt_layout(u) for u in n.up],
                    dec=[(self._conv_layout(b[0]), self._conv_layout(b[2])) for b in n.dec],
                    out=self._conv_layout(n.out))

    # ---- kernel calls ---------------------------------------------------
    def _stream(self):
        import torch
        return torch.cuda.current_stream().cuda_stream

    def _empty(self, *shape, f32=False):
        import torch
        return torch.empty(*shape, device=self.device,
                           dtype=torch.float32 if f32 else torch.bfloat16)

    def _conv(self, x, in_stride, cin, wt, bias, out, out_stride, out_off, cout, bhw, ks=3,
              mode=0, gelu=False, pre=None, dz=None, dz_stride=0, f32=False):
        b, h, w = bhw
        self.check(self.lib.psv_conv_bf16(
            x.data_ptr(), in_stride, wt.data_ptr(), 0 if bias is None else bias.data_ptr(),
            out.data_ptr(), int(f32), out_stride, out_off,
            0 if pre is None else pre.data_ptr(), 0 if dz is None else dz.data_ptr(),
            dz_stride, b, h, w, cin, cout, ks, mode, int(gelu), self._stream()))
        return out

    def _wgrad(self, dy, dy_stride, gather, dy_ch, x, x_stride, x_ch, ks, bhw, dw):
        """Weight gradient into dw; returns the bias gradient."""
        b, h, w = bhw
        db = self._empty(4 * dy_ch if gather else dy_ch, f32=True)
        self.check(self.lib.psv_conv_wgrad_bf16(
            dy.data_ptr(), dy_stride, int(gather), dy_ch, x.data_ptr(), x_stride, x_ch, ks,
            b, h, w, dw.data_ptr(), db.data_ptr(), self.ws.data_ptr(), self.ws.numel(),
            self._stream()))
        return db.view(4, dy_ch).sum(0) if gather else db

    # ---- network --------------------------------------------------------
    def _block_forward(self, x, cin_stride, cin, L, bhw):
        c1, c2 = L
        b, h, w = bhw
        z1, a1 = self._empty(b, h, w, c1["co"]), self._empty(b, h, w, c1["co"])
        self._conv(x, cin_stride, cin, c1["wf"], c1["b"], a1, c1["co"], 0, c1["co"], bhw,
                   gelu=True, pre=z1)
        z2, a2 = self._empty(b, h, w, c2["co"]), self._empty(b, h, w, c2["co"])
        self._conv(a1, c1["co"], c1["co"], c2["wf"], c2["b"], a2, c2["co"], 0, c2["co"], bhw,
                   gelu=True, pre=z2)
        return dict(x=x, xs=cin_stride, xc=cin, z1=z1, a1=a1, z2=z2, a2=a2, bhw=bhw)

    def forward(self, x, L, keep=True):
        """x: (B, h, w, 8) bfloat16 NHWC input (h, w multiples of 16).
        Returns the fp32 (B, h, w, 3) output and the saved activations."""
        b, h, w, c8 = x.shape
        saved = dict(enc=[], dec=[], up_in=[], cat=[])
        a, stride, ch = x, c8, c8
        for lvl in range(3):
            s = self._block_forward(a, stride, ch, L["enc"][lvl], (b, h, w))
            saved["enc"].append(s)
            c = L["enc"][lvl][1]["co"]
            pooled = self._empty(b, h // 2, w // 2, c)
            self.check(self.lib.psv_maxpool2_bf16(s["a2"].data_ptr(), pooled.data_ptr(),
                                                  b * h, w, c, self._stream()))
            a, stride, ch, h, w = pooled, c, c, h // 2, w // 2
        s = self._block_forward(a, stride, ch, L["mid"], (b, h, w))
        saved["mid"] = s
        a = s["a2"]
        for lvl in range(3):
            up, skip = L["up"][lvl], saved["enc"][2 - lvl]
            cu, cs = up["co"], skip["a2"].shape[-1]
            cat = self._empty(b, 2 * h, 2 * w, cu + cs)
            self._conv(a, up["ci"], up["ci"], up["wf"], up["b"], cat, cu + cs, 0, cu,
                       (b, h, w), ks=1, mode=1)
            cat[..., cu:] = skip["a2"]
            saved["up_in"].append(a)
            saved["cat"].append(cat)
            h, w = 2 * h, 2 * w
            s = self._block_forward(cat, cu + cs, cu + cs, L["dec"][lvl], (b, h, w))
            saved["dec"].append(s)
            a = s["a2"]
        o = L["out"]
        y = self._empty(b, h, w, o["co"], f32=True)
        self._conv(a, o["ci"], o["ci"], o["wf"], o["b"], y, o["co"], 0, o["co"], (b, h, w),
                   ks=1, f32=True)
        return y, saved

    def _block_backward(self, s, L, dz2, grads, dx_f32=False, need_dx=True):
        """dz2: gradient at the block's second pre-activation.  Adds the
        weight/bias gradients to `grads`; returns the gradient at the block
        input (fp32 or bf16), or None."""
        c1, c2 = L
        b, h, w = s["bhw"]
        dw2 = self._empty(c2["co"], 9 * c2["c8"], f32=True)
        grads[c2["mod"]] = (dw2, self._wgrad(dz2, c2["co"], 0, c2["co"], s["a1"], c1["co"],
                                             c2["c8"], 3, s["bhw"], dw2))
        dz1 = self._empty(b, h, w, c1["co"])
        self._conv(dz2, c2["co"], c2["o8"], c2["wd"], None, dz1, c1["co"], 0, c2["ci"],
                   s["bhw"], dz=s["z1"], dz_stride=c1["co"])
        dw1 = self._empty(c1["co"], 9 * c1["c8"], f32=True)
        grads[c1["mod"]] = (dw1, self._wgrad(dz1, c1["co"], 0, c1["co"], s["x"], s["xs"],
                                             c1["c8"], 3, s["bhw"], dw1))
        if not need_dx:
            return None
        dx = self._empty(b, h, w, c1["ci"], f32=dx_f32)
        self._conv(dz1, c1["co"], c1["o8"], c1["wd"], None, dx, c1["ci"], 0, c1["ci"],
                   s["bhw"], f32=dx_f32)
        return dx

    def backward(self, dy, saved, L):
        """dy: fp32 (B, h, w, 3) gradient at the output.  Returns
        {module: (dW in kernel layout, db)}."""
        import torch
        grads = {}
        b, h, w, _ = dy.shape
        o = L["out"]
        dyo = torch.zeros(b, h, w, o["o8"], device=self.device, dtype=torch.bfloat16)
        dyo[..., :o["co"]] = dy
        last = saved["dec"][2]
        dwo = self._empty(o["co"], o["c8"], f32=True)
        grads[o["mod"]] = (dwo, self._wgrad(dyo, o["o8"], 0, o["co"], last["a2"], o["ci"],
                                            o["c8"], 1, (b, h, w), dwo))
        dz2 = self._empty(b, h, w, o["ci"])
        self._conv(dyo, o["o8"], o["o8"], o["wd"], None, dz2, o["ci"], 0, o["ci"], (b, h, w),
                   ks=1, dz=last["z2"], dz_stride=o["ci"])
        dskips = [None] * 3
        for lvl in (2, 1, 0):
            s, up = saved["dec"][lvl], L["up"][lvl]
            dcat = self._block_backward(s, L["dec"][lvl], dz2, grads)
            bb, hh, ww = s["bhw"]
            cu, ccat = up["co"], dcat.shape[-1]
            dskips[2 - lvl] = (dcat, ccat, cu)
            lo = (bb, hh // 2, ww // 2)
            dwu = self._empty(4 * cu, up["ci"], f32=True)
            grads[up["mod"]] = (dwu, self._wgrad(dcat, ccat, 1, cu, saved["up_in"][lvl],
                                                 up["ci"], up["ci"], 1, lo, dwu))
            prev = saved["mid"] if lvl == 0 else saved["dec"][lvl - 1]
            dz2 = self._empty(bb, hh // 2, ww // 2, up["ci"])
            self._conv(dcat, ccat, cu, up["wd"], None, dz2, up["ci"], 0, up["ci"], lo, ks=1,
                       mode=2, dz=prev["z2"], dz_stride=up["ci"])
        dpool = self._block_backward(saved["mid"], L["mid"], dz2, grads, dx_f32=True)
        for lvl in (2, 1, 0):
            s = saved["enc"][lvl]
            bb, hh, ww = s["bhw"]
            c = s["a2"].shape[-1]
            dcat, ccat, cu = dskips[lvl]
            dz2 = self._empty(bb, hh, ww, c)
            self.check(self.lib.psv_maxpool2_backward_bf16(
                s["a2"].data_ptr(), s["z2"].data_ptr(), dpool.data_ptr(), dcat.data_ptr(),
                ccat, cu, dz2.data_ptr(), bb, hh, ww, c, self._stream()))
            dpool = self._block_backward(s, L["enc"][lvl], dz2, grads, dx_f32=True,
                                         need_dx=lvl > 0)
        return grads

    @staticmethod
    def _to_param_grads(L, grads):
        """Kernel-layout gradients -> (module, dW shaped like the weight, db)."""
        out = []
        for mod, (dw, db) in grads.items():
            wsh = mod.weight.shape
            if isinstance(mod, type(L["up"][0]["mod"])):
                ci, co = wsh[:2]
                g = dw.view(2, 2, co, ci).permute(3, 2, 0, 1)
            else:
                co, ci, k, _ = wsh
                g = dw.view(co, k, k, -1)[..., :ci].permute(0, 3, 1, 2)
            out.append((mod, g.contiguous(), db))
        return out

    def prepare(self, x):
        """(B, 3, nx, nz) fp32 tensor -> (B, h, w, 8) bfloat16 NHWC padded."""
        import torch
        b, _, nx, nz = x.shape
        h, w = nx + (-nx) % 16, nz + (-nz) % 16
        xx = torch.zeros(b, h, w, 8, device=self.device, dtype=torch.bfloat16)
        xx[:, :nx, :nz, :3] = x.permute(0, 2, 3, 1)
        return xx

    def loss_and_grads(self, x, y, mask):
        """Masked mean-squared-error loss and the parameter gradients (module,
        dW, db)."""
        import torch
        L = self._layouts()
        nx, nz = x.shape[-2:]
        out, saved = self.forward(self.prepare(x), L)
        pred = out[:, :nx, :nz, :].permute(0, 3, 1, 2).detach().requires_grad_(True)
        loss = (((pred - y) * mask) ** 2).mean() / (mask.mean() * 1e-3)
        loss.backward()
        dy = torch.zeros_like(out)
        dy[:, :nx, :nz, :] = pred.grad.permute(0, 2, 3, 1)
        grads = self.backward(dy, saved, L)
        return loss.detach(), self._to_param_grads(L, grads)

    def step(self, x, y, mask):
        """One AdamW step on a batch; returns the loss (a tensor)."""
        loss, grads = self.loss_and_grads(x, y, mask)
        for mod, gw, gb in grads:
            mod.weight.grad = gw
            mod.bias.grad = gb
        self.opt.step()
        return loss


def train_hip(net, grads, perts, epochs=20, batch=16, lr=2e-3, device="cuda", log=None):
    """train() with the forward and backward passes on HipUNetTrainer's
    kernels (same data order, augmentation, loss and optimiser)."""
    import torch
    tr = HipUNetTrainer(net, lr=lr, epochs=epochs, device=device)
    mask = torch.tensor(kernel_mask(grads.shape[-2:]), dtype=torch.float32, device=device)
    x_all = torch.tensor(normalise_gradient(grads), device=device)
    y_all = torch.tensor(perts, device=device)
    n = x_all.shape[0]
    gen = torch.Generator(device="cpu").manual_seed(0)
    history = []
    for ep in range(epochs):
        te = time.perf_counter()
        perm = torch.randperm(n, generator=gen)
        tot = torch.zeros((), device=device)
        for i in range(0, n, batch):
            idx = perm[i:i + batch].to(device)
            x, y = x_all[idx], y_all[idx]
            if torch.rand(1, generator=gen).item() < 0.5:  # mirror in x
                x, y = torch.flip(x, dims=[2]), torch.flip(y, dims=[2])
            tot += tr.step(x, y, mask) * x.shape[0]
        tr.sched.step()
        history.append(float(tot) / n)
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {history[-1]:.5f} "
                f"({time.perf_counter() - te:.1f} s)")
    return history
