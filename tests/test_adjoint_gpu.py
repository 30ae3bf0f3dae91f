"""The adjoint raymarch on the GPU (the differentiable configuration: a
128^3 float32 normalized density built from the 10,000-strand style, a
512x512 view, 256 steps, L2 loss against a render of the volume with +10%
Gaussian density noise):
  * the density and parameter gradients against the float64 analytic
    gradient (relative L2 error <= 1e-3);
  * central differences (eps 1e-3) on 1,000 random voxels that carry
    gradient, with the piecewise-constant decisions held at their base values
    (float64 re-evaluation of the touched pixels), allowing 2% outliers where
    a quotient straddles a clamp;
  * a torch.autograd.Function over the C ABI (test harness only): backward
    reproduces the adjoint and a few descent steps on the parameters lower
    the loss."""
import numpy as np
import pytest
import torch

from oracle import raymarch as OR
from vkhr_amd import params, synth
from vkhr_amd import device as D

pytestmark = pytest.mark.gpu

W = H = 512
STEPS = 256.0
RES = (128, 128, 128)


class Problem:
    def __init__(self, device):
        s = synth.strands_10k()
        b, vol = D.style_volume(s, RES, device, segments=False, want_f32=True)
        self.bounds = b
        self.dens = vol["density_f32"]
        self.tan = vol["tangents"]
        self.sc = synth.make_scene(b, W, H)
        self.cam, self.light = params.camera(self.sc), params.light(self.sc)
        self.sp = params.strand(b, RES)
        self.rp = params.render(raycast_steps=STEPS)
        self.diff = D.diff_params(0.3, 11.0, 80.0)
        g = torch.Generator(device="cpu").manual_seed(0)
        noise = torch.randn(self.dens.shape, generator=g).to(device)
        self.noisy = torch.clamp(self.dens * (1.0 + 0.1 * noise), min=0.0)
        self.target, _ = self.forward(self.noisy)

    def vol(self, dens):
        return dict(density_f32=dens, tangents=self.tan)

    def forward(self, dens, diff=None):
        return D.raymarch_record(self.vol(dens), RES, self.cam, self.light, self.sp, self.rp,
                                 diff if diff is not None else self.diff, W, H)

    def loss(self, rgba):
        return float(((rgba.double() - self.target.double()) ** 2).sum().item())


@pytest.fixture(scope="module")
def prob(device):
    return Problem(device)


def test_gradient_matches_float64_adjoint(prob):
    rgba, rec = prob.forward(prob.dens)
    d_rgba = 2.0 * (rgba - prob.target)
    gd, gp = D.raymarch_adjoint(prob.vol(prob.dens), RES, prob.cam, prob.light, prob.sp, prob.rp, prob.diff, W, H,
                                rec, d_rgba)
    diff = np.array([0.3, 11.0, 80.0], np.float32)
    _, dd, dp = OR.raymarch_grad(prob.tan.cpu().numpy(), prob.dens.cpu().numpy(), RES, prob.cam, prob.light,
                                 prob.sp, prob.rp, diff, W, H, d_rgba=d_rgba.reshape(-1, 4).cpu().numpy())
    g = gd.double().cpu().numpy()
    rel = np.linalg.norm(g - dd) / np.linalg.norm(dd)
    print("density grad rel L2", rel, "max abs", np.abs(g - dd).max(), "max", np.abs(dd).max())
    assert rel <= 1e-3
    assert np.abs(g - dd).max() <= 1e-3 * np.abs(dd).max()
    gpn = gp.double().cpu().numpy()
    print("params", gpn, dp)
    assert np.allclose(gpn, dp, rtol=1e-3, atol=1e-6 * np.abs(dp).max())


def test_central_differences_on_1000_voxels(prob):
    """The forward is piecewise smooth: its isosurface/fragment decisions are
    piecewise constant in the density, and a +-1e-3 change of one voxel flips
    them for some pixels.  The gradient is that of the smooth piece, so the
    difference quotient is taken with the decisions held at their base values:
    the pixels a perturbation touches (found with the GPU forward) are
    re-evaluated by the float64 oracle with forced decisions."""
    rgba, rec = prob.forward(prob.dens)
    d_rgba = 2.0 * (rgba - prob.target)
    gd, _ = D.raymarch_adjoint(prob.vol(prob.dens), RES, prob.cam, prob.light, prob.sp, prob.rp, prob.diff, W, H,
                               rec, d_rgba)
    base = D.raymarch(prob.vol(prob.dens), RES, prob.cam, prob.light, prob.sp, prob.rp, W, H, want_info=True)
    assert torch.equal(base["rgba"], rgba)
    info = base["info"].reshape(-1, 4).cpu().numpy()
    target = prob.target.reshape(-1, 4).double().cpu().numpy()
    tan = prob.tan.cpu().numpy()
    g = gd.cpu().numpy()
    gmax = np.abs(g).max()
    cand = np.nonzero(np.abs(g) > 1e-2 * gmax)[0]
    rng = np.random.default_rng(11)
    picks = rng.choice(cand, 1000, replace=False)
    eps = 1e-3
    dens = prob.dens.clone()
    flat = rgba.reshape(-1, 4)
    bad, worst = 0, []
    for v in picks:
        b0 = float(dens[v])
        outs = []
        for s in (+1, -1):
            dens[v] = b0 + s * eps
            o, _ = prob.forward(dens)
            outs.append(dens.cpu().numpy())
            if s == 1:
                changed = (o.reshape(-1, 4) != flat).any(dim=1)
            else:
                changed |= (o.reshape(-1, 4) != flat).any(dim=1)
        dens[v] = b0
        pix = torch.nonzero(changed).reshape(-1).cpu().numpy().astype(np.int64)
        if len(pix) == 0:
            fd = 0.0
        else:
            forced = info[pix].astype(np.int32)
            lp = OR.raymarch(None, tan, RES, prob.cam, prob.light, prob.sp, prob.rp, W, H, pixels=pix,
                             density_f32=outs[0], forced=forced)["rgba"]
            lm = OR.raymarch(None, tan, RES, prob.cam, prob.light, prob.sp, prob.rp, W, H, pixels=pix,
                             density_f32=outs[1], forced=forced)["rgba"]
            dp = float(np.float32(b0 + eps)) - float(np.float32(b0 - eps))
            fd = float((((lp - target[pix]) ** 2) - ((lm - target[pix]) ** 2)).sum()) / dp
        if abs(fd - g[v]) > 1e-2 * abs(g[v]) + 1e-3 * gmax:
            bad += 1
            worst.append((int(v), fd, float(g[v])))
    print("FD outliers", bad, worst[:5])
    assert bad <= 20


class RaymarchFn(torch.autograd.Function):
    """Differentiable raymarch over the C ABI (test harness only)."""

    @staticmethod
    def forward(ctx, density, theta, prob):
        diff = D.diff_params(float(theta[0]), float(theta[1]), float(theta[2]))
        rgba, rec = prob.forward(density.detach().contiguous(), diff)
        ctx.save_for_backward(density)
        ctx.rec, ctx.diff, ctx.prob = rec, diff, prob
        return rgba

    @staticmethod
    def backward(ctx, grad_rgba):
        (density,) = ctx.saved_tensors
        p = ctx.prob
        gd, gp = D.raymarch_adjoint(p.vol(density.detach().contiguous()), RES, p.cam, p.light, p.sp, p.rp, ctx.diff,
                                    W, H, ctx.rec, grad_rgba.contiguous())
        return gd, gp, None


def test_autograd_function_and_descent(prob):
    density = prob.dens.clone().requires_grad_(True)
    theta = torch.tensor([0.3, 11.0, 80.0], device=density.device, requires_grad=True)
    rgba = RaymarchFn.apply(density, theta, prob)
    loss = ((rgba - prob.target) ** 2).sum()
    loss.backward()
    rgba2, rec = prob.forward(prob.dens)
    gd, gp = D.raymarch_adjoint(prob.vol(prob.dens), RES, prob.cam, prob.light, prob.sp, prob.rp, prob.diff, W, H,
                                rec, 2.0 * (rgba2 - prob.target))
    assert torch.allclose(density.grad, gd, rtol=1e-4, atol=1e-6 * float(gd.abs().max()))
    assert torch.allclose(theta.grad, gp, rtol=1e-4)
    # The parameters enter no isosurface/fragment decision, so the loss is
    # smooth in them: a few normalized descent steps toward a target rendered
    # at other parameter values lower it.
    target, _ = prob.forward(prob.dens, D.diff_params(0.36, 9.0, 60.0))
    th = torch.tensor([0.3, 11.0, 80.0], device=density.device)
    lr = torch.tensor([0.006, 0.4, 4.0], device=density.device)
    losses = []
    for _ in range(6):
        t = th.clone().requires_grad_(True)
        out = RaymarchFn.apply(prob.dens, t, prob)
        l = ((out - target) ** 2).sum()
        l.backward()
        losses.append(float(l.detach()))
        with torch.no_grad():
            th = th - lr * torch.sign(t.grad)
    print("descent", losses, th.tolist())
    assert losses[-1] < 0.5 * losses[0]
