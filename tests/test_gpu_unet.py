"""Learned warm start: the BF16 MFMA kernels against a plain PyTorch FP32 CPU
reference of the same operations.

Tolerances: activations and gradients are stored in bf16 (8-bit mantissa,
relative rounding 2^-9 = 2e-3 per store); each kernel is compared at 1e-2 of
the reference's max magnitude, the whole network's weight gradients at 5e-2
(rounding compounds over 15 layers) against a reference that rounds to bf16
at the same points."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from jsolver_amd import warmstart as ws

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _bf(x):
    return x.to(torch.bfloat16).float()


def _rel(a, b):
    return float((a - b).abs().max() / max(float(b.abs().max()), 1e-30))


@pytest.mark.parametrize("cin,cout,k,relu", [(4, 16, 3, 1), (16, 32, 3, 1), (48, 16, 3, 0),
                                             (64, 64, 3, 1), (16, 4, 1, 0), (96, 32, 3, 1)])
def test_conv_fwd_dgrad_wgrad(cin, cout, k, relu):
    g = torch.Generator().manual_seed(cin * 100 + cout)
    N, H, W = 2, 12, 20
    L = ws.Layer("t", cin, cout, k, relu, DEV, g)
    L.b.copy_(torch.randn(cout, generator=g).to(DEV) * 0.1)
    x = _bf(torch.randn(N, H, W, cin, generator=g))
    net = ws.UNet.__new__(ws.UNet)
    net.dev = DEV
    y16, y32 = net._conv(L, x.to(torch.bfloat16).to(DEV).contiguous(), N, H, W, out32=True)
    wt = L.torch_weight().cpu()
    ref = F.conv2d(x.permute(0, 3, 1, 2), wt, L.b.cpu(), padding=k // 2).permute(0, 2, 3, 1)
    if relu:
        ref = ref.clamp_min(0)
    assert _rel(y32.cpu(), ref) < 1e-5          # fp32 output: accumulation order only
    assert _rel(y16.float().cpu(), ref) < 1e-2
    # backward against autograd
    dy = _bf(torch.randn(N, H, W, cout, generator=g))
    xr = x.permute(0, 3, 1, 2).clone().requires_grad_(True)
    wr = wt.clone().requires_grad_(True)
    br = L.b.cpu().clone().requires_grad_(True)
    out = F.conv2d(xr, wr, br, padding=k // 2)
    if relu:
        out = out.clamp_min(0)
    out.backward(dy.permute(0, 3, 1, 2))
    dx = net._conv_bwd(L, x.to(torch.bfloat16).to(DEV).contiguous(), y16,
                       dy.to(torch.bfloat16).to(DEV).contiguous(), N, H, W)
    fan = k * k * cin
    gw = L.gw[:, :fan].reshape(cout, k, k, cin).permute(0, 3, 1, 2).cpu()
    assert _rel(gw, wr.grad) < 1e-2
    assert _rel(L.gb.cpu(), br.grad) < 1e-2
    assert _rel(dx.float().cpu(), xr.grad.permute(0, 2, 3, 1)) < 1e-2


def test_pool_upcat_fwd_bwd():
    g = torch.Generator().manual_seed(3)
    N, H, W, C1, C2 = 2, 8, 12, 16, 8
    net = ws.UNet.__new__(ws.UNet)
    net.dev = DEV
    x = _bf(torch.randn(N, H, W, C1, generator=g))
    p = net._pool(x.to(torch.bfloat16).to(DEV), N, H, W, C1).float().cpu()
    ref = F.avg_pool2d(x.permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1)
    assert _rel(p, ref) < 1e-2
    lo = _bf(torch.randn(N, H // 2, W // 2, C1, generator=g))
    sk = _bf(torch.randn(N, H, W, C2, generator=g))
    u = net._upcat(lo.to(torch.bfloat16).to(DEV), sk.to(torch.bfloat16).to(DEV), N, H, W, C1, C2).float().cpu()
    ref = torch.cat([F.interpolate(lo.permute(0, 3, 1, 2), scale_factor=2, mode="nearest"),
                     sk.permute(0, 3, 1, 2)], 1).permute(0, 2, 3, 1)
    assert torch.equal(u, ref)
    dy = _bf(torch.randn(N, H, W, C1 + C2, generator=g))
    dlo, dsk = net._upcat_bwd(dy.to(torch.bfloat16).to(DEV), N, H, W, C1, C2)
    ref_lo = F.avg_pool2d(dy[..., :C1].permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1) * 4
    assert _rel(dlo.float().cpu(), ref_lo) < 1e-2
    assert torch.equal(dsk.float().cpu(), dy[..., C1:])


class _Round(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return _bf(x)

    @staticmethod
    def backward(ctx, g):
        return _bf(g)


def _torch_unet(net, x):
    """CPU FP32 reference of UNet.forward with bf16 rounding where the HIP path
    stores activations (and, in backward, gradients)."""
    Ls = net.layers
    R = _Round.apply
    ws_ = {n: L.torch_weight().cpu().requires_grad_(True) for n, L in Ls.items()}
    bs_ = {n: L.b.cpu().clone().requires_grad_(True) for n, L in Ls.items()}

    def conv(n, t):
        L = Ls[n]
        y = F.conv2d(t, ws_[n], bs_[n], padding=L.k // 2)
        return R(y.clamp_min(0)) if L.relu else y

    t = x.permute(0, 3, 1, 2)
    e1 = conv("e1b", conv("e1a", t))
    e2 = conv("e2b", conv("e2a", R(F.avg_pool2d(e1, 2))))
    e3 = conv("e3b", conv("e3a", R(F.avg_pool2d(e2, 2))))
    b = conv("b2", conv("b1", R(F.avg_pool2d(e3, 2))))
    up = lambda z: F.interpolate(z, scale_factor=2, mode="nearest")
    d3 = conv("d3b", conv("d3a", torch.cat([up(b), e3], 1)))
    d2 = conv("d2b", conv("d2a", torch.cat([up(d3), e2], 1)))
    d1 = conv("d1b", conv("d1a", torch.cat([up(d2), e1], 1)))
    out = conv("out", R(d1))
    return out[:, 0], ws_, bs_


def test_unet_forward_backward_vs_torch():
    net = ws.UNet(seed=5, device=DEV)
    g = torch.Generator().manual_seed(9)
    # non-zero biases so every term of the backward is exercised
    for L in net.layers.values():
        L.b.copy_(0.05 * torch.randn(L.cout, generator=g).to(DEV))
    N, H, W = 2, 32, 24
    x = _bf(torch.randn(N, H, W, 4, generator=g))
    target = torch.randn(N, H, W, generator=g)
    out = net.forward(x.to(torch.bfloat16).to(DEV).contiguous(), keep=True)
    ref, ws_, bs_ = _torch_unet(net, x)
    assert _rel(out.cpu(), ref.detach()) < 2e-2
    diff = out - target.to(DEV)
    net.backward((2.0 / diff.numel()) * diff)
    loss = ((ref - target) ** 2).mean()
    loss.backward()
    for n, L in net.layers.items():
        fan = L.k * L.k * L.cin
        gw = L.gw[:, :fan].reshape(L.cout, L.k, L.k, L.cin).permute(0, 3, 1, 2).cpu()
        assert _rel(gw, ws_[n].grad) < 5e-2, n
        assert _rel(L.gb.cpu(), bs_[n].grad) < 5e-2, n


def test_training_reduces_loss():
    from jsolver_amd.warmstart_train import train
    net = ws.UNet(seed=1, device=DEV)
    g = torch.Generator().manual_seed(2)
    X = torch.randn(8, 32, 32, 4, generator=g)
    # a local target the network can represent: a 5x5 box average of the
    # rhs channel times exp(-log D_xx)
    Y = F.avg_pool2d(X[..., 3][:, None], 5, stride=1, padding=2)[:, 0] * torch.exp(-0.5 * X[..., 0])
    X, Y = X.to(DEV), Y.to(DEV)
    net.out_scale = float(Y.std())
    hist = train(net, X, Y, epochs=60, batch=4, lr=2e-3, log=lambda *_: None)
    assert hist[-1] < 0.5 * hist[0], hist


def test_smooth_guess_matches_oracle():
    """The guess post-smoothing of the warm-start evaluation is `sweeps` RBGS
    sweeps of the fine operator, problem by problem (bit-exact)."""
    from jsolver_amd.warmstart_train import make_problems, smooth_guess
    from oracle import gmg
    m, B, sweeps = 32, 3, 3
    lam, dxx, dyy, rhs = make_problems(B, m, 4, DEV)
    g = torch.Generator(device=DEV).manual_seed(1)
    phi0 = torch.randn(rhs.shape, generator=g, device=DEV, dtype=torch.float64)
    phi0[:, :2] = 0
    phi0[:, -2:] = 0
    phi0[:, :, :2] = 0
    phi0[:, :, -2:] = 0
    got = smooth_guess(phi0, rhs, lam, dxx, dyy, sweeps).cpu().numpy()
    og = gmg.OracleGrid(m)
    for b in range(B):
        op = gmg.build_operator(2, (m, m), (og.idel, og.idel), lam[b].cpu().numpy(),
                                [dxx[b].cpu().numpy(), dyy[b].cpu().numpy()])
        want = phi0[b].cpu().numpy()
        for _ in range(sweeps):
            want = gmg.smooth(op, 0, want, rhs[b].cpu().numpy())
        np.testing.assert_array_equal(got[b], want)
