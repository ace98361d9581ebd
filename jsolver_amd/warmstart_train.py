"""Training data, training loop and evaluation of the learned warm start.

Data: 2D anisotropic problems from the coefficient generator with correlation
length ~ U[0.02, 0.1], anisotropy ratio r ~ U[1, 100], lambda ~ U[0, 10] per
cell and rhs = sin(2 pi x) sin(3 pi y) + N(0, 0.1^2); D_xx = exp(GRF),
D_yy = r exp(GRF).  Fields are drawn on the GPU (torch's FFT as plumbing,
seeded per batch) and solved by the project's own batched FP64 multigrid
(solve_batched_2d) to the relative residual `label_rtol`; those solutions are
the training targets.

    python -m jsolver_amd.warmstart_train --samples 20000 --m 256 --epochs 4 \
        --out unet.pt --eval-m 512 --eval-batch 64
"""
import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from .grid_1d import grid_1D
from .solver_2d import solve_batched_2d
from .warmstart import UNet, _chk, _lib_unet, _p, _sp, features


def grf_batch(B, n, h, corr, gen, dev):
    """B Gaussian random fields on (n, n), each with its own correlation length
    corr[b]: periodic Gaussian filter exp(-2 pi^2 l^2 |k|^2), standardised."""
    noise = torch.randn((B, n, n), generator=gen, device=dev, dtype=torch.float64)
    f = torch.fft.rfft2(noise)
    ky = torch.fft.fftfreq(n, d=h, device=dev, dtype=torch.float64)
    kx = torch.fft.rfftfreq(n, d=h, device=dev, dtype=torch.float64)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    filt = torch.exp(-2.0 * math.pi ** 2 * corr[:, None, None] ** 2 * k2[None])
    g = torch.fft.irfft2(f * filt, s=(n, n))
    g = g - g.mean(dim=(1, 2), keepdim=True)
    return g / g.std(dim=(1, 2), keepdim=True)


def make_problems(B, m, seed, dev):
    """One batch of problems on an m x m grid: (lam, dxx, dyy, rhs) each
    (B, m+3, m+3) float64 on the GPU."""
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    g = grid_1D(m)
    n = m + 3
    corr = 0.02 + 0.08 * torch.rand(B, generator=gen, device=dev, dtype=torch.float64)
    r = 1.0 + 99.0 * torch.rand(B, generator=gen, device=dev, dtype=torch.float64)
    dxx = torch.exp(grf_batch(B, n, g.del_, corr, gen, dev))
    dyy = r[:, None, None] * torch.exp(grf_batch(B, n, g.del_, corr, gen, dev))
    lam = 10.0 * torch.rand((B, n, n), generator=gen, device=dev, dtype=torch.float64)
    c = torch.tensor(g.centers, device=dev)
    base = torch.sin(2 * math.pi * c[None, :]) * torch.sin(3 * math.pi * c[:, None])
    rhs = base[None] + 0.1 * torch.randn((B, n, n), generator=gen, device=dev, dtype=torch.float64)
    return lam, dxx, dyy, rhs


def generate(samples, m, batch, seed, label_rtol, dev, log=print, gamma=1, max_cycles=50000):
    """Solved training set, kept on the GPU: features (S, m, m, 4) fp32 raw and
    targets (S, m, m) fp32 (phi on nodes 1..m).  Problems whose solve does
    not stay finite are dropped: with the reference's non-Galerkin coarse
    operators a W-cycle (gamma = 2) diverges on some strongly anisotropic
    fields that V-cycles solve, so labels use V-cycles by default."""
    g = grid_1D(m)
    X, Y, counts = [], [], []
    dropped = 0
    t0 = time.perf_counter()
    for i in range(0, samples, batch):
        B = min(batch, samples - i)
        lam, dxx, dyy, rhs = make_problems(B, m, seed + i, dev)
        phi, cnt, stopv = solve_batched_2d(g, g, torch.zeros_like(rhs), rhs, lam, dxx, dyy, gamma=gamma,
                                           stop="rel_residual", rtol=label_rtol, max_iters_outer=max_cycles)
        ok = torch.isfinite(phi).flatten(1).all(1) & torch.as_tensor(stopv <= label_rtol, device=dev)
        dropped += int((~ok).sum())
        sl = (ok, slice(1, -2), slice(1, -2))
        X.append(torch.stack([torch.log(dxx[sl]), torch.log(dyy[sl]), lam[sl], rhs[sl]], -1).float())
        Y.append(phi[sl].float())
        counts.append(cnt[ok.cpu().numpy()])
        log(f"[data] {i + B}/{samples} solved ({time.perf_counter() - t0:.1f}s, "
            f"mean cycles {np.mean(cnt):.1f}, dropped {dropped})")
    return torch.cat(X), torch.cat(Y), np.concatenate(counts)


def train(net, X, Y, epochs, batch, lr, log=print, seed=0, on_epoch=None):
    """Adam on MSE of the normalised output; all kernels hand-written.
    on_epoch(epoch, train_loss) runs after every epoch (learning curves)."""
    S = X.shape[0]
    gen = torch.Generator(device=X.device)
    gen.manual_seed(seed)
    steps_per_epoch = S // batch
    total = epochs * steps_per_epoch
    step = 0
    t0 = time.perf_counter()
    hist = []
    shift = net.in_shift.to(X.device)
    scale = net.in_scale.to(X.device)
    for ep in range(epochs):
        perm = torch.randperm(S, generator=gen, device=X.device)
        run = 0.0
        for j in range(steps_per_epoch):
            idx = perm[j * batch:(j + 1) * batch]
            xb = ((X[idx] - shift) / scale).to(torch.bfloat16).contiguous()
            yb = Y[idx] / net.out_scale
            out = net.forward(xb, keep=True)
            diff = out - yb
            loss = float((diff * diff).mean())
            dout = (2.0 / diff.numel()) * diff
            net.backward(dout)
            cur_lr = lr * 0.5 * (1 + math.cos(math.pi * step / max(1, total)))
            net.adam_step(lr=cur_lr)
            run += loss
            step += 1
            if step % 50 == 0:
                log(f"[train] epoch {ep} step {step}/{total} loss {run / (j + 1):.4e} "
                    f"({time.perf_counter() - t0:.1f}s)")
        hist.append(run / max(1, steps_per_epoch))
        if on_epoch is not None:
            on_epoch(ep, hist[-1])
    return hist


def fit_normalisation(net, X, Y):
    flat = X.reshape(-1, 4)
    net.in_shift = flat.mean(0).cpu()
    net.in_scale = flat.std(0).clamp_min(1e-6).cpu()
    net.out_scale = float(Y.std())


def guess_phi(net, lam, dxx, dyy, rhs, m_train):
    """Initial phi for (B, m+3, m+3) problems.  A grid finer than the training
    grid is predicted on the training resolution (inputs averaged 2x2 per
    level, i.e. the same physical scales the network saw) and the prediction
    is interpolated bilinearly to the fine nodes."""
    B, n, _ = rhs.shape
    m = n - 3
    if m == m_train:
        out = net.forward(features(net, lam, dxx, dyy, rhs)).double() * net.out_scale
    else:
        f = int(m // m_train)
        sl = (slice(None), slice(1, -2), slice(1, -2))
        raw = torch.stack([torch.log(dxx[sl]), torch.log(dyy[sl]), lam[sl], rhs[sl]], 1)  # B,4,m,m
        coarse = torch.nn.functional.avg_pool2d(raw, f)
        x = coarse.permute(0, 2, 3, 1).float()
        x = ((x - net.in_shift.to(x.device)) / net.in_scale.to(x.device)).to(torch.bfloat16).contiguous()
        oc = net.forward(x).double() * net.out_scale                           # B, m_train, m_train
        out = torch.nn.functional.interpolate(oc[:, None], size=(m, m), mode="bilinear",
                                              align_corners=False)[:, 0]
    phi = torch.zeros_like(rhs)
    phi[:, 2:-2, 2:-2] = out[:, 1:, 1:]
    return phi


def smooth_guess(phi0, rhs, lam, dxx, dyy, sweeps):
    """`sweeps` FP64 RBGS sweeps of the fine-level operator applied to a batch
    of guesses (B, m+3, m+3): they damp the guess's high-frequency error,
    whose h^-2 weighted residual would otherwise dominate ||rhs - L phi0||."""
    from jsolver_amd import _lib, _solver
    from jsolver_amd._device import ptr, stream_ptr
    B, n, _ = rhs.shape
    g = grid_1D(n - 3)
    plan = _solver.get_plan(2, (g.mx, g.mx), [g.idel, g.idel], "fp64", B)
    plan.set_operator(lam, [dxx, dyy])
    lib = _lib.load()
    phi = phi0.contiguous().clone()
    r = rhs.contiguous()
    _lib.check(lib.jsv_level_upload(plan.h, _lib.JSV_FP64, 0, 0, ptr(phi), stream_ptr()))
    _lib.check(lib.jsv_level_upload(plan.h, _lib.JSV_FP64, 0, 1, ptr(r), stream_ptr()))
    _lib.check(lib.jsv_smooth(plan.h, _lib.JSV_FP64, 0, int(sweeps), stream_ptr()))
    _lib.check(lib.jsv_level_download(plan.h, _lib.JSV_FP64, 0, 0, ptr(phi), stream_ptr()))
    return phi


def evaluate(net, m, B, seed, m_train, dev, rtol=1e-10, gamma=1, log=print, max_cycles=50000,
             guess_sweeps=4, criterion="rhs"):
    """Cycles and wall time to rtol from a zero start and from the network's
    guess after `guess_sweeps` RBGS sweeps, with zero-start fallback per
    problem.  criterion="rhs": fall back when ||rhs - L phi0|| > ||rhs|| (the
    residual of the zero start).  criterion="cycle": run one cycle from the guess
    and one from zero and fall back when the guess's residual after it is the
    larger; that cycle pair is counted in the warm-start cycles.  Problems on
    which either solve does not converge are excluded from the cycle
    statistics and counted."""
    g = grid_1D(m)
    lam, dxx, dyy, rhs = make_problems(B, m, seed, dev)
    kw = dict(gamma=gamma, stop="rel_residual", rtol=rtol, max_iters_outer=max_cycles)
    solve_batched_2d(g, g, torch.zeros_like(rhs), rhs, lam, dxx, dyy, fixed=True, max_iters_outer=1, gamma=gamma)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, c0, r0 = solve_batched_2d(g, g, torch.zeros_like(rhs), rhs, lam, dxx, dyy, **kw)
    torch.cuda.synchronize()
    t_zero = time.perf_counter() - t0
    t0 = time.perf_counter()
    phi0 = guess_phi(net, lam, dxx, dyy, rhs, m_train)
    if guess_sweeps > 0:
        phi0 = smooth_guess(phi0, rhs, lam, dxx, dyy, guess_sweeps)
    extra = 0
    if criterion == "rhs":
        # ||rhs - L phi0|| vs ||rhs|| per problem
        _, _, rel0 = solve_batched_2d(g, g, phi0, rhs, lam, dxx, dyy, fixed=True, max_iters_outer=0,
                                      stop="rel_residual", rtol=rtol, precision="mixed")
        use = torch.tensor(rel0 <= 1.0, device=dev)
    else:
        # residuals after one cycle from the guess and from zero
        one = dict(fixed=True, max_iters_outer=1, gamma=gamma)
        pg, _, _ = solve_batched_2d(g, g, phi0, rhs, lam, dxx, dyy, **one)
        pz, _, _ = solve_batched_2d(g, g, torch.zeros_like(rhs), rhs, lam, dxx, dyy, **one)
        res0 = dict(fixed=True, max_iters_outer=0, stop="rel_residual", rtol=rtol, precision="mixed")
        _, _, rg = solve_batched_2d(g, g, pg, rhs, lam, dxx, dyy, **res0)
        _, _, rz = solve_batched_2d(g, g, pz, rhs, lam, dxx, dyy, **res0)
        rel0 = rg / np.maximum(rz, 1e-300)
        use = torch.tensor(rg <= rz, device=dev)
        extra = 2
    phi_start = torch.where(use[:, None, None], phi0, torch.zeros_like(phi0))
    _, c1, r1 = solve_batched_2d(g, g, phi_start, rhs, lam, dxx, dyy, **kw)
    c1 = np.asarray(c1) + extra
    torch.cuda.synchronize()
    t_warm = time.perf_counter() - t0
    good = np.isfinite(r0) & (r0 <= rtol) & np.isfinite(r1) & (r1 <= rtol)
    c0_all, c1_all = np.asarray(c0), np.asarray(c1)
    c0, c1 = c0_all[good], c1_all[good]
    res = {"m": m, "batch": B, "gamma": gamma, "rtol": rtol, "guess_sweeps": guess_sweeps, "criterion": criterion,
           "unconverged_excluded": int((~good).sum()),
           "cycles_zero_mean": float(np.mean(c0)), "cycles_warm_mean": float(np.mean(c1)),
           "cycle_reduction": float(1.0 - np.sum(c1) / np.sum(c0)),
           "per_problem_reduction_median": float(np.median(1.0 - c1 / c0)),
           "time_zero_s": t_zero, "time_warm_s": t_warm,
           "guess_criterion_ratio_median": float(np.median(rel0)),
           "fallbacks": int((~use).sum().item()),
           "max_final_rel_residual": float(max(np.max(r0[good]), np.max(r1[good]))),
           "per_problem": {"criterion_ratio": [float(v) for v in np.asarray(rel0)],
                           "used": [bool(v) for v in use.cpu().numpy()],
                           "converged": [bool(v) for v in good],
                           "cycles_zero": [int(v) for v in np.asarray(c0_all)],
                           "cycles_warm": [int(v) for v in np.asarray(c1_all)]}}
    log(f"[eval] {json.dumps({k: v for k, v in res.items() if k != 'per_problem'})}")
    return res


def _rel_err(a, b):
    """Per-problem ||a - b|| / ||b|| over the unknowns of (B, m+3, m+3) arrays."""
    sl = (slice(None), slice(2, -2), slice(2, -2))
    d = (a[sl] - b[sl]).flatten(1).norm(dim=1)
    return (d / b[sl].flatten(1).norm(dim=1).clamp_min(1e-300)).cpu().numpy()


def _criteria(g, phi0, rhs, lam, dxx, dyy, gamma, rtol):
    """The two fallback criteria of a batch of starts: ||rhs - L phi0|| / ||rhs||
    ("rhs") and the relative residual after one cycle from phi0 ("cycle")."""
    res0 = dict(fixed=True, max_iters_outer=0, stop="rel_residual", rtol=rtol, precision="mixed")
    _, _, r_rhs = solve_batched_2d(g, g, phi0, rhs, lam, dxx, dyy, **res0)
    p1, _, _ = solve_batched_2d(g, g, phi0, rhs, lam, dxx, dyy, fixed=True, max_iters_outer=1, gamma=gamma)
    _, _, r_cyc = solve_batched_2d(g, g, p1, rhs, lam, dxx, dyy, **res0)
    return np.asarray(r_rhs), np.asarray(r_cyc)


def diagnose(net, m, B, seed, m_train, dev, rtol=1e-10, gamma=1, log=print, max_cycles=50000,
             n_after=50, guess_sweeps=4):
    """What a warm start can gain on this problem class, and why the fallback
    rule decides as it does.  For the network's guess (raw and after
    `guess_sweeps` RBGS sweeps) and for reference guesses made from the
    converged solution phi* (phi* rounded to bf16: the best a bf16-output
    network can give; phi* plus a smooth relative error of 1e-3, 1e-2, 1e-1),
    per problem:
      * both fallback criteria (relative to the zero start's: > 1 means the
        guess falls back);
      * cycles to rtol when the guess is used regardless of the criteria;
      * the error ||phi_N - phi*|| / ||phi*|| after n_after cycles, against
        the zero start's (the error left in the slowly converging modes)."""
    g = grid_1D(m)
    lam, dxx, dyy, rhs = make_problems(B, m, seed, dev)
    kw = dict(gamma=gamma, stop="rel_residual", rtol=rtol, max_iters_outer=max_cycles)
    zero = torch.zeros_like(rhs)
    star, _, rs = solve_batched_2d(g, g, zero, rhs, lam, dxx, dyy, gamma=gamma, stop="rel_residual",
                                   rtol=min(rtol, 1e-13), max_iters_outer=max_cycles)
    _, c0, r0 = solve_batched_2d(g, g, zero, rhs, lam, dxx, dyy, **kw)
    pz, _, _ = solve_batched_2d(g, g, zero, rhs, lam, dxx, dyy, fixed=True, max_iters_outer=n_after, gamma=gamma)
    e_zero = _rel_err(pz, star)
    crit_zero = _criteria(g, zero, rhs, lam, dxx, dyy, gamma, rtol)
    guesses = {}
    phi_net = guess_phi(net, lam, dxx, dyy, rhs, m_train)
    guesses["net"] = phi_net
    if guess_sweeps > 0:
        guesses[f"net_sweeps{guess_sweeps}"] = smooth_guess(phi_net, rhs, lam, dxx, dyy, guess_sweeps)
    guesses["exact_bf16"] = star.to(torch.bfloat16).double()
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed) + 17)
    pert = grf_batch(B, m + 3, g.del_, torch.full((B,), 0.1, device=dev, dtype=torch.float64), gen, dev)
    sl = (slice(None), slice(2, -2), slice(2, -2))
    rms = star[sl].flatten(1).norm(dim=1) / math.sqrt((m - 1) ** 2)
    for eps in (1e-3, 1e-2, 1e-1):
        gs = star + eps * rms[:, None, None] * pert
        gz = torch.zeros_like(gs)
        gz[sl] = gs[sl]
        guesses[f"exact_smooth_{eps:g}"] = gz
    out = {"m": m, "batch": B, "gamma": gamma, "rtol": rtol, "n_after": n_after,
           "cycles_zero_mean": float(np.mean(c0)), "error_after_n_zero_median": float(np.median(e_zero)),
           "unconverged_zero": int(np.sum(~(np.asarray(r0) <= rtol))), "guesses": {}}
    for name, phi0 in guesses.items():
        crit = _criteria(g, phi0, rhs, lam, dxx, dyy, gamma, rtol)
        _, c1, r1 = solve_batched_2d(g, g, phi0, rhs, lam, dxx, dyy, **kw)
        pn, _, _ = solve_batched_2d(g, g, phi0, rhs, lam, dxx, dyy, fixed=True, max_iters_outer=n_after,
                                    gamma=gamma)
        ok = (np.asarray(r0) <= rtol) & (np.asarray(r1) <= rtol)
        c0k, c1k = np.asarray(c0)[ok], np.asarray(c1)[ok]
        ratio_rhs = crit[0] / np.maximum(crit_zero[0], 1e-300)
        ratio_cyc = crit[1] / np.maximum(crit_zero[1], 1e-300)
        e_g = _rel_err(pn, star)
        out["guesses"][name] = {
            "guess_rel_error_median": float(np.median(_rel_err(phi0, star))),
            "rhs_criterion_ratio_median": float(np.median(ratio_rhs)),
            "cycle_criterion_ratio_median": float(np.median(ratio_cyc)),
            "fallbacks_rhs": int(np.sum(ratio_rhs > 1.0)),
            "fallbacks_cycle": int(np.sum(ratio_cyc > 1.0)),
            "cycles_used_anyway_mean": float(np.mean(c1k)) if c1k.size else None,
            "cycle_reduction_used_anyway": float(1.0 - np.sum(c1k) / np.sum(c0k)) if c0k.size else None,
            "error_after_n_median": float(np.median(e_g)),
            "error_after_n_vs_zero_median": float(np.median(e_g / np.maximum(e_zero, 1e-300))),
            "unconverged": int(np.sum(~ok)),
        }
        log(f"[diag] {name}: {json.dumps(out['guesses'][name])}")
    return out


class CurveProbe:
    """Learning-curve measurements after every epoch: normalised MSE on
    held-out samples, and on a held-out batch of problems at the training
    resolution the cycles to rtol from the network's guess used regardless of
    the fallback rule, and its error after n_after cycles against the zero
    start's."""

    def __init__(self, m, B, seed, dev, rtol, gamma, n_after, label_rtol, max_cycles, heldout, log):
        self.g = grid_1D(m)
        self.m, self.dev, self.rtol, self.gamma, self.n_after = m, dev, rtol, gamma, n_after
        self.max_cycles, self.log = max_cycles, log
        self.prob = make_problems(B, m, seed, dev)
        lam, dxx, dyy, rhs = self.prob
        zero = torch.zeros_like(rhs)
        self.star, _, _ = solve_batched_2d(self.g, self.g, zero, rhs, lam, dxx, dyy, gamma=gamma,
                                           stop="rel_residual", rtol=1e-13, max_iters_outer=max_cycles)
        _, self.c0, _ = solve_batched_2d(self.g, self.g, zero, rhs, lam, dxx, dyy, gamma=gamma,
                                         stop="rel_residual", rtol=rtol, max_iters_outer=max_cycles)
        pz, _, _ = solve_batched_2d(self.g, self.g, zero, rhs, lam, dxx, dyy, fixed=True,
                                    max_iters_outer=n_after, gamma=gamma)
        self.e_zero = _rel_err(pz, self.star)
        self.Xh, self.Yh, _ = generate(heldout, m, 64, seed + 1, label_rtol, dev, lambda *_: None, gamma,
                                       max_cycles)
        self.points = []

    def __call__(self, net, epoch, train_loss):
        lam, dxx, dyy, rhs = self.prob
        mse, n = 0.0, 0
        for i in range(0, self.Xh.shape[0], 16):
            x = ((self.Xh[i:i + 16] - net.in_shift.to(self.dev)) / net.in_scale.to(self.dev)).to(torch.bfloat16)
            out = net.forward(x.contiguous())
            y = self.Yh[i:i + 16] / net.out_scale
            mse += float(((out - y) ** 2).sum())
            n += y.numel()
        phi0 = guess_phi(net, lam, dxx, dyy, rhs, self.m)
        _, c1, r1 = solve_batched_2d(self.g, self.g, phi0, rhs, lam, dxx, dyy, gamma=self.gamma,
                                     stop="rel_residual", rtol=self.rtol, max_iters_outer=self.max_cycles)
        pn, _, _ = solve_batched_2d(self.g, self.g, phi0, rhs, lam, dxx, dyy, fixed=True,
                                    max_iters_outer=self.n_after, gamma=self.gamma)
        e = _rel_err(pn, self.star)
        pt = {"epoch": epoch + 1, "train_loss": train_loss, "heldout_nmse": mse / n,
              "guess_rel_error_median": float(np.median(_rel_err(phi0, self.star))),
              "cycles_zero_mean": float(np.mean(self.c0)), "cycles_guess_used_anyway_mean": float(np.mean(c1)),
              "cycle_reduction_used_anyway": float(1.0 - np.sum(c1) / np.sum(self.c0)),
              "error_after_n_vs_zero_median": float(np.median(e / np.maximum(self.e_zero, 1e-300)))}
        self.points.append(pt)
        self.log(f"[curve] {json.dumps(pt)}")


def time_convs(m, batch, dev, reps=10, log=print):
    """Per-layer time and achieved TFLOP/s of the network's convolutions on a
    (batch, m, m) input, with HIP events on the launch stream: forward,
    input gradient (dgrad, the forward kernel on flipped weights) and weight
    gradient (wgrad: the MFMA pixel reduction, the chunk sum and the bias
    sum).  Flops: 2 * pixels * Cout * k^2 * Cin per pass."""
    from .warmstart import LAYERS
    net = UNet(seed=0, device=dev)
    res = {"e1": m, "e2": m // 2, "e3": m // 4, "b1": m // 8, "b2": m // 8, "d3": m // 4, "d2": m // 2,
           "d1": m, "ou": m}
    rows = []
    for (name, cin, cout, k, relu) in LAYERS:
        L = net.layers[name]
        H = W = res[name[:2]]
        x = (torch.randn((batch, H, W, cin), device=dev) * 0.5).to(torch.bfloat16).contiguous()
        dy = (torch.randn((batch, H, W, cout), device=dev) * 0.5).to(torch.bfloat16).contiguous()
        flops = 2.0 * batch * H * W * cout * k * k * cin

        def timed(fn):
            fn()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / 1e3 / reps

        t_f = timed(lambda: net._conv(L, x, batch, H, W))
        lib = _lib_unet()
        M = batch * H * W
        chunks = max(1, min(256, M // 4096))
        part = torch.empty(chunks * L.cout * (L.kp + 1), dtype=torch.float32, device=dev)

        def wgrad():
            _chk(lib.jsv_conv_wgrad(_p(x), _p(dy), _p(L.gw), _p(L.gb), _p(part), chunks, batch, H, W, L.cin,
                                    L.cout, L.k, L.kp, _sp()))
        t_w = timed(wgrad)
        wd, kpd = L.dgrad_weight16()

        def dgrad():
            dx = torch.empty((batch, H, W, cin), dtype=torch.bfloat16, device=dev)
            _chk(lib.jsv_conv_fwd(_p(dy), _p(wd), None, _p(dx), None, batch, H, W, cout, cin, k, kpd, 0, _sp()))
        t_d = timed(dgrad)
        row = {"layer": name, "hw": H, "cin": cin, "cout": cout, "k": k, "gflop": flops / 1e9,
               "fwd_us": t_f * 1e6, "fwd_tflops": flops / t_f / 1e12,
               "dgrad_us": t_d * 1e6, "dgrad_tflops": flops / t_d / 1e12,
               "wgrad_us": t_w * 1e6, "wgrad_tflops": flops / t_w / 1e12}
        rows.append(row)
        log(f"[convs] {json.dumps(row)}")
    tot = {kind: sum(r[f"{kind}_us"] for r in rows) for kind in ("fwd", "dgrad", "wgrad")}
    gf = sum(r["gflop"] for r in rows)
    summary = {kind: {"us": tot[kind], "tflops": gf * 1e9 / (tot[kind] * 1e-6) / 1e12} for kind in tot}
    return {"m": m, "batch": batch, "layers": rows, "total": summary}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--m", type=int, default=256)
    ap.add_argument("--gen-batch", type=int, default=64)
    ap.add_argument("--label-rtol", type=float, default=1e-8)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--eval-m", type=int, default=512)
    ap.add_argument("--eval-batch", type=int, default=64)
    ap.add_argument("--gamma", type=int, default=1)
    ap.add_argument("--max-cycles", type=int, default=50000)
    ap.add_argument("--guess-sweeps", type=int, default=4)
    ap.add_argument("--curve", type=int, default=0,
                    help="problems of the per-epoch learning-curve batch (0: no curve)")
    ap.add_argument("--curve-heldout", type=int, default=256, help="held-out samples for the curve's MSE")
    ap.add_argument("--diagnose", type=int, default=0,
                    help="problems at --eval-m for the gain diagnosis (0: none)")
    ap.add_argument("--time-convs", action="store_true",
                    help="time every convolution (forward, dgrad, wgrad) on a --batch x --m input and exit")
    ap.add_argument("--out", default=None)
    ap.add_argument("--eval-only", default=None, help="state file of a trained network: evaluate it only")
    a = ap.parse_args(argv)
    dev = torch.device("cuda", 0)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    def evaluate_all(net):
        ev = {}
        for criterion in ("rhs", "cycle"):
            for sw in sorted({0, a.guess_sweeps}):
                ev[f"{criterion}_sweeps{sw}"] = evaluate(net, a.eval_m, a.eval_batch, 777_000, a.m, dev,
                                                         gamma=a.gamma, log=log, max_cycles=a.max_cycles,
                                                         guess_sweeps=sw, criterion=criterion)
        return ev

    if a.time_convs:
        print(json.dumps(time_convs(a.m, a.batch, dev, log=log)), flush=True)
        return
    if a.eval_only:
        net = UNet(seed=a.seed, device=dev)
        net.load_state_dict(torch.load(a.eval_only, weights_only=True))
        out = {"eval": evaluate_all(net)}
        if a.diagnose:
            out["diagnosis"] = diagnose(net, a.eval_m, a.diagnose, 999_000, a.m, dev, gamma=a.gamma, log=log,
                                        max_cycles=a.max_cycles, guess_sweeps=a.guess_sweeps)
        print(json.dumps(out), flush=True)
        return
    t0 = time.perf_counter()
    X, Y, counts = generate(a.samples, a.m, a.gen_batch, a.seed, a.label_rtol, dev, log, a.gamma, a.max_cycles)
    t_data = time.perf_counter() - t0
    X_kept = X.shape[0]
    net = UNet(seed=a.seed, device=dev)
    fit_normalisation(net, X, Y)
    probe = None
    if a.curve:
        probe = CurveProbe(a.m, a.curve, 555_000, dev, 1e-10, a.gamma, 50, a.label_rtol, a.max_cycles,
                           a.curve_heldout, log)
    t0 = time.perf_counter()
    hist = train(net, X, Y, a.epochs, a.batch, a.lr, log,
                 on_epoch=(lambda ep, loss: probe(net, ep, loss)) if probe else None)
    t_train = time.perf_counter() - t0
    if a.out:
        torch.save(net.state_dict(), a.out)
    del X, Y
    torch.cuda.empty_cache()
    out = {"samples": a.samples, "kept": int(X_kept), "m": a.m, "gamma": a.gamma, "epochs": a.epochs,
           "data_s": t_data, "train_s": t_train, "label_cycles_mean": float(np.mean(counts)),
           "loss_per_epoch": hist}
    if probe is not None:
        out["curve"] = probe.points
    if a.eval_batch > 0:
        out["eval"] = evaluate_all(net)
    if a.diagnose:
        out["diagnosis"] = diagnose(net, a.eval_m, a.diagnose, 999_000, a.m, dev, gamma=a.gamma, log=log,
                                    max_cycles=a.max_cycles, guess_sweeps=a.guess_sweeps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
