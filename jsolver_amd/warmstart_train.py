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
from .warmstart import UNet, features


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


def train(net, X, Y, epochs, batch, lr, log=print, seed=0):
    """Adam on MSE of the normalised output; all kernels hand-written."""
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
    c0, c1 = np.asarray(c0)[good], np.asarray(c1)[good]
    res = {"m": m, "batch": B, "gamma": gamma, "rtol": rtol, "guess_sweeps": guess_sweeps, "criterion": criterion,
           "unconverged_excluded": int((~good).sum()),
           "cycles_zero_mean": float(np.mean(c0)), "cycles_warm_mean": float(np.mean(c1)),
           "cycle_reduction": float(1.0 - np.sum(c1) / np.sum(c0)),
           "per_problem_reduction_median": float(np.median(1.0 - c1 / c0)),
           "time_zero_s": t_zero, "time_warm_s": t_warm,
           "guess_criterion_ratio_median": float(np.median(rel0)),
           "fallbacks": int((~use).sum().item()),
           "max_final_rel_residual": float(max(np.max(r0[good]), np.max(r1[good])))}
    log(f"[eval] {json.dumps(res)}")
    return res


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
    ap.add_argument("--out", default=None)
    ap.add_argument("--eval-only", default=None, help="state file of a trained network: evaluate it only")
    a = ap.parse_args(argv)
    dev = torch.device("cuda", 0)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)
    if a.eval_only:
        net = UNet(seed=a.seed, device=dev)
        net.load_state_dict(torch.load(a.eval_only, weights_only=True))
        out = {}
        for criterion in ("rhs", "cycle"):
            for sw in sorted({0, a.guess_sweeps}):
                out[f"{criterion}_sweeps{sw}"] = evaluate(net, a.eval_m, a.eval_batch, 777_000, a.m, dev, gamma=a.gamma,
                                                      log=log, max_cycles=a.max_cycles, guess_sweeps=sw,
                                                      criterion=criterion)
        print(json.dumps(out), flush=True)
        return
    t0 = time.perf_counter()
    X, Y, counts = generate(a.samples, a.m, a.gen_batch, a.seed, a.label_rtol, dev, log, a.gamma, a.max_cycles)
    t_data = time.perf_counter() - t0
    X_kept = X.shape[0]
    net = UNet(seed=a.seed, device=dev)
    fit_normalisation(net, X, Y)
    t0 = time.perf_counter()
    hist = train(net, X, Y, a.epochs, a.batch, a.lr, log)
    t_train = time.perf_counter() - t0
    if a.out:
        torch.save(net.state_dict(), a.out)
    del X, Y
    torch.cuda.empty_cache()
    ev = {}
    for sw in sorted({0, a.guess_sweeps}):
        ev[f"sweeps{sw}"] = evaluate(net, a.eval_m, a.eval_batch, 777_000, a.m, dev, gamma=a.gamma, log=log,
                                     max_cycles=a.max_cycles, guess_sweeps=sw)
    ev_train_res = evaluate(net, a.m, a.eval_batch, 888_000, a.m, dev, gamma=a.gamma, log=log,
                            max_cycles=a.max_cycles, guess_sweeps=a.guess_sweeps)
    print(json.dumps({"samples": a.samples, "kept": int(X_kept), "m": a.m, "gamma": a.gamma,
                      "data_s": t_data, "train_s": t_train,
                      "label_cycles_mean": float(np.mean(counts)), "loss_per_epoch": hist,
                      "eval": ev, "eval_train_resolution": ev_train_res}), flush=True)


if __name__ == "__main__":
    main()
