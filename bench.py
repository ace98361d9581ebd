#!/usr/bin/env python3
"""Benchmark: time-to-solution at relative residual 1e-10 on the 3D 256^3
anisotropic diffusion-absorption problem, FP32-inner/FP64-refined cycles vs
plain FP64 cycles, on one MI355X; plus the HBM bandwidth of the fine-level
RBGS sweep against the 8 TB/s peak.

A "step" is one complete solve (mixed precision, V-cycles, dynamic stop at
||rhs - L phi||_2 / ||rhs||_2 <= 1e-10) from phi = 0, with every input already
resident in HBM.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--m 256] [--cycle V|F] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

METRIC = "time-to-solution at rel. residual 1e-10 (3D 256^3), mixed vs FP64; HBM GB/s"
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=256)
    ap.add_argument("--cycle", choices=["V", "F"], default="V")
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--inner-cycles", type=int, default=0,
                    help="FP32 cycles per refinement step (0: predicted from the contraction)")
    ap.add_argument("--sweeps", type=int, default=20, help="sweeps per timed block of the roofline timing")
    ap.add_argument("--sweep-blocks", type=int, default=7,
                    help="timed blocks per level; the median block gives the sweep time")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-cycles", type=int, default=8,
                    help="minimum V-cycles of the C oracle timed for the CPU baseline (more are run "
                         "until about 10 s of CPU work)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only set up the operator and time the per-level sweeps (for profiling)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def roofline(a, grids, m):
    """Per-level RBGS sweep time with HIP events on the stream the kernels run
    on (median of several blocks of back-to-back sweeps; consecutive sweeps
    of a level march in opposite z directions, as in a cycle's post- and
    pre-smoothing); the fine level's FP32 sweep is the roofline kernel.
    Algorithmic bytes per sweep and unknown: 7 words (phi read + write, rhs,
    3 face arrays, the diagonal)."""
    import torch

    from jsolver_amd import _lib, _solver
    from jsolver_amd._device import stream_ptr
    lib = _lib.load()
    plan = _solver.get_plan(3, (m, m, m), [g.idel for g in grids], "mixed")
    per_level, roof = [], None
    for prec_name, prec, w in (("fp32", _lib.JSV_MIXED, 4), ("fp64", _lib.JSV_FP64, 8)):
        for level in range(plan.levels - 1):
            mm = m >> level
            n_unk = (mm - 1) ** 3
            sweeps = a.sweeps if level < 3 else 50
            _lib.check(lib.jsv_smooth(plan.h, prec, level, 2, stream_ptr()))
            # median over blocks of `sweeps` back-to-back sweeps
            blocks = []
            for _ in range(max(1, a.sweep_blocks)):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                _lib.check(lib.jsv_smooth(plan.h, prec, level, sweeps, stream_ptr()))
                e1.record()
                torch.cuda.synchronize()
                blocks.append(e0.elapsed_time(e1) / 1e3 / sweeps)
            t_sweep = sorted(blocks)[len(blocks) // 2]
            gbs = 7 * w * n_unk / t_sweep / 1e9
            per_level.append({"prec": prec_name, "level": level, "m": mm,
                              "us_per_sweep": round(t_sweep * 1e6, 2), "GBps": round(gbs, 1)})
            if level == 0 and prec_name == "fp32":
                roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": round(gbs / PEAK_HBM_GBS, 4), "traffic": _traffic(),
                        "kernel": "k_rbsweep3p<float>: one fused out-of-place RBGS sweep (red + black "
                                  "halves, one launch; two column pairs per thread in 10-row tiles of "
                                  "768 threads), fine level",
                        "algorithmic_bytes_per_sweep": 7 * w * n_unk,
                        "us_per_sweep": round(t_sweep * 1e6, 2)}
    log("per-level sweeps: " + ", ".join(f"{p['prec']}/L{p['level']}:{p['GBps']}GB/s" for p in per_level))
    if roof is not None:
        roof["stream_GBps"] = _stream_bandwidth()
    return roof, per_level


def _stream_bandwidth():
    """Bandwidth this GPU sustains on the library's own streaming kernel
    (jsv_stream_bench: 16-byte non-temporal loads and stores over 256 MiB
    arrays): a copy (1 read : 1 write) and the RBGS sweep's ratio (6 reads :
    1 write).  Context for the roofline fraction, whose peak is the published
    8 TB/s."""
    import torch

    from jsolver_amd import _lib
    from jsolver_amd._device import ptr, stream_ptr
    lib = _lib.load()
    n = 1 << 26
    out = {}
    for k in (1, 6):
        src = torch.ones(k * n, dtype=torch.float32, device="cuda")
        dst = torch.empty(n, dtype=torch.float32, device="cuda")
        _lib.check(lib.jsv_stream_bench(ptr(src), k, n, ptr(dst), 0, stream_ptr()))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            _lib.check(lib.jsv_stream_bench(ptr(src), k, n, ptr(dst), 0, stream_ptr()))
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / 1e3 / 10
        out[f"read{k}_write1"] = round(4 * n * (k + 1) / t / 1e9, 1)
        del src, dst
    return out


def _traffic():
    """HBM bytes per launch of the fine-level FP32 sweep kernel measured with
    rocprofv3 PMC (FETCH_SIZE + WRITE_SIZE, each scaled by its reading on a
    copy kernel of known size), recorded in
    profiles/traffic_rbsweep3p_fp32_256.json; None when absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "traffic_rbsweep3p_fp32_256.json")
    try:
        with open(path) as f:
            return json.load(f)["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def _host_cpus():
    """Threads for the CPU baseline: every CPU this process may run on
    (sched affinity), capped by OMP_NUM_THREADS where the host sets it; plus
    the host's CPU count and model for the record."""
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    except ValueError:
        cap = 0
    threads = max(1, min(share, cap) if cap > 0 else share)
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return threads, {"cpu_count": os.cpu_count(), "affinity": share, "omp_num_threads": cap or None,
                     "cpu_model": model}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    from jsolver_amd import fields
    from jsolver_amd.solver_3d import solve_3d

    m = a.m
    t0 = time.perf_counter()
    grids, rhs, lam, dxx, dyy, dzz = fields.anisotropic_3d(m)
    log(f"inputs generated in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda", 0)
    rhs_d, lam_d, dx_d, dy_d, dz_d = (torch.from_numpy(v).to(dev) for v in (rhs, lam, dxx, dyy, dzz))
    phi0 = torch.zeros_like(rhs_d)
    unknowns = (m - 1) ** 3
    kw = dict(stop="rel_residual", rtol=a.rtol, fcycle=(a.cycle == "F"))

    def run(precision, check_every=1):
        return solve_3d(*grids, phi0, rhs_d, lam_d, dx_d, dy_d, dz_d, precision=precision,
                        inner_cycles=a.inner_cycles, check_every=check_every, return_info=True, **kw)

    if a.roofline_only:
        solve_3d(*grids, phi0, rhs_d, lam_d, dx_d, dy_d, dz_d, precision="mixed", fixed=True,
                 max_iters_outer=1)
        roof, per_level = roofline(a, grids, m)
        print(json.dumps({"metric": METRIC, "roofline": roof, "levels": per_level}), flush=True)
        return

    # ---- FP64 reference solve -------------------------------------------
    for _ in range(max(1, a.warmup)):
        run("fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phi64, n64, info64 = run("fp64")
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    log(f"fp64: {n64} cycles, {t64 * 1e3:.1f} ms, rel res {info64['rel_residual']:.3e}")
    # FP64 with the residual checks spaced like the mixed solve's refinement steps
    run("fp64", check_every=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, n64a, info64a = run("fp64", check_every=0)
    torch.cuda.synchronize()
    t64a = time.perf_counter() - t0
    log(f"fp64 (adaptive checks): {n64a} cycles, {t64a * 1e3:.1f} ms")
    # FP64 with the per-cycle test read by the host after every cycle (the
    # device-resident test switched off), for comparison
    from jsolver_amd import _lib, _solver
    plan64 = _solver.get_plan(3, [m, m, m], [g.idel for g in grids], "fp64")
    lib = _lib.load()
    _lib.check(lib.jsv_plan_set_tuning(plan64.h, _lib.JSV_TUNE_DEVICE_STOP, 0))
    try:
        run("fp64")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, n64h, info64h = run("fp64")
        torch.cuda.synchronize()
        t64h = time.perf_counter() - t0
    finally:
        _lib.check(lib.jsv_plan_set_tuning(plan64.h, _lib.JSV_TUNE_DEVICE_STOP, 1))
    log(f"fp64 (host reads every test): {n64h} cycles, {t64h * 1e3:.1f} ms")

    # ---- mixed precision: the timed steps ----------------------------------
    for _ in range(a.warmup):
        run("mixed")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        phim, nm, infom = run("mixed")
    torch.cuda.synchronize()
    tm = (time.perf_counter() - t0) / a.steps
    log(f"mixed: {nm} fp32 cycles in {infom['outer_steps']} refinements, {tm * 1e3:.1f} ms, "
        f"rel res {infom['rel_residual']:.3e}")
    err_vs_fp64 = float(torch.linalg.norm(phim - phi64) / torch.linalg.norm(phi64))

    roof, per_level = roofline(a, grids, m)

    # ---- CPU baseline: the oracle's C restatement (FP64, OpenMP), bounded sample
    cpu = None
    if not a.no_cpu:
        try:
            sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
            from oracle import gmg, gmg_c
            threads, host = _host_cpus()
            t0 = time.perf_counter()
            op = gmg.build_operator(3, (m, m, m), tuple(g.idel for g in grids), lam, [dxx, dyy, dzz])
            t_setup = time.perf_counter() - t0
            cfg = gmg.Config(fixed=True)
            phi = np.zeros(rhs.shape)
            done, t_run = 0, 0.0
            while done < a.cpu_cycles or (t_run < 10.0 and done < 200):
                t0 = time.perf_counter()
                phi = gmg_c.cycles3d(op, cfg, phi, rhs, 1, threads=threads)
                t_run += time.perf_counter() - t0
                done += 1
            t_cyc = t_run / done
            cpu = {"value": round(t_setup + t_cyc * n64, 2), "unit": "s", "cores": threads, "kind": "port",
                   "extrapolated": True, "host": host,
                   "sample": f"the oracle's C restatement (FP64, OpenMP, bit-identical to the NumPy oracle): "
                             f"NumPy operator setup ({t_setup:.1f}s) + {done} V-cycles at {m}^3 timed "
                             f"({t_cyc:.3f}s/cycle on {threads} threads: every CPU this process may use), "
                             f"extrapolated to the FP64 GPU cycle count {n64}"}
            log(f"cpu oracle (C, {threads} threads): {t_cyc:.3f}s/cycle")
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "s", "cores": 1, "kind": "port", "sample": f"failed: {e}"}

    out = {
        "metric": METRIC,
        "value": round(tm, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(tm * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-refined",
        "data": "synthetic (seeded log-normal GRF coefficients, corr. length 0.05)",
        "config": {"workload": f"3D anisotropic diffusion-absorption {m}^3, {a.cycle}-cycles, "
                               f"dynamic to rel. residual {a.rtol:g}",
                   "unknowns": unknowns, "cycle": a.cycle, "inner_cycles": a.inner_cycles},
        "fp64": {"time_to_solution_s": round(t64, 5), "cycles": n64,
                 "rel_residual": info64["rel_residual"],
                 "residual_checked": "every cycle, on the device (read back once per block of cycles)",
                 "time_vs_adaptive_checks": round(t64 / t64a, 4)},
        "fp64_host_checks": {"time_to_solution_s": round(t64h, 5), "cycles": n64h,
                             "rel_residual": info64h["rel_residual"],
                             "residual_checked": "every cycle, read by the host after each cycle"},
        "fp64_adaptive_checks": {"time_to_solution_s": round(t64a, 5), "cycles": n64a,
                                 "rel_residual": info64a["rel_residual"],
                                 "speedup_mixed_vs_this": round(t64a / tm, 3)},
        "mixed": {"time_to_solution_s": round(tm, 5), "fp32_cycles": nm,
                  "refinements": infom["outer_steps"], "rel_residual": infom["rel_residual"],
                  "rel_error_vs_fp64": err_vs_fp64},
        "speedup_mixed_vs_fp64": round(t64 / tm, 3),
        "cycles_per_s_fp64": round(n64 / t64, 2),
        "unknowns_per_s_fp64": round(unknowns * n64 / t64, 1),
        "roofline": roof,
        "levels": per_level,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
