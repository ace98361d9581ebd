"""The learned warm start on the C5 workload, reduced in sample counts only:
the network is trained by the project's own kernels on 256 problems at 256^2
solved by the project's FP64 solver (the C5 coefficient generator:
correlation length ~ U[0.02, 0.1], anisotropy r ~ U[1, 100], lambda ~
U[0, 10]), then 8 problems at 512^2 are solved to relative residual 1e-10
from a zero start and from the network's guess, under both fallback
criteria.  Checked: every solve converges; a problem falls back exactly when
its criterion says the guess is worse; a problem that falls back takes the
zero start's cycle count (plus the two criterion cycles of the "cycle"
rule).  The cycles zero -> warm are printed: what the trained network gains
over the method without it (the full-size record, 20,000 samples and 64
problems, is committed under profiles/)."""
import json

import numpy as np
import pytest
import torch

from jsolver_amd import warmstart_train as wt
from jsolver_amd.warmstart import UNet

pytestmark = pytest.mark.gpu


def test_reduced_c5_train_and_evaluate():
    dev = torch.device("cuda", 0)
    X, Y, counts = wt.generate(256, 256, 64, 4242, 1e-8, dev, log=lambda *_: None)
    assert X.shape[1:] == (256, 256, 4) and X.shape[0] >= 250
    net = UNet(seed=1, device=dev)
    wt.fit_normalisation(net, X, Y)
    hist = wt.train(net, X, Y, 6, 16, 2e-3, log=lambda *_: None)
    assert hist[-1] < 0.7 * hist[0], hist
    del X, Y
    torch.cuda.empty_cache()
    summary = {}
    for criterion in ("rhs", "cycle"):
        res = wt.evaluate(net, 512, 8, 777_000, 256, dev, criterion=criterion, guess_sweeps=4,
                          log=lambda *_: None)
        pp = res["per_problem"]
        assert all(pp["converged"]), pp
        assert res["max_final_rel_residual"] <= 1e-10
        extra = 2 if criterion == "cycle" else 0
        for ratio, used, c0, c1 in zip(pp["criterion_ratio"], pp["used"], pp["cycles_zero"], pp["cycles_warm"]):
            assert used == (ratio <= 1.0), (ratio, used)
            if not used:
                assert c1 == c0 + extra, (c0, c1)
        summary[criterion] = {"cycles_zero_mean": res["cycles_zero_mean"],
                              "cycles_warm_mean": res["cycles_warm_mean"],
                              "cycle_reduction": res["cycle_reduction"], "fallbacks": res["fallbacks"],
                              "criterion_ratio_median": res["guess_criterion_ratio_median"]}
    print("C5 reduced:", json.dumps({"loss_per_epoch": hist, "eval_512": summary}))
