"""Multiscale FWI driver: the control flow of src/main.c:275-806 on the batched device path.

Per frequency stage: C-PML table and Ricker wavelet of the stage frequency, observed data modelled from the
true model (main.c:323-366), then iterations of {gradient + FP64 misfit (main.c:396-538); steepest descent on
the first iteration (main.c:539-568); L-BFGS direction + backtracking line search afterwards (main.c:569-800)}.
A stage ends after `iterations` or when the line search fails more than 5 times in the stage; the next
stage starts from the current model (main.c:383-387, 803-805).

Run as a module for a small end-to-end inversion:
    python -m fwi_amd.driver --workload plain200 --iterations 4
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import time

import numpy as np
import torch

from . import models as M
from .optim import LBFGS, SD_STEP, LineSearch, clamp_model
from .survey import Plan, cpml_table, pre_variables


@dataclasses.dataclass
class StageLog:
    f: float
    misfits: list
    trials: list
    ended_by_line_search: bool
    seconds: float


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def run(w: M.Workload, stages=(None,), iterations: int = 100, memory: int = 5, it_lbfgs: int = 1,
        grad_mode: int = 1, verbose: bool = True, bounds=None, **plan_kw):
    """Invert `w` from its initial model; stages are Ricker frequencies (None = the workload's).
    bounds = (vmin, vmax) in m/s clamps every updated model (fwi_model_clamp); None keeps the reference's
    update, whose only bound is v_hor >= v_nmo.
    Returns (final (v_hor, v_po, v_nmo) device tensors, list of StageLog)."""
    bound = (lambda: clamp_model(model, *bounds)) if bounds else (lambda: None)
    true_sq = pre_variables(*(_dev(x) for x in w.true_model))
    model = list(pre_variables(*(_dev(x) for x in w.init_model)))
    logs = []
    for f in stages:
        f = w.f if f is None else f
        t0 = time.perf_counter()
        p = dataclasses.replace(w.cpml_params, f=f)
        src = M.ricker(w.nt, f, w.dt)
        plan = Plan(w.nx, w.nz, w.nt, w.dh, w.dt, w.cpml, w.shots, src, cpml_table(p), grad_mode=grad_mode,
                    **plan_kw)
        obs = plan.forward(true_sq)
        if verbose:
            torch.cuda.synchronize()
            print(json.dumps({"stage_f": f, "observed_data_s": round(time.perf_counter() - t0, 3)}), flush=True)
        opt = LBFGS(w.nx, w.nz, memory)
        ls = LineSearch()
        misfits, trials = [], []
        ended = False
        for n in range(iterations):
            phi, grad, _ = plan.gradient(model, obs)
            misfits.append(phi)
            if verbose:
                print(json.dumps({"stage_f": f, "iteration": n, "misfit": phi}), flush=True)
            if n < it_lbfgs:
                opt.sd_step(model, grad, SD_STEP)
                bound()
                continue
            opt.direction(model, grad)

            def evaluate(alpha):
                opt.trial(model, alpha)
                bound()
                m = plan.misfit(model, obs)
                if verbose:
                    print(json.dumps({"stage_f": f, "iteration": n, "trial_alpha": alpha, "trial_misfit": m}),
                          flush=True)
                return m

            acc, nt_ = ls.run(phi, evaluate, lambda: opt.restore(model))
            trials.append(nt_)
            if acc is None:
                ended = True
                break
        torch.cuda.synchronize()
        logs.append(StageLog(f, misfits, trials, ended, time.perf_counter() - t0))
        plan.close()
        opt.close()
    return model, logs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="plain200")
    ap.add_argument("--iterations", type=int, default=5)
    ap.add_argument("--stages", default="", help="comma-separated Ricker frequencies (Hz)")
    ap.add_argument("--shots", type=int, default=0)
    ap.add_argument("--nt", type=int, default=0)
    ap.add_argument("--out", default="")
    ap.add_argument("--bounds", default="", help="vmin,vmax in m/s: clamp every updated model")
    args = ap.parse_args()
    w = M.workload(args.workload, nshots=args.shots or None, nt=args.nt or None)
    stages = tuple(float(x) for x in args.stages.split(",")) if args.stages else (None,)
    bounds = tuple(float(x) for x in args.bounds.split(",")) if args.bounds else None
    model, logs = run(w, stages=stages, iterations=args.iterations, bounds=bounds)
    for lg in logs:
        print(json.dumps(dataclasses.asdict(lg)))
    if args.out:
        from . import io

        io.write_result(args.out, *(t.cpu().numpy() for t in model))


if __name__ == "__main__":
    main()
