"""GPU: the multiscale drivers (src/main.c:275-806) against the CPU composition of their pinned pieces.

Both drivers -- the Python one (fwi_amd/driver.py) and the C one (fwi_amd/driver/fwi_main.c, main.c's control flow
on the C ABI) -- run two frequency stages of three iterations each: per stage the C-PML table and wavelet of its
frequency, observed data from the true model, a fresh L-BFGS history and line search; the first iteration a
steepest-descent step, the others an L-BFGS direction and the backtracking line search.  The CPU composition runs
the same schedule from the oracle's gradient and forward (bit-exact with the device's, tested elsewhere), the FP64
L-BFGS oracle and the line-search state machine.  The misfit sequences (every gradient's and, for the C driver, every
line-search trial's), the trial counts and the stage ends agree -- gradient misfits to 1e-9, trial misfits to 1e-7,
the final models to 1e-6 (the device's FP64 reductions add in another order, which can move a model value by an
ulp).  On these small cases the
first L-BFGS step of each stage, built from one curvature pair, fails six trials and ends the stage, as the
reference's rules have it (main.c:773-784), so each stage runs two gradients.
"""
import dataclasses
import math
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle_lib as O
from helpers import oracle_gathers, oracle_gradient, say, small_workload
from fwi_amd import models as M

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def cpu_inversion(w, stages, iterations):
    """The drivers' schedule on the CPU: returns (misfits of every gradient, trial misfits of every line search,
    final squared-velocity model, trials per line search, stages ended by the line search)."""
    from lbfgs_oracle import LbfgsOracle

    from fwi_amd.optim import SD_STEP, LineSearch

    true_sq = O.pre_variables(*w.true_model)
    model = list(O.pre_variables(*w.init_model))
    misfits, trials_all, counts, ended = [], [], [], []
    for f in stages:
        p = dataclasses.replace(w.cpml_params, f=f)
        tab = O.cpml_table(p)
        src = M.ricker(w.nt, f, w.dt)
        obs = oracle_gathers(w, "f32i", true_sq, tab, src)
        orc = LbfgsOracle(5)
        ls = LineSearch()
        ended.append(False)
        for n in range(iterations):
            phi, G, _ = oracle_gradient(w, "f32i", True, tuple(model), tab, src, obs)
            misfits.append(phi)
            if n < 1:
                model = orc.sd_step(model, [G[k] for k in range(3)], SD_STEP)
                continue
            orc.direction(model, [G[k] for k in range(3)])
            state = {}

            def evaluate(alpha):
                m = orc.trial(alpha)
                state["m"] = m
                syn = oracle_gathers(w, "f32i", tuple(m), tab, src)
                t = sum(0.5 * float(np.sum((s.astype(np.float64) - o.astype(np.float64)) ** 2))
                        for s, o in zip(syn, obs))
                trials_all.append(t)
                return t

            acc, ntr = ls.run(phi, evaluate, lambda: state.update(m=[x.copy() for x in orc.back_m]))
            counts.append(ntr)
            model = state["m"]
            if acc is None:
                ended[-1] = True
                break
    return misfits, trials_all, model, counts, ended


def _close(a, b, rel):
    return abs(a - b) <= rel * max(abs(a), abs(b))


def test_python_driver_matches_cpu_composition():
    import torch

    from fwi_amd import driver

    w = small_workload(nshots=2, nt=240, f=14.0)
    stages = (14.0, 17.0)
    say("python driver: 2 stages x 3 iterations vs the CPU composition")
    model, logs = driver.run(w, stages=stages, iterations=3, verbose=False)
    torch.cuda.synchronize()
    got = [phi for lg in logs for phi in lg.misfits]
    ref, _, ref_model, ref_counts, ref_ended = cpu_inversion(w, stages, 3)
    print("misfits device", got, [lg.trials for lg in logs], [lg.ended_by_line_search for lg in logs])
    print("misfits cpu   ", ref, ref_counts, ref_ended)
    assert len(got) == len(ref) >= 4
    assert all(_close(a, b, 1e-9) for a, b in zip(got, ref)), (got, ref)
    assert [t for lg in logs for t in lg.trials] == ref_counts
    assert [lg.ended_by_line_search for lg in logs] == ref_ended
    assert got[1] < got[0]  # the steepest-descent step descends
    for k in range(3):
        a = model[k].cpu().numpy().astype(np.float64)
        b = ref_model[k].astype(np.float64)
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max(), k


def test_c_driver_matches_cpu_composition():
    """fwi_main on model files: the reference's geometry (nshots+1 sources on row sz-1 = 4 at CPML + i*space,
    receivers over the live columns), its wavelet and C-PML parameters, stages f, f+3."""
    exe = os.path.join(ROOT, "fwi_amd", "fwi_main")
    assert os.path.exists(exe), "fwi_main not built"
    nx, nz, cpml, dh, dt, tend, f0, nshots = 100, 60, 12, 10.0, 1e-3, 0.2, 12.0, 2
    nt = int(math.ceil(float(np.float32(tend) / np.float32(dt))))
    space = int(math.floor(float(np.float32(nx - 2 * cpml - 1) / np.float32(nshots))))
    shots = [M.Shot(x0=0, nx=nx, isx=cpml + i * space - 1, isz=4, irz=4, ir0=cpml, nrec=nx - 2 * cpml)
             for i in range(nshots + 1)]
    vt = M.gradient_model(nx, nz, seed=3, cpml=cpml, sigma=6.0)
    vi = M.smooth_model(*vt, width=8.0)
    w = M.Workload("c_driver", nx, nz, dh, dt, nt, f0, cpml, vt, vi, shots)
    say(f"C driver: {nshots + 1} shots, nt={nt}, stages {f0}, {f0 + 3} Hz x 3 iterations vs the CPU composition")
    with tempfile.TemporaryDirectory() as d:
        for name, arr in zip(("vp_ori", "vh_ori", "vn_ori", "vp_ini", "vh_ini", "vn_ini"), (*vt, *vi)):
            np.ascontiguousarray(arr, np.float32).tofile(os.path.join(d, name + ".bin"))
        cmd = [exe, "--dir", d, "--nx", str(nx), "--nz", str(nz), "--dh", str(dh), "--dt", str(dt), "--tend", str(tend),
               "--cpml", str(cpml), "--nshots", str(nshots), "--f", str(f0), "--stages", "2", "--iterations", "3"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        print(r.stdout[-3000:])
        assert r.returncode == 0, r.stderr
        got = [float(l.split(":")[1]) for l in r.stdout.splitlines() if l.startswith("Error total")]
        got_trials = [float(l.split(":")[1]) for l in r.stdout.splitlines() if l.startswith("trial misfit")]
        out = [np.fromfile(os.path.join(d, n + ".bin"), np.float32).reshape(nz, nx) for n in ("v_hor", "v_po", "v_nmo")]
    ref, ref_trials, ref_model, _, _ = cpu_inversion(w, (f0, f0 + 3.0), 3)
    print("misfits C driver", got, got_trials)
    print("misfits cpu     ", ref, ref_trials)
    assert len(got) == len(ref) >= 4
    assert all(_close(a, b, 1e-9) for a, b in zip(got, ref)), (got, ref)
    # a trial model is m - alpha r rounded to float: the FP64 direction's last bits (device reductions against numpy)
    # can move a cell by an ulp, so the trial misfits agree to 1e-7 (measured 2e-9), the gradient misfits to 1e-9
    assert len(got_trials) == len(ref_trials) and all(_close(a, b, 1e-7) for a, b in zip(got_trials, ref_trials))
    for k in range(3):
        b = ref_model[k].astype(np.float64)
        assert np.abs(out[k].astype(np.float64) - b).max() <= 1e-6 * np.abs(b).max(), k
