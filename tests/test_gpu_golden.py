"""The HIP path against the committed golden fixtures (tests/golden/golden_small64.npz): forward gathers and
exact and reference-compatible gradients bit for bit, misfits, and the device optimizer's steepest-descent
and L-BFGS steps, on the fixture's inputs."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from helpers import rel_l2, say
from fwi_amd import models as M

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "golden_small64.npz")


@pytest.fixture(scope="module")
def torch():
    import torch as t

    if not t.cuda.is_available():
        pytest.skip("no GPU")
    return t


@pytest.fixture(scope="module")
def gold():
    with open(FIXTURE.replace(".npz", ".json")) as fh:
        meta = json.load(fh)
    return np.load(FIXTURE, allow_pickle=False), meta


def _plan(meta, g, **kw):
    from fwi_amd.survey import Plan

    s = meta["shot"]
    shot = M.Shot(x0=0, nx=meta["nx"], isx=s["isx"], isz=s["isz"], irz=s["irz"], ir0=s["ir0"], nrec=s["nrec"])
    return Plan(meta["nx"], meta["nz"], meta["nt"], meta["dh"], meta["dt"], meta["cpml"], [shot], g["ricker"],
                (g["cpml_a"], g["cpml_b"]), **kw)


def _model(torch, g, which):
    return tuple(torch.from_numpy(np.ascontiguousarray(g[f"sq_{which}_{n}"])).cuda() for n in ("vh", "vp", "vn"))


@pytest.mark.parametrize("arith,key", [(1, "gather_mix"), (3, "gather_f32i")])
def test_forward_matches_golden(torch, gold, arith, key):
    g, meta = gold
    say(f"golden small64 forward arith={arith}")
    plan = _plan(meta, g, arith=arith)
    got = plan.forward(_model(torch, g, "true")).cpu().numpy().reshape(meta["nt"], -1)
    err = rel_l2(got, g[key])
    print(f"{key}: rel-L2 {err:.3e}, max|diff| {np.abs(got - g[key]).max():.3e}")
    assert np.array_equal(got, g[key])


@pytest.mark.parametrize("exact,arith,key", [(1, 3, "exact_f32i"), (0, 1, "ref_mix")])
def test_gradient_matches_golden(torch, gold, exact, arith, key):
    g, meta = gold
    say(f"golden small64 gradient exact={exact} arith={arith}")
    plan = _plan(meta, g, arith=arith, grad_mode=exact)
    obs = torch.from_numpy(np.ascontiguousarray(g["obs"]).reshape(-1)).cuda()
    phi, G, _ = plan.gradient(_model(torch, g, "init"), obs)
    G = G.cpu().numpy()
    ref = g[f"grad_{key}"]
    assert abs(phi - float(g[f"misfit_{key}"])) <= 1e-9 * float(g[f"misfit_{key}"])
    for k in range(3):
        e = rel_l2(G[k], ref[k])
        print(f"grad {key}[{k}]: rel-L2 {e:.3e}, max|diff| {np.abs(G[k] - ref[k]).max():.3e}")
        assert np.array_equal(G[k], ref[k].astype(np.float64))  # one shot: the FP64 sum of one FP32 slot
    if exact:  # FP32 increment form against the FP64 exact gradient: the 1e-4 target
        for k in range(3):
            assert rel_l2(G[k], g["grad_exact_f64"][k]) < 1e-4


def test_optimizer_steps_match_golden(torch, gold):
    from fwi_amd.optim import LBFGS

    g, meta = gold
    say("golden small64 device steepest-descent and L-BFGS steps")
    nx, nz = meta["nx"], meta["nz"]
    opt = LBFGS(nx, nz, meta["lbfgs_memory"])
    dev = [t.clone() for t in _model(torch, g, "init")]
    opt.sd_step(dev, torch.from_numpy(np.ascontiguousarray(g["grad_exact_f32i"], np.float64)).cuda(),
                tuple(meta["sd_alpha"]))
    torch.cuda.synchronize()
    for k in range(3):  # norms: FP64 tree sums on the device, pairwise sums in numpy (<= 1 ulp apart)
        np.testing.assert_allclose(dev[k].cpu().numpy(), g["sd_model"][k], rtol=1e-6)
    opt.direction(dev, torch.from_numpy(np.ascontiguousarray(g["lbfgs_grad"], np.float64)).cuda())
    r = opt.get_direction().cpu().numpy()
    for k in range(3):
        assert rel_l2(r[k], g["lbfgs_dir"][k]) < 1e-10, k
    opt.trial(dev, (meta["lbfgs_trial_alpha"],) * 3)
    torch.cuda.synchronize()
    for k in range(3):
        np.testing.assert_allclose(dev[k].cpu().numpy(), g["lbfgs_trial"][k], rtol=1e-6)
