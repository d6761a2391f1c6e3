"""GPU: device optimizer (fwi_lbfgs_*) vs the FP64 numpy restatement of main.c:539-800; autograd binding;
the multiscale drivers on a small inversion."""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import oracle_gathers, oracle_gradient, oracle_setup, rel_l2, say, small_workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from lbfgs_oracle import LbfgsOracle  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    return t


@pytest.mark.parametrize("nx,nz,m,its", [(40, 30, 3, 7), (1600, 400, 5, 7), (4000, 1000, 5, 3)])
def test_lbfgs_iterations_match_numpy_oracle(torch, nx, nz, m, its):
    """SD step then L-BFGS iterations (with memory 3 over 7 iterations the history ring wraps twice):
    direction, trial and restore against the FP64 oracle, on a small grid and on the survey1600 and large4000
    grids, whose FP64 dot products and norms reduce over 0.6 and 4 million cells per block."""
    from fwi_amd.optim import LBFGS

    say(f"device L-BFGS vs numpy oracle, {nx}x{nz}, m={m}, {its} iterations")
    rng = np.random.default_rng(5)
    model_h = [(2.5e6 + 1e5 * rng.standard_normal((nz, nx))).astype(np.float32) for _ in range(3)]
    model_h[0] = np.maximum(model_h[0], model_h[2])
    opt = LBFGS(nx, nz, m)
    orc = LbfgsOracle(m)
    dev = [torch.from_numpy(x.copy()).cuda() for x in model_h]
    g = rng.standard_normal((3, nz, nx))
    orc_model = orc.sd_step(model_h, list(g), (3e5, 2e5, 2.5e5))
    opt.sd_step(dev, torch.from_numpy(g).cuda(), (3e5, 2e5, 2.5e5))
    torch.cuda.synchronize()
    for k in range(3):  # norms are FP64 tree sums on the device, pairwise sums in numpy: <= 1 ulp apart
        np.testing.assert_allclose(dev[k].cpu().numpy(), orc_model[k], rtol=1e-6)
    cur = [d.cpu().numpy() for d in dev]
    for it in range(its):
        # a quadratic-like gradient so that curvature pairs are positive
        g = np.stack([1e-6 * (c.astype(np.float64) - 2.4e6) + 0.01 * rng.standard_normal((nz, nx)) for c in cur])
        r_ref = orc.direction(cur, list(g))
        opt.direction(dev, torch.from_numpy(g).cuda())
        r = opt.get_direction().cpu().numpy()
        for k in range(3):
            assert rel_l2(r[k], r_ref[k]) < 1e-10, (it, k)
        alpha = (0.5 ** it,) * 3
        trial_ref = orc.trial(alpha)
        opt.trial(dev, alpha)
        torch.cuda.synchronize()
        got = [d.cpu().numpy() for d in dev]
        for k in range(3):
            np.testing.assert_allclose(got[k], trial_ref[k], rtol=1e-6)
        cur = got
        assert opt.pairs == min(it + 1, m)
    opt.restore(dev)
    torch.cuda.synchronize()
    for k in range(3):
        np.testing.assert_array_equal(dev[k].cpu().numpy(), orc.back_m[k])


def test_autograd_matches_gradient_and_fd(torch):
    """torch.autograd through the HIP forward and exact adjoint: (1) the misfit gradient equals
    fwi_gradient + chain rule bit for bit; (2) it agrees with the FP64 oracle gradient; (3) a central
    finite difference of the FP32 misfit along a smooth direction agrees to 1e-2."""
    from fwi_amd import models as M
    from fwi_amd.autograd import misfit
    from fwi_amd.survey import Plan, chain_rule, pre_variables

    say("autograd misfit gradient vs fwi_gradient, oracle FP64 and FD")
    w = small_workload(nshots=2, nt=240)
    tab, src, sq_true, sq_init = oracle_setup(w, "f32")
    obs_np = np.concatenate([o.reshape(-1) for o in oracle_gathers(w, "f32", sq_true, tab, src)])
    obs = torch.from_numpy(obs_np).cuda()
    plan = Plan(w.nx, w.nz, w.nt, w.dh, w.dt, w.cpml, w.shots, src, tab)
    vp, eps, dl = (torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda().requires_grad_(True)
                   for x in w.init_model)
    J = misfit(vp, eps, dl, plan, obs)
    J.backward()
    phi, gsq, _ = plan.gradient(pre_variables(vp.detach(), eps.detach(), dl.detach()), obs)
    g3 = chain_rule(vp.detach(), eps.detach(), dl.detach(), gsq).float()
    for got, ref in zip((vp.grad, eps.grad, dl.grad), g3):
        assert torch.equal(got, ref)
    assert abs(J.item() - phi) <= 1e-9 * phi
    # FP64 oracle gradient with respect to (vp, eps, delta)
    import oracle_lib as O

    tab64, src64, st64, si64 = oracle_setup(w, "f64")
    obs64 = [o.astype(np.float64) for o in oracle_gathers(w, "f32", sq_true, tab, src)]
    _, G64, _ = oracle_gradient(w, "f64", True, si64, tab64, src64.astype(np.float64), obs64)
    r64 = O.chain_rule(*[x.astype(np.float64) for x in w.init_model], G64[0], G64[1], G64[2])
    errs = [rel_l2(a.cpu().numpy(), b) for a, b in zip((vp.grad, eps.grad, dl.grad), r64)]
    print(f"autograd vs FP64 oracle (vp, eps, delta): {errs}")
    assert max(errs) < 5e-3
    # finite difference along a smooth direction of vp
    d = torch.from_numpy(M.smooth_perturbation(w.nx, w.nz, seed=9, sigma=6.0)[0].astype(np.float32)).cuda() * 20.0
    h = 1.0
    with torch.no_grad():
        Jp = misfit(vp + h * d, eps, dl, plan, obs).item()
        Jm = misfit(vp - h * d, eps, dl, plan, obs).item()
    fd = (Jp - Jm) / (2 * h)
    pred = torch.sum(vp.grad.double() * d.double()).item()
    print(f"FD {fd:.6e} vs adjoint {pred:.6e}")
    assert abs(fd - pred) <= 1e-2 * abs(pred)


def test_python_driver_reduces_misfit(torch):
    from fwi_amd import driver

    say("python multiscale driver, small workload, 4 iterations")
    w = small_workload(nshots=2, nt=240)
    model, logs = driver.run(w, stages=(None,), iterations=4, verbose=False)
    m = logs[0].misfits
    print("misfits", m)
    assert len(m) >= 2 and m[-1] < m[0]


def test_c_driver_runs_reference_flow(torch, tmp_path):
    """fwi_main (main.c's control flow on the C ABI) on raw model files of a small workload."""
    from fwi_amd import io as fio
    from fwi_amd.build import DRIVER_BIN

    say("C driver on small raw model files, 2 stages x 3 iterations")
    assert os.path.exists(DRIVER_BIN), "build the C driver first"
    w = small_workload(nx=120, nz=80, nt=240)
    fio.write_models(str(tmp_path), w.true_model, w.init_model)
    cmd = [DRIVER_BIN, "--dir", str(tmp_path), "--nx", "120", "--nz", "80", "--dh", "10", "--dt", "1e-3",
           "--tend", "0.24", "--cpml", "12", "--nshots", "3", "--f", "10", "--stages", "2", "--iterations", "3"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    errs = [float(l.split(":")[1]) for l in r.stdout.splitlines() if l.startswith("Error total")]
    assert len(errs) >= 4 and errs[1] < errs[0]
    for name in ("v_hor.bin", "v_po.bin", "v_nmo.bin"):
        a = fio.read_raw(os.path.join(tmp_path, name), 120, 80)
        assert np.isfinite(a).all()


def test_model_clamp_matches_numpy(torch):
    """fwi_model_clamp: each squared velocity to [vmin^2, vmax^2], then v_hor >= v_nmo."""
    from fwi_amd.optim import clamp_model

    rng = np.random.default_rng(2)
    m = [(rng.standard_normal((30, 40)) * 1e7 + 5e6).astype(np.float32) for _ in range(3)]
    dev = [torch.from_numpy(x.copy()).cuda() for x in m]
    clamp_model(dev, 1400.0, 5000.0)
    torch.cuda.synchronize()
    lo, hi = np.float32(1400.0 ** 2), np.float32(5000.0 ** 2)
    c = [np.minimum(np.maximum(x, lo), hi) for x in m]
    c[0] = np.where(c[0] < c[2], c[2], c[0])
    for k in range(3):
        np.testing.assert_array_equal(dev[k].cpu().numpy(), c[k])


@pytest.mark.parametrize("n", [1, 1000, 640000 + 17])
def test_dot3_matches_numpy(torch, n):
    """fwi_dot3: the three FP64 block dot products the optimizer uses (main.c:576-676 in FP64), against numpy,
    for one element per block, a small block and a block of the survey1600 size plus a ragged tail."""
    import ctypes

    from fwi_amd import _lib

    say(f"fwi_dot3 vs numpy, n={n}")
    rng = np.random.default_rng(n)
    a = rng.standard_normal((3, n))
    b = rng.standard_normal((3, n))
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = (ctypes.c_double * 3)()
    _lib.check(_lib.load().fwi_dot3(da.data_ptr(), db.data_ptr(), n, out, None), "fwi_dot3")
    ref = [float(np.dot(a[k], b[k])) for k in range(3)]
    scale = [float(np.dot(np.abs(a[k]), np.abs(b[k]))) for k in range(3)]
    for k in range(3):
        assert abs(out[k] - ref[k]) <= 1e-13 * scale[k], (k, out[k], ref[k])
