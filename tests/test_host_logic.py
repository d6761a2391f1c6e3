"""Host logic on the CPU: line-search state machine, L-BFGS oracle pins, file formats, survey geometry."""
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from lbfgs_oracle import LbfgsOracle, dense_bfgs_direction, two_loop  # noqa: E402

from fwi_amd import io as fio  # noqa: E402
from fwi_amd import models as M  # noqa: E402
from fwi_amd.optim import LineSearch  # noqa: E402


# ------------------------------------------------------------------------------------------------
# Line search (main.c:690-786) on scripted misfits
# ------------------------------------------------------------------------------------------------

def _script(values):
    it = iter(values)
    seen = []

    def evaluate(alpha):
        seen.append(alpha)
        return next(it)

    return evaluate, seen


def test_line_search_accepts_first_decrease():
    ls = LineSearch()
    ev, seen = _script([9.0])
    phi, trials = ls.run(10.0, ev, lambda: pytest.fail("no restore"))
    assert phi == 9.0 and trials == 1 and seen == [(1.0, 1.0, 1.0)] and ls.count == 0


def test_line_search_halves_on_increase_and_nan_and_keeps_state():
    ls = LineSearch()
    ev, seen = _script([11.0, float("nan"), 10.0])
    phi, trials = ls.run(10.0, ev, lambda: pytest.fail("no restore"))
    assert phi == 10.0 and trials == 3  # equal misfit is accepted (loop runs while trial > phi_n)
    assert seen == [(1.0,) * 3, (0.5,) * 3, (0.25,) * 3]
    assert ls.count == 2
    # alpha and count persist into the next iteration of the stage (main.c:306-311)
    ev2, seen2 = _script([5.0])
    ls.run(6.0, ev2, lambda: None)
    assert seen2 == [(0.25,) * 3] and ls.count == 2


def test_line_search_aborts_after_six_failures_cumulative():
    ls = LineSearch()
    restored = []
    ev, seen = _script([20.0] * 10)
    phi, trials = ls.run(10.0, ev, lambda: restored.append(1))
    assert phi is None and trials == 6 and restored == [1] and ls.count == 6
    assert seen[-1] == (1.0 / 32,) * 3


def test_line_search_nan_counts_as_failure():
    ls = LineSearch()
    ev, _ = _script([float("nan")] * 6)
    restored = []
    phi, trials = ls.run(1e40, ev, lambda: restored.append(1))
    # NaN becomes 1e30 <= 1e40 and is accepted, with alpha already halved (the reference's order)
    assert phi == 1e30 and trials == 1 and ls.alpha == [0.5] * 3 and ls.count == 1


# ------------------------------------------------------------------------------------------------
# L-BFGS oracle pinned against dense BFGS
# ------------------------------------------------------------------------------------------------

def _pairs(rng, n, k):
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)  # SPD: every pair has positive curvature
    out = []
    for _ in range(k):
        s = rng.standard_normal(n)
        out.append((s, A @ s))
    return out


@pytest.mark.parametrize("k", [1, 3, 5])
def test_two_loop_equals_dense_bfgs(k):
    """The two-loop recursion equals H g for the dense BFGS inverse update with H0 = gamma I, and H maps
    the newest y to the newest s (secant identity)."""
    rng = np.random.default_rng(k)
    n = 24
    pairs = _pairs(rng, n, k)
    g = rng.standard_normal(n)
    r = two_loop([p[0] for p in pairs], [p[1] for p in pairs], g)
    np.testing.assert_allclose(r, dense_bfgs_direction(pairs, g), rtol=1e-9, atol=1e-12)
    sec = two_loop([p[0] for p in pairs], [p[1] for p in pairs], pairs[-1][1])
    np.testing.assert_allclose(sec, pairs[-1][0], rtol=1e-9, atol=1e-12)
    # a planted error (gamma from the oldest pair) breaks the equality once there are two pairs
    if k > 1:
        bad = dense_bfgs_direction(pairs[::-1], g)
        assert np.abs(bad - r).max() > 1e-6 * np.abs(r).max()


def test_lbfgs_history_is_fifo_of_m():
    orc = LbfgsOracle(m=2)
    rng = np.random.default_rng(0)
    n = 40
    model = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
    grad = [rng.standard_normal(n) for _ in range(3)]
    orc.sd_step([x.reshape(20, 2) for x in model], [g.reshape(20, 2) for g in grad], (1.0, 1.0, 1.0))
    for it in range(4):
        m2 = [(x + rng.standard_normal(n).astype(np.float32) * 0.1).reshape(20, 2) for x in model]
        g2 = [(g + rng.standard_normal(n)).reshape(20, 2) for g in grad]
        orc.direction(m2, g2)
        assert len(orc.s) == min(it + 1, 2)


# ------------------------------------------------------------------------------------------------
# Files and geometry
# ------------------------------------------------------------------------------------------------

def test_raw_model_round_trip(tmp_path):
    nx, nz = 30, 20
    rng = np.random.default_rng(1)
    true = tuple(rng.random((nz, nx)).astype(np.float32) for _ in range(3))
    init = tuple(rng.random((nz, nx)).astype(np.float32) for _ in range(3))
    fio.write_models(str(tmp_path), true, init)
    t2, i2 = fio.read_models(str(tmp_path), nx, nz)
    for a, b in zip(true + init, t2 + i2):
        np.testing.assert_array_equal(a, b)
    # raw layout: x fastest, little-endian float32 (main.c:101-122)
    raw = np.fromfile(os.path.join(tmp_path, "vp_ori.bin"), "<f4")
    assert raw[1] == true[0][0, 1] and raw[nx] == true[0][1, 0]
    with pytest.raises(ValueError):
        fio.read_raw(os.path.join(tmp_path, "vp_ori.bin"), nx + 1, nz)


def test_squared_velocities_invert_pre_variables():
    vp = np.full((4, 5), 2500.0, np.float32)
    eps = np.full((4, 5), 0.2, np.float32)
    dl = np.full((4, 5), 0.1, np.float32)
    v_po = vp.astype(np.float64) ** 2
    v_hor = v_po * (1 + 2 * eps)
    v_nmo = v_po * (1 + 2 * dl)
    a, b, c = fio.squared_to_vti(v_hor, v_po, v_nmo)
    np.testing.assert_allclose(a, vp, rtol=1e-6)
    np.testing.assert_allclose(b, eps, rtol=1e-5)
    np.testing.assert_allclose(c, dl, rtol=1e-5)


def test_workload_geometry_rules():
    w = M.workload("plain200")
    s = w.shots[0]
    assert (s.isx, s.isz) == (100, 6) and s.nrec == 160 and s.ir0 == 20
    w2 = M.workload("survey1600", nt=10)
    assert len(w2.shots) == 64 and w2.shots[0].nrec == 1560
    xs = [s.isx for s in w2.shots]
    assert xs[0] == 20 and xs[-1] == 1579 and all(b > a for a, b in zip(xs, xs[1:]))
    w3 = M.workload("uneven1600", nt=10)
    assert len(w3.shots) == 96
    for s in w3.shots:
        assert 300 + 40 <= s.nx <= 1300 + 40 and 200 <= s.nrec <= 1200 or s.nrec == s.nx - 40
        assert s.x0 + s.nx <= 1600 and s.ir0 >= 20 and s.ir0 + s.nrec <= s.nx - 20
        assert 20 <= s.isx < s.nx - 20


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says how shots would shard, before any
    device work (so it runs here without a GPU)."""
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and "all-reduce" in r.stderr
    assert r.stdout == ""


def test_gradient_groups_rule():
    """The gradient-group rule (its restatement in tests/helpers.py, which the GPU tests hold the plan to): every
    shot in exactly one group of at most four, leaders in window order; identical windows group in runs of four
    (survey1600); on the ragged survey the rule leaves fewer placeholder waves than runs of four in window order."""
    from helpers import _placeholder_waves, gradient_groups

    from fwi_amd import models as M

    full = M.workload("survey1600")
    assert gradient_groups(full.shots, 4) == [list(range(k, k + 4)) for k in range(0, 64, 4)]
    w = M.workload("uneven1600")
    sh = w.shots
    groups = gradient_groups(sh, 4)
    assert sorted(i for g in groups for i in g) == list(range(len(sh)))
    assert all(1 <= len(g) <= 4 for g in groups)
    key = lambda i: (sh[i].x0, sh[i].nx, i)
    assert [key(g[0]) for g in groups] == sorted(key(g[0]) for g in groups)
    order = sorted(range(len(sh)), key=key)
    runs = [order[k:k + 4] for k in range(0, len(order), 4)]
    waves = lambda gs: sum(len(g) * len(range(min(sh[i].x0 for i in g), max(sh[i].x0 + sh[i].nx for i in g), 64))
                           for g in gs)
    share = sum(_placeholder_waves(sh, g) for g in groups) / waves(groups)
    share_runs = sum(_placeholder_waves(sh, g) for g in runs) / waves(runs)
    print(f"placeholder waves: {share:.3f} of the grouped waves, runs of four {share_runs:.3f}")
    assert share < 0.6 * share_runs
    # batches group separately
    assert all(max(g) < 48 or min(g) >= 48 for g in gradient_groups(sh, 4, max_batch=48))
