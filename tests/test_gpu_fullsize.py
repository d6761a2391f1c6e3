"""GPU checks at the benchmark's own size: survey1600, 1600 x 400 cells, 64 full-width shots, 4000 steps.

The oracle cannot run 64 shots x 4000 steps in seconds, so the full size is covered three ways:
  * three shots of the full 64-shot batch (the first, a middle and the last, whose fields lie gigabytes into
    the workspace) against the oracle's per-shot forward, bit for bit, on a 300-step time axis;
  * one full-width shot's exact gradient against the oracle, bit for bit, on a 400-step time axis with
    recomputation from checkpoints;
  * size-independent properties of the full 4000-step, 64-shot gradient in the benchmark's configuration
    (default options, captured graphs, the tuner's tiling, trailing segments kept to the memory budget):
    a source delayed by m steps delays every gather by m steps bit for bit, and a second plan with another
    checkpoint interval, one kept segment, eager launches, another tiling and single steps instead of step pairs
    returns the same misfit and the same gradient bits (recomputation and the schedule change nothing).
"""
import numpy as np
import pytest

from helpers import oracle_gathers, oracle_gradient, oracle_setup, rel_l2, say
from fwi_amd import models as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t

    return t


def _plan(w, tab, src, **kw):
    from fwi_amd.survey import Plan

    return Plan(w.nx, w.nz, w.nt, w.dh, w.dt, w.cpml, w.shots, src, tab, **kw)


def _dev_model(torch, sq):
    return tuple(torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda() for x in sq)


def test_survey1600_batch_gathers_match_oracle(torch):
    w = M.workload("survey1600", nt=300)
    say("survey1600 forward, 64 shots x 300 steps, shots 0/31/63 against the oracle")
    tab, src, sq_true, _ = oracle_setup(w, "f32i")
    plan = _plan(w, tab, src, mem_budget=40e9)
    try:
        g = plan.split_gathers(plan.forward(_dev_model(torch, sq_true)))
        torch.cuda.synchronize()
        for i in (0, 31, 63):
            sub = M.Workload(w.name, w.nx, w.nz, w.dh, w.dt, w.nt, w.f, w.cpml, w.true_model, w.init_model,
                             [w.shots[i]])
            ref = oracle_gathers(sub, "f32i", sq_true, tab, src)[0]
            got = g[i].cpu().numpy()
            print(f"shot {i} (x={w.shots[i].isx}): rel-L2 {rel_l2(got, ref):.3e}")
            assert np.array_equal(got, ref), i
    finally:
        plan.close()


def test_full_width_shot_gradient_matches_oracle(torch):
    w = M.workload("survey1600", nshots=1, nt=400)
    say("survey1600 grid, one full-width shot, 400 steps: exact gradient against the oracle")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = oracle_gathers(w, "f32i", sq_true, tab, src)
    phi_ref, G_ref, _ = oracle_gradient(w, "f32i", True, sq_init, tab, src, obs)
    plan = _plan(w, tab, src, store_segments=1)
    try:
        assert plan.kept_steps < w.nt
        phi, G, _ = plan.gradient(_dev_model(torch, sq_init), torch.from_numpy(obs[0].reshape(-1)).cuda())
        torch.cuda.synchronize()
        G = G.cpu().numpy()
        assert abs(phi - phi_ref) <= 1e-9 * phi_ref
        for k in range(3):
            print(f"grad[{k}] rel-L2 {rel_l2(G[k], G_ref[k]):.3e}")
            assert np.array_equal(G[k], G_ref[k]), k
    finally:
        plan.close()


def test_survey1600_full_gradient_properties(torch):
    w = M.workload("survey1600")
    say("survey1600 at full size: 64 shots x 4000 steps; time shift and schedule invariance")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    true_d, init_d = _dev_model(torch, sq_true), _dev_model(torch, sq_init)
    plan = _plan(w, tab, src)  # the benchmark's configuration
    try:
        obs = plan.forward(true_d)
        phi_a, G_a, _ = plan.gradient(init_d, obs)
        torch.cuda.synchronize()
        assert np.isfinite(phi_a) and phi_a > 0 and torch.isfinite(G_a).all().item()
        kept, tiling, plan_pairs = plan.kept_steps, (plan.rows_per_item, plan.item_order, plan.adj_tiling), plan.pair_rows
        print(f"misfit {phi_a:.9e}; kept steps {kept}, K {plan.ckpt_interval}, tiling {tiling}")
        # time shift: the whole survey's gathers move by m steps, nothing else changes
        m = 173
        late = np.zeros_like(src)
        late[m:] = src[:w.nt - m]
        plan.set_wavelet(late)
        shifted = plan.forward(true_d)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(plan.split_gathers(obs), plan.split_gathers(shifted))):
            assert not b[:m].any().item(), i
            assert torch.equal(b[m:], a[:w.nt - m]), i
        del shifted
    finally:
        plan.close()
    other = _plan(w, tab, src, ckpt_interval=50, store_segments=1, use_graph=0, rows_per_item=32, item_order=0,
                  step_pairs=0)
    try:
        assert other.kept_steps != kept
        phi_b, G_b, _ = other.gradient(init_d, obs)
        torch.cuda.synchronize()
        assert plan_pairs > 0 and other.pair_rows == 0
        print(f"second plan: K {other.ckpt_interval}, kept {other.kept_steps}, single steps, misfit {phi_b:.9e}")
        assert phi_b == phi_a
        assert torch.equal(G_b, G_a)
    finally:
        other.close()


def test_plain200_full_gradient_matches_oracle(torch):
    """The reference's own case at its full size (200 x 200, one shot, 1000 steps), both gradient modes."""
    w = M.workload("plain200")
    say("plain200 at full size: gradient (exact and reference-compatible) against the oracle")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = oracle_gathers(w, "f32i", sq_true, tab, src)
    for exact in (True, False):
        phi_ref, G_ref, _ = oracle_gradient(w, "f32i", exact, sq_init, tab, src, obs)
        plan = _plan(w, tab, src, grad_mode=1 if exact else 0)
        try:
            phi, G, _ = plan.gradient(_dev_model(torch, sq_init), torch.from_numpy(obs[0].reshape(-1)).cuda())
            torch.cuda.synchronize()
            G = G.cpu().numpy()
            assert abs(phi - phi_ref) <= 1e-9 * phi_ref
            for k in range(3):
                print(f"exact={exact} grad[{k}] rel-L2 {rel_l2(G[k], G_ref[k]):.3e}")
                assert np.array_equal(G[k], G_ref[k]), (exact, k)
        finally:
            plan.close()


def test_uneven1600_ragged_batch(torch):
    """The ragged survey's 96 windows (343-1340 columns) in one batch on a 600-step axis: the narrowest, the
    widest and a middle shot's gathers equal the oracle's bit for bit, and the grouped gradient over the union
    windows is the same for another checkpoint interval, kept-segment count, launch path and tiling."""
    w = M.workload("uneven1600", nt=600)
    widths = [s.nx for s in w.shots]
    pick = sorted({int(np.argmin(widths)), int(np.argmax(widths)), len(w.shots) // 2})
    say(f"uneven1600, 96 ragged shots x 600 steps: shots {pick} against the oracle, schedule invariance")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    init_d = _dev_model(torch, sq_init)
    plan = _plan(w, tab, src, mem_budget=60e9)
    try:
        obs = plan.forward(_dev_model(torch, sq_true))
        torch.cuda.synchronize()
        g = plan.split_gathers(obs)
        for i in pick:
            s = w.shots[i]
            sub = M.Workload(w.name, w.nx, w.nz, w.dh, w.dt, w.nt, w.f, w.cpml, w.true_model, w.init_model, [s])
            ref = oracle_gathers(sub, "f32i", sq_true, tab, src)[0]
            got = g[i].cpu().numpy()
            print(f"shot {i} (window {s.x0}+{s.nx}, {s.nrec} receivers): rel-L2 {rel_l2(got, ref):.3e}")
            assert np.array_equal(got, ref), i
        phi_a, G_a, _ = plan.gradient(init_d, obs)
        torch.cuda.synchronize()
        assert plan.adj_tiling[2]  # grouped
    finally:
        plan.close()
    other = _plan(w, tab, src, ckpt_interval=31, store_segments=1, use_graph=0, rows_per_item=24, item_order=2,
                  mem_budget=60e9)
    try:
        phi_b, G_b, _ = other.gradient(init_d, obs)
        torch.cuda.synchronize()
        print(f"misfit {phi_a:.9e} / {phi_b:.9e}")
        assert phi_b == phi_a
        assert torch.equal(G_b, G_a)
    finally:
        other.close()


@pytest.mark.parametrize("fmt", [1, 2])
def test_uneven1600_16bit_checkpoints(torch, fmt):
    """16-bit checkpoints (fmt 1 BF16, 2 scaled FP16) on the 96 ragged windows, 600 steps: the result does not
    depend on the tiling or the launch path; the rounding moves the gradient away from the FP32-checkpoint one
    (which proves the 16-bit path ran), by less than 5e-2 (BF16) and 1e-2 (FP16) relative L2."""
    w = M.workload("uneven1600", nt=600)
    say(f"uneven1600 x 600 steps, 16-bit checkpoints fmt={fmt}: schedule invariance and distance to FP32")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    init_d = _dev_model(torch, sq_init)
    runs = []
    # the FP32-checkpoint plan without the FP64 band, as the 16-bit plans run
    for kw in ({}, {"use_graph": 0, "rows_per_item": 24, "item_order": 2},
               {"ckpt_format": 0, "store_segments": 1, "surface_rows": -1}):
        opts = {"ckpt_format": fmt, "ckpt_interval": 31, "mem_budget": 60e9}
        opts.update(kw)
        plan = _plan(w, tab, src, **opts)
        try:
            if not runs:
                obs = plan.forward(_dev_model(torch, sq_true))
            runs.append(plan.gradient(init_d, obs)[:2])
            torch.cuda.synchronize()
        finally:
            plan.close()
    (phi_a, G_a), (phi_b, G_b), (phi_32, G_32) = runs
    assert phi_b == phi_a and torch.equal(G_b, G_a)
    assert phi_32 == phi_a  # the misfit comes from the forward, which checkpoints do not touch
    errs = [rel_l2(G_a[k].cpu().numpy(), G_32[k].cpu().numpy()) for k in range(3)]
    print(f"fmt {fmt}: gradient vs FP32 checkpoints rel-L2 {errs}")
    assert 0 < max(errs) < (5e-2 if fmt == 1 else 1e-2)


def test_survey_scale_fp32_gradient_error_against_fp64(torch):
    """The default FP32 arithmetic (increment-form leapfrog in the forward and on the near-surface rows of the
    adjoint, rows 4..23 of the forward in FP64, one gradient group of four with its FP32 in-group sum) against the
    FP64 oracle at the survey's own scale: four full-width shots of survey1600 (0, 21, 42, 63; the first source on
    the C-PML's edge), 8 Hz, dt = 0.8 ms, all 4000 steps, on the FP64 oracle's observed data.  The summed gradient
    is within 1e-4 (relative L2) of the FP64 one on the full grid, both against FP64 on the same float32 model
    values (the arithmetic's own error) and against FP64 on the double model (which adds the model's float32
    rounding); the rows the reference's update uses (iz > 15, include/parallelcalls.h:6) likewise."""
    w = M.workload("survey1600")
    w.shots = [w.shots[i] for i in (0, 21, 42, 63)]
    say("survey-scale FP32 gradient vs FP64 oracle: 4 full-width shots of survey1600, 4000 steps")
    tab64, src, sq_true64, sq_init64 = oracle_setup(w, "f64")
    src64 = src.astype(np.float64)
    obs64 = oracle_gathers(w, "f64", sq_true64, tab64, src64)
    tab, src32, _, sq_init = oracle_setup(w, "f32i")
    plan = _plan(w, tab, src32)
    try:
        assert plan.surface_rows == 24 and plan.adj_tiling[2]
        obs = torch.from_numpy(np.concatenate([o.astype(np.float32).reshape(-1) for o in obs64])).cuda()
        _, G, _ = plan.gradient(_dev_model(torch, sq_init), obs)
        torch.cuda.synchronize()
        G = G.cpu().numpy()
        del obs
    finally:
        plan.close()
    upd = np.arange(w.nz) > 15
    same = tuple(x.astype(np.float64) for x in sq_init)
    for name, model in (("same float32 model", same), ("double model", sq_init64)):
        say(f"FP64 oracle gradient, {name}")
        _, ref, _ = oracle_gradient(w, "f64", True, model, tab64, src64, obs64, group=1)
        e_all = [rel_l2(G[k], ref[k]) for k in range(3)]
        e_upd = [rel_l2(G[k][upd], ref[k][upd]) for k in range(3)]
        print(f"vs FP64, {name}: full grid {['%.2e' % e for e in e_all]}; rows iz>15 {['%.2e' % e for e in e_upd]}")
        assert max(e_all) < 1e-4, name
        assert max(e_upd) < 1e-4, name
