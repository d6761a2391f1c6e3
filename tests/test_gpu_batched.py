"""GPU parity of the batched FWI path (include/fwi.h) against the oracle.

Each arithmetic (FP32, MIXED, FP32D, FP32I) is compared with the oracle's instantiation of the same name:
same expressions, same order, so gathers and gradients must be equal bit for bit (the rel-L2 is printed
beside); the FP64 misfit is held to 1e-9 relative, since the device adds its per-block partial sums in
another order than the oracle's sequential sum.  Different checkpoint intervals and batch splits must give
identical results, because recomputation replays the same deterministic kernels.  The error of the FP32 path against the FP64
oracle is measured and bounded separately (the FP32 rounding floor of the leapfrog: a few 1e-4 in the
two-level form, ~2e-5 in the default increment form).
"""
import numpy as np
import pytest

import oracle_lib as O
from helpers import oracle_gathers, oracle_gradient, oracle_setup, rel_l2, say, small_workload
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


def _cat(gs):
    return np.concatenate([g.reshape(-1) for g in gs]).astype(np.float32)


@pytest.mark.parametrize("arith,mode", [(0, "f32"), (1, "mix"), (2, "f32d"), (3, "f32i")])
def test_forward_plain200_matches_oracle(torch, arith, mode):
    w = M.workload("plain200")
    say(f"batched forward plain200 arith={mode}")
    tab, src, sq_true, _ = oracle_setup(w, mode)
    plan = _plan(w, tab, src, arith=arith)
    g = plan.forward(_dev_model(torch, sq_true))
    torch.cuda.synchronize()
    ref = oracle_gathers(w, mode, sq_true, tab, src)[0]
    got = g.cpu().numpy().reshape(w.nt, -1)
    err = rel_l2(got, ref)
    print(f"forward plain200 {mode}: rel-L2 {err:.3e}, max|diff| {np.abs(got - ref).max():.3e}")
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("arith,mode", [(0, "f32"), (1, "mix"), (2, "f32d"), (3, "f32i")])
def test_gradient_matches_oracle(torch, exact, arith, mode):
    w = small_workload(nshots=2)
    say(f"batched gradient small 96x80 nt=240 exact={exact} arith={mode}")
    tab, src, sq_true, sq_init = oracle_setup(w, mode)
    obs = oracle_gathers(w, mode, sq_true, tab, src)
    phi_ref, G_ref, syn_ref = oracle_gradient(w, mode, exact, sq_init, tab, src, obs)
    plan = _plan(w, tab, src, arith=arith, grad_mode=1 if exact else 0)
    obs_d = torch.from_numpy(_cat(obs)).cuda()
    phi, G, syn = plan.gradient(_dev_model(torch, sq_init), obs_d, want_syn=True)
    torch.cuda.synchronize()
    G = G.cpu().numpy()
    es = rel_l2(syn.cpu().numpy(), _cat(syn_ref))
    print(f"syn rel-L2 {es:.3e}; misfit {phi:.9e} vs {phi_ref:.9e}")
    assert np.array_equal(syn.cpu().numpy(), _cat(syn_ref))
    assert abs(phi - phi_ref) <= 1e-9 * phi_ref
    for k, name in enumerate(("v_hor", "v_po", "v_nmo")):
        e = rel_l2(G[k], G_ref[k])
        print(f"grad {name}: rel-L2 {e:.3e}")
        assert np.array_equal(G[k], G_ref[k])


@pytest.mark.parametrize("cpml", [5, 12])
def test_odd_geometries_match_oracle(torch, cpml):
    """Windows that exercise the kernels' choice between the pipelined interior bodies and the general body and
    the cuts of the row runs: narrower than one 64-column band, exactly one interior band wide, one column more,
    a source next to a C-PML strip, a short last row run (nz - 8 not a multiple of the rows per item), with the
    thinnest C-PML the plan accepts (one cell past the rim) or a thicker one; gathers, misfit and both
    gradients against the oracle."""
    nx, nz, nt = 230, 45, 90
    w = small_workload(nx=nx, nz=nz, nt=nt, cpml=cpml, nshots=1)
    e = max(cpml, 4)
    w.shots = [M.Shot(x0=0, nx=40, isx=20, isz=6, irz=6, ir0=cpml, nrec=40 - 2 * cpml),
               M.Shot(x0=10, nx=64 + 2 * e, isx=e + 1, isz=7, irz=6, ir0=cpml, nrec=64 + 2 * e - 2 * cpml),
               M.Shot(x0=100, nx=65 + 2 * e, isx=40, isz=9, irz=5, ir0=cpml + 3, nrec=30),
               M.Shot(x0=0, nx=nx, isx=nx - cpml - 1, isz=6, irz=6, ir0=cpml, nrec=nx - 2 * cpml)]
    say(f"odd geometries: cpml={cpml}, nz={nz}, windows 40 / {64 + 2 * e} / {65 + 2 * e} / {nx}")
    for exact in (True, False):
        tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
        obs = oracle_gathers(w, "f32i", sq_true, tab, src)
        phi_ref, G_ref, syn_ref = oracle_gradient(w, "f32i", exact, sq_init, tab, src, obs)
        plan = _plan(w, tab, src, arith=3, grad_mode=1 if exact else 0, ckpt_interval=13, store_segments=1)
        g = plan.forward(_dev_model(torch, sq_true))
        phi, G, syn = plan.gradient(_dev_model(torch, sq_init), torch.from_numpy(_cat(obs)).cuda(), want_syn=True)
        torch.cuda.synchronize()
        assert np.array_equal(g.cpu().numpy(), _cat(obs))
        assert np.array_equal(syn.cpu().numpy(), _cat(syn_ref))
        assert abs(phi - phi_ref) <= 1e-9 * phi_ref
        G = G.cpu().numpy()
        for k in range(3):
            e_k = rel_l2(G[k], G_ref[k])
            print(f"exact={exact} grad[{k}] rel-L2 {e_k:.3e}")
            assert np.array_equal(G[k], G_ref[k])


def test_tiling_invariance(torch):
    """Work-item shape and order (rows per item, item order, as the tuner would choose them) change only the
    schedule: gathers, misfit and both gradients are bit-identical for every combination, on ragged windows
    with C-PML cuts inside the row runs."""
    w = small_workload(nx=240, nz=70, nt=80, cpml=12, nshots=3)
    w.shots = [M.Shot(x0=0, nx=150, isx=70, isz=6, irz=6, ir0=12, nrec=126),
               M.Shot(x0=20, nx=220, isx=100, isz=7, irz=6, ir0=12, nrec=196),
               M.Shot(x0=170, nx=70, isx=30, isz=6, irz=8, ir0=14, nrec=40)]
    say("tiling invariance: rows {4, 8, 16, 24} x orders {0, 1, 2, 3}, ragged windows")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = torch.from_numpy(_cat(oracle_gathers(w, "f32i", sq_true, tab, src))).cuda()
    model = _dev_model(torch, sq_init)
    true_model = _dev_model(torch, sq_true)
    for exact in (1, 0):
        ref = None
        for rows in (4, 8, 16, 24):
            for order in (0, 1, 2, 3):
                plan = _plan(w, tab, src, grad_mode=exact, rows_per_item=rows, item_order=order,
                             ckpt_interval=9, store_segments=1)
                g = plan.forward(true_model)
                phi, G, syn = plan.gradient(model, obs, want_syn=True)
                torch.cuda.synchronize()
                out = (g.cpu().numpy(), phi, G.cpu().numpy(), syn.cpu().numpy())
                if ref is None:
                    ref = out
                    continue
                assert np.array_equal(out[0], ref[0]), (exact, rows, order)
                assert out[1] == ref[1], (exact, rows, order)
                assert np.array_equal(out[2], ref[2]), (exact, rows, order)
                assert np.array_equal(out[3], ref[3]), (exact, rows, order)


def test_shared_model_tiles_invariance(torch):
    """Item order 4 (four shots with the same window per workgroup, the model rows of interior tiles shared
    through LDS) changes only the schedule: on eight full-width shots its gathers, misfit and gradient equal
    those of order 1 bit for bit, and the gradient equals the oracle's; so does an adjoint tiling pinned apart
    from the forward's (adj_rows_per_item, adj_item_order)."""
    w = small_workload(nx=200, nz=72, nt=90, cpml=12, nshots=8)
    say("shared-model tiles: order 4 vs order 1, 8 full-width shots")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs_np = oracle_gathers(w, "f32i", sq_true, tab, src)
    obs = torch.from_numpy(_cat(obs_np)).cuda()
    model = _dev_model(torch, sq_init)
    outs = []
    for order, adj in ((1, {}), (4, {}), (4, {"adj_rows_per_item": 8, "adj_item_order": 2})):
        plan = _plan(w, tab, src, rows_per_item=16, item_order=order, ckpt_interval=9, store_segments=1, **adj)
        try:
            g = plan.forward(_dev_model(torch, sq_true))
            if adj:
                assert plan.adj_tiling == (8, 2, True) and (plan.rows_per_item, plan.item_order) == (16, 4)
            phi, G, syn = plan.gradient(model, obs, want_syn=True)
            torch.cuda.synchronize()
            outs.append((g.cpu().numpy(), phi, G.cpu().numpy(), syn.cpu().numpy()))
        finally:
            plan.close()
    a, b, c = outs
    for x in (b, c):
        assert np.array_equal(a[0], x[0]) and a[1] == x[1]
        assert np.array_equal(a[2], x[2]) and np.array_equal(a[3], x[3])
    assert np.array_equal(b[0], _cat(obs_np))
    phi_ref, G_ref, _ = oracle_gradient(w, "f32i", True, sq_init, tab, src, obs_np)
    for k in range(3):
        assert np.array_equal(b[2][k], G_ref[k])


@pytest.mark.parametrize("rows", [-1, 6, 7, 30])
def test_surface_band_matches_oracle(torch, rows):
    """The FP64 near-surface band of the increment form at several depths -- none, ending on the source and
    receiver row 6 (both in the float rows), just below it (both in the band), deep -- gives the oracle's
    gathers and exact gradient bit for bit, with checkpoints every 9 steps (the band state is checkpointed and
    recomputed), ragged windows and the side C-PML crossing the band; the reference-compatible gradient too."""
    w = small_workload(nx=150, nz=72, nt=100, cpml=12, nshots=3)
    w.shots = [M.Shot(x0=0, nx=110, isx=50, isz=6, irz=6, ir0=12, nrec=86),
               M.Shot(x0=20, nx=130, isx=70, isz=6, irz=6, ir0=14, nrec=100),
               M.Shot(x0=90, nx=60, isx=30, isz=6, irz=6, ir0=12, nrec=36)]
    say(f"surface band rows={rows}: 3 ragged shots, K=9")
    with O.surface_rows(rows):
        tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
        obs = oracle_gathers(w, "f32i", sq_true, tab, src)
        phi_ref, G_ref, syn_ref = oracle_gradient(w, "f32i", True, sq_init, tab, src, obs)
        _, Gr_ref, _ = oracle_gradient(w, "f32i", False, sq_init, tab, src, obs)
    obs_d = torch.from_numpy(_cat(obs)).cuda()
    model = _dev_model(torch, sq_init)
    for exact, ref in ((1, G_ref), (0, Gr_ref)):
        plan = _plan(w, tab, src, ckpt_interval=9, store_segments=1, surface_rows=rows, grad_mode=exact)
        try:
            assert plan.surface_rows == (0 if rows < 0 else rows)
            g = plan.forward(_dev_model(torch, sq_true))
            phi, G, syn = plan.gradient(model, obs_d, want_syn=True)
            torch.cuda.synchronize()
            assert np.array_equal(g.cpu().numpy(), _cat(obs))
            if exact:
                assert np.array_equal(syn.cpu().numpy(), _cat(syn_ref))
                assert abs(phi - phi_ref) <= 1e-9 * phi_ref
            G = G.cpu().numpy()
            for k in range(3):
                assert np.array_equal(G[k], ref[k]), (exact, k, rel_l2(G[k], ref[k]))
        finally:
            plan.close()


def test_surface_band_option_checks(torch):
    """An explicit band needs the increment form and FP32 checkpoints, and must end above the
    bottom C-PML; the default switches itself off where it cannot run.  Out-of-range work-item tilings are refused."""
    w = small_workload(nshots=1, nt=20)
    tab, src, _, _ = oracle_setup(w, "f32i")
    for kw in ({"surface_rows": 24, "ckpt_format": 1}, {"surface_rows": 24, "arith": 0}, {"surface_rows": 4},
               {"surface_rows": w.nz - w.cpml - 3}):
        with pytest.raises(RuntimeError, match="surface_rows"):
            _plan(w, tab, src, **kw)
    for kw, rows in (({}, 24), ({"step_pairs": 1}, 24), ({"ckpt_format": 2}, 0), ({"arith": 1}, 0)):
        plan = _plan(w, tab, src, **kw)
        assert plan.surface_rows == rows, kw
        plan.close()
    for kw in ({"item_order": 5}, {"adj_item_order": 4}, {"adj_rows_per_item": -2}, {"step_pairs": -1},
               {"adj_step_pairs": -1}):
        with pytest.raises(RuntimeError, match="item_order"):
            _plan(w, tab, src, **kw)


def test_checkpoint_interval_and_batching_invariance(torch):
    """Recomputation from checkpoints replays the same kernels: every interval K (including storing all
    steps) gives bit-identical gradients, in the two-level and the increment form; batch splits and group
    sizes that form the same gradient groups give bit-identical gradients, and the misfit (per-shot
    reductions added in shot order) is identical for every split."""
    from helpers import gradient_groups

    w = small_workload(nshots=3, nt=200)
    say("checkpoint interval / batch split / gradient group invariance small nt=200")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32")
    obs = torch.from_numpy(_cat(oracle_gathers(w, "f32", sq_true, tab, src))).cuda()
    model = _dev_model(torch, sq_init)
    ref, phis = {}, {}
    # (K, max_batch, grad_group, store_segments); store_segments 1 recomputes every segment but the last,
    # 0 keeps as many as memory allows (here all of them)
    cases = [(0, 0, 0, 1), (1, 0, 0, 1), (7, 0, 0, 1), (200, 0, 0, 1), (13, 1, 0, 1), (0, 2, 0, 1), (0, 0, 1, 1),
             (7, 0, 2, 1), (0, 2, 1, 1), (7, 0, 0, 0), (7, 0, 0, 3), (13, 1, 0, 2)]
    for K, mb, grp, keep in cases:
        for exact in (1, 0):
            for arith in (3, 0):
                plan = _plan(w, tab, src, ckpt_interval=K, max_batch=mb, grad_mode=exact, arith=arith,
                             grad_group=grp, store_segments=keep)
                if keep == 1 and K in (7, 13):
                    assert plan.kept_steps < w.nt // 2
                assert plan.batch_size == (mb or 3) and plan.batch_count == -(-3 // (mb or 3))
                phi, G, _ = plan.gradient(model, obs)
                torch.cuda.synchronize()
                G = G.cpu().numpy()
                sm = plan.shot_misfits()
                assert phi == sum(sm.tolist())  # per-shot misfits added in shot order
                groups = tuple(map(tuple, gradient_groups(w.shots, (grp or 4) if exact else 1, mb)))
                key = (exact, arith, groups)
                phis.setdefault((exact, arith), phi)
                assert phi == phis[(exact, arith)], (K, mb, grp, keep, exact, arith)
                if key not in ref:
                    ref[key] = G
                else:
                    assert np.array_equal(G, ref[key]), (K, mb, grp, keep, exact, arith)
    # the exact mode met each of the three group structures at least twice
    assert sum(1 for k in ref if k[0] == 1 and k[1] == 3) == 3


@pytest.mark.parametrize("layout", ["six_full", "mixed_windows", "per_shot"])
def test_gradient_groups_match_oracle(torch, layout):
    """Exact gradient with gradient groups: six full-width shots form groups of 4 and 2 (two placeholder
    waves); with one narrower window among them the groups span the union of their windows; grad_group=1
    keeps one accumulator per shot.  Each against the oracle with the same groups (bit-level bar)."""
    from helpers import gradient_groups

    w = small_workload(nshots=6, nt=160)
    grp = 1 if layout == "per_shot" else 4
    if layout == "mixed_windows":
        s = w.shots[3]
        w.shots[3] = M.Shot(x0=8, nx=w.nx - 8, isx=s.isx - 8, isz=s.isz, irz=s.irz, ir0=12, nrec=w.nx - 8 - 24)
    say(f"gradient groups {layout}: {gradient_groups(w.shots, grp)}")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = oracle_gathers(w, "f32i", sq_true, tab, src)
    phi_ref, G_ref, _ = oracle_gradient(w, "f32i", True, sq_init, tab, src, obs, group=grp)
    plan = _plan(w, tab, src, grad_group=grp)
    phi, G, _ = plan.gradient(_dev_model(torch, sq_init), torch.from_numpy(_cat(obs)).cuda())
    torch.cuda.synchronize()
    print("adjoint tiling", plan.adj_tiling)
    assert plan.adj_tiling[2] == (layout != "per_shot")
    G = G.cpu().numpy()
    assert abs(phi - phi_ref) <= 1e-9 * phi_ref
    for k in range(3):
        e = rel_l2(G[k], G_ref[k])
        print(f"groups {layout} grad[{k}] rel-L2 {e:.3e}")
        assert np.array_equal(G[k], G_ref[k])


def test_ragged_window_groups_match_oracle(torch):
    """Eight ragged windows: the plan's gradient groups follow the placeholder-minimising rule (here not runs of
    four in window order), the same groups as the oracle's restatement of the rule, so the grouped gradient equals
    the oracle's bit for bit."""
    from helpers import _placeholder_waves, gradient_groups

    w = small_workload(nx=200, nz=72, nt=100, cpml=12, nshots=8)
    wins = [(0, 200), (0, 80), (4, 196), (60, 140), (0, 90), (100, 100), (10, 150), (120, 80)]
    w.shots = [M.Shot(x0=x0, nx=wd, isx=wd // 2, isz=6, irz=6, ir0=12, nrec=wd - 24) for x0, wd in wins]
    groups = gradient_groups(w.shots, 4)
    order = sorted(range(8), key=lambda i: (w.shots[i].x0, w.shots[i].nx, i))
    assert groups != [order[:4], order[4:]]
    say(f"ragged groups {groups}, placeholder waves {[_placeholder_waves(w.shots, g) for g in groups]}")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = oracle_gathers(w, "f32i", sq_true, tab, src)
    phi_ref, G_ref, _ = oracle_gradient(w, "f32i", True, sq_init, tab, src, obs)
    plan = _plan(w, tab, src)
    try:
        phi, G, _ = plan.gradient(_dev_model(torch, sq_init), torch.from_numpy(_cat(obs)).cuda())
        torch.cuda.synchronize()
        assert plan.adj_tiling[2]
        G = G.cpu().numpy()
    finally:
        plan.close()
    assert abs(phi - phi_ref) <= 1e-9 * phi_ref
    for k in range(3):
        assert np.array_equal(G[k], G_ref[k]), k


@pytest.mark.parametrize("fmt", [1, 2])
def test_16bit_checkpoints_match_oracle(torch, fmt):
    """BF16 (fmt 1) and scaled FP16 (fmt 2) checkpoints every 16 steps: the recomputed segments start from
    the rounded states exactly as the oracle's emulation has them (bit-level bar), with a narrower window
    in the batch so the strip packing covers two window widths; the FP32-checkpoint gradient differs by the
    rounding's effect, which is printed."""
    w = small_workload(nshots=3, nt=160)
    s = w.shots[2]
    w.shots[2] = M.Shot(x0=8, nx=w.nx - 8, isx=s.isx - 8, isz=s.isz, irz=s.irz, ir0=12, nrec=w.nx - 8 - 24)
    say(f"16-bit checkpoints fmt={fmt} K=16 nt=160, 3 shots")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = oracle_gathers(w, "f32i", sq_true, tab, src)
    phi_ref, G_ref, _ = oracle_gradient(w, "f32i", True, sq_init, tab, src, obs, ckpt=(16, fmt))
    obs_d = torch.from_numpy(_cat(obs)).cuda()
    model = _dev_model(torch, sq_init)
    plan = _plan(w, tab, src, ckpt_interval=16, ckpt_format=fmt)
    assert plan.kept_steps == 16
    phi, G, _ = plan.gradient(model, obs_d)
    plan32 = _plan(w, tab, src, ckpt_interval=16, store_segments=1)
    _, G32, _ = plan32.gradient(model, obs_d)
    torch.cuda.synchronize()
    G, G32 = G.cpu().numpy(), G32.cpu().numpy()
    assert plan.workspace_bytes < plan32.workspace_bytes
    assert abs(phi - phi_ref) <= 1e-9 * phi_ref
    for k in range(3):
        e = rel_l2(G[k], G_ref[k])
        print(f"ckpt fmt {fmt} grad[{k}] vs oracle rel-L2 {e:.3e}, vs FP32 checkpoints {rel_l2(G[k], G32[k]):.3e}")
        assert np.array_equal(G[k], G_ref[k])


def test_ragged_windows_match_oracle(torch):
    """Shots with their own windowed subdomains and receiver spreads in one launch."""
    nx, nz, nt, cpml = 220, 72, 200, 12
    w = small_workload(nx=nx, nz=nz, nt=nt, cpml=cpml, nshots=2)
    w.shots = [M.Shot(x0=0, nx=120, isx=60, isz=6, irz=6, ir0=20, nrec=70),
               M.Shot(x0=50, nx=170, isx=90, isz=7, irz=6, ir0=cpml, nrec=170 - 2 * cpml),
               M.Shot(x0=140, nx=80, isx=30, isz=6, irz=8, ir0=30, nrec=20)]
    say("ragged windows 3 shots widths 120/170/80")
    for exact in (True, False):
        tab, src, sq_true, sq_init = oracle_setup(w, "f32")
        obs = oracle_gathers(w, "f32", sq_true, tab, src)
        phi_ref, G_ref, _ = oracle_gradient(w, "f32i", exact, sq_init, tab, src, obs)
        plan = _plan(w, tab, src, grad_mode=1 if exact else 0)  # default arithmetic: FP32 increment form
        phi, G, _ = plan.gradient(_dev_model(torch, sq_init), torch.from_numpy(_cat(obs)).cuda())
        torch.cuda.synchronize()
        G = G.cpu().numpy()
        assert abs(phi - phi_ref) <= 1e-9 * phi_ref
        for k in range(3):
            e = rel_l2(G[k], G_ref[k])
            print(f"ragged exact={exact} grad[{k}] rel-L2 {e:.3e}")
            assert np.array_equal(G[k], G_ref[k])


def test_misfit_matches_gradient_misfit(torch):
    w = small_workload(nshots=2, nt=160)
    say("misfit-only path vs gradient misfit")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32")
    obs = torch.from_numpy(_cat(oracle_gathers(w, "f32", sq_true, tab, src))).cuda()
    plan = _plan(w, tab, src)
    model = _dev_model(torch, sq_init)
    phi_m = plan.misfit(model, obs)
    phi_g, _, _ = plan.gradient(model, obs)
    assert phi_m == phi_g


@pytest.mark.parametrize("arith,bound,fine", [(0, 2e-3, False), (2, 2e-3, False), (3, 3e-4, False),
                                              (0, 2e-3, True), (3, 1e-4, True)])
def test_fp32_error_against_fp64_oracle(torch, arith, bound, fine):
    """FP32 batched gradient vs the FP64 oracle (exact mode).  The two-level leapfrog's round-off
    (amplified by the time recursion) reaches the residual syn - obs, where the direct wave cancels, so
    its gradient error is a few 1e-4; the increment form (arith 3, the default) is 5-20x closer.  `fine`:
    8 Hz at dt = 0.8 ms (156 samples per period, the survey's sampling), 1000 steps, 1 shot."""
    w = small_workload(nshots=1, nt=1000, f=8.0, dt=0.8e-3) if fine else small_workload(nshots=2)
    say(f"fp32 gradient vs fp64 oracle {'fine' if fine else 'small'}, arith={arith}")
    tab64, src, sq_true64, sq_init64 = oracle_setup(w, "f64")
    obs64 = oracle_gathers(w, "f64", sq_true64, tab64, src.astype(np.float64))
    phi64, G64, _ = oracle_gradient(w, "f64", True, sq_init64, tab64, src.astype(np.float64), obs64)
    tab, src32, _, sq_init = oracle_setup(w, "f32")
    plan = _plan(w, tab, src32, arith=arith)
    phi, G, _ = plan.gradient(_dev_model(torch, sq_init), torch.from_numpy(_cat(obs64)).cuda())
    torch.cuda.synchronize()
    G = G.cpu().numpy()
    errs = [rel_l2(G[k], G64[k]) for k in range(3)]
    print(f"arith {arith} vs FP64: misfit rel {abs(phi - phi64) / phi64:.3e}, grad rel-L2 {errs}")
    assert max(errs) < bound


def test_adjoint128_checkpoint28_fp32(torch):
    """The adjoint-check configuration on the GPU (128x128, 2 shots, nt = 800, 12 Hz, FP32 increment form,
    checkpoints every 28 steps + C-PML strips): the gradient equals the oracle's at the bit level, its error
    against the FP64 oracle is stated and bounded, and it agrees with central differences of the GPU's own
    FP32 misfit along a smooth perturbation of (vp, eps, delta)."""
    from fwi_amd.survey import chain_rule, pre_variables

    w = M.workload("adjoint128")
    say("adjoint128: K=28 FP32 vs oracle f32i, FP64 and FD")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = oracle_gathers(w, "f32i", sq_true, tab, src)
    phi_ref, G_ref, _ = oracle_gradient(w, "f32i", True, sq_init, tab, src, obs)
    plan = _plan(w, tab, src, ckpt_interval=28, store_segments=1)
    assert plan.ckpt_interval == 28 and plan.kept_steps == 800 - 28 * 28
    obs_d = torch.from_numpy(_cat(obs)).cuda()
    phi, G, _ = plan.gradient(_dev_model(torch, sq_init), obs_d)
    torch.cuda.synchronize()
    Gn = G.cpu().numpy()
    for k in range(3):
        assert np.array_equal(Gn[k], G_ref[k])
    # FP64 oracle on the same observed data
    tab64, src64, _, si64 = oracle_setup(w, "f64")
    _, G64, _ = oracle_gradient(w, "f64", True, si64, tab64, src64.astype(np.float64), [o.astype(np.float64) for o in obs])
    e64 = [rel_l2(Gn[k], G64[k]) for k in range(3)]
    print(f"adjoint128 FP32 (K=28) vs FP64 gradient rel-L2 (v_hor, v_po, v_nmo): {e64}")
    assert max(e64) < 3e-4
    # central differences of the FP32 misfit, h = 1e-1 .. 1e-3 of a perturbation of unit size
    m0 = [torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda() for x in w.init_model]
    dm = M.smooth_perturbation(w.nx, w.nz, seed=11, sigma=6.0)
    d = [torch.from_numpy(np.ascontiguousarray(x * s, np.float32)).cuda() for x, s in zip(dm, (50.0, 0.01, 0.01))]
    gv = chain_rule(*m0, G)
    pred = sum(torch.sum(g * dd.double()).item() for g, dd in zip(gv, d))
    errs = []
    for h in (1e-1, 3e-2, 1e-2):
        Jp = plan.misfit(pre_variables(*[a + h * b for a, b in zip(m0, d)]), obs_d)
        Jm = plan.misfit(pre_variables(*[a - h * b for a, b in zip(m0, d)]), obs_d)
        errs.append(abs((Jp - Jm) / (2 * h) - pred) / abs(pred))
    print(f"adjoint128 FP32 FD relative errors (h = 1e-1, 3e-2, 1e-2): {errs}")
    assert min(errs) < 2e-3


def _pair_workloads():
    """Ragged windows for the step-pair checks: a core several tiles wide, one 32 columns wide, one window exactly
    2 cpml + 8 wide (no core), a source and a receiver row inside the core; nz chosen for cores 28, 1 and 4 rows tall
    under the FP64 band (thin cores make the parts of the shell meet)."""
    out = []
    for nz, cpml in ((72, 12), (45, 12), (41, 5)):
        w = small_workload(nx=420, nz=nz, nt=150, cpml=cpml, nshots=4)
        zc = min(nz - cpml - 5, 30)  # a row inside the core without the band
        w.shots = [M.Shot(x0=0, nx=300, isx=60, isz=6, irz=6, ir0=20, nrec=250),
                   M.Shot(x0=30, nx=170, isx=90, isz=zc, irz=zc - 2, ir0=12, nrec=170 - 24),
                   M.Shot(x0=150, nx=2 * cpml + 8 + 32, isx=20, isz=6, irz=8, ir0=14, nrec=16),
                   M.Shot(x0=420 - 2 * cpml - 8, nx=2 * cpml + 8, isx=cpml + 4, isz=zc, irz=6, ir0=6,
                          nrec=2 * cpml - 4)]
        out.append(w)
    return out


@pytest.mark.parametrize("geometry", [0, 1, 2])
def test_step_pairs_identical_to_single_steps(torch, geometry):
    """Step pairs (the core advanced two steps per launch, the default) against single steps everywhere: gathers,
    misfit and exact and reference-compatible gradients bit for bit, with and without the FP64 near-surface band, an
    odd checkpoint interval (pairs stop at checkpoints and at the first kept step), recomputed segments of odd
    length, ragged windows, a source and a receiver row inside the core, core tiles cut at several row-run lengths
    and cores 1 and 4 rows tall."""
    w = _pair_workloads()[geometry]
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = torch.from_numpy(_cat(oracle_gathers(w, "f32i", sq_true, tab, src))).cuda()
    model = _dev_model(torch, sq_init)
    for surface, rows in ((0, 16), (-1, 24), (0, 9)):
        for exact in (1, 0):
            say(f"step pairs vs single steps: nz={w.nz} cpml={w.cpml}, 4 ragged shots, K=7, 1 kept segment, "
                f"surface_rows={surface}, pair rows {rows}, exact={exact}")
            res = []
            for pairs in (rows, 0):
                plan = _plan(w, tab, src, ckpt_interval=7, store_segments=1, grad_mode=exact, step_pairs=pairs,
                             surface_rows=surface, rows_per_item=8)
                assert (plan.pair_rows > 0) == bool(pairs)
                assert (plan.surface_rows > 0) == (surface == 0 and w.nz - w.cpml - 4 >= 24)
                g = plan.forward(model)
                phi, G, syn = plan.gradient(model, obs, want_syn=True)
                torch.cuda.synchronize()
                res.append((g.cpu().numpy(), phi, G.cpu().numpy(), syn.cpu().numpy()))
                plan.close()
            (g1, p1, G1, s1), (g0, p0, G0, s0) = res
            assert np.array_equal(g1, g0) and np.array_equal(s1, s0)
            assert p1 == p0
            assert np.array_equal(G1, G0), exact


def _adj_pair_workloads():
    """Geometries for the adjoint step pairs, whose core is each gradient group's common core (columns
    [cpml+8, nx-cpml-8) of every member's window, rows [max(4, band)+4, nz-cpml-8)): four ragged windows in one group
    with overlapping cores (placeholder waves where a member's window misses a band of the union) and receiver rows
    inside the core; five full-width shots (a group of one: three placeholder waves in its core tiles); a core four
    rows tall without the FP64 band."""
    out = []
    w = small_workload(nx=320, nz=72, nt=120, cpml=12, nshots=4)
    w.shots = [M.Shot(x0=0, nx=300, isx=60, isz=6, irz=6, ir0=20, nrec=250),
               M.Shot(x0=20, nx=280, isx=90, isz=30, irz=30, ir0=12, nrec=250),
               M.Shot(x0=10, nx=250, isx=40, isz=10, irz=10, ir0=14, nrec=200),
               M.Shot(x0=40, nx=200, isx=100, isz=6, irz=40, ir0=20, nrec=150)]
    out.append(w)
    w = small_workload(nx=200, nz=60, nt=110, cpml=10, nshots=5)
    for k, s in enumerate(w.shots):
        w.shots[k] = M.Shot(x0=0, nx=200, isx=s.isx, isz=s.isz, irz=(6, 20, 9, 6, 30)[k], ir0=10, nrec=180)
    out.append(w)
    out.append(small_workload(nx=240, nz=28, nt=100, cpml=8, nshots=4))
    return out


@pytest.mark.parametrize("geometry", [0, 1, 2])
def test_adjoint_step_pairs_identical_to_single_steps(torch, geometry):
    """Adjoint step pairs (each gradient group's core advanced two adjoint steps per launch, the gradient slot updated
    once per pair) against single adjoint steps: misfit and exact gradient bit for bit, with and without the FP64
    near-surface band, an odd checkpoint interval (pairs stop where step k-1's imaging terms lie in another
    recomputed segment), core tiles of several row-run lengths; and against the oracle."""
    w = _adj_pair_workloads()[geometry]
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs_np = oracle_gathers(w, "f32i", sq_true, tab, src)
    obs = torch.from_numpy(_cat(obs_np)).cuda()
    model = _dev_model(torch, sq_init)
    for surface, rows in ((0, 16), (-1, 9), (-1, 32)):
        say(f"adjoint step pairs vs single steps: {w.nx}x{w.nz} cpml={w.cpml}, {len(w.shots)} shots, K=7, "
            f"1 kept segment, surface_rows={surface}, pair rows {rows}")
        res = []
        for apairs in (rows, 0):
            plan = _plan(w, tab, src, ckpt_interval=7, store_segments=1, adj_step_pairs=apairs,
                         surface_rows=surface, rows_per_item=8)
            band = plan.surface_rows
            core = w.nz - w.cpml - 8 > max(4, band) + 4
            assert (plan.adj_pair_rows > 0) == (bool(apairs) and core), (plan.adj_pair_rows, band)
            phi, G, syn = plan.gradient(model, obs, want_syn=True)
            torch.cuda.synchronize()
            res.append((phi, G.cpu().numpy(), syn.cpu().numpy()))
            plan.close()
        (p1, G1, s1), (p0, G0, s0) = res
        assert p1 == p0 and np.array_equal(s1, s0)
        assert np.array_equal(G1, G0), (surface, rows)
        if surface == 0:
            phi_ref, G_ref, _ = oracle_gradient(w, "f32i", True, sq_init, tab, src, obs_np)
            for k in range(3):
                assert np.array_equal(G1[k], G_ref[k]), k


def test_entry_points_refuse_null_and_pageable_arrays(torch):
    """With a live plan, a null array and a pageable host array (which a kernel would fault on) are refused with
    a message naming the argument, before any kernel runs; the plan works normally afterwards (at the true
    model the misfit and gradient are exactly zero)."""
    import ctypes

    from fwi_amd import _lib

    w = small_workload(nshots=1, nt=40)
    say("argument checks: null and pageable host arrays refused before any kernel")
    tab, src, sq_true, _ = oracle_setup(w, "f32i")
    plan = _plan(w, tab, src)
    try:
        L = _lib.load()
        msg = lambda: L.fwi_error_message().decode()
        model = _dev_model(torch, sq_true)
        ptrs = [t.data_ptr() for t in model]
        g = torch.empty(plan.gather_size, dtype=torch.float32, device="cuda")
        host_f = np.zeros(1 << 20, np.float32)  # 4 MiB of pageable memory, classified as unregistered
        host_d = np.zeros(1 << 20, np.float64)
        assert L.fwi_forward(plan._h, ptrs[0], ptrs[1], ptrs[2], None, None) == 1
        assert msg() == "fwi_forward: gathers is null"
        assert L.fwi_forward(plan._h, host_f.ctypes.data, ptrs[1], ptrs[2], g.data_ptr(), None) == 1
        assert msg().startswith("fwi_forward: v_hor is host memory the GPU cannot address"), msg()
        obs = plan.forward(model)
        m = ctypes.c_double()
        assert L.fwi_gradient(plan._h, *ptrs, obs.data_ptr(), host_d.ctypes.data, ctypes.byref(m), None, None) == 1
        assert msg().startswith("fwi_gradient: grad is host memory"), msg()
        phi, G, _ = plan.gradient(model, obs)
        torch.cuda.synchronize()
        assert phi == 0.0 and not G.any().item()
    finally:
        plan.close()


def test_plan_retargets_to_another_stage_frequency(torch):
    """fwi_plan_set_wavelet + fwi_plan_set_cpml move a plan to another stage's frequency (a multiscale
    inversion's next stage): its gathers then equal those of a plan created at that frequency, bit for bit."""
    import dataclasses

    w = small_workload(nshots=2, nt=120, f=10.0)
    say("plan retargeted from 10 Hz to 15 Hz: wavelet and C-PML table replaced")
    tab10, src10, sq_true, _ = oracle_setup(w, "f32i")
    p15 = dataclasses.replace(w.cpml_params, f=15.0)
    tab15, src15 = O.cpml_table(p15), O.ricker(w.nt, 15.0, w.dt)
    assert not np.array_equal(tab15[0], tab10[0])
    model = _dev_model(torch, sq_true)
    moved = _plan(w, tab10, src10)
    fresh = _plan(w, tab15, src15)
    try:
        g10 = moved.forward(model).clone()
        moved.set_wavelet(src15)
        moved.set_cpml(tab15)
        g_moved = moved.forward(model)
        g_fresh = fresh.forward(model)
        torch.cuda.synchronize()
        assert not torch.equal(g10, g_fresh)
        assert torch.equal(g_moved, g_fresh)
        # the same on a non-blocking side stream, with the retarget issued while a forward is still in flight
        # there: the tables are replaced only after it has finished
        moved.set_wavelet(src10)
        moved.set_cpml(tab10)
        side = torch.cuda.Stream()
        with torch.cuda.stream(side):
            moved.forward(model)
            g10s = moved.forward(model).clone()
            moved.set_wavelet(src15)
            moved.set_cpml(tab15)
            g15s = moved.forward(model).clone()
        side.synchronize()
        assert torch.equal(g10s, g10)
        assert torch.equal(g15s, g_fresh)
    finally:
        moved.close()
        fresh.close()
