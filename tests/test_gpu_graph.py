"""GPU: captured HIP graphs (use_graph=1, the default) against eager launches (use_graph=0).

Each entry point's device work per batch is captured on the first call with a set of buffers and replayed
on later calls.  The test replays several times with the program's own eager work and synchronisations in
between, changes the model in place between replays, and checks every replay bit for bit against the eager
plan.  Two batches with checkpoints exercise the cross-batch accumulation and the recomputation inside the
graph."""
import numpy as np
import pytest

from helpers import oracle_gathers, oracle_setup, say, small_workload

pytestmark = pytest.mark.gpu


def test_graph_replays_match_eager():
    import torch

    from fwi_amd.survey import Plan

    w = small_workload(nshots=3, nt=160)
    say("graph replays vs eager: 3 shots in 2 batches, K=16, gradient/misfit/forward/gradient_src")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = torch.from_numpy(np.concatenate([o.reshape(-1) for o in oracle_gathers(w, "f32i", sq_true, tab, src)])).cuda()
    mk = lambda g: Plan(w.nx, w.nz, w.nt, w.dh, w.dt, w.cpml, w.shots, src, tab, ckpt_interval=16, max_batch=2,
                        use_graph=g, store_segments=1)
    pg, pe = mk(1), mk(0)
    model = [torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda() for x in sq_init]
    out = torch.empty((3, w.nz, w.nx), dtype=torch.float64, device="cuda")
    fo = torch.empty(pg.gather_size, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(0)
    adj = torch.linspace(-1, 1, pg.gather_size, device="cuda")
    for rep in range(4):
        phi_g, g_g, _ = pg.gradient(model, obs, out=out)
        g_g = g_g.clone()
        torch.cuda.synchronize()
        phi_e, g_e, _ = pe.gradient(model, obs)
        torch.cuda.synchronize()
        assert phi_g == phi_e, rep
        assert torch.equal(g_g, g_e), rep
        assert pg.misfit(model, obs) == pe.misfit(model, obs)
        assert torch.equal(pg.forward(model, out=fo), pe.forward(model))
        assert torch.equal(pg.gradient_src(model, adj)[0], pe.gradient_src(model, adj)[0])
        # eager work of the program between replays, then the model changes in place (same addresses)
        torch.cuda.synchronize()
        with torch.no_grad():
            for t in model:
                t.mul_(1.0 + 1e-3 * float(rng.standard_normal()))
        torch.cuda.synchronize()
    captures, replays = pg.graph_stats()
    print("graph captures, replays:", captures, replays)
    # 2 batches x 4 entry points captured on the first pass, replayed on the three others
    assert captures == 8 and replays == 3 * 8
    assert pe.graph_stats() == (0, 0)


def test_graph_cache_eviction():
    """More buffer sets than the plan keeps graphs for (8): forwards into 11 output buffers, twice over, evict
    and recapture least recently used graphs, and every result still equals the eager plan's."""
    import torch

    from fwi_amd.survey import Plan

    w = small_workload(nshots=2, nt=60)
    say("graph cache eviction: 11 output buffers, 2 rounds")
    tab, src, sq_true, _ = oracle_setup(w, "f32i")
    mk = lambda g: Plan(w.nx, w.nz, w.nt, w.dh, w.dt, w.cpml, w.shots, src, tab, use_graph=g)
    pg, pe = mk(1), mk(0)
    model = [torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda() for x in sq_true]
    ref = pe.forward(model)
    outs = [torch.empty(pg.gather_size, dtype=torch.float32, device="cuda") for _ in range(11)]
    for rnd in range(2):
        for k, o in enumerate(outs):
            o.fill_(float("nan"))
            pg.forward(model, out=o)
        torch.cuda.synchronize()
        for k, o in enumerate(outs):
            assert torch.equal(o, ref), (rnd, k)
    captures, replays = pg.graph_stats()
    print("graph captures, replays:", captures, replays)
    # one batch: 11 distinct keys per round and a cache of 8, visited in order, so every call recaptures
    assert captures == 22 and replays == 0


def test_long_capture_is_cut_into_bounded_graphs():
    """A gradient of 3000 steps (3000 forward, ~2900 recomputed, 3000 adjoint steps) is captured as consecutive
    graphs of at most 4096 steps each, replayed in order: three replays with the model changed in place in
    between equal the eager plan bit for bit, and no graph holds more than 4096 steps' kernels."""
    import torch

    from fwi_amd.survey import Plan

    w = small_workload(nshots=2, nt=3000, nx=64, nz=48, cpml=10)
    say("long capture: 3000 steps, graphs of at most 4096 steps, 3 replays vs eager")
    tab, src, sq_true, sq_init = oracle_setup(w, "f32i")
    obs = torch.from_numpy(np.concatenate([o.reshape(-1) for o in oracle_gathers(w, "f32i", sq_true, tab, src)])).cuda()
    mk = lambda g: Plan(w.nx, w.nz, w.nt, w.dh, w.dt, w.cpml, w.shots, src, tab, use_graph=g, store_segments=1)
    pg, pe = mk(1), mk(0)
    model = [torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda() for x in sq_init]
    for rep in range(3):
        phi_g, g_g, _ = pg.gradient(model, obs)
        g_g = g_g.clone()
        torch.cuda.synchronize()
        phi_e, g_e, _ = pe.gradient(model, obs)
        torch.cuda.synchronize()
        assert phi_g == phi_e and torch.equal(g_g, g_e), rep
        with torch.no_grad():
            for t in model:
                t.mul_(1.0 + 1e-3 * (rep + 1))
        torch.cuda.synchronize()
    largest, total = pg.graph_nodes()
    steps = 3000 + (3000 - pg.kept_steps) + 3000
    print(f"graph nodes: largest {largest}, capture total {total} for {steps} steps")
    assert pg.graph_stats() == (1, 2)
    assert total > largest  # the capture was cut
    assert largest <= 4 * 4096 + 64  # 2-4 kernels per step, plus the checkpoint copies of a graph
