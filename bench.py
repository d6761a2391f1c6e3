#!/usr/bin/env python3
"""Benchmark: one full FWI gradient of the 1600x400, 64-shot survey per step on one MI355X.

A "step" is one gradient (all shots: forward with checkpoints, residual, FP64 misfit, adjoint with
recomputation, imaging, FP64 cross-shot sum) of the batched path.  value = Gcell-updates/s (fwd+adj) =
2 * sum_shots(nx*nz) * nt / seconds per gradient; recomputation is not counted as work.

Beside it, on the same workload:
  * plain per-shot GPU path: the reference's per-shot flow on the legacy C ABI (launchPropagation with a
    pinned-host snapshot every step, launchBackPropagation), timed on one shot for a bounded number of
    steps and scaled to the full survey (it is linear in shots and steps);
  * roofline of the dominant kernel (the adjoint step): algorithmic bytes per launch over its average
    duration, measured with HIP events on the stream the kernel runs on during the timed region;
  * cpu_baseline: the CPU oracle (f32i, exact gradient, OpenMP) on one shot for a bounded number of steps.
  * the device optimizer of one iteration (steepest-descent step, L-BFGS direction, trial), timed apart.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="survey1600")
    ap.add_argument("--shots", type=int, default=0, help="shots (0 = the workload's own count)")
    ap.add_argument("--nt", type=int, default=0)
    ap.add_argument("--mode", choices=["exact", "refcompat"], default="exact")
    ap.add_argument("--ckpt", type=int, default=0)
    ap.add_argument("--ckpt-format", type=int, default=0, help="checkpoints: 0 FP32, 1 BF16, 2 FP16 (scaled)")
    ap.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying captured graphs")
    ap.add_argument("--keep", type=int, default=0, help="trailing segments kept from the forward (0 = memory allows)")
    ap.add_argument("--grad-group", type=int, default=0, help="shots per gradient group (0 = 4)")
    ap.add_argument("--pairs", action="store_true", help="fused core step pairs (increment form)")
    ap.add_argument("--arith", type=int, default=3,
                    help="0 FP32 two-level, 1 mixed (reference), 2 FP32 difference stencil, 3 FP32 increment form")
    ap.add_argument("--rows", type=int, default=0, help="rows per work item (0 = tuned)")
    ap.add_argument("--order", type=int, default=-1, help="work-item order (-1 = tuned)")
    ap.add_argument("--plain-steps", type=int, default=400)
    ap.add_argument("--cpu-steps", type=int, default=4000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-stride", type=int, default=4)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device-to-device (a kernel of known traffic for PMC scaling)")
    return ap.parse_args()


def plain_path_seconds(w, model_sq_dev, src_dev, ab, nsteps):
    """Reference flow on the legacy ABI for one shot and `nsteps` steps: forward with a pinned snapshot
    per step (launchPropagation), residual, adjoint with the snapshot copy back (launchBackPropagation)."""
    import numpy as np
    import torch

    from fwi_amd import _lib

    L = _lib.load()
    nx, nz, cpml = w.nx, w.nz, w.cpml
    n = nx * nz
    nt = nsteps
    shot = w.shots[len(w.shots) // 2]
    z = lambda k: torch.zeros(k, dtype=torch.float32, device="cuda")
    names = ("p2", "r2", "psx", "psz", "zx", "zz", "atp", "atf", "PF", "aux1", "aux2", "aux3", "aux4", "p2cx", "p2cz",
             "r2dx", "r2cz", "psixcx", "psixcz", "atenxcx", "atenxcz", "atenzdx", "atenzcz", "gcx", "gcz", "gdx")
    f = {k: z(n) for k in names}
    ax, bx, az, bz = z(nx), z(nx), z(nz), z(nz)
    p = w.cpml_params
    L.CPML_x_h(ax.data_ptr(), bx.data_ptr(), cpml, p.vmax, nx, p.dt, p.dh, p.f, p.d0, p.L)
    L.CPML_z_h(az.data_ptr(), bz.data_ptr(), cpml, p.vmax, nz, p.dt, p.dh, p.f, p.d0, p.L)
    vh, vp, vn = model_sq_dev
    Pa, Pb, Pre2 = z(2 * n), z(2 * n), z(2 * n)
    PFh = torch.empty(2 * nt * n, dtype=torch.float32, pin_memory=True)
    shot1 = z(nx * nt)
    shot2 = z(nx * nt)
    Res = z(nx * nt)
    dh = np.float32(w.dh)
    dh2 = dh * dh
    dt = np.float32(w.dt)
    C = np.float32(dt * dt) / dh2
    s1 = _lib.create_blocking_stream()
    s2 = _lib.create_blocking_stream()
    ptr = lambda k: f[k].data_ptr()
    sx, sz = shot.isx + 1, shot.isz + 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Ppre, Ppast = Pa, Pb
    for it in range(nt):
        L.launchPropagation(s1, s2, PFh.data_ptr() + 8 * it * n, ptr("PF"), Ppast.data_ptr(), Ppre.data_ptr(),
                            Ppast.data_ptr() + 4 * n, Ppre.data_ptr() + 4 * n, ptr("p2"), ptr("r2"), dh2, nx, nz,
                            ax.data_ptr(), bx.data_ptr(), az.data_ptr(), bz.data_ptr(), ptr("psx"), ptr("psz"), dh,
                            cpml, ptr("zx"), ptr("zz"), ptr("atp"), ptr("atf"), vh.data_ptr(), vp.data_ptr(),
                            vn.data_ptr(), src_dev.data_ptr(), shot2.data_ptr(), C, sx, sz, it)
        Ppre, Ppast = Ppast, Ppre
    L.res_h(Res.data_ptr(), shot1.data_ptr(), shot2.data_ptr(), nx, nt)
    torch.cuda.synchronize()
    Pa.zero_()
    Pb.zero_()
    Ppre, Ppast = Pa, Pb
    for it in range(nt):
        L.launchBackPropagation(s1, s2, nx, nz, nt, PFh.data_ptr() + 8 * (nt - it - 1) * n, Ppre.data_ptr(),
                                Ppast.data_ptr(), Ppre.data_ptr() + 4 * n, Ppast.data_ptr() + 4 * n, ptr("aux1"),
                                ptr("aux2"), ptr("aux3"), ptr("aux4"), vh.data_ptr(), vp.data_ptr(), vn.data_ptr(),
                                ptr("p2cx"), ptr("p2cz"), ptr("r2dx"), ptr("r2cz"), dh2, Pre2.data_ptr(),
                                Pre2.data_ptr() + 4 * n, ax.data_ptr(), bx.data_ptr(), az.data_ptr(), bz.data_ptr(),
                                ptr("psixcx"), ptr("psixcz"), ptr("psx"), ptr("psz"), dh, cpml, ptr("zx"), ptr("zz"),
                                ptr("atf"), ptr("atp"), ptr("atenxcx"), ptr("atenxcz"), ptr("atenzdx"),
                                ptr("atenzcz"), Res.data_ptr(), C, ptr("PF"), it, sx, sz, dt, ptr("gcx"), ptr("gcz"),
                                ptr("gdx"))
        Ppre, Ppast = Ppast, Ppre
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    _lib.check_last("plain path")
    _lib.destroy_stream(s1)
    _lib.destroy_stream(s2)
    del PFh
    return sec


def cpu_baseline(w, nsteps):
    """Oracle (f32i arithmetic, exact gradient, OpenMP over all host threads) on one shot of the workload
    for nsteps steps: forward + adjoint + imaging.  Returns (Gcell-updates/s, threads, seconds)."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    shot = w.shots[len(w.shots) // 2]
    tab = O.cpml_table(w.cpml_params)
    src = O.ricker(nsteps, w.f, w.dt)
    sq = O.pre_variables(*[np.ascontiguousarray(x[:, shot.x0:shot.x0 + shot.nx]) for x in w.init_model])
    obs = np.zeros((nsteps, shot.nrec), np.float32)
    t0 = time.perf_counter()
    O.gradient("f32i", True, sq, tab, src, obs, shot.nx, w.nz, nsteps, w.cpml, w.dh, w.dt, shot)
    sec = time.perf_counter() - t0
    return 2.0 * shot.nx * w.nz * nsteps / sec / 1e9, threads, sec


def optimizer_ms(w, model, grad):
    """The device optimizer of one inversion iteration (src/main.c:539-800) on this model and gradient, timed
    apart from the gradient as the reference's flow separates them: a steepest-descent step, an L-BFGS
    direction from one curvature pair (FP64 two-loop recursion with device dot products) and one trial update,
    in milliseconds each.  The model is a copy; the second gradient of the pair is a scaled copy of the first
    (the timing does not depend on the values)."""
    import torch

    from fwi_amd.optim import LBFGS, SD_STEP

    m = [t.clone() for t in model]
    opt = LBFGS(w.nx, w.nz, 5)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return round((time.perf_counter() - t0) * 1e3, 3)

    try:
        out = {"sd_step_ms": timed(lambda: opt.sd_step(m, grad, SD_STEP))}
        g2 = grad * 1.5
        out["lbfgs_direction_ms"] = timed(lambda: opt.direction(m, g2))
        out["trial_ms"] = timed(lambda: opt.trial(m, (1.0, 1.0, 1.0)))
    finally:
        opt.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported.  Shots are independent, so "
              "several GPUs would each take a share of the shots and add their FP64 gradients and misfits "
              "with one all-reduce per gradient (3*nx*nz + 1 doubles).", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from fwi_amd import models as M
    from fwi_amd.survey import Plan, cpml_table, pre_variables

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.calibrate_copy:
        a = torch.ones(256 * 2 ** 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read + 1 GiB written")
        del a, b
    w = M.workload(args.workload, nshots=args.shots or None, nt=args.nt or None)
    nt = w.nt
    log(f"workload {w.name}: {w.nx}x{w.nz}, {len(w.shots)} shots, nt={nt}, dt={w.dt}, f={w.f}")
    ab = cpml_table(w.cpml_params)
    src = M.ricker(nt, w.f, w.dt)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    true_sq = pre_variables(*(dev(x) for x in w.true_model))
    init_sq = pre_variables(*(dev(x) for x in w.init_model))
    grad_mode = 1 if args.mode == "exact" else 0
    plan = Plan(w.nx, w.nz, nt, w.dh, w.dt, w.cpml, w.shots, src, ab, grad_mode=grad_mode, ckpt_interval=args.ckpt,
                rows_per_item=args.rows, item_order=args.order, arith=args.arith,
                ckpt_format=args.ckpt_format, use_graph=0 if args.no_graph else 1, store_segments=args.keep,
                grad_group=args.grad_group, step_pairs=1 if args.pairs else 0)
    log(f"plan: batch {plan.batch_size} shots, K={plan.ckpt_interval}, {plan.kept_steps} steps kept from the forward, "
        f"workspace {plan.workspace_bytes / 1e9:.1f} GB")
    t0 = time.perf_counter()
    obs = plan.forward(true_sq)
    torch.cuda.synchronize()
    log(f"observed data modelled in {time.perf_counter() - t0:.2f} s (includes tile tuning)")
    for i in range(args.warmup):
        t0 = time.perf_counter()
        plan.gradient(init_sq, obs)
        torch.cuda.synchronize()
        log(f"warmup gradient {i}: {time.perf_counter() - t0:.3f} s")
    torch.cuda.synchronize()
    times = []
    phi = None
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        phi, grad, _ = plan.gradient(init_sq, obs)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"gradient {i}: {times[-1]:.3f} s, misfit {phi:.6e}")
    total = time.perf_counter() - t_all
    sec = total / args.steps
    ncell, nstrip = plan.cells()
    work = 2.0 * ncell * nt
    gcells = work / sec / 1e9
    assert np.isfinite(phi) and torch.isfinite(grad).all().item()
    graphs = plan.graph_stats()
    log(f"graph captures {graphs[0]}, replays {graphs[1]}")
    # kernel durations: one more gradient with every timing_stride-th launch bracketed by HIP events (the
    # instrumented call launches eagerly; its kernels are the ones the timed calls replayed)
    sec_instr = None
    if args.timing_stride > 0:
        plan.set_timing(args.timing_stride)
        t0 = time.perf_counter()
        phi_i, _, _ = plan.gradient(init_sq, obs)
        torch.cuda.synchronize()
        sec_instr = time.perf_counter() - t0
        assert phi_i == phi
        log(f"instrumented eager gradient: {sec_instr:.3f} s")

    # roofline of the adjoint step kernel (dominant): algorithmic bytes per launch / mean event time
    adj_ms, adj_n = plan.kernel_time(2)
    fwd_ms, fwd_n = plan.kernel_time(0)
    rec_ms, rec_n = plan.kernel_time(1)
    nmodel = w.nx * w.nz
    adj_bytes = 56.0 * ncell + 12.0 * nmodel
    adj_avg = adj_ms / max(adj_n, 1) / 1e3
    achieved = adj_bytes / adj_avg / 1e9 if adj_n else None
    fwd_bytes = 24.0 * ncell + 12.0 * nmodel
    fwd_avg = fwd_ms / max(fwd_n, 1) / 1e3
    if adj_n:
        log(f"adjoint step: {adj_n} timed launches, mean {adj_avg * 1e6:.1f} us -> {achieved:.0f} GB/s algorithmic")
    pair_ms, pair_n = plan.kernel_time(3)
    pairs_ms, pairs_n = plan.kernel_time(4)
    if pair_n or pairs_n:
        log(f"fused core pair: {pair_n} timed launches, mean {pair_ms / max(pair_n, 1) * 1e3:.1f} us; storing pair "
            f"{pairs_n} launches, mean {pairs_ms / max(pairs_n, 1) * 1e3:.1f} us (rows per item {plan.pair_rows})")
    if fwd_n:
        log(f"forward step: {fwd_n} timed launches, mean {fwd_avg * 1e6:.1f} us -> "
            f"{fwd_bytes / max(fwd_avg, 1e-12) / 1e9:.0f} GB/s; storing step mean "
            f"{rec_ms / max(rec_n, 1) * 1e3:.1f} us")

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "fused_steps_survey1600_hbm_traffic.json")
    kname = "k_adj_ex_main" if grad_mode == 1 else "k_adj_ref_main"
    if os.path.exists(tfile) and w.name == "survey1600" and len(w.shots) == 64:
        with open(tfile) as fh:
            traffic = json.load(fh)["bytes_per_launch"].get(kname)

    opt_ms = optimizer_ms(w, init_sq, grad)
    log(f"optimizer: SD step {opt_ms['sd_step_ms']} ms, L-BFGS direction {opt_ms['lbfgs_direction_ms']} ms, "
        f"trial {opt_ms['trial_ms']} ms")

    plain = None
    if not args.no_plain:
        ns = min(args.plain_steps, nt)
        # the snapshot buffer of the plain path is 2*ns*nx*nz floats of pinned host memory
        s_plain = plain_path_seconds(w, init_sq, torch.from_numpy(src[:ns].copy()).cuda(), ab, ns)
        plain = s_plain * (nt / ns) * len(w.shots)
        log(f"plain per-shot path: {s_plain:.3f} s for 1 shot x {ns} steps -> {plain:.1f} s per gradient")

    cpu = None
    if not args.no_cpu:
        ns = min(args.cpu_steps, nt)
        cg, threads, csec = cpu_baseline(w, ns)
        cpu = {"value": round(cg, 4), "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"oracle f32i exact gradient, 1 shot of {w.name} ({w.nx}x{w.nz}), {ns} of {nt} steps, "
                         f"fwd+adj, {csec:.1f} s"}
        log(f"cpu baseline: {cg:.3f} Gcell/s on {threads} threads ({csec:.1f} s)")

    nrecs = [sh.nrec for sh in w.shots]
    widths = [sh.nx for sh in w.shots]
    spread = (f"{nrecs[0]} receivers" if min(nrecs) == max(nrecs) else f"{min(nrecs)}-{max(nrecs)} receivers") + (
        "" if min(widths) == max(widths) == w.nx else f", windows of {min(widths)}-{max(widths)} columns")
    line = {
        "metric": "Gcell-updates/s fwd+adj; s per FWI gradient, batched vs plain, 1 MI355X",
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded layered VTI model with faults, Ricker source)",
        "config": {
            "workload": f"{w.name}: {w.nx}x{w.nz} dh={w.dh} m, {len(w.shots)} shots, {spread}, "
                        f"nt={nt}, dt={w.dt} s, {w.f} Hz Ricker, cpml={w.cpml}",
            "grad_mode": args.mode,
            "arith": {0: "fp32", 1: "mixed", 2: "fp32-diff", 3: "fp32-increment"}[args.arith],
            "ckpt_interval": plan.ckpt_interval,
            "kept_steps": plan.kept_steps,
            "step_pair_rows": plan.pair_rows,
            "ckpt_format": {0: "fp32", 1: "bf16-scaled", 2: "fp16-scaled"}[args.ckpt_format],
            "workspace_gb": round(plan.workspace_bytes / 1e9, 2),
            "batch_shots": plan.batch_size,
            "rows_per_item": plan.rows_per_item,
            "item_order": plan.item_order,
            "adjoint_tiling": dict(zip(("rows_per_item", "item_order", "grouped"), plan.adj_tiling)),
            "s_per_gradient_batched": round(sec, 4),
            "s_per_gradient_median": round(float(np.median(times)), 4),
            "optimizer_ms": opt_ms,
            "graphs": {"captures": graphs[0], "replays": graphs[1]},
            "s_per_gradient_eager_instrumented": round(sec_instr, 4) if sec_instr else None,
            "s_per_gradient_plain": round(plain, 3) if plain else None,
            "speedup_vs_plain": round(plain / sec, 2) if plain else None,
            "misfit": phi,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic,
            "bytes_per_launch": adj_bytes,
            "mean_launch_us": round(adj_avg * 1e6, 2) if adj_n else None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
