#!/usr/bin/env python3
"""Benchmark: one full FWI gradient of the 1600x400, 64-shot survey per step on one MI355X.

A "step" is one gradient (all shots: forward with checkpoints, residual, FP64 misfit, adjoint with
recomputation, imaging, FP64 cross-shot sum) of the batched path.  value = Gcell-updates/s (fwd+adj) =
2 * sum_shots(nx*nz) * nt / seconds per gradient; recomputation is not counted as work.

Beside it, on the same workload:
  * plain per-shot GPU path: the reference's per-shot flow on the legacy C ABI (launchPropagation with a
    pinned-host snapshot every step, launchBackPropagation), timed over every shot and step of the gradient
    when that is at most PLAIN_FULL_CELL_STEPS cell-steps (survey1600: about 75 s), else on a sample of shots and
    steps scaled to the survey.  Short samples overstate it: this host-bound flow runs its steps faster the
    longer it runs (survey1600: 0.51 ms per step for one shot, 0.29 ms over all 64 shots);
  * roofline of every step kernel and of the dominant one (the largest share of the gradient's kernel time):
    algorithmic bytes per launch (the survey's count, and this design's minimum) over its average duration,
    measured with HIP events on the stream the kernel runs on during an instrumented gradient, beside the
    counter-measured HBM bytes of the same kernel when profiles/ holds them for these kernel sources (the run
    then uses the work-item tiling those counters were measured with, unless --retune or a tiling is given);
  * the same gradient in the reference's arithmetic (FP64 stencil sums), without the FP64 near-surface band,
    and with single steps only (no step pairs), each timed apart;
  * cpu_baseline: the CPU oracle (f32i, exact gradient, OpenMP) on one shot for a bounded number of steps, and
    the reference's arithmetic and gradient on one thread, with the CPU model.
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
PLAIN_FULL_CELL_STEPS = 2e11  # largest plain-path gradient timed in full (survey1600: 1.64e11)


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
    ap.add_argument("--pairs", type=int, default=1,
                    help="step pairs: the core two steps per launch (1 = rows tuned, k > 1 = k rows per core item, "
                         "0 = single steps)")
    ap.add_argument("--adj-pairs", type=int, default=1,
                    help="adjoint step pairs: each gradient group's core two adjoint steps per launch (1 = rows tuned "
                         "and kept where faster than single steps, k > 1 = k rows per core item, 0 = single steps)")
    ap.add_argument("--arith", type=int, default=3,
                    help="0 FP32 two-level, 1 mixed (reference), 2 FP32 difference stencil, 3 FP32 increment form")
    ap.add_argument("--rows", type=int, default=0, help="rows per work item (0 = tuned)")
    ap.add_argument("--order", type=int, default=-1, help="work-item order (-1 = tuned)")
    ap.add_argument("--adj-rows", type=int, default=0, help="rows per adjoint work item (0 = --rows or tuned)")
    ap.add_argument("--adj-order", type=int, default=-1, help="adjoint work-item order (-1 = from --order or tuned)")
    ap.add_argument("--retune", action="store_true",
                    help="let the tuner choose the tiling even where the counter profile names one")
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="steps per shot of the plain path (0: all, or 1000 when the full gradient is too large)")
    ap.add_argument("--plain-shots", type=int, default=0,
                    help="shots the plain path runs, the rest extrapolated (0: all, or 1 when too large)")
    ap.add_argument("--cpu-steps", type=int, default=4000)
    ap.add_argument("--cpu1-steps", type=int, default=200, help="steps of the single-thread reference-arithmetic leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the reference-arithmetic (mixed) GPU gradient")
    ap.add_argument("--no-band-leg", action="store_true", help="skip the gradient timed without the FP64 band")
    ap.add_argument("--no-single-leg", action="store_true", help="skip the gradient timed with single steps only")
    ap.add_argument("--surface-rows", type=int, default=0, help="FP64 near-surface band rows (0 default, -1 none)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--timing-stride", type=int, default=4)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device-to-device (a kernel of known traffic for PMC scaling)")
    return ap.parse_args()


def plain_path_seconds(w, model_sq_dev, src_dev, ab, nsteps, nshots=1):
    """Reference flow on the legacy ABI for `nshots` shots one after the other (spread over the survey; the middle
    one when 1) and `nsteps` steps each: the shot's fields cleared, forward with a pinned snapshot per step
    (launchPropagation), residual, adjoint with the snapshot copy back (launchBackPropagation)."""
    import numpy as np
    import torch

    from fwi_amd import _lib

    L = _lib.load()
    nx, nz, cpml = w.nx, w.nz, w.cpml
    n = nx * nz
    nt = nsteps
    picks = [w.shots[len(w.shots) // 2]] if nshots <= 1 else \
        [w.shots[(k * len(w.shots)) // nshots] for k in range(min(nshots, len(w.shots)))]
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
    per_shot = [k for k in names if k not in ("gcx", "gcz", "gdx")]  # the gradient sums over the shots
    def run_shots(shots):
        for shot in shots:
            sx, sz = shot.isx + 1, shot.isz + 1
            for k in per_shot:
                f[k].zero_()
            for t in (Pa, Pb, Pre2):
                t.zero_()
            Ppre, Ppast = Pa, Pb
            for it in range(nt):
                L.launchPropagation(
                    s1, s2, PFh.data_ptr() + 8 * it * n, ptr("PF"), Ppast.data_ptr(), Ppre.data_ptr(),
                    Ppast.data_ptr() + 4 * n, Ppre.data_ptr() + 4 * n, ptr("p2"), ptr("r2"), dh2, nx, nz,
                    ax.data_ptr(), bx.data_ptr(), az.data_ptr(), bz.data_ptr(), ptr("psx"), ptr("psz"), dh, cpml,
                    ptr("zx"), ptr("zz"), ptr("atp"), ptr("atf"), vh.data_ptr(), vp.data_ptr(), vn.data_ptr(),
                    src_dev.data_ptr(), shot2.data_ptr(), C, sx, sz, it)
                Ppre, Ppast = Ppast, Ppre
            L.res_h(Res.data_ptr(), shot1.data_ptr(), shot2.data_ptr(), nx, nt)
            torch.cuda.synchronize()
            Pa.zero_()
            Pb.zero_()
            Ppre, Ppast = Pa, Pb
            for it in range(nt):
                L.launchBackPropagation(
                    s1, s2, nx, nz, nt, PFh.data_ptr() + 8 * (nt - it - 1) * n, Ppre.data_ptr(), Ppast.data_ptr(),
                    Ppre.data_ptr() + 4 * n, Ppast.data_ptr() + 4 * n, ptr("aux1"), ptr("aux2"), ptr("aux3"),
                    ptr("aux4"), vh.data_ptr(), vp.data_ptr(), vn.data_ptr(), ptr("p2cx"), ptr("p2cz"), ptr("r2dx"),
                    ptr("r2cz"), dh2, Pre2.data_ptr(), Pre2.data_ptr() + 4 * n, ax.data_ptr(), bx.data_ptr(),
                    az.data_ptr(), bz.data_ptr(), ptr("psixcx"), ptr("psixcz"), ptr("psx"), ptr("psz"), dh, cpml,
                    ptr("zx"), ptr("zz"), ptr("atf"), ptr("atp"), ptr("atenxcx"), ptr("atenxcz"), ptr("atenzdx"),
                    ptr("atenzcz"), Res.data_ptr(), C, ptr("PF"), it, sx, sz, dt, ptr("gcx"), ptr("gcz"), ptr("gdx"))
                Ppre, Ppast = Ppast, Ppre

    # one untimed pass over the first shot: a first pass takes about twice as long (first launches, first copies
    # into the pinned snapshot buffer), and the timed shots stand for shots that run after it
    run_shots(picks[:1])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_shots(picks)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    _lib.check_last("plain path")
    _lib.destroy_stream(s1)
    _lib.destroy_stream(s2)
    del PFh
    return sec


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, nsteps, mode="f32i", exact=True, threads=None):
    """The oracle on one shot of the workload for nsteps steps (forward + adjoint + imaging): by default the
    f32i arithmetic's exact gradient on all OpenMP threads; mode "mix" with exact=False is the reference's
    arithmetic and gradient.  threads: OpenMP threads for this call (None: the runtime's default).  Returns
    (Gcell-updates/s, threads used, seconds)."""
    import ctypes

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    omp = ctypes.CDLL("libgomp.so.1")
    before = omp.omp_get_max_threads()
    if threads:
        omp.omp_set_num_threads(threads)
    used = omp.omp_get_max_threads()
    try:
        shot = w.shots[len(w.shots) // 2]
        tab = O.cpml_table(w.cpml_params)
        src = O.ricker(nsteps, w.f, w.dt)
        sq = O.pre_variables(*[np.ascontiguousarray(x[:, shot.x0:shot.x0 + shot.nx]) for x in w.init_model])
        obs = np.zeros((nsteps, shot.nrec), np.float32)
        t0 = time.perf_counter()
        O.gradient(mode, exact, sq, tab, src, obs, shot.nx, w.nz, nsteps, w.cpml, w.dh, w.dt, shot)
        sec = time.perf_counter() - t0
    finally:
        omp.omp_set_num_threads(before)
    return 2.0 * shot.nx * w.nz * nsteps / sec / 1e9, used, sec


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


# Step launches of the batched gradient: (timing kind, kernel name, what it does).  With step pairs (the default)
# a pair of forward steps is two launches: k_fwd_pair (two steps of every window's core, one step of the FP64
# near-surface band, the C-PML shell and the core's cells next to it) and k_fwd_main completing the pair (the band's
# and the shell's second step); single forward steps (where a pair would cross a checkpoint, or without pairs) are
# one k_fwd_main launch with the band.  Launches per gradient are counted by the library.
STEP_KERNELS = [
    (0, "k_fwd_main<0>", "forward step (one launch with the FP64 band)"),
    (1, "k_fwd_main<1>", "forward step storing the imaging terms (one launch with the FP64 band)"),
    (2, "k_adj_ex_main", "exact adjoint step with imaging"),
    (3, "k_fwd_pair<0>", "step pair: two steps of the core, one of the band and the shell"),
    (4, "k_fwd_pair<1>", "step pair storing the imaging terms: two steps of the core, one of the band and the shell"),
    (8, "k_fwd_main<0> (pair)", "the band's and the shell's second step of a step pair"),
    (9, "k_fwd_main<1> (pair)", "the band's and the shell's second step of a step pair, storing"),
    (5, "k_adj_pair", "adjoint step pair: two adjoint steps of every gradient group's core, one of the shell"),
    (7, "k_adj_shell", "the shell's second adjoint step of an adjoint step pair"),
]


def kernel_bytes(kind, N, Nb, Nc, Nm, Ncore, grouped, inc, Nacore=0):
    """Bytes per launch of a step launch: (the survey's algorithmic count, this design's minimum).

    Survey count per forward step (FP32): 24 B per cell + 32 per C-PML cell + 12 per model cell (read once per
    batch); adjoint + imaging 56 B per cell + 32 per C-PML cell + 12 per model cell.  A step-pair launch advances the Ncore core cells two steps and the other N - Ncore cells one
    step, the launch completing the pair those N - Ncore cells one more step: together they count two steps; an
    adjoint step-pair launch likewise advances the Nacore cells of the gradient groups' cores two adjoint steps.
    Design minimum: the increment form reads P, Q, F, R and writes them (32 B per cell and step or per core cell and
    step pair; + 8 B per cell and step when storing the imaging terms); the FP64 band's Nb cells move two FP64 levels
    of P and F (48 B) + their float rounding (8 B) + the model (12 B, per shot); the adjoint reads mu and mu^{k+2} (or
    D) and the imaging terms (24 B), writes mu (and D) and read-modify-writes the FP32 gradient slot (24 B per cell,
    6 with gradient groups of four), and reads and writes the two C-PML memory variables of its direction on each
    C-PML cell (16 B; both directions on the corners)."""
    st = 8.0 if kind in (1, 4, 9) else 0.0
    band = (48.0 + 8.0 + 12.0 + st) * Nb

    def shell(cells):  # one step of `cells` cells, the band's among them
        return (32.0 + st) * (cells - Nb) + band + 32.0 * Nc + 12.0 * Nm

    if kind in (0, 1):
        survey = 24.0 * N + 32.0 * Nc + 12.0 * Nm
        design = shell(N)
    elif kind in (3, 4):
        survey = 24.0 * (2 * Ncore + (N - Ncore)) + 32.0 * Nc + 12.0 * Nm
        design = (32.0 + 2 * st) * Ncore + shell(N - Ncore)
    elif kind in (8, 9):
        survey = 24.0 * (N - Ncore) + 32.0 * Nc + 12.0 * Nm
        design = shell(N - Ncore)
    else:
        adj_shell = (24.0 + (16.0 if inc else 8.0) + (6.0 if grouped else 24.0))
        if kind == 5:  # core: mu^{k+1}, mu^{k+2} read, mu^k, mu^{k-1} written, both steps' terms, one slot update
            survey = 56.0 * (2 * Nacore + (N - Nacore)) + 32.0 * Nc + 12.0 * Nm
            design = (48.0 + (6.0 if grouped else 24.0)) * Nacore + adj_shell * (N - Nacore) + 16.0 * Nc + 12.0 * Nm
        elif kind == 7:
            survey = 56.0 * (N - Nacore) + 32.0 * Nc + 12.0 * Nm
            design = adj_shell * (N - Nacore) + 16.0 * Nc + 12.0 * Nm
        else:
            survey = 56.0 * N + 32.0 * Nc + 12.0 * Nm
            design = adj_shell * N + 16.0 * Nc + 12.0 * Nm
    return survey, design


def load_traffic(w, nshots, tiling):
    """Counter-measured HBM bytes per launch of each kernel (profiles/fused_steps_survey1600_hbm_traffic.json), if it
    was measured on this workload with the kernel sources in this tree (its source hash) and the work-item tiling
    this plan runs (the tiling decides the halo and model re-reads); else None and why."""
    from fwi_amd.pmc import source_hash

    tfile = os.path.join(ROOT, "profiles", "fused_steps_survey1600_hbm_traffic.json")
    if not (w.name == "survey1600" and nshots == 64):
        return None, "counters measured on survey1600 only"
    if not os.path.exists(tfile):
        return None, "no counter profile"
    with open(tfile) as fh:
        t = json.load(fh)
    if t.get("source_hash") != source_hash():
        return None, f"counter profile is of other kernel sources ({t.get('source_hash')} vs {source_hash()})"
    if tiling is not None and t.get("tiling") != tiling:
        return None, f"counter profile is of another work-item tiling ({t.get('tiling')} vs {tiling})"
    return t["kernels"], "profiles/fused_steps_survey1600_hbm_traffic.json (source hash " + t["source_hash"] + ")"


def profiled_tiling(w, args):
    """The work-item tiling of the committed counter profile, when this run measures its workload with its kernel
    sources and no tiling was asked for: the bench then runs that tiling (the tuner's candidates near the top
    time within about 1 % of each other, so its own choice varies from run to run) and the counters describe the
    kernels it times."""
    if (args.rows, args.order, args.adj_rows, args.adj_order) != (0, -1, 0, -1) or args.retune:
        return None
    t, _ = load_traffic(w, len(w.shots), None)
    if t is None:
        return None
    with open(os.path.join(ROOT, "profiles", "fused_steps_survey1600_hbm_traffic.json")) as fh:
        return json.load(fh).get("tiling")


def tiling_of(plan):
    """The plan's work-item tiling as the bench line reports it."""
    r, o, (ar, ao, g) = plan.rows_per_item, plan.item_order, plan.adj_tiling
    t = {"rows_per_item": r, "item_order": o,
         "adjoint_tiling": {"rows_per_item": ar, "item_order": ao, "grouped": bool(g)}}
    if plan.pair_rows:
        t["pair_rows"] = plan.pair_rows
    if plan.adj_pair_rows:
        t["adj_pair_rows"] = plan.adj_pair_rows
    return t


TRAFFIC_KEYS = {"k_fwd_main<0>": ("k_fwd_main<0, float",), "k_fwd_main<1>": ("k_fwd_main<1, float",),
                "k_fwd_main<0> (pair)": ("k_fwd_main<0, float",), "k_fwd_main<1> (pair)": ("k_fwd_main<1, float",),
                "k_fwd_pair<0>": ("k_fwd_pair<0>",), "k_fwd_pair<1>": ("k_fwd_pair<1>",),
                "k_adj_ex_main": ("k_adj_ex_main<float",), "k_adj_pair": ("k_adj_pair",),
                "k_adj_shell": ("k_adj_shell",)}


def traffic_of(traffic, name):
    """Counter bytes (read + written) per launch of a step kernel, summed over the kernels it stands for."""
    if not traffic:
        return None
    parts = [v for k, v in traffic.items() for p in TRAFFIC_KEYS.get(name, ()) if k.startswith(p)]
    if not parts:
        return None
    return float(sum(v["read_bytes"] + v["write_bytes"] for v in parts))


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
    pinned = profiled_tiling(w, args)
    if pinned:
        args.rows, args.order = pinned["rows_per_item"], pinned["item_order"]
        args.adj_rows, args.adj_order = pinned["adjoint_tiling"]["rows_per_item"], pinned["adjoint_tiling"]["item_order"]
        if args.pairs == 1 and pinned.get("pair_rows"):
            args.pairs = pinned["pair_rows"]
        if args.adj_pairs == 1:
            args.adj_pairs = pinned.get("adj_pair_rows", 0)
        log(f"tiling of the counter profile: forward {args.rows} rows order {args.order}, adjoint {args.adj_rows} rows "
            f"order {args.adj_order}")
    plan = Plan(w.nx, w.nz, nt, w.dh, w.dt, w.cpml, w.shots, src, ab, grad_mode=grad_mode, ckpt_interval=args.ckpt,
                rows_per_item=args.rows, item_order=args.order, adj_rows_per_item=args.adj_rows,
                adj_item_order=args.adj_order, arith=args.arith,
                ckpt_format=args.ckpt_format, use_graph=0 if args.no_graph else 1, store_segments=args.keep,
                grad_group=args.grad_group, step_pairs=args.pairs, surface_rows=args.surface_rows,
                adj_step_pairs=args.adj_pairs)
    log(f"plan: batch {plan.batch_size} shots, K={plan.ckpt_interval}, {plan.kept_steps} steps kept from the forward, "
        f"workspace {plan.workspace_bytes / 1e9:.1f} GB, FP64 band rows 4..{plan.surface_rows - 1}")
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
    gnodes = plan.graph_nodes()
    log(f"graph captures {graphs[0]}, replays {graphs[1]}; kernel nodes: largest graph {gnodes[0]}, the gradient's "
        f"capture {gnodes[1]}")
    # kernel durations: one more gradient with every timing_stride-th launch of each step kernel bracketed by HIP
    # events on the stream it runs on (the instrumented call launches eagerly; its kernels are the ones the timed
    # calls replayed)
    sec_instr = None
    kernels = []
    if args.timing_stride > 0:
        plan.set_timing(args.timing_stride)
        t0 = time.perf_counter()
        phi_i, _, _ = plan.gradient(init_sq, obs)
        torch.cuda.synchronize()
        sec_instr = time.perf_counter() - t0
        assert phi_i == phi
        log(f"instrumented eager gradient: {sec_instr:.3f} s")
        nmodel = w.nx * w.nz
        band = plan.surface_rows
        nband = sum((band - 4) * s.nx for s in w.shots) if band else 0
        ncore = plan.core_cells()
        nacore = plan.adj_core_cells()
        nbat = plan.batch_count  # batches of a gradient, each with the same launches
        grouped = bool(plan.adj_tiling[2])
        inc = args.arith == 3
        traffic, tnote = load_traffic(w, len(w.shots), tiling_of(plan))
        for kind, name, what in STEP_KERNELS:
            ms, n = plan.kernel_time(kind)
            if not n:
                continue
            mean = ms / n / 1e3
            # a launch covers one batch of shots: the mean launch moves the cells of the mean batch
            survey, design = kernel_bytes(kind, ncell / nbat, nband / nbat, nstrip / nbat, nmodel, ncore / nbat,
                                          grouped, inc, nacore / nbat)
            per_grad = plan.kernel_launches(kind)
            tb = traffic_of(traffic, name)
            k = {"kernel": name, "does": what, "launches_per_gradient": per_grad, "mean_launch_us": round(mean * 1e6, 1),
                 "s_per_gradient": round(per_grad * mean, 4),
                 "bytes_survey": survey, "frac_survey": round(survey / mean / 1e9 / HBM_PEAK_GBS, 4),
                 "bytes_design": design, "frac_design": round(design / mean / 1e9 / HBM_PEAK_GBS, 4),
                 "traffic": tb, "traffic_frac": round(tb / mean / 1e9 / HBM_PEAK_GBS, 4) if tb else None}
            kernels.append(k)
            log(f"{name}: {n} timed launches, mean {mean * 1e6:.1f} us, {per_grad} per gradient "
                f"({k['s_per_gradient']:.3f} s); frac of 8 TB/s: survey bytes {k['frac_survey']:.3f}, design bytes "
                f"{k['frac_design']:.3f}, counters {k['traffic_frac']}")
        ksum = sum(k["s_per_gradient"] for k in kernels)
        for k in kernels:
            k["share"] = round(k["s_per_gradient"] / ksum, 4) if ksum else None
        log(f"step kernels {ksum:.3f} s of the {sec_instr:.3f} s instrumented gradient; traffic: {tnote}")
    dom = max(kernels, key=lambda k: k["s_per_gradient"]) if kernels else None

    opt_ms = optimizer_ms(w, init_sq, grad)
    log(f"optimizer: SD step {opt_ms['sd_step_ms']} ms, L-BFGS direction {opt_ms['lbfgs_direction_ms']} ms, "
        f"trial {opt_ms['trial_ms']} ms")
    surface_rows = plan.surface_rows
    kept, K, wsb, bsz = plan.kept_steps, plan.ckpt_interval, plan.workspace_bytes, plan.batch_size
    tiling = tiling_of(plan)
    plan.close()
    del plan, grad
    torch.cuda.empty_cache()

    # the reference's arithmetic (double-literal stencil sums) on the same batched path, timed apart
    sec_mixed = None
    if not args.no_mixed:
        pm = Plan(w.nx, w.nz, nt, w.dh, w.dt, w.cpml, w.shots, src, ab, grad_mode=grad_mode, arith=1)
        try:
            pm.gradient(init_sq, obs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pm.gradient(init_sq, obs)
            torch.cuda.synchronize()
            sec_mixed = time.perf_counter() - t0
        finally:
            pm.close()
        log(f"reference arithmetic (mixed: FP64 stencil sums), same path: {sec_mixed:.3f} s per gradient")

    # what the FP64 near-surface band costs: the same gradient with every row in FP32
    sec_noband = None
    if surface_rows > 0 and not args.no_band_leg:
        pb = Plan(w.nx, w.nz, nt, w.dh, w.dt, w.cpml, w.shots, src, ab, grad_mode=grad_mode, arith=args.arith,
                  ckpt_interval=args.ckpt, ckpt_format=args.ckpt_format, grad_group=args.grad_group,
                  surface_rows=-1)
        try:
            pb.gradient(init_sq, obs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pb.gradient(init_sq, obs)
            torch.cuda.synchronize()
            sec_noband = time.perf_counter() - t0
        finally:
            pb.close()
        log(f"without the FP64 surface band: {sec_noband:.3f} s per gradient")

    # the same gradient with single steps only (no temporal blocking), timed apart
    sec_single = None
    if args.pairs and args.arith == 3 and not args.no_single_leg:
        p1 = Plan(w.nx, w.nz, nt, w.dh, w.dt, w.cpml, w.shots, src, ab, grad_mode=grad_mode, arith=args.arith,
                  ckpt_interval=args.ckpt, ckpt_format=args.ckpt_format, grad_group=args.grad_group,
                  surface_rows=args.surface_rows, step_pairs=0)
        try:
            p1.gradient(init_sq, obs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            p1.gradient(init_sq, obs)
            torch.cuda.synchronize()
            sec_single = time.perf_counter() - t0
        finally:
            p1.close()
        log(f"single steps only: {sec_single:.3f} s per gradient")

    plain = None
    if not args.no_plain:
        # the snapshot buffer of the plain path is 2*pns*nx*nz floats of pinned host memory (20.5 GB for survey1600)
        full = len(w.shots) * nt * w.nx * w.nz <= PLAIN_FULL_CELL_STEPS
        pns = min(args.plain_steps or (nt if full else 1000), nt)
        pnsh = max(1, min(args.plain_shots or (len(w.shots) if full else 1), len(w.shots)))
        s_plain = plain_path_seconds(w, init_sq, torch.from_numpy(src[:pns].copy()).cuda(), ab, pns, pnsh)
        plain = s_plain * (nt / pns) * (len(w.shots) / pnsh)
        log(f"plain per-shot path: {s_plain:.3f} s for {pnsh} shot(s) x {pns} steps -> {plain:.1f} s per gradient")

    cpu = None
    if not args.no_cpu:
        ns = min(args.cpu_steps, nt)
        cg, threads, csec = cpu_baseline(w, ns)
        cpu = {"value": round(cg, 4), "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"oracle f32i exact gradient on {threads} OpenMP threads, 1 shot of {w.name} "
                         f"({w.nx}x{w.nz}), {ns} of {nt} steps, fwd+adj, {csec:.1f} s",
               "cpu_model": cpu_model()}
        log(f"cpu baseline: {cg:.3f} Gcell/s on {threads} threads ({csec:.1f} s)")
        ns1 = min(args.cpu1_steps, nt)
        c1, _, c1sec = cpu_baseline(w, ns1, mode="mix", exact=False, threads=1)
        cpu["single_thread_reference_arith"] = {
            "value": round(c1, 4), "unit": "Gcell-updates/s", "cores": 1,
            "sample": f"oracle in the reference's arithmetic (FP64 stencil sums, FP32 storage) and gradient "
                      f"formula, 1 thread, 1 shot, {ns1} of {nt} steps, fwd+adj, {c1sec:.1f} s"}
        log(f"cpu baseline, reference arithmetic, 1 thread: {c1:.4f} Gcell/s ({c1sec:.1f} s)")

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
            "fp64_surface_rows": surface_rows,
            "ckpt_interval": K,
            "kept_steps": kept,
            "step_pairs": bool(tiling.get("pair_rows")),
            "adj_step_pairs": bool(tiling.get("adj_pair_rows")),
            "ckpt_format": {0: "fp32", 1: "bf16-scaled", 2: "fp16-scaled"}[args.ckpt_format],
            "workspace_gb": round(wsb / 1e9, 2),
            "batch_shots": bsz,
            **tiling,
            "tiling_from": "counter profile" if pinned else ("options" if (args.rows or args.order >= 0 or args.adj_rows or args.adj_order >= 0) else "tuner"),
            "s_per_gradient_batched": round(sec, 4),
            "s_per_gradient_median": round(float(np.median(times)), 4),
            "s_per_gradient_reference_arith": round(sec_mixed, 4) if sec_mixed else None,
            "s_per_gradient_without_surface_band": round(sec_noband, 4) if sec_noband else None,
            "s_per_gradient_single_steps": round(sec_single, 4) if sec_single else None,
            "optimizer_ms": opt_ms,
            "graphs": {"captures": graphs[0], "replays": graphs[1], "nodes_largest_graph": gnodes[0],
                       "nodes_gradient": gnodes[1]},
            "s_per_gradient_eager_instrumented": round(sec_instr, 4) if sec_instr else None,
            "s_per_gradient_plain": round(plain, 3) if plain else None,
            "plain_timed": f"{pnsh} shots x {pns} steps" if plain else None,
            "speedup_vs_plain": round(plain / sec, 2) if plain else None,
            "misfit": phi,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom["kernel"] if dom else None,
            "achieved": round(dom["bytes_survey"] / dom["mean_launch_us"] / 1e3, 1) if dom else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": dom["frac_survey"] if dom else None,
            "traffic": dom["traffic"] if dom else None,
            "bytes_per_launch": dom["bytes_survey"] if dom else None,
            "mean_launch_us": dom["mean_launch_us"] if dom else None,
            "share_of_step_kernels": dom["share"] if dom else None,
            "kernels": kernels,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
