"""Shared test helpers: workload setup on the CPU oracle and on the GPU."""
from __future__ import annotations

import itertools
import sys

import numpy as np

import oracle_lib as O
from fwi_amd import models as M


def say(msg: str):
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def small_workload(nx=96, nz=80, nt=240, cpml=12, f=14.0, nshots=2, seed=0, dt=1e-3, dh=10.0):
    """A quick workload of the plain200 kind at reduced size (same model family, same geometry rules)."""
    vt = M.gradient_model(nx, nz, seed=seed, cpml=cpml, sigma=8.0)
    vi = M.smooth_model(*vt, width=10.0)
    if nshots == 1:
        shots = [M.reference_shot(nx, cpml, nx // 2, M.depth_row(2))]
    else:
        shots = M.even_shots(nx, cpml, nshots, M.depth_row(2))
    return M.Workload("small", nx, nz, dh, dt, nt, f, cpml, vt, vi, shots)


def oracle_setup(w, mode="f32"):
    """Oracle-side tables, wavelet, squared-velocity models (true, initial) of a workload."""
    T = np.float64 if mode == "f64" else np.float32
    p = w.cpml_params
    ta, tb = O.cpml_table(p)
    src = O.ricker(w.nt, w.f, w.dt)
    sq_true = O.pre_variables(*[x.astype(T) for x in w.true_model], dtype=T)
    sq_init = O.pre_variables(*[x.astype(T) for x in w.init_model], dtype=T)
    return (ta, tb), src, sq_true, sq_init


def window_model(sq, shot):
    return tuple(np.ascontiguousarray(x[:, shot.x0:shot.x0 + shot.nx]) for x in sq)


def oracle_gathers(w, mode, sq, tab, src):
    out = []
    for s in w.shots:
        out.append(O.forward(mode, window_model(sq, s), tab, src, s.nx, w.nz, w.nt, w.cpml, w.dh, w.dt, s))
    return out


GROUP_LOOK = 12  # ungrouped shots a group leader chooses its members among (plan.hip: kGroupLook)


def _placeholder_waves(shots, members):
    """Waves of a grouped workgroup whose shot's window misses the 64-column band of the group's union window."""
    ux0 = min(shots[i].x0 for i in members)
    ux1 = max(shots[i].x0 + shots[i].nx for i in members)
    return sum(1 for c in range(ux0, ux1, 64) for i in members
               if shots[i].x0 >= c + 64 or shots[i].x0 + shots[i].nx <= c)


def gradient_groups(shots, group=4, max_batch=0):
    """The batched path's gradient groups (plan.hip: group_shots): within each batch (of `max_batch` shots, 0 = one
    batch) the shots ordered by window (x0, width, then index); each group is led by the first ungrouped shot and
    completed by the up to group-1 shots among the next GROUP_LOOK ungrouped ones whose union window leaves the
    fewest placeholder waves, the first such choice in lexicographic order.  Member lists in wave order."""
    out, b0 = [], 0
    n = len(shots)
    while b0 < n:
        b1 = n if max_batch <= 0 else min(n, b0 + max_batch)
        rest = sorted(range(b0, b1), key=lambda i: (shots[i].x0, shots[i].nx, i))
        while rest:
            nc = min(GROUP_LOOK, len(rest) - 1)
            k = min(group - 1, nc)
            best, bw = (), None
            for pick in itertools.combinations(range(1, nc + 1), k):
                w = _placeholder_waves(shots, [rest[0]] + [rest[j] for j in pick])
                if bw is None or w < bw:
                    best, bw = pick, w
            out.append([rest[0]] + [rest[j] for j in best])
            rest = [r for j, r in enumerate(rest) if j and j not in best]
        b0 = b1
    return out


def oracle_gradient(w, mode, exact, sq, tab, src, obs_list, group=4, max_batch=0, ckpt=None):
    """Sum over shots (FP64) of the oracle gradients, placed into model columns; total FP64 misfit.  In exact
    mode shots are accumulated in the batched path's gradient groups (group=1: one accumulator per shot), whose
    slots are added in the order of their leaders within each batch; ckpt=(K, fmt) emulates 16-bit checkpoints
    every K steps (fmt 1 BF16, 2 FP16)."""
    T = np.float64 if mode == "f64" else np.float32
    G = np.zeros((3, w.nz, w.nx))
    phi = 0.0
    if exact and (group > 1 or ckpt is not None):
        groups = gradient_groups(w.shots, group, max_batch)
        results = {}
        mis_all = {}
        syns = [None] * len(w.shots)
        for members in groups:
            mis, sy, ux0, gh, gp, gn = O.group_gradient(mode, sq, tab, src.astype(T), [obs_list[i] for i in members],
                                                        w.nx, w.nz, w.nt, w.cpml, w.dh, w.dt,
                                                        [w.shots[i] for i in members], ckpt=ckpt)
            results[members[0]] = (ux0, (gh, gp, gn))
            for i, v, y in zip(members, mis, sy):
                mis_all[i] = v
                syns[i] = y
        for i in range(len(w.shots)):  # misfits in shot order, slots in leader order (per batch)
            phi += mis_all[i]
        b0 = 0
        while b0 < len(w.shots):  # each batch's slots summed first, then added (the device's order)
            b1 = len(w.shots) if max_batch <= 0 else min(len(w.shots), b0 + max_batch)
            Gb = np.zeros_like(G)
            for lead in range(b0, b1):
                if lead in results:
                    ux0, gs = results[lead]
                    for k, g in enumerate(gs):
                        Gb[k][:, ux0:ux0 + g.shape[1]] += g.astype(np.float64)
            G += Gb
            b0 = b1
        return phi, G, syns
    syns = []
    for s, obs in zip(w.shots, obs_list):
        ph, _, syn, gh, gp, gn = O.gradient(mode, exact, window_model(sq, s), tab, src.astype(T), obs, s.nx, w.nz, w.nt,
                                            w.cpml, w.dh, w.dt, s)
        phi += ph
        syns.append(syn)
        for k, g in enumerate((gh, gp, gn)):
            G[k][:, s.x0:s.x0 + s.nx] += g.astype(np.float64)
    return phi, G, syns
