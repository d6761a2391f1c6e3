"""ctypes loader for the CPU oracle (oracle/fwi_oracle.cpp) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "build", "libfwi_oracle.so")

_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    L.orc_ricker.argtypes = [_F, _i, _f, _f]
    L.orc_ricker_f64.argtypes = [_D, _i, _d, _d]
    L.orc_cpml_table.argtypes = [_D, _D, _i, _d, _d, _d, _d, _d, _d, _i]
    L.orc_pre_variables.argtypes = [_F] * 6 + [_i, _i]
    L.orc_pre_variables_f64.argtypes = [_D] * 6 + [_i, _i]
    L.orc_SecondDerivate_P.argtypes = [_i, _F, _F, _F, _F, _f, _i, _i]
    L.orc_PSI_F.argtypes = [_i] + [_F] * 8 + [_i, _i, _f, _i]
    L.orc_ZETA_F.argtypes = [_i] + [_F] * 12 + [_i, _i, _f, _i]
    L.orc_propagator_F.argtypes = [_F] * 14 + [_f, _f, _i, _i, _i, _i, _i, _i]
    L.orc_res.argtypes = [_F, _F, _F, _i, _i]
    L.orc_entryPoint.argtypes = [_F] * 9 + [_i, _i]
    L.orc_secondDerivate.argtypes = [_i] + [_F] * 8 + [_f, _i, _i]
    L.orc_PSI_R.argtypes = [_i] + [_F] * 12 + [_i, _i, _f, _i]
    L.orc_ZETA_R.argtypes = [_i] + [_F] * 20 + [_i, _i, _f, _i]
    L.orc_retroPropagation.argtypes = [_F] * 5 + [_f, _f, _i, _i, _i, _i, _i, _i] + [_F] * 8
    L.orc_GRADIENT.argtypes = [_i] + [_F] * 7 + [_f, _f, _i, _i]
    L.orc_NORMAL.argtypes = [_F] * 3 + [_i, _i, _f, _f, _f]
    L.orc_UpdateModel.argtypes = [_F] * 6 + [_f, _f, _f, _i, _i]
    L.orc_compute_sy.argtypes = [_F] * 18 + [_i, _i, _i]
    L.orc_compute_qi.argtypes = [_F] * 6 + [_f, _f, _f, _i, _i, _i]
    L.orc_compute_ri.argtypes = [_F] * 6 + [_f, _f, _f, _i, _i]
    L.orc_update_r.argtypes = [_F] * 6 + [_f] * 6 + [_i, _i, _i]
    L.orc_update_LBFGS.argtypes = [_F] * 9 + [_i, _i, _f, _f, _f] + [_i] * 5
    for suf, P in (("mix", _F), ("f32", _F), ("f64", _D), ("f32d", _F), ("f32i", _F)):
        fn = getattr(L, f"orc_forward_{suf}")
        fn.argtypes = [P] * 6 + [_i, _i, _i, _i, _d, _d, _i, _i, _i, _i, _i, P]
        fn.restype = None
        fn = getattr(L, f"orc_gradient_{suf}")
        fn.argtypes = [_i] + [P] * 7 + [_i, _i, _i, _i, _d, _d, _i, _i, _i, _i, _i, P, P, P, P, _D]
        fn.restype = _d
        fn = getattr(L, f"orc_group_gradient_{suf}")
        fn.argtypes = [P] * 6 + [_i, _i, _i, _i, _d, _d, _i, ctypes.c_void_p, P, P, P, P, P, _D]
        fn.restype = None
    for suf, P in (("mix", _F), ("f64", _D)):
        fn = getattr(L, f"orc_gradient_states_{suf}")
        fn.argtypes = [_i] + [P] * 7 + [_i, _i, _i, _i, _d, _d, _i, _i, _i, _i, _i, P, P, P, P, P, P]
        fn.restype = _d
    L.orc_group_gradient_ck_f32i.argtypes = [_F] * 6 + [_i, _i, _i, _i, _d, _d, _i, ctypes.c_void_p] + [_F] * 5 + [
        _D, _i, _i]
    L.orc_group_gradient_ck_f32i.restype = None
    L.orc_ck_round.argtypes = [_F, ctypes.c_long, _i]
    L.orc_ck_round.restype = None
    L.orc_set_surface_rows.argtypes = [_i]
    L.orc_set_surface_rows.restype = None
    L.orc_surface_rows.argtypes = [_i, _i]
    L.orc_surface_rows.restype = _i
    L.orc_step_fwd_f64.argtypes = [_D] * 5 + [_i, _i, _i, _d, _d, _D, _D]
    L.orc_step_adj_f64.argtypes = [_D] * 5 + [_i, _i, _i, _d, _d, _D, _D]


class surface_rows:
    """Within the block, the f32i arithmetic's FP64 near-surface band is k rows (0: the library's default, -1:
    none), as a plan's surface_rows option sets it; the default is restored on exit."""

    def __init__(self, k: int):
        self.k = k

    def __enter__(self):
        lib().orc_set_surface_rows(self.k)
        return self

    def __exit__(self, *exc):
        lib().orc_set_surface_rows(0)
        return False


def ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    if a.dtype == np.float32:
        return a.ctypes.data_as(_F)
    if a.dtype == np.float64:
        return a.ctypes.data_as(_D)
    raise TypeError(a.dtype)


# ------------------------------------------------------------------------------------------------
# Convenience wrappers
# ------------------------------------------------------------------------------------------------

def cpml_table(p, round_b=True):
    """(a, b) C-PML table of p.cpml entries in FP64, indexed by distance from the edge.  The parameters are
    the reference's float values (main.c:38-52) promoted to double."""
    a = np.zeros(p.cpml)
    b = np.zeros(p.cpml)
    f32 = lambda v: float(np.float32(v))
    lib().orc_cpml_table(ptr(a), ptr(b), p.cpml, f32(p.vmax), f32(p.dt), f32(p.dh), f32(p.f), f32(p.d0), f32(p.L),
                         1 if round_b else 0)
    return a, b


def ricker(nt, f, dt):
    s = np.zeros(nt, np.float32)
    lib().orc_ricker(ptr(s), nt, f, dt)
    return s


def pre_variables(vp, eps, dl, dtype=np.float32):
    vp, eps, dl = (np.ascontiguousarray(x, dtype) for x in (vp, eps, dl))
    nz, nx = vp.shape
    vh = np.empty_like(vp)
    vpo = np.empty_like(vp)
    vn = np.empty_like(vp)
    fn = lib().orc_pre_variables if dtype == np.float32 else lib().orc_pre_variables_f64
    fn(ptr(vp), ptr(eps), ptr(dl), ptr(vh), ptr(vpo), ptr(vn), nx, nz)
    return vh, vpo, vn


def _dtype_of(mode):
    return np.float64 if mode == "f64" else np.float32


def forward(mode, model_sq, tab, src, nx, nz, nt, cpml, dh, dt, shot):
    """Gather [nt, nrec] of one shot.  model_sq = (v_hor, v_po, v_nmo); tab = (a, b) C-PML table."""
    T = _dtype_of(mode)
    vh, vp, vn = (np.ascontiguousarray(x, T) for x in model_sq)
    ta, tb = (np.ascontiguousarray(x, T) for x in tab)
    s = np.ascontiguousarray(src, T)
    g = np.zeros((nt, shot.nrec), T)
    getattr(lib(), f"orc_forward_{mode}")(ptr(vh), ptr(vp), ptr(vn), ptr(ta), ptr(tb), ptr(s), nx, nz, nt, cpml,
                                           dh, dt, shot.isx, shot.isz, shot.irz, shot.ir0, shot.nrec, ptr(g))
    return g


def gradient(mode, exact, model_sq, tab, src, obs, nx, nz, nt, cpml, dh, dt, shot):
    """One shot's gradient.  Returns (misfit_f64, misfit_ref32, syn[nt,nrec], g_hor, g_po, g_nmo)."""
    T = _dtype_of(mode)
    vh, vp, vn = (np.ascontiguousarray(x, T) for x in model_sq)
    ta, tb = (np.ascontiguousarray(x, T) for x in tab)
    s = np.ascontiguousarray(src, T)
    o = np.ascontiguousarray(obs, T)
    syn = np.zeros((nt, shot.nrec), T)
    gh = np.zeros((nz, nx), T)
    gp = np.zeros((nz, nx), T)
    gn = np.zeros((nz, nx), T)
    m32 = np.zeros(1)
    phi = getattr(lib(), f"orc_gradient_{mode}")(1 if exact else 0, ptr(vh), ptr(vp), ptr(vn), ptr(ta), ptr(tb),
                                                 ptr(s), ptr(o), nx, nz, nt, cpml, dh, dt, shot.isx, shot.isz,
                                                 shot.irz, shot.ir0, shot.nrec, ptr(syn), ptr(gh), ptr(gp), ptr(gn),
                                                 ptr(m32))
    return phi, float(m32[0]), syn, gh, gp, gn


def gradient_states(mode, exact, model_sq, tab, src, obs, nx, nz, nt, cpml, dh, dt, shot):
    """gradient() that also returns the fields it went through (mode "mix" or "f64"): fwd[n] = (P^n, F^n), the
    state forward step n starts from, n = 0..nt, shape [nt+1, 2, nz, nx]; adj[k-1] = the adjoint (P, F) fields
    paired with k = nt..1, shape [nt, 2, nz, nx]: mu^k of the exact adjoint, or the reference-compatible
    adjoint's fields after its step nt - k.  Returns (misfit_f64, syn, g_hor, g_po, g_nmo, fwd, adj)."""
    T = _dtype_of(mode)
    vh, vp, vn = (np.ascontiguousarray(x, T) for x in model_sq)
    ta, tb = (np.ascontiguousarray(x, T) for x in tab)
    s = np.ascontiguousarray(src, T)
    o = np.ascontiguousarray(obs, T)
    syn = np.zeros((nt, shot.nrec), T)
    gh, gp, gn = (np.zeros((nz, nx), T) for _ in range(3))
    fwd = np.zeros((nt + 1, 2, nz, nx), T)
    adj = np.zeros((nt, 2, nz, nx), T)
    phi = getattr(lib(), f"orc_gradient_states_{mode}")(
        1 if exact else 0, ptr(vh), ptr(vp), ptr(vn), ptr(ta), ptr(tb), ptr(s), ptr(o), nx, nz, nt, cpml, dh, dt,
        shot.isx, shot.isz, shot.irz, shot.ir0, shot.nrec, ptr(syn), ptr(gh), ptr(gp), ptr(gn), ptr(fwd), ptr(adj))
    return phi, syn, gh, gp, gn, fwd, adj


def group_gradient(mode, model_sq, tab, src, obs_list, nxm, nz, nt, cpml, dh, dt, shots, ckpt=None):
    """Exact gradient of a gradient group (1..4 shots, each with its own window of the nxm-wide model; imaging
    terms added in member order per step over the union of the windows).  Returns (per-shot FP64 misfits,
    syn list [nt,nrec], union x0, g_hor, g_po, g_nmo over [nz, union width])."""
    T = _dtype_of(mode)
    vh, vp, vn = (np.ascontiguousarray(x, T) for x in model_sq)
    assert vh.shape == (nz, nxm)
    ta, tb = (np.ascontiguousarray(x, T) for x in tab)
    s = np.ascontiguousarray(src, T)
    geo = np.array([[sh.x0, sh.nx, sh.isx, sh.isz, sh.irz, sh.ir0, sh.nrec] for sh in shots], np.int32).reshape(-1)
    obs = np.ascontiguousarray(np.concatenate([np.asarray(o, T).reshape(-1) for o in obs_list]), T)
    syn = np.zeros_like(obs)
    ux0 = min(sh.x0 for sh in shots)
    unx = max(sh.x0 + sh.nx for sh in shots) - ux0
    gh, gp, gn = (np.zeros((nz, unx), T) for _ in range(3))
    mis = np.zeros(len(shots))
    if ckpt is not None and ckpt[1] != 0:
        # checkpoints every ckpt[0] steps in format ckpt[1] (1 = BF16, 2 = FP16), f32i only
        assert mode == "f32i"
        lib().orc_group_gradient_ck_f32i(ptr(vh), ptr(vp), ptr(vn), ptr(ta), ptr(tb), ptr(s), nxm, nz, nt, cpml, dh,
                                         dt, len(shots), geo.ctypes.data, ptr(obs), ptr(syn), ptr(gh), ptr(gp), ptr(gn),
                                         ptr(mis), int(ckpt[0]), int(ckpt[1]))
    else:
        getattr(lib(), f"orc_group_gradient_{mode}")(ptr(vh), ptr(vp), ptr(vn), ptr(ta), ptr(tb), ptr(s), nxm, nz, nt,
                                                     cpml, dh, dt, len(shots), geo.ctypes.data, ptr(obs), ptr(syn),
                                                     ptr(gh), ptr(gp), ptr(gn), ptr(mis))
    syns, o = [], 0
    for sh in shots:
        syns.append(syn[o:o + nt * sh.nrec].reshape(nt, sh.nrec))
        o += nt * sh.nrec
    return mis, syns, ux0, gh, gp, gn


def chain_rule(vp, eps, dl, g_hor, g_po, g_nmo):
    """Gradients with respect to (vp, epsilon, delta) from those with respect to the squared velocities
    v_po = vp^2, v_hor = vp^2 (1+2 eps), v_nmo = vp^2 (1+2 delta)."""
    vp = vp.astype(np.float64)
    gv = 2 * vp * (g_po + (1 + 2 * eps) * g_hor + (1 + 2 * dl) * g_nmo)
    ge = 2 * vp * vp * g_hor
    gd = 2 * vp * vp * g_nmo
    return gv, ge, gd
