"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests, fixtures or golden vectors, so each oracle function is checked here against
an independent fact: polynomial exactness and the Nyquist symbol of the stencils, the C-PML profile shapes
recovered from the tables, closed-form properties of the Ricker wavelet, the isotropic reduction, the
analytic 2-D Green's function, linearity, time-shift invariance and source-receiver reciprocity of the
forward, a dot-product test, finite-difference directional derivatives and a per-cell brute-force
finite-difference gradient of the exact adjoint, and the relation between the reference-compatible and exact
gradients.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from fwi_amd import models as M


def _f(a):
    return np.ascontiguousarray(a, np.float32)


# ------------------------------------------------------------------------------------------------
# Stencils (parallelcalls.cu:57, :127)
# ------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("mode", [0, 1])
def test_d2_polynomial_exactness(mode):
    """8th-order D2 is exact on polynomials of degree <= 9 (error term ~ h^8 u^(10))."""
    nx, nz, h = 40, 24, 0.5
    x = (np.arange(nx) - nx / 2) * h / 10.0
    for deg in range(0, 10):
        u = np.tile(x ** deg, (nz, 1))
        P = _f(u)
        F = _f(u.T[:nz, :nx] if False else u)
        p2 = np.zeros((nz, nx), np.float32)
        r2 = np.zeros((nz, nx), np.float32)
        O.lib().orc_SecondDerivate_P(mode, O.ptr(P), O.ptr(F), O.ptr(p2), O.ptr(r2), np.float32(h * h), nx, nz)
        # d2/dx2 of (ix*h/10 - c)^deg with respect to the grid coordinate ix*h
        exact = deg * (deg - 1) * x ** max(deg - 2, 0) / 100.0 if deg >= 2 else 0 * x
        got = p2[10, 4:nx - 4]
        ref = np.broadcast_to(exact, (nx,))[4:nx - 4]
        scale = max(np.abs(u).max() / (h * h), 1.0)
        assert np.abs(got - ref).max() <= 2e-5 * scale, (deg, np.abs(got - ref).max())


def test_d2_nyquist_symbol():
    """On (-1)^ix the D2 stencil returns (-205/72 - 2*8/5 - 2/5 - 2*8/315 - 2/560)/h^2 = -6.50159/h^2."""
    nx, nz = 32, 16
    u = np.tile((-1.0) ** np.arange(nx), (nz, 1))
    p2 = np.zeros((nz, nx), np.float32)
    r2 = np.zeros((nz, nx), np.float32)
    O.lib().orc_SecondDerivate_P(0, O.ptr(_f(u)), O.ptr(_f(u)), O.ptr(p2), O.ptr(r2), np.float32(1.0), nx, nz)
    assert abs(p2[8, 10] - (-6.501587301587)) < 1e-5
    # z-derivative of a field constant in z is zero
    assert np.abs(r2[4:-4, 4:-4]).max() < 1e-12


@pytest.mark.parametrize("mode", [0, 1])
def test_d1_polynomial_exactness(mode):
    """8th-order D1 (via PSI_F with a=1, b=0 so psi = D1 P) is exact on degree <= 8."""
    nx, nz, cpml, h = 40, 40, 40, 1.0
    x = (np.arange(nx) - nx / 2) / 8.0
    ones = np.ones(nx, np.float32)
    zeros = np.zeros(nx, np.float32)
    for deg in range(0, 9):
        P = _f(np.tile(x ** deg, (nz, 1)))
        F = _f(np.tile((x ** deg)[:, None], (1, nx)))  # same profile along z
        psx = np.zeros((nz, nx), np.float32)
        psz = np.zeros((nz, nx), np.float32)
        O.lib().orc_PSI_F(mode, O.ptr(P), O.ptr(F), O.ptr(ones), O.ptr(zeros), O.ptr(ones), O.ptr(zeros),
                          O.ptr(psx), O.ptr(psz), nx, nz, np.float32(h), cpml)
        exact = deg * x ** max(deg - 1, 0) / 8.0 if deg >= 1 else 0 * x
        ref = np.broadcast_to(exact, (nx,))[4:nx - 4]
        tol = 2e-5 * max(np.abs(P).max(), 1.0)
        assert np.abs(psx[12, 4:nx - 4] - ref).max() <= tol, deg
        assert np.abs(psz[4:nx - 4, 12] - ref).max() <= tol, deg


# ------------------------------------------------------------------------------------------------
# C-PML table (parallelcalls.cu:18-46)
# ------------------------------------------------------------------------------------------------

def test_cpml_table_profiles():
    """Invert the table: d+alpha = -ln(b)/dt and d = a (d+alpha)/(b-1).  The damping must be quadratic in
    the PML depth, alpha linear (zero at the outer edge), and the outer edge gives a = b - 1."""
    p = M.CpmlParams(cpml=20, dh=10.0, dt=1e-3, f=10.0)
    a, b = O.cpml_table(p, round_b=False)
    f32 = lambda v: float(np.float32(v))  # the reference's parameters are floats (main.c:38-52)
    assert np.all(b > 0) and np.all(b < 1) and np.all(a <= 0)
    s = -np.log(b) / f32(p.dt)
    d = a * s / (b - 1)
    al = s - d
    depth = (p.cpml - np.arange(p.cpml)) * f32(p.dh) / f32(p.L)
    np.testing.assert_allclose(d, f32(p.d0) * f32(p.vmax) * depth ** 2, rtol=1e-9)
    np.testing.assert_allclose(al, math.pi * f32(p.f) * (1 - depth), rtol=1e-9, atol=1e-9)
    assert abs(a[0] - (b[0] - 1)) < 1e-12
    # theoretical reflection coefficient of the quadratic profile: exp(-2/3 d0 L) = R (at unit speed)
    assert abs(math.exp(-2.0 / 3.0 * p.d0 * p.L) - p.R) / p.R < 1e-5


# ------------------------------------------------------------------------------------------------
# Ricker (main.c:328-333)
# ------------------------------------------------------------------------------------------------

def test_ricker_closed_form():
    nt, f, dt = 1000, 10.0, 1e-3
    s = O.ricker(nt, f, dt).astype(np.float64)
    # peak 1 at t = 1/f, i.e. (it+1)*dt = 0.1 -> it = 99
    assert int(np.argmax(s)) == 99 and abs(s[99] - 1.0) < 1e-6
    # zero crossings at pi f |t - 1/f| = 1/sqrt(2)
    tz = 1 / f + 1 / (math.sqrt(2) * math.pi * f)
    iz = int(round(tz / dt)) - 1
    assert abs(s[iz]) < 0.05 and s[iz - 3] * s[iz + 3] < 0
    # zero mean, symmetry about the peak
    assert abs(s.sum() * dt) < 1e-4
    np.testing.assert_allclose(s[99 - 50:99], s[100:150][::-1], atol=1e-5)
    np.testing.assert_array_equal(O.ricker(nt, f, dt), M.ricker(nt, f, dt))


# ------------------------------------------------------------------------------------------------
# Forward
# ------------------------------------------------------------------------------------------------

def _homog(nx, nz, v, eps=0.0, dl=0.0):
    return (np.full((nz, nx), v, np.float32), np.full((nz, nx), eps, np.float32), np.full((nz, nx), dl, np.float32))


@pytest.mark.parametrize("mode", ["mix", "f32", "f64", "f32d", "f32i"])
def test_isotropic_reduction(mode):
    """epsilon = delta = 0 gives v_hor = v_nmo = v_po, so P and F stay identical: the gather is 2P."""
    nx, nz, nt, cpml = 64, 56, 200, 12
    p = M.CpmlParams(cpml=cpml, dh=10.0, dt=1e-3, f=12.0)
    T = np.float64 if mode == "f64" else np.float32
    vp, eps, dl = M.gradient_model(nx, nz, seed=1, cpml=cpml, sigma=6.0)
    sq = O.pre_variables(vp.astype(T), 0 * eps.astype(T), 0 * dl.astype(T), dtype=T)
    assert np.array_equal(sq[0], sq[1]) and np.array_equal(sq[2], sq[1])
    ta, tb = O.cpml_table(p)
    src = O.ricker(nt, p.f, p.dt)
    shot = M.reference_shot(nx, cpml, 30, M.depth_row(2))
    g = O.forward(mode, sq, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt, shot)
    assert np.isfinite(g).all() and np.abs(g).max() > 0.1


def _green2d_trace(r, v, t, q):
    """Analytic P(r, t) = int q(tau) G(r, t - tau) dtau for P_tt = v^2 lap P + q(t) delta(x), with
    G = H(vt - r) / (2 pi v sqrt(v^2 t^2 - r^2)); substitution t - tau = (r/v) cosh w removes the
    singularity: P = (1 / 2 pi v^2) int_0^{acosh(vt/r)} q(t - (r/v) cosh w) dw."""
    out = np.zeros_like(t)
    for k, tk in enumerate(t):
        if v * tk <= r:
            continue
        W = math.acosh(v * tk / r)
        w = np.linspace(0.0, W, 4001)
        out[k] = np.trapezoid(q(tk - (r / v) * np.cosh(w)), w) / (2 * math.pi * v * v)
    return out


def test_forward_matches_2d_green_function():
    """Homogeneous isotropic medium: the FP64 forward trace at 400 m offset agrees with the analytic 2-D
    Green's function convolved with the discrete source (correlation > 0.999, amplitude within 2 %)."""
    nx = nz = 200
    cpml, dh, dt, f, v, nt = 20, 10.0, 1e-3, 8.0, 2500.0, 520
    p = M.CpmlParams(cpml=cpml, dh=dh, dt=dt, f=f)
    sq = O.pre_variables(*[x.astype(np.float64) for x in _homog(nx, nz, v)], dtype=np.float64)
    ta, tb = O.cpml_table(p)
    src = O.ricker(nt, f, dt).astype(np.float64)
    isx = isz = 100
    shot = M.Shot(0, nx, isx, isz, isz, cpml, nx - 2 * cpml)
    g = O.forward("f64", sq, (ta, tb), src, nx, nz, nt, cpml, dh, dt, shot)
    off = 40
    trace = g[:, isx + off - cpml]
    # source term: step n adds s[n] to the cell, i.e. q(n dt) = s[n] dh^2 / dt^2 with s[n] = ricker((n+1) dt)
    def q(tau):
        tt = tau + dt
        a = math.pi * f * (tt - 1 / f)
        val = (1 - 2 * a * a) * np.exp(-a * a)
        return np.where(tau >= 0, val, 0.0) * dh * dh / (dt * dt)
    t = (np.arange(nt) + 1) * dt
    ana = 2.0 * _green2d_trace(off * dh, v, t, q)  # gather = P + F = 2P
    corr = np.dot(trace, ana) / (np.linalg.norm(trace) * np.linalg.norm(ana))
    ratio = np.linalg.norm(trace) / np.linalg.norm(ana)
    assert corr > 0.999, corr
    assert abs(ratio - 1) < 0.02, ratio


def test_forward_linear_in_source():
    nx, nz, nt, cpml = 56, 48, 160, 10
    p = M.CpmlParams(cpml=cpml, dh=10.0, dt=1e-3, f=12.0)
    vp, eps, dl = M.gradient_model(nx, nz, seed=3, cpml=cpml, sigma=5.0)
    sq = O.pre_variables(vp.astype(np.float64), eps.astype(np.float64), dl.astype(np.float64), dtype=np.float64)
    ta, tb = O.cpml_table(p)
    rng = np.random.default_rng(0)
    s1 = rng.standard_normal(nt)
    s2 = rng.standard_normal(nt)
    shot = M.reference_shot(nx, cpml, 25, M.depth_row(3))
    args = ((ta, tb),)
    g1 = O.forward("f64", sq, *args, s1, nx, nz, nt, cpml, p.dh, p.dt, shot)
    g2 = O.forward("f64", sq, *args, s2, nx, nz, nt, cpml, p.dh, p.dt, shot)
    g12 = O.forward("f64", sq, *args, 2 * s1 - 3 * s2, nx, nz, nt, cpml, p.dh, p.dt, shot)
    np.testing.assert_allclose(g12, 2 * g1 - 3 * g2, rtol=0, atol=1e-10 * np.abs(g12).max())


@pytest.mark.parametrize("mode", ["mix", "f32", "f64", "f32d", "f32i"])
def test_forward_time_shift_invariance(mode):
    """The step has no explicit time dependence: a source delayed by m steps leaves every field exactly zero
    for m steps, then repeats the undelayed run, so the gather is the undelayed gather delayed by m steps,
    bit for bit, in every arithmetic (a wrong source index, a lagged receiver read or state that is not reset
    would break it)."""
    nx, nz, nt, cpml, m = 60, 52, 180, 10, 37
    p = M.CpmlParams(cpml=cpml, dh=10.0, dt=1e-3, f=12.0)
    T = np.float64 if mode == "f64" else np.float32
    vp, eps, dl = M.gradient_model(nx, nz, seed=4, cpml=cpml, sigma=5.0)
    sq = O.pre_variables(vp.astype(T), eps.astype(T), dl.astype(T), dtype=T)
    ta, tb = O.cpml_table(p)
    src = O.ricker(nt, p.f, p.dt).astype(T)
    late = np.zeros_like(src)
    late[m:] = src[:nt - m]
    shot = M.reference_shot(nx, cpml, 27, M.depth_row(3))
    g = O.forward(mode, sq, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt, shot)
    gl = O.forward(mode, sq, (ta, tb), late, nx, nz, nt, cpml, p.dh, p.dt, shot)
    assert not gl[:m].any()
    assert np.array_equal(gl[m:], g[:nt - m])
    assert np.abs(g[:nt - m]).max() > 0.01


def test_forward_source_receiver_reciprocity():
    """Isotropic (epsilon = delta = 0, so P = F) heterogeneous medium, FP64.  The step is
    P^{n+1} - 2 P^n + P^{n-1} = C v L P^n + s_n e_A with L the symmetric undivided 8th-order Laplacian on the
    live cells (zero outside) and v the squared velocity, i.e. (1/v) D_tt P - C L P = s_n e_A / v_A with a
    symmetric spatial operator; its Green's function is symmetric, so v_A P_B(source at A) = v_B P_A(source at B).
    The run ends before either wave reaches the C-PML strips, whose discrete update is not symmetric.  A swapped
    grid index, a transposed stencil or a source scaled by the local velocity would break the identity."""
    nx = nz = 220
    cpml, nt = 10, 240
    p = M.CpmlParams(cpml=cpml, dh=10.0, dt=1e-3, f=20.0)
    vp, _, _ = M.gradient_model(nx, nz, seed=5, cpml=cpml, sigma=12.0)
    zero = np.zeros_like(vp, np.float64)
    sq = O.pre_variables(vp.astype(np.float64), zero, zero, dtype=np.float64)
    v = sq[1]
    ta, tb = O.cpml_table(p)
    src = O.ricker(nt, p.f, p.dt).astype(np.float64)
    (xa, za), (xb, zb) = (100, 100), (125, 115)
    # every wave stays inside: 85 cells to the nearest strip, at most 0.35 cells per step for 240 steps
    assert np.sqrt(v.max()) * p.dt / p.dh * nt < min(xa - cpml, nx - cpml - xb, nz - cpml - zb)
    ga = O.forward("f64", sq, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt,
                   M.Shot(x0=0, nx=nx, isx=xa, isz=za, irz=zb, ir0=xb, nrec=1))[:, 0]
    gb = O.forward("f64", sq, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt,
                   M.Shot(x0=0, nx=nx, isx=xb, isz=zb, irz=za, ir0=xa, nrec=1))[:, 0]
    lhs, rhs = v[za, xa] * ga, v[zb, xb] * gb
    assert np.abs(lhs).max() > 1e-3 * np.abs(v[za, xa] * src).max()  # the wave has arrived
    np.testing.assert_allclose(lhs, rhs, rtol=0, atol=1e-9 * np.abs(lhs).max())
    # without the velocity weights the two traces differ (v_A != v_B here)
    assert abs(v[za, xa] / v[zb, xb] - 1) > 0.05
    assert np.abs(ga - gb).max() > 1e-3 * np.abs(ga).max()


# ------------------------------------------------------------------------------------------------
# Exact adjoint and gradients
# ------------------------------------------------------------------------------------------------

def _small_setup(nx=48, nz=40, cpml=10, nt=300, seed=2):
    p = M.CpmlParams(cpml=cpml, dh=10.0, dt=1e-3, f=12.0)
    ta, tb = O.cpml_table(p)
    vp, eps, dl = M.gradient_model(nx, nz, seed=seed, cpml=cpml, sigma=5.0)
    vt = M.gradient_model(nx, nz, seed=seed + 3, cpml=cpml, sigma=5.0)
    return p, ta, tb, (vp, eps, dl), vt


def test_exact_adjoint_dot_product():
    """<M x, y> = <x, M^T y> for one linear forward step including all C-PML memory variables (FP64)."""
    nx, nz, cpml = 48, 40, 10
    p, ta, tb, m, _ = _small_setup(nx, nz, cpml)
    sq = O.pre_variables(*[x.astype(np.float64) for x in m], dtype=np.float64)
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz))
    om = (X >= 4) & (X < nx - 4) & (Z >= 4) & (Z < nz - 4)
    chix = om & ((X < cpml) | (X > nx - cpml - 1))
    chiz = om & (Z > nz - cpml - 1)
    rng = np.random.default_rng(7)

    def rnd():
        masks = [om] * 4 + [chix, chix, chiz, chiz]
        return np.ascontiguousarray(np.stack([rng.standard_normal((nz, nx)) * mk for mk in masks]))

    x, y = rnd(), rnd()
    Mx = np.zeros_like(x)
    MTy = np.zeros_like(y)
    L = O.lib()
    L.orc_step_fwd_f64(O.ptr(sq[0]), O.ptr(sq[1]), O.ptr(sq[2]), O.ptr(ta), O.ptr(tb), nx, nz, cpml, p.dh, p.dt,
                       O.ptr(x), O.ptr(Mx))
    L.orc_step_adj_f64(O.ptr(sq[0]), O.ptr(sq[1]), O.ptr(sq[2]), O.ptr(ta), O.ptr(tb), nx, nz, cpml, p.dh, p.dt,
                       O.ptr(y), O.ptr(MTy))
    a, b = np.sum(Mx * y), np.sum(x * MTy)
    assert abs(a - b) / abs(a) < 1e-12
    # a planted error in the transpose (C-PML memory adjoint dropped) must fail the same test
    y2 = y.copy()
    y2[4:] = 0
    MTy2 = np.zeros_like(y2)
    L.orc_step_adj_f64(O.ptr(sq[0]), O.ptr(sq[1]), O.ptr(sq[2]), O.ptr(ta), O.ptr(tb), nx, nz, cpml, p.dh, p.dt,
                       O.ptr(y2), O.ptr(MTy2))
    assert abs(np.sum(x * MTy2) - b) / abs(b) > 1e-6


def _fd_check(exact_mode=True):
    nx, nz, cpml, nt = 48, 40, 10, 300
    p, ta, tb, m, vt = _small_setup(nx, nz, cpml, nt)
    src = O.ricker(nt, p.f, p.dt).astype(np.float64)
    shot = M.reference_shot(nx, cpml, 24, M.depth_row(2))
    sqt = O.pre_variables(*[x.astype(np.float64) for x in vt], dtype=np.float64)
    obs = O.forward("f64", sqt, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt, shot)
    m0 = [x.astype(np.float64) for x in m]

    def J(mm):
        s = O.pre_variables(*mm, dtype=np.float64)
        g = O.forward("f64", s, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt, shot)
        return 0.5 * np.sum((g - obs) ** 2)

    sq = O.pre_variables(*m0, dtype=np.float64)
    phi, _, _, gh, gp, gn = O.gradient("f64", exact_mode, sq, (ta, tb), src, obs, nx, nz, nt, cpml, p.dh, p.dt, shot)
    return J, m0, phi, O.chain_rule(*m0, gh, gp, gn), (nx, nz)


def test_exact_gradient_fd_directional_derivatives():
    """Central finite differences of the FP64 misfit along smooth directions of vp, epsilon and delta
    converge to <grad, dm> with O(h^2) until round-off."""
    J, m0, phi, grads, (nx, nz) = _fd_check(True)
    assert abs(phi - J(m0)) <= 1e-12 * phi
    dm = M.smooth_perturbation(nx, nz, seed=4, sigma=3.0)
    scale = [300.0, 0.05, 0.05]
    for comp in range(3):
        d = [np.zeros((nz, nx)) for _ in range(3)]
        d[comp] = dm[comp] * scale[comp]
        pred = sum(np.sum(g * dd) for g, dd in zip(grads, d))
        errs = []
        for h in (1e-2, 1e-3, 1e-4):
            fd = (J([a + h * b for a, b in zip(m0, d)]) - J([a - h * b for a, b in zip(m0, d)])) / (2 * h)
            errs.append(abs(fd - pred) / abs(pred))
        assert errs[2] < 1e-7, (comp, errs)
        assert errs[0] / errs[1] > 50, (comp, errs)  # second-order convergence


def test_exact_gradient_brute_force_per_cell():
    """Central differences of the FP64 misfit with respect to every cell of v_hor, v_po and v_nmo on a 24 x 24
    grid (rim, C-PML strips and interior; 3456 forward runs) equal the exact gradient cell by cell: a smooth
    direction cannot see a gradient shifted by one cell or a term assigned to the wrong field, this does."""
    nx = nz = 24
    cpml, nt = 6, 80
    p = M.CpmlParams(cpml=cpml, dh=10.0, dt=1e-3, f=25.0)
    ta, tb = O.cpml_table(p)
    src = O.ricker(nt, p.f, p.dt).astype(np.float64)
    shot = M.reference_shot(nx, cpml, 11, M.depth_row(2))
    Z = np.arange(nz, dtype=np.float64)[:, None] * np.ones((1, nx))

    def model(seed):  # VTI with depth trends and seeded cell-scale noise, so no two cells are alike
        r = np.random.default_rng(seed)
        vp = 2000 + 40 * Z + 60 * r.standard_normal((nz, nx))
        return O.pre_variables(vp, 0.1 + 0.004 * Z + 0.01 * r.standard_normal((nz, nx)),
                               0.05 + 0.002 * Z + 0.005 * r.standard_normal((nz, nx)), dtype=np.float64)

    sqt, sq0 = model(8), model(9)
    obs = O.forward("f64", sqt, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt, shot)

    def J(model):
        g = O.forward("f64", model, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt, shot)
        return 0.5 * np.sum((g - obs) ** 2)

    _, _, _, gh, gp, gn = O.gradient("f64", True, sq0, (ta, tb), src, obs, nx, nz, nt, cpml, p.dh, p.dt, shot)
    h = 1e-5
    for f, grad in enumerate((gh, gp, gn)):
        fd = np.zeros((nz, nx))
        for iz in range(nz):
            for ix in range(nx):
                d = h * sq0[f][iz, ix]
                up = [m.copy() for m in sq0]
                dn = [m.copy() for m in sq0]
                up[f][iz, ix] += d
                dn[f][iz, ix] -= d
                fd[iz, ix] = (J(up) - J(dn)) / (2 * d)
        # compare the sensitivities to a relative change of each cell
        scale = np.abs(grad * sq0[f]).max()
        assert scale > 0
        np.testing.assert_allclose(fd * sq0[f], grad * sq0[f], rtol=0, atol=1e-6 * scale)
        # the rim is never updated, so neither side sees it
        assert not grad[:4].any() and not grad[-4:].any() and not grad[:, :4].any() and not grad[:, -4:].any()


def test_refcompat_gradient_relates_to_exact():
    """The reference's gradient is the exact one scaled by 1/dt with a one-step lag of the forward field:
    away from the C-PML the two agree in shape (correlation > 0.99) and roughly in size."""
    nx, nz, cpml, nt = 48, 40, 10, 300
    p, ta, tb, m, vt = _small_setup(nx, nz, cpml, nt)
    src = O.ricker(nt, p.f, p.dt).astype(np.float64)
    shot = M.reference_shot(nx, cpml, 24, M.depth_row(2))
    sqt = O.pre_variables(*[x.astype(np.float64) for x in vt], dtype=np.float64)
    obs = O.forward("f64", sqt, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt, shot)
    sq = O.pre_variables(*[x.astype(np.float64) for x in m], dtype=np.float64)
    _, _, _, eh, ep, en = O.gradient("f64", True, sq, (ta, tb), src, obs, nx, nz, nt, cpml, p.dh, p.dt, shot)
    _, _, _, rh, rp, rn = O.gradient("f64", False, sq, (ta, tb), src, obs, nx, nz, nt, cpml, p.dh, p.dt, shot)
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz))
    inner = (X >= cpml + 4) & (X < nx - cpml - 4) & (Z >= 4) & (Z < nz - cpml - 4)
    for e, r in ((eh, rh), (ep, rp), (en, rn)):
        a, b = e[inner], r[inner] * p.dt
        assert np.corrcoef(a, b)[0, 1] > 0.99
        assert 0.7 < np.linalg.norm(b) / np.linalg.norm(a) < 1.4


# ------------------------------------------------------------------------------------------------
# Reference-compatible gradient (serialcalls.cu:55-90, main.c:422-492, parallelcalls.cu:144-259)
# ------------------------------------------------------------------------------------------------

_W2 = (-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0)  # Fornberg 8th-order D2 weights


def _np_d2(u, axis, h2):
    """8th-order second derivative along x (axis 1) or z (axis 0) on the interior, zero on the 4-cell rim,
    written from the Fornberg weights in FP64 (independent of the oracle's stencil code)."""
    nz, nx = u.shape
    c = _W2[0] * u[4:nz - 4, 4:nx - 4]
    for k in range(1, 5):
        if axis == 1:
            c = c + _W2[k] * (u[4:nz - 4, 4 + k:nx - 4 + k] + u[4:nz - 4, 4 - k:nx - 4 - k])
        else:
            c = c + _W2[k] * (u[4 + k:nz - 4 + k, 4:nx - 4] + u[4 - k:nz - 4 - k, 4:nx - 4])
    out = np.zeros_like(u, dtype=np.float64)
    out[4:nz - 4, 4:nx - 4] = c / h2
    return out


def _refcompat_case(mode, cpml_on, nt=100, nx=40, nz=36, cpml=8):
    """A small VTI shot (true and initial models differ) in `mode` ("f64" or "mix"); C-PML tables zeroed when
    not cpml_on.  Returns the parameters and the oracle's states of both gradient modes."""
    p = M.CpmlParams(cpml=cpml, dh=10.0, dt=1e-3, f=20.0)
    ta, tb = O.cpml_table(p)
    if not cpml_on:
        ta, tb = ta * 0, tb * 0
    T = np.float64 if mode == "f64" else np.float32
    src = O.ricker(nt, p.f, p.dt).astype(T)
    shot = M.reference_shot(nx, cpml, 17, M.depth_row(2))
    sqt = O.pre_variables(*M.gradient_model(nx, nz, seed=5, cpml=cpml, sigma=4.0), dtype=T)
    sq = O.pre_variables(*M.gradient_model(nx, nz, seed=2, cpml=cpml, sigma=4.0), dtype=T)
    obs = O.forward(mode, sqt, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt, shot)
    out = {}
    for exact in (True, False):
        out[exact] = O.gradient_states(mode, exact, sq, (ta, tb), src, obs, nx, nz, nt, cpml, p.dh, p.dt, shot)
    return p, (ta, tb), src, shot, sq, obs, out


def test_refcompat_adjoint_equals_exact_adjoint_without_cpml():
    """With the C-PML tables zeroed, the reference's adjoint recursion (entryPoint, secondDerivate, PSI_R, ZETA_R,
    retroPropagation; parallelcalls.cu:144-243) is the transpose of the forward step: the fields after its step
    it equal the exact discrete adjoint mu^(nt-it), which the dot-product, FD and per-cell brute-force tests
    pin.  A field taken one step off (the pairing a lag error would produce) differs at the percent level."""
    p, tab, src, shot, sq, obs, out = _refcompat_case("f64", cpml_on=False)
    mu = out[True][6]
    lam = out[False][6]
    scale = np.abs(mu).max()
    assert scale > 0
    assert np.abs(lam - mu).max() <= 1e-11 * scale
    assert np.abs(lam[1:] - mu[:-1]).max() > 1e-2 * scale


def test_refcompat_gradient_is_the_lagged_dt_scaled_imaging():
    """The reference's gradient (GRADIENT after each retroPropagation, with the snapshot PF_h[nt-it-1] that
    launchPropagation copied from P_past before its update; serialcalls.cu:60,76,88) is
        g_hor = dt sum_k lamP^k D2x P^(k-2),  g_nmo = dt sum_k lamF^k D2x P^(k-2),
        g_po  = dt sum_k (lamP^k + lamF^k) D2z F^(k-2)
    with lam^k the adjoint fields paired with k (the exact mu^k when the C-PML is off) and P^(-1) = 0.  It is
    assembled here in numpy (its own D2) from the oracle's FP64 forward fields and adjoint fields, with and
    without C-PML; the exact pairing (lag 0), a lag of 2 and a dt^2 scale each miss by orders of magnitude."""
    for cpml_on in (False, True):
        p, tab, src, shot, sq, obs, out = _refcompat_case("f64", cpml_on=cpml_on)
        nt = out[False][5].shape[0] - 1
        fwd = out[False][5]
        lam = out[True][6] if not cpml_on else out[False][6]
        dt, h2 = p.dt, p.dh * p.dh
        ref = out[False][2:5]  # g_hor, g_po, g_nmo of the reference-compatible mode

        def assemble(lag, scale):
            gh, gp, gn = (np.zeros_like(ref[0]) for _ in range(3))
            for k in range(1, nt + 1):
                j = k - 1 - lag
                if j < 0:
                    continue
                dxp = _np_d2(fwd[j, 0], 1, h2)
                dzf = _np_d2(fwd[j, 1], 0, h2)
                lp, lf = lam[k - 1, 0], lam[k - 1, 1]
                gh += scale * dxp * lp
                gn += scale * dxp * lf
                gp += scale * dzf * (lp + lf)
            return gh, gp, gn

        def err(g):
            return max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(g, ref))

        assert err(assemble(1, dt)) < 1e-10, cpml_on
        assert err(assemble(0, dt)) > 5e-2, cpml_on
        assert err(assemble(2, dt)) > 5e-2, cpml_on
        assert err(assemble(1, dt * dt)) > 0.5, cpml_on


def test_refcompat_step_composition_with_cpml_bit_exact():
    """With C-PML on, in the reference's arithmetic (mix: float storage, double-literal stencil sums): one
    reference adjoint step is the forward routines applied to the auxiliary fields -- aux1 = v_hor lP, aux2 =
    v_po lP, aux3 = v_nmo lF, aux4 = v_po lF (entryPoint); D2x aux1, D2z aux2, D2x aux3, D2z aux4 by
    SecondDerivate_P; PSI_R = PSI_F on (aux1, aux2) and on (aux3, aux4); ZETA_R = ZETA_F on (p2cx, p2cz, psi1,
    psi2) and on (r2dx, r2cz, psi3, psi4) -- then the retroPropagation leapfrog with the residual injected at
    time nt-it-1, and GRADIENT with the forward snapshot u^(nt-it-2).  Composed here step by step from the
    oracle's pinned per-kernel forward routines, it reproduces the oracle's reference-compatible adjoint fields
    and gradient bit for bit; a snapshot one step off or the v_po/v_hor products swapped do not."""
    nt = 80
    p, (ta, tb), src, shot, sq, obs, out = _refcompat_case("mix", cpml_on=True, nt=nt)
    nx, nz, cpml = 40, 36, 8
    L = O.lib()
    vh, vp, vn = (np.ascontiguousarray(x, np.float32) for x in sq)
    fwd = out[False][5]
    syn = out[False][1]
    res = (syn - obs).astype(np.float32)
    # C-PML coefficients per column / row (a, b = 0 off the strips; no C-PML at the top)
    ax, bx, az, bz = (np.zeros(n, np.float32) for n in (nx, nx, nz, nz))
    for ix in range(nx):
        e = ix if ix < cpml else (nx - 1 - ix if ix > nx - cpml - 1 else -1)
        if e >= 0:
            ax[ix], bx[ix] = ta[e], tb[e]
    for iz in range(nz - cpml, nz):
        az[iz], bz[iz] = ta[nz - 1 - iz], tb[nz - 1 - iz]
    dh = np.float32(p.dh)
    dt = np.float32(p.dt)
    dh2 = dh * dh
    C = (dt * dt) / dh2
    Z = lambda: np.zeros((nz, nx), np.float32)

    def run(snap_lag=2, swap=False):
        lP, lF, lPm, lFm = Z(), Z(), Z(), Z()
        psi = [Z() for _ in range(4)]
        zeta = [Z() for _ in range(4)]
        gh, gp, gn = Z(), Z(), Z()
        states = np.zeros((nt, 2, nz, nx), np.float32)
        for it in range(nt):
            a1 = (vp if swap else vh) * lP
            a2 = vp * lP
            a3 = vn * lF
            a4 = vp * lF
            p2cx, p2cz, r2dx, r2cz = Z(), Z(), Z(), Z()
            L.orc_SecondDerivate_P(0, O.ptr(a1), O.ptr(a2), O.ptr(p2cx), O.ptr(p2cz), dh2, nx, nz)
            L.orc_SecondDerivate_P(0, O.ptr(a3), O.ptr(a4), O.ptr(r2dx), O.ptr(r2cz), dh2, nx, nz)
            L.orc_PSI_F(0, O.ptr(a1), O.ptr(a2), O.ptr(ax), O.ptr(bx), O.ptr(az), O.ptr(bz), O.ptr(psi[0]),
                        O.ptr(psi[1]), nx, nz, dh, cpml)
            L.orc_PSI_F(0, O.ptr(a3), O.ptr(a4), O.ptr(ax), O.ptr(bx), O.ptr(az), O.ptr(bz), O.ptr(psi[2]),
                        O.ptr(psi[3]), nx, nz, dh, cpml)
            at = [Z() for _ in range(4)]
            L.orc_ZETA_F(0, O.ptr(p2cx), O.ptr(p2cz), O.ptr(ax), O.ptr(bx), O.ptr(az), O.ptr(bz), O.ptr(psi[0]),
                         O.ptr(psi[1]), O.ptr(zeta[0]), O.ptr(zeta[1]), O.ptr(at[0]), O.ptr(at[1]), nx, nz, dh, cpml)
            L.orc_ZETA_F(0, O.ptr(r2dx), O.ptr(r2cz), O.ptr(ax), O.ptr(bx), O.ptr(az), O.ptr(bz), O.ptr(psi[2]),
                         O.ptr(psi[3]), O.ptr(zeta[2]), O.ptr(zeta[3]), O.ptr(at[2]), O.ptr(at[3]), nx, nz, dh, cpml)
            i = (slice(4, nz - 4), slice(4, nx - 4))
            nP, nF = lPm.copy(), lFm.copy()
            nP[i] = np.float32(2) * lP[i] - lPm[i] + (dh2 * C) * (((p2cx[i] + at[0][i]) + r2dx[i]) + at[2][i])
            nF[i] = np.float32(2) * lF[i] - lFm[i] + (dh2 * C) * (((p2cz[i] + r2cz[i]) + at[1][i]) + at[3][i])
            r = res[nt - it - 1]
            nP[shot.irz, shot.ir0:shot.ir0 + shot.nrec] += r
            nF[shot.irz, shot.ir0:shot.ir0 + shot.nrec] += r
            states[nt - it - 1] = (nP, nF)
            j = nt - it - snap_lag
            Pf, Ff = (fwd[j, 0], fwd[j, 1]) if j >= 0 else (Z(), Z())
            d2p, d2f = Z(), Z()
            L.orc_SecondDerivate_P(0, O.ptr(np.ascontiguousarray(Pf)), O.ptr(np.ascontiguousarray(Ff)), O.ptr(d2p),
                                   O.ptr(d2f), dh2, nx, nz)
            gh[i] = gh[i] + (dt * d2p[i]) * nP[i]
            gp[i] = gp[i] + (dt * d2f[i]) * (nP[i] + nF[i])
            gn[i] = gn[i] + (dt * d2p[i]) * nF[i]
            lPm, lFm, lP, lF = lP, lF, nP, nF
        return states, (gh, gp, gn)

    states, g = run()
    ref_states = out[False][6]
    ref_g = out[False][2:5]
    assert np.array_equal(states, ref_states)
    for a, b in zip(g, ref_g):
        assert np.abs(b).max() > 0
        assert np.array_equal(a, b)
    _, g_lag = run(snap_lag=1)
    assert not np.array_equal(g_lag[0], ref_g[0])
    s_swap, _ = run(swap=True)
    assert not np.array_equal(s_swap, ref_states)


def test_misfit_reference_float_sum_close_to_fp64():
    J, m0, phi, _, _ = _fd_check(True)
    nx, nz, cpml, nt = 48, 40, 10, 300
    p, ta, tb, m, vt = _small_setup(nx, nz, cpml, nt)
    src = O.ricker(nt, p.f, p.dt)
    shot = M.reference_shot(nx, cpml, 24, M.depth_row(2))
    sqt = O.pre_variables(*vt)
    obs = O.forward("mix", sqt, (ta, tb), src, nx, nz, nt, cpml, p.dh, p.dt, shot)
    sq = O.pre_variables(*m)
    phi64, phi32, syn, *_ = O.gradient("mix", False, sq, (ta, tb), src, obs, nx, nz, nt, cpml, p.dh, p.dt, shot)
    res = syn.astype(np.float64) - obs.astype(np.float64)
    assert abs(phi64 - 0.5 * np.sum(res ** 2)) < 1e-9 * phi64
    assert abs(phi32 - phi64) < 1e-4 * phi64


def test_f32d_forward_is_closer_to_fp64():
    """The difference-form FP32 arithmetic computes the same operator (weights sum to zero) without the
    cancellation of the reference form: in a homogeneous isotropic medium its gather differs from the FP64
    forward by < 5e-6 relative L2, at least 4x less than the reference FP32 expression order."""
    w = M.workload("plain200")
    tab = O.cpml_table(w.cpml_params)
    src = O.ricker(w.nt, w.f, w.dt)
    model = (np.full((w.nz, w.nx), 2500.0, np.float32), np.zeros((w.nz, w.nx), np.float32),
             np.zeros((w.nz, w.nx), np.float32))
    shot = M.Shot(0, w.nx, 100, 100, 100, w.cpml, w.nx - 2 * w.cpml)
    out = {}
    for mode in ("f64", "f32", "f32d"):
        T = np.float64 if mode == "f64" else np.float32
        sq = O.pre_variables(*[x.astype(T) for x in model], dtype=T)
        out[mode] = O.forward(mode, sq, tab, src, w.nx, w.nz, w.nt, w.cpml, w.dh, w.dt, shot).astype(np.float64)
    e = lambda a: np.linalg.norm(a - out["f64"]) / np.linalg.norm(out["f64"])
    assert e(out["f32d"]) < 5e-6
    assert e(out["f32d"]) * 4 < e(out["f32"])


def test_f32i_increment_form_tracks_fp64():
    """The increment form computes the same recursion as the two-level leapfrog; on a finely sampled run
    (8 Hz at dt = 0.8 ms: 156 samples per period, VTI model, 1000 steps) its gather and exact gradient stay
    far closer to FP64 than the two-level FP32 form, whose rounding the time recursion amplifies."""
    from helpers import small_workload, oracle_setup, oracle_gathers, oracle_gradient, rel_l2

    w = small_workload(nshots=1, nt=1000, f=8.0, dt=0.8e-3)
    tab, src, sq_true, sq_init = oracle_setup(w, "f32")
    obs = oracle_gathers(w, "f32", sq_true, tab, src)
    tab64, src64, _, si64 = oracle_setup(w, "f64")
    g64 = oracle_gathers(w, "f64", si64, tab64, src64)[0]
    _, G64, _ = oracle_gradient(w, "f64", True, si64, tab64, src64, [o.astype(np.float64) for o in obs])
    err = {}
    for mode in ("f32", "f32i"):
        g = oracle_gathers(w, mode, sq_init, tab, src)[0]
        _, G, _ = oracle_gradient(w, mode, True, sq_init, tab, src, obs)
        err[mode] = (rel_l2(g, g64), [rel_l2(G[k], G64[k]) for k in range(3)])
    print(err)
    assert err["f32i"][0] < 1e-5 and err["f32i"][0] * 10 < err["f32"][0]
    assert max(err["f32i"][1]) < 1e-4
    for k in range(3):
        assert err["f32i"][1][k] * 5 < err["f32"][1][k]


def test_group_gradient_equals_per_shot_sum():
    """Gradient groups (shots whose per-step imaging terms are added over the union of their windows before
    one update; here one window is narrower) compute the same sum as one accumulator per shot: equal to 1e-14
    in FP64 and to FP32 rounding in f32i; a group of one is the per-shot function bit for bit."""
    from helpers import oracle_gathers, oracle_gradient, oracle_setup, rel_l2, small_workload

    w = small_workload(nshots=3, nt=120)
    s = w.shots[1]  # one narrower window: the group's slot spans the union of the windows
    w.shots[1] = M.Shot(x0=8, nx=w.nx - 20, isx=s.isx - 8, isz=s.isz, irz=s.irz, ir0=12, nrec=w.nx - 20 - 24)
    for mode, tol in (("f64", 1e-14), ("f32i", 1e-6)):
        tab, src, st, si = oracle_setup(w, mode)
        obs = oracle_gathers(w, mode, st, tab, src)
        p1, G1, s1 = oracle_gradient(w, mode, True, si, tab, src, obs, group=1)
        p4, G4, s4 = oracle_gradient(w, mode, True, si, tab, src, obs, group=4)
        assert p1 == p4
        for a, b in zip(s1, s4):
            assert np.array_equal(a, b)
        for k in range(3):
            assert rel_l2(G4[k], G1[k]) < tol
        # a planted mistake: one member's terms counted twice moves the result by O(1/3)
        G_bad = G4 + G1 / 3.0
        assert rel_l2(G_bad[0], G1[0]) > 0.1
    tab, src, st, si = oracle_setup(w, "f32i")
    obs = oracle_gathers(w, "f32i", st, tab, src)
    s0 = w.shots[0]
    mis, _, ux0, gh, gp, gn = O.group_gradient("f32i", si, tab, src, [obs[0]], w.nx, w.nz, w.nt, w.cpml, w.dh, w.dt,
                                               [s0])
    assert ux0 == s0.x0 and gh.shape == (w.nz, s0.nx)
    ph, _, _, h1, p1_, n1 = O.gradient("f32i", True, tuple(np.ascontiguousarray(x[:, s0.x0:s0.x0 + s0.nx]) for x in si),
                                       tab, src, obs[0], s0.nx, w.nz, w.nt, w.cpml, w.dh, w.dt, s0)
    assert mis[0] == ph and np.array_equal(gh, h1) and np.array_equal(gp, p1_) and np.array_equal(gn, n1)


def test_group_fp32_sum_is_a_small_share_of_the_fp32_error():
    """The gradient groups add four shots' imaging terms in FP32 before the FP32 slot update (the cross-shot sum over
    slots is FP64).  Against one FP32 accumulator per shot summed in FP64 this moves the gradient by less than 5 % of
    the f32i gradient's own distance to FP64 -- the in-group FP32 sum is not what limits the accuracy."""
    from helpers import oracle_gathers, oracle_gradient, oracle_setup, rel_l2, small_workload

    w = small_workload(nshots=4, nt=300, f=12.0)
    tab64, src, st64, si64 = oracle_setup(w, "f64")
    obs64 = oracle_gathers(w, "f64", st64, tab64, src.astype(np.float64))
    _, G64, _ = oracle_gradient(w, "f64", True, si64, tab64, src.astype(np.float64), obs64)
    tab, src32, _, si = oracle_setup(w, "f32i")
    obs = [o.astype(np.float32) for o in obs64]
    _, G1, _ = oracle_gradient(w, "f32i", True, si, tab, src32, obs, group=1)
    _, G4, _ = oracle_gradient(w, "f32i", True, si, tab, src32, obs, group=4)
    for k in range(3):
        e_group = rel_l2(G4[k], G1[k])
        e_fp32 = rel_l2(G1[k], G64[k])
        print(f"grad[{k}]: in-group FP32 sum {e_group:.2e}, FP32 against FP64 {e_fp32:.2e}")
        assert e_group < 0.05 * e_fp32


def test_16bit_checkpoint_rounding_matches_torch_and_numpy():
    """The oracle's 16-bit checkpoint rounding (exact power-of-two scale from the max-abs, then round to
    nearest even) equals torch's float32 -> bfloat16 and numpy's float32 -> float16 conversions of the scaled
    field, for fields of very different magnitudes, with zeros, ties and FP16 subnormals."""
    import torch

    rng = np.random.default_rng(3)
    for scale in (1e-30, 1e-3, 1.0, 7.5e4, 3e30):
        x = (rng.standard_normal(20000) * scale).astype(np.float32)
        x *= np.float32(10.0) ** rng.integers(-9, 1, x.size).astype(np.float32)  # down into FP16 subnormals
        x[:10] = 0.0
        x[10:20] = np.float32(scale) * np.float32(1 + 2 ** -8)   # BF16 ties
        x[20:30] = np.float32(scale) * np.float32(1 + 2 ** -11)  # FP16 ties
        m = np.abs(x).max()
        e = max(int(np.floor(np.log2(m))), -126)
        y = x.copy()
        O.lib().orc_ck_round(O.ptr(y), y.size, 1)
        ref = torch.from_numpy(np.ldexp(x, -e).astype(np.float32)).to(torch.bfloat16).float().numpy()
        assert np.array_equal(y, np.ldexp(ref, e).astype(np.float32)), scale
        nz = x != 0
        assert np.max(np.abs(y[nz] - x[nz]) / np.abs(x[nz])) <= 2.0 ** -8
        y = x.copy()
        O.lib().orc_ck_round(O.ptr(y), y.size, 2)
        ref = np.ldexp(np.ldexp(x, 14 - e).astype(np.float32).astype(np.float16).astype(np.float32), e - 14)
        assert np.array_equal(y, ref.astype(np.float32)), scale
        big = np.abs(x) >= m * 2.0 ** -28
        assert np.max(np.abs(y[big] - x[big]) / np.abs(x[big])) <= 2.0 ** -10


def test_16bit_checkpoint_gradient_close_to_fp32_checkpoints():
    """Emulated 16-bit checkpoints move the exact gradient (BF16 about 1e-2, FP16 about 8x less) and
    change nothing when there is no intermediate checkpoint (K >= nt)."""
    from helpers import oracle_gathers, oracle_gradient, oracle_setup, rel_l2, small_workload

    w = small_workload(nshots=2, nt=160)
    tab, src, st, si = oracle_setup(w, "f32i")
    obs = oracle_gathers(w, "f32i", st, tab, src)
    with O.surface_rows(-1):  # 16-bit checkpoints run without the FP64 band
        p0, G0, _ = oracle_gradient(w, "f32i", True, si, tab, src, obs)
    errs = {}
    for fmt in (1, 2):
        p1, G1, _ = oracle_gradient(w, "f32i", True, si, tab, src, obs, ckpt=(16, fmt))
        p2, G2, _ = oracle_gradient(w, "f32i", True, si, tab, src, obs, ckpt=(160, fmt))
        assert p0 == p1 == p2  # the forward itself never reads a checkpoint
        errs[fmt] = [rel_l2(G1[k], G0[k]) for k in range(3)]
        assert all(np.array_equal(G2[k], G0[k]) for k in range(3))
    print("16-bit checkpoints vs fp32 checkpoints (bf16, fp16):", errs)
    assert 0 < max(errs[1]) < 3e-2 and 0 < max(errs[2]) < 5e-3
    assert max(errs[2]) * 4 < max(errs[1])


def test_adjoint128_fd_fp64_over_h_range():
    """The adjoint-check configuration (128x128, dh 10 m, 2 shots, 88 live receivers each, nt = 800, dt 1 ms,
    12 Hz): central finite differences of the FP64 misfit along a seeded smooth perturbation of vp, epsilon and
    delta at once, h = 1e-2 ... 1e-6, against <grad, dm> of the exact FP64 gradient (both shots)."""
    from helpers import oracle_gathers, oracle_gradient, oracle_setup

    w = M.workload("adjoint128")
    assert (w.nx, w.nz, w.nt, w.f, len(w.shots)) == (128, 128, 800, 12.0, 2) and w.shots[0].nrec == 88
    tab, src, sq_true, sq_init = oracle_setup(w, "f64")
    src = src.astype(np.float64)
    obs = oracle_gathers(w, "f64", sq_true, tab, src)
    m0 = [x.astype(np.float64) for x in w.init_model]
    dm = M.smooth_perturbation(w.nx, w.nz, seed=11, sigma=6.0)
    d = [dm[0] * 50.0, dm[1] * 0.01, dm[2] * 0.01]

    def J(mm):
        sq = O.pre_variables(*mm, dtype=np.float64)
        return sum(0.5 * np.sum((g - o) ** 2) for g, o in zip(oracle_gathers(w, "f64", sq, tab, src), obs))

    phi, G, _ = oracle_gradient(w, "f64", True, O.pre_variables(*m0, dtype=np.float64), tab, src, obs)
    gv = O.chain_rule(*m0, G[0], G[1], G[2])
    pred = sum(np.sum(g * dd) for g, dd in zip(gv, d))
    errs = []
    for h in (1e-2, 1e-3, 1e-4, 1e-5, 1e-6):
        fd = (J([a + h * b for a, b in zip(m0, d)]) - J([a - h * b for a, b in zip(m0, d)])) / (2 * h)
        errs.append(abs(fd - pred) / abs(pred))
    print("adjoint128 FP64 FD relative errors, h = 1e-2..1e-6:", errs)
    assert errs[0] / errs[1] > 50 and errs[1] / errs[2] > 50  # O(h^2) until round-off
    assert min(errs) < 1e-9
