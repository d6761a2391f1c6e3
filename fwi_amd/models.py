"""Seeded synthetic VTI models, source wavelet, C-PML parameters and survey geometry.

The reference reads its models from raw files that are not shipped (src/main.c:101-122), so every
workload here is generated from a seed.  Models are (vp [m/s], epsilon, delta) arrays of shape (nz, nx),
row-major with x fastest -- the reference's I(ix,iz) = ix + nx*iz layout (include/parallelcalls.h:5).

Depth mapping of "z" in workload descriptions: the reference never updates a 4-cell rim, so a source
placed literally at row z would sit in that frozen rim.  Here a depth z (in cells below the top of the
live grid) maps to row iz = 4 + z.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

RIM = 4  # stencil radius; rows/columns 0..3 and n-4..n-1 are never updated


def depth_row(z: int) -> int:
    """Row index of a depth given in cells below the top of the live grid."""
    return RIM + int(z)


# --------------------------------------------------------------------------------------------
# Wavelet and C-PML parameters (src/main.c:44-55, :328-333)
# --------------------------------------------------------------------------------------------

def ricker(nt: int, f: float, dt: float) -> np.ndarray:
    """Ricker wavelet as in src/main.c:328-333: t accumulates in float32 (t += dt), delay 1/f."""
    out = np.empty(nt, np.float32)
    t = np.float32(0.0)
    f32 = np.float32(f)
    dt32 = np.float32(dt)
    for it in range(nt):
        t = np.float32(t + dt32)
        a = math.pi * float(f32) * float(np.float32(t - np.float32(1.0) / f32))
        out[it] = np.float32((1.0 - 2.0 * a * a) * math.exp(-a * a))
    return out


@dataclasses.dataclass(frozen=True)
class CpmlParams:
    """C-PML parameters of src/main.c:48-52: R = 1e-6, L = cpml*dh, d0 = -3 ln R / (2L), Vmax = 6000."""

    cpml: int
    dh: float
    dt: float
    f: float
    vmax: float = 6000.0
    R: float = 1e-6

    @property
    def L(self) -> float:
        return float(np.float32(self.cpml * np.float32(self.dh)))

    @property
    def d0(self) -> float:
        # float d0 = -3*log(R)/(2*L) with float R and L (main.c:48-50)
        return float(np.float32(-3.0 * math.log(float(np.float32(self.R))) / (2.0 * self.L)))


# --------------------------------------------------------------------------------------------
# Models
# --------------------------------------------------------------------------------------------

def _gaussian_blur(a: np.ndarray, sigma: float) -> np.ndarray:
    from scipy.ndimage import gaussian_filter

    return gaussian_filter(a, sigma=sigma, mode="nearest")


def gradient_model(nx: int, nz: int, seed: int = 0, anomalies: int = 3, amp: float = 0.08,
                   sigma: float = 15.0, cpml: int = 20):
    """Small-grid VTI model: vp linear in depth 1500->3500 m/s, epsilon 0->0.2, delta 0->0.1, times
    `anomalies` seeded Gaussian perturbations (+-amp, width sigma cells) centred in the non-C-PML interior."""
    rng = np.random.default_rng(seed)
    zz = np.arange(nz, dtype=np.float64)[:, None] / max(nz - 1, 1)
    vp = 1500.0 + 2000.0 * zz * np.ones((1, nx))
    eps = 0.2 * zz * np.ones((1, nx))
    dl = 0.1 * zz * np.ones((1, nx))
    X, Z = np.meshgrid(np.arange(nx), np.arange(nz))
    pert = np.zeros((nz, nx))
    lo = cpml + 5
    for _ in range(anomalies):
        cx = rng.uniform(lo, nx - 1 - lo)
        cz = rng.uniform(RIM + 10, nz - 1 - lo)
        sgn = 1.0 if rng.random() < 0.5 else -1.0
        pert += sgn * amp * np.exp(-((X - cx) ** 2 + (Z - cz) ** 2) / (2 * sigma ** 2))
    vp = vp * (1 + pert)
    eps = eps * (1 + pert)
    dl = dl * (1 + pert)
    return vp.astype(np.float32), eps.astype(np.float32), dl.astype(np.float32)


def smooth_model(vp, eps, dl, width: float = 20.0):
    """Initial model: the true model smoothed by a Gaussian of `width`/2 cells standard deviation."""
    s = width / 2.0
    return (_gaussian_blur(vp.astype(np.float64), s).astype(np.float32),
            _gaussian_blur(eps.astype(np.float64), s).astype(np.float32),
            _gaussian_blur(dl.astype(np.float64), s).astype(np.float32))


def layered_model(nx: int, nz: int, seed: int = 0, layers: int = 12, faults: int = 3,
                  vp_range=(1500.0, 4500.0), eps_max: float = 0.25, delta_max: float = 0.12,
                  salt: bool = False):
    """Seeded layered VTI model with dipping fault offsets (and optionally a salt body with vp 4500,
    epsilon = delta = 0).  Layer properties increase with depth; epsilon >= delta everywhere."""
    rng = np.random.default_rng(seed)
    top = RIM + max(4, nz // 40)
    cuts = np.sort(rng.uniform(top, nz - 8, size=layers - 1))
    vps = np.linspace(vp_range[0], vp_range[1], layers) * (1 + rng.uniform(-0.03, 0.03, layers))
    vps[0] = vp_range[0]
    epss = np.linspace(0.0, eps_max, layers) * rng.uniform(0.85, 1.0, layers)
    dls = np.minimum(np.linspace(0.0, delta_max, layers) * rng.uniform(0.85, 1.0, layers), epss)
    epss[0] = dls[0] = 0.0
    X, Z = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(nz, dtype=np.float64))
    shift = np.zeros_like(Z)
    for _ in range(faults):
        x0 = rng.uniform(0.15 * nx, 0.85 * nx)
        dip = math.radians(rng.uniform(55, 75))
        throw = rng.uniform(0.02, 0.05) * nz * (1 if rng.random() < 0.5 else -1)
        xf = x0 + (Z - top) / math.tan(dip)
        shift += np.where((X > xf) & (Z > top), throw, 0.0)
    zeff = Z - shift
    idx = np.searchsorted(cuts, zeff)
    idx = np.clip(idx, 0, layers - 1)
    vp = vps[idx]
    eps = epss[idx]
    dl = dls[idx]
    if salt:
        cx, cz = 0.55 * nx, 0.55 * nz
        rx, rz = 0.12 * nx, 0.18 * nz
        body = ((X - cx) / rx) ** 2 + ((Z - cz) / rz) ** 2 + 0.15 * np.sin(X / (0.03 * nx)) < 1.0
        vp = np.where(body, 4500.0, vp)
        eps = np.where(body, 0.0, eps)
        dl = np.where(body, 0.0, dl)
    return vp.astype(np.float32), eps.astype(np.float32), dl.astype(np.float32)


def smooth_perturbation(nx: int, nz: int, seed: int = 1, sigma: float = 6.0):
    """Seeded smooth random direction (Gaussian-filtered noise) for (vp, epsilon, delta), unit max-norm."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(3):
        a = _gaussian_blur(rng.standard_normal((nz, nx)), sigma)
        out.append(a / np.abs(a).max())
    return out


# --------------------------------------------------------------------------------------------
# Geometry
# --------------------------------------------------------------------------------------------

@dataclasses.dataclass
class Shot:
    """One shot in window-local coordinates.

    The shot's grid is the model columns [x0, x0 + nx) (a windowed subdomain with its own C-PML on the
    window's left/right edges).  Source at (isx, isz); receivers on row irz at columns [ir0, ir0+nrec).
    """

    x0: int
    nx: int
    isx: int
    isz: int
    irz: int
    ir0: int
    nrec: int


def reference_shot(nx: int, cpml: int, ix: int, iz: int) -> Shot:
    """Full-width shot with the reference's receiver spread: row of the source, live columns
    cpml..nx-cpml-1 (src/parallelcalls.cu:113)."""
    return Shot(x0=0, nx=nx, isx=ix, isz=iz, irz=iz, ir0=cpml, nrec=nx - 2 * cpml)


def even_shots(nx: int, cpml: int, nshots: int, iz: int):
    """nshots sources evenly spaced along row iz between the C-PML strips, full-width windows."""
    lo, hi = cpml, nx - cpml - 1
    if nshots == 1:
        xs = [(lo + hi) // 2]
    else:
        xs = [lo + int(round(i * (hi - lo) / (nshots - 1))) for i in range(nshots)]
    return [reference_shot(nx, cpml, x, iz) for x in xs]


def uneven_shots(nx: int, cpml: int, nshots: int, iz: int, seed: int = 0,
                 spread=(200, 1200), window=(300, 1300)):
    """Seeded ragged survey: per-shot windowed subdomain (live width in `window` plus 2*cpml) and a
    receiver spread of `spread` traces centred on the source inside the window's live columns."""
    rng = np.random.default_rng(seed)
    shots = []
    for _ in range(nshots):
        live = int(rng.integers(window[0], window[1] + 1))
        wn = min(live + 2 * cpml, nx)
        live = wn - 2 * cpml
        x0 = int(rng.integers(0, nx - wn + 1))
        isx = cpml + int(rng.integers(0, live))
        nrec = min(int(rng.integers(spread[0], spread[1] + 1)), live)
        ir0 = min(max(cpml, isx - nrec // 2), cpml + live - nrec)
        shots.append(Shot(x0=x0, nx=wn, isx=isx, isz=iz, irz=iz, ir0=ir0, nrec=nrec))
    return shots


# --------------------------------------------------------------------------------------------
# Named workloads
# --------------------------------------------------------------------------------------------

@dataclasses.dataclass
class Workload:
    name: str
    nx: int
    nz: int
    dh: float
    dt: float
    nt: int
    f: float
    cpml: int
    true_model: tuple
    init_model: tuple
    shots: list

    @property
    def cpml_params(self) -> CpmlParams:
        return CpmlParams(cpml=self.cpml, dh=self.dh, dt=self.dt, f=self.f)


def workload(name: str, seed: int = 0, nshots: int | None = None, nt: int | None = None) -> Workload:
    """Build a named workload.

    plain200  : 200x200, dh=10, 20-cell C-PML, 1 shot at (z=2, x=100), 10 Hz, nt=1000, dt=1 ms
    adjoint128: 128x128, dh=10, 2 shots (x=43, 85), 12 Hz, nt=800, dt=1 ms
    survey1600: 1600x400 layered+faults, 64 shots on the top, 8 Hz, nt=4000, dt=0.8 ms
    uneven1600: same model, 96 shots with ragged windows (300-1300 live columns) and spreads (200-1200)
    large4000 : 4000x1000 layered+faults+salt, 256 shots, nt=6000, dt=0.8 ms, 3 Hz first stage
    """
    cpml = 20
    if name == "plain200":
        nx = nz = 200
        vt = gradient_model(nx, nz, seed=seed, cpml=cpml)
        vi = smooth_model(*vt)
        shots = [reference_shot(nx, cpml, 100, depth_row(2))]
        return Workload(name, nx, nz, 10.0, 1e-3, nt or 1000, 10.0, cpml, vt, vi, shots)
    if name == "adjoint128":
        nx = nz = 128
        vt = gradient_model(nx, nz, seed=seed, cpml=cpml, sigma=10.0)
        vi = smooth_model(*vt)
        shots = [reference_shot(nx, cpml, x, depth_row(2)) for x in (43, 85)]
        return Workload(name, nx, nz, 10.0, 1e-3, nt or 800, 12.0, cpml, vt, vi, shots)
    if name in ("survey1600", "uneven1600"):
        nx, nz = 1600, 400
        vt = layered_model(nx, nz, seed=seed)
        vi = smooth_model(*vt, width=40.0)
        if name == "survey1600":
            shots = even_shots(nx, cpml, nshots or 64, depth_row(2))
        else:
            shots = uneven_shots(nx, cpml, nshots or 96, depth_row(2), seed=seed + 7)
        return Workload(name, nx, nz, 10.0, 8e-4, nt or 4000, 8.0, cpml, vt, vi, shots)
    if name == "large4000":
        nx, nz = 4000, 1000
        vt = layered_model(nx, nz, seed=seed, salt=True)
        vi = smooth_model(*vt, width=60.0)
        shots = even_shots(nx, cpml, nshots or 256, depth_row(2))
        return Workload(name, nx, nz, 10.0, 8e-4, nt or 6000, 3.0, cpml, vt, vi, shots)
    raise ValueError(f"unknown workload {name!r}")
