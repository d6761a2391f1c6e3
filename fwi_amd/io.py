"""Model and gather files in the reference's raw format (src/main.c:101-122, :823-834).

Raw little-endian float32, nx*nz values, x fastest.  The reference reads the true model from
vp_ori.bin (vp), vh_ori.bin (epsilon), vn_ori.bin (delta) and the initial model from the *_ini.bin files,
and at the end writes the squared velocities v_hor, v_po, v_nmo back into vh_ini.bin, vp_ini.bin,
vn_ini.bin -- files that its next run reads as vp, epsilon, delta.  `write_result` keeps explicit names
(v_hor.bin, v_po.bin, v_nmo.bin plus vp.bin, epsilon.bin, delta.bin); `write_reference_outputs`
reproduces the reference's naming when a drop-in is wanted.
"""
from __future__ import annotations

import os

import numpy as np


def read_raw(path: str, nx: int, nz: int) -> np.ndarray:
    a = np.fromfile(path, dtype="<f4")
    if a.size != nx * nz:
        raise ValueError(f"{path}: {a.size} floats, expected nx*nz = {nx * nz}")
    return a.reshape(nz, nx)


def write_raw(path: str, a: np.ndarray):
    np.ascontiguousarray(a, dtype="<f4").tofile(path)


def read_models(directory: str, nx: int, nz: int):
    """((vp, eps, delta) true, (vp, eps, delta) initial) from the reference's six input files."""
    r = lambda name: read_raw(os.path.join(directory, name), nx, nz)
    true = (r("vp_ori.bin"), r("vh_ori.bin"), r("vn_ori.bin"))
    init = (r("vp_ini.bin"), r("vh_ini.bin"), r("vn_ini.bin"))
    return true, init


def write_models(directory: str, true, init):
    os.makedirs(directory, exist_ok=True)
    for name, a in zip(("vp_ori.bin", "vh_ori.bin", "vn_ori.bin", "vp_ini.bin", "vh_ini.bin", "vn_ini.bin"),
                       (*true, *init)):
        write_raw(os.path.join(directory, name), a)


def squared_to_vti(v_hor, v_po, v_nmo):
    """(vp, epsilon, delta) from the squared velocities (inverse of pre_variables)."""
    v_po = np.asarray(v_po, np.float64)
    vp = np.sqrt(np.maximum(v_po, 0.0))
    eps = (np.asarray(v_hor, np.float64) / v_po - 1.0) / 2.0
    dl = (np.asarray(v_nmo, np.float64) / v_po - 1.0) / 2.0
    return vp.astype(np.float32), eps.astype(np.float32), dl.astype(np.float32)


def write_result(directory: str, v_hor, v_po, v_nmo):
    os.makedirs(directory, exist_ok=True)
    for name, a in (("v_hor.bin", v_hor), ("v_po.bin", v_po), ("v_nmo.bin", v_nmo)):
        write_raw(os.path.join(directory, name), a)
    for name, a in zip(("vp.bin", "epsilon.bin", "delta.bin"), squared_to_vti(v_hor, v_po, v_nmo)):
        write_raw(os.path.join(directory, name), a)


def write_reference_outputs(directory: str, v_hor, v_po, v_nmo):
    """The reference's output naming (main.c:823-834): squared velocities in vh_ini / vp_ini / vn_ini."""
    os.makedirs(directory, exist_ok=True)
    write_raw(os.path.join(directory, "vh_ini.bin"), v_hor)
    write_raw(os.path.join(directory, "vp_ini.bin"), v_po)
    write_raw(os.path.join(directory, "vn_ini.bin"), v_nmo)


def write_gathers(path: str, gathers: np.ndarray):
    """Survey gathers [shot][it][receiver] as raw float32."""
    write_raw(path, gathers)
