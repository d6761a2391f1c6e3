"""Write the golden fixtures of the FWI gradient path from the CPU oracle (test infrastructure).

The reference ships no tests, fixtures or known-answer vectors, so these fixtures are written by this
project's oracle (oracle/fwi_oracle.cpp, oracle/lbfgs_oracle.py), whose functions are pinned by
tests/test_oracle_pins.py and tests/test_host_logic.py.  They freeze the oracle's outputs on one small seeded
case so that any later change to the oracle or to the device path is caught without re-deriving them:

  grid 64 x 64, dh = 10 m, cpml = 12, dt = 1 ms, nt = 200, 15 Hz Ricker, one shot at x = 32 on row 6, receivers
  on row 6 over the live columns; true model of the plain200 family (seed 0), initial model its smooth version.

Contents (golden_small64.npz, read with numpy.load(allow_pickle=False)):
  ricker                    the wavelet (float32, nt)
  cpml_a, cpml_b            the C-PML table (float64, cpml)
  sq_true_{vh,vp,vn}        pre_variables of the true model (float32)
  sq_init_{vh,vp,vn}        pre_variables of the initial model (float32)
  obs                       observed gather: the FP64 forward of the true model, rounded to float32 [nt, nrec]
  gather_{mix,f32i,f64}     forward gathers of the true model in the reference's mixed, FP32 increment and
                            FP64 arithmetic (f32i: increment-form leapfrog in the forward and the exact adjoint,
                            rows 4..23 of the forward in FP64 -- the library's default)
  grad_exact_{f32i,f64}     exact gradient (v_hor, v_po, v_nmo) [3, nz, nx] at the initial model against obs
                            (float32 and float64: each mode's own storage type)
  grad_ref_mix              reference-compatible gradient in the reference's arithmetic (float32)
  misfit_{exact_f32i,exact_f64,ref_mix}   the FP64 misfits
  sd_model                  one steepest-descent step (alpha = 3e5, 2e5, 2.5e5) from the initial model along
                            grad_exact_f32i (float32 [3, nz, nx])
  lbfgs_grad                exact f32i gradient at sd_model
  lbfgs_dir, lbfgs_trial    the L-BFGS direction (k = 0, one curvature pair) and the trial model at alpha = 1

Run:  python tests/golden/make_fixtures.py   (needs the oracle library, built by `make -C oracle`)
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402
from fwi_amd import models as M  # noqa: E402

OUT = os.path.join(HERE, "golden_small64.npz")
META = os.path.join(HERE, "golden_small64.json")
SEED = 0
SD_ALPHA = (3e5, 2e5, 2.5e5)


def workload():
    nx = nz = 64
    cpml = 12
    vt = M.gradient_model(nx, nz, seed=SEED, cpml=cpml, sigma=6.0)
    vi = M.smooth_model(*vt, width=8.0)
    shots = [M.reference_shot(nx, cpml, 32, M.depth_row(2))]
    return M.Workload("golden_small64", nx, nz, 10.0, 1e-3, 200, 15.0, cpml, vt, vi, shots)


def compute():
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
    from lbfgs_oracle import LbfgsOracle

    w = workload()
    sh = w.shots[0]
    tab = O.cpml_table(w.cpml_params)
    src = O.ricker(w.nt, w.f, w.dt)
    out = {"ricker": src, "cpml_a": tab[0], "cpml_b": tab[1]}
    sq = {}
    for mode, T in (("f32", np.float32), ("f64", np.float64)):
        sq[mode] = (O.pre_variables(*[x.astype(T) for x in w.true_model], dtype=T),
                    O.pre_variables(*[x.astype(T) for x in w.init_model], dtype=T))
    for name, arr in zip(("vh", "vp", "vn"), sq["f32"][0]):
        out[f"sq_true_{name}"] = arr
    for name, arr in zip(("vh", "vp", "vn"), sq["f32"][1]):
        out[f"sq_init_{name}"] = arr

    def run_forward(mode, model):
        return O.forward(mode, model, tab, src, w.nx, w.nz, w.nt, w.cpml, w.dh, w.dt, sh)

    def run_gradient(mode, exact, model, obs):
        ph, _, _, gh, gp, gn = O.gradient(mode, exact, model, tab, src, obs, w.nx, w.nz, w.nt, w.cpml, w.dh, w.dt,
                                          sh)
        return ph, np.stack([gh, gp, gn])

    obs = run_forward("f64", sq["f64"][0]).astype(np.float32)
    out["obs"] = obs
    out["gather_mix"] = run_forward("mix", sq["f32"][0])
    out["gather_f32i"] = run_forward("f32i", sq["f32"][0])
    out["gather_f64"] = run_forward("f64", sq["f64"][0])
    ph, g = run_gradient("f32i", True, sq["f32"][1], obs)
    out["grad_exact_f32i"], out["misfit_exact_f32i"] = g, np.float64(ph)
    ph, g = run_gradient("f64", True, sq["f64"][1], obs.astype(np.float64))
    out["grad_exact_f64"], out["misfit_exact_f64"] = g, np.float64(ph)
    ph, g = run_gradient("mix", False, sq["f32"][1], obs)
    out["grad_ref_mix"], out["misfit_ref_mix"] = g, np.float64(ph)

    orc = LbfgsOracle(5)
    init = [np.ascontiguousarray(x) for x in sq["f32"][1]]
    sd = orc.sd_step(init, list(out["grad_exact_f32i"]), SD_ALPHA)
    out["sd_model"] = np.stack(sd)
    _, g2 = run_gradient("f32i", True, tuple(np.ascontiguousarray(x) for x in sd), obs)
    out["lbfgs_grad"] = g2
    r = orc.direction(sd, list(g2))
    out["lbfgs_dir"] = np.stack(r)
    out["lbfgs_trial"] = np.stack(orc.trial((1.0, 1.0, 1.0)))
    meta = {"seed": SEED, "nx": w.nx, "nz": w.nz, "dh": w.dh, "dt": w.dt, "nt": w.nt, "f": w.f, "cpml": w.cpml,
            "shot": {"isx": sh.isx, "isz": sh.isz, "irz": sh.irz, "ir0": sh.ir0, "nrec": sh.nrec},
            "sd_alpha": SD_ALPHA, "lbfgs_trial_alpha": 1.0, "lbfgs_memory": 5,
            "obs": "FP64 oracle forward of the true model, rounded to float32",
            "f32i_surface_rows": int(O.lib().orc_surface_rows(w.nz, w.cpml))}
    return out, meta


def main():
    out, meta = compute()
    np.savez_compressed(OUT, **out)
    with open(META, "w") as fh:
        json.dump(meta, fh, indent=1)
    print(f"wrote {OUT} ({os.path.getsize(OUT) // 1024} KiB) and {META}")


if __name__ == "__main__":
    main()
