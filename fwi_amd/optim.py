"""Steepest descent, L-BFGS and the line search of src/main.c:539-800 on the device optimizer.

`LBFGS` wraps fwi_lbfgs_* (include/fwi.h): FP64 two-loop recursion with device dot products.
`LineSearch` is the reference's backtracking state machine (main.c:690-786) on the library's host-only
fwi_line_search_* functions, which the C driver runs too, so both drivers' behaviour is tested on scripted
misfits without a GPU:
  * a trial is accepted when its misfit <= the misfit of the current iteration;
  * otherwise, or when the misfit is NaN (then taken as 1e30), every step length is halved and `count`
    grows; when count > 5 the backed-up model is restored and the stage ends;
  * step lengths start at 1 per stage and are never reset between iterations, `count` accumulates over
    the stage (main.c:306-311).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .survey import _stream_ptr

SD_STEP = (3e5, 2e5, 2.5e5)  # v_hor, v_po, v_nmo step lengths of the first iteration (main.c:567)


def clamp_model(model, vmin: float, vmax: float, stream=None):
    """Physical bounds on the squared velocities (v_hor, v_po, v_nmo): each to [vmin^2, vmax^2], then
    v_hor >= v_nmo (fwi_model_clamp; the reference has only the last bound)."""
    for t in model:
        assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
    _lib.check(_lib.load().fwi_model_clamp(*[t.data_ptr() for t in model], model[0].numel(), float(vmin) ** 2,
                                           float(vmax) ** 2, _stream_ptr(stream)), "fwi_model_clamp")


class LBFGS:
    """Device optimizer state for the blocks (v_hor, v_po, v_nmo) of an nz x nx model, memory m."""

    def __init__(self, nx: int, nz: int, m: int = 5):
        h = ctypes.c_void_p()
        _lib.check(_lib.load().fwi_lbfgs_create(nx, nz, m, ctypes.byref(h)), "fwi_lbfgs_create")
        self._h = h
        self.nx, self.nz, self.m = nx, nz, m

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().fwi_lbfgs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ptrs(model):
        for t in model:
            assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
        return [t.data_ptr() for t in model]

    def sd_step(self, model, grad: torch.Tensor, alpha=SD_STEP, stream=None):
        a = (ctypes.c_double * 3)(*alpha)
        _lib.check(_lib.load().fwi_lbfgs_sd_step(self._h, *self._ptrs(model), grad.contiguous().data_ptr(), a,
                                                 _stream_ptr(stream)), "fwi_lbfgs_sd_step")

    def direction(self, model, grad: torch.Tensor, stream=None):
        _lib.check(_lib.load().fwi_lbfgs_direction(self._h, *self._ptrs(model), grad.contiguous().data_ptr(),
                                                   _stream_ptr(stream)), "fwi_lbfgs_direction")

    def trial(self, model, alpha, stream=None):
        a = (ctypes.c_double * 3)(*alpha)
        _lib.check(_lib.load().fwi_lbfgs_trial(self._h, *self._ptrs(model), a, _stream_ptr(stream)),
                   "fwi_lbfgs_trial")

    def restore(self, model, stream=None):
        _lib.check(_lib.load().fwi_lbfgs_restore(self._h, *self._ptrs(model), _stream_ptr(stream)),
                   "fwi_lbfgs_restore")

    def get_direction(self, stream=None) -> torch.Tensor:
        out = torch.empty((3, self.nz, self.nx), dtype=torch.float64, device="cuda")
        _lib.check(_lib.load().fwi_lbfgs_get_direction(self._h, out.data_ptr(), _stream_ptr(stream)),
                   "fwi_lbfgs_get_direction")
        return out

    @property
    def pairs(self) -> int:
        return int(_lib.load().fwi_lbfgs_pairs(self._h))


class LineSearch:
    """Backtracking line search of main.c:690-786 (per stage state: step lengths and failure count), on the
    library's state machine (fwi_line_search_init / fwi_line_search_step, host code that the C driver runs too)."""

    def __init__(self, max_failures: int = 5):
        self._L = _lib.load()
        self._ls = _lib.FwiLineSearch()
        self._L.fwi_line_search_init(ctypes.byref(self._ls))
        self._ls.max_failures = max_failures

    @property
    def alpha(self):
        return [float(a) for a in self._ls.alpha]

    @property
    def count(self) -> int:
        return int(self._ls.count)

    def run(self, phi_n: float, evaluate, restore):
        """evaluate(alpha) sets the trial model and returns its misfit; restore() restores the backed-up
        model.  Returns (accepted misfit or None when the stage ends, number of trials)."""
        trials = 0
        while True:
            phi = evaluate(tuple(self.alpha))
            trials += 1
            t = ctypes.c_double(float("nan") if phi is None else float(phi))
            r = self._L.fwi_line_search_step(ctypes.byref(self._ls), float(phi_n), ctypes.byref(t))
            if r < 0:
                raise RuntimeError(self._L.fwi_error_message().decode())
            if r == 2:
                restore()
                return None, trials
            if r == 0:
                return t.value, trials
