"""Host-side mirror of the batched FWI gradient path (include/fwi.h) on torch device tensors.

`Plan` wraps fwi_plan_*: it owns the device workspace of a survey and exposes the operations the
reference's driver performs per FWI iteration (src/main.c:334-366 observed data, :396-538 gradient,
:695-737 line-search misfit) on all shots at once.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from .models import CpmlParams, Shot

GRAD_REFCOMPAT = 0
GRAD_EXACT = 1
ARITH_FP32 = 0
ARITH_MIXED = 1
ARITH_FP32D = 2
ARITH_FP32I = 3  # default: FP32 with the forward leapfrog in increment form
CKPT_FP32 = 0
CKPT_BF16 = 1  # 16-bit checkpoints, power-of-two scale per field, shot and snapshot
CKPT_FP16 = 2


def cpml_table(p: CpmlParams):
    """C-PML (a, b) tables by distance from the edge (cpml entries, float32), evaluated by the library."""
    L = _lib.load()
    a = (ctypes.c_float * p.cpml)()
    b = (ctypes.c_float * p.cpml)()
    L.fwi_cpml_table(p.cpml, p.vmax, p.dt, p.dh, p.f, p.d0, p.L, a, b)
    return np.frombuffer(a, np.float32).copy(), np.frombuffer(b, np.float32).copy()


def pre_variables(vp: torch.Tensor, eps: torch.Tensor, delta: torch.Tensor):
    """(v_hor, v_po, v_nmo) = (vp^2 (1+2 eps), vp^2, vp^2 (1+2 delta)) on the device (pre_variables_h)."""
    L = _lib.load()
    nz, nx = vp.shape
    vp, eps, delta = (t.contiguous().float() for t in (vp, eps, delta))
    vh = torch.empty_like(vp)
    vpo = torch.empty_like(vp)
    vn = torch.empty_like(vp)
    L.pre_variables_h(vp.data_ptr(), eps.data_ptr(), delta.data_ptr(), vh.data_ptr(), vpo.data_ptr(), vn.data_ptr(),
                      nx, nz)
    _lib.check_last("pre_variables_h")
    return vh, vpo, vn


def _stream_ptr(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class Plan:
    """A survey prepared for batched forward / misfit / gradient evaluation on one GPU.

    Parameters mirror fwi_survey / fwi_options.  `shots` are window-local Shot records.
    """

    def __init__(self, nx: int, nz: int, nt: int, dh: float, dt: float, cpml: int, shots: Sequence[Shot],
                 wavelet: np.ndarray, cpml_ab, grad_mode: int = GRAD_EXACT, arith: int = ARITH_FP32I,
                 ckpt_interval: int = 0, max_batch: int = 0, mem_budget: float = 0.0, rows_per_item: int = 0,
                 item_order: int = -1, grad_group: int = 0, ckpt_format: int = 0, use_graph: int = 1,
                 store_segments: int = 0, step_pairs: int = 1, surface_rows: int = 0, adj_rows_per_item: int = 0,
                 adj_item_order: int = -1, adj_step_pairs: int = 1):
        L = _lib.load()
        self.nx, self.nz, self.nt, self.dh, self.dt, self.cpml = nx, nz, nt, dh, dt, cpml
        self.shots = list(shots)
        self._shot_arr = (_lib.FwiShot * len(self.shots))(
            *[_lib.FwiShot(s.x0, s.nx, s.isx, s.isz, s.irz, s.ir0, s.nrec) for s in self.shots])
        self._wav = np.ascontiguousarray(wavelet, np.float32)
        assert self._wav.size == nt
        ta, tb = (np.ascontiguousarray(x, np.float32) for x in cpml_ab)
        assert ta.size == cpml and tb.size == cpml
        self._ta, self._tb = ta, tb
        fp = ctypes.POINTER(ctypes.c_float)
        sv = _lib.FwiSurvey(nx, nz, nt, cpml, dh, dt, len(self.shots), self._shot_arr,
                            self._wav.ctypes.data_as(fp), ta.ctypes.data_as(fp), tb.ctypes.data_as(fp))
        opt = _lib.FwiOptions()
        L.fwi_options_default(ctypes.byref(opt))
        opt.grad_mode = grad_mode
        opt.arith = arith
        opt.ckpt_interval = ckpt_interval
        opt.max_batch = max_batch
        opt.mem_budget = mem_budget
        opt.rows_per_item = rows_per_item
        opt.item_order = item_order
        opt.grad_group = grad_group
        opt.ckpt_format = ckpt_format
        opt.use_graph = use_graph
        opt.store_segments = store_segments
        opt.step_pairs = step_pairs
        opt.surface_rows = surface_rows
        opt.adj_rows_per_item = adj_rows_per_item
        opt.adj_item_order = adj_item_order
        opt.adj_step_pairs = adj_step_pairs
        h = ctypes.c_void_p()
        _lib.check(L.fwi_plan_create(ctypes.byref(sv), ctypes.byref(opt), ctypes.byref(h)), "fwi_plan_create")
        self._h = h
        self._bufs = {}
        self.grad_mode = grad_mode
        self.gather_size = int(L.fwi_plan_gather_size(h))
        offs = np.cumsum([0] + [nt * s.nrec for s in self.shots])
        self.gather_offsets = offs

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().fwi_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def batch_size(self) -> int:
        return int(_lib.load().fwi_plan_batch_size(self._h))

    @property
    def batch_count(self) -> int:
        return int(_lib.load().fwi_plan_batch_count(self._h))

    @property
    def workspace_bytes(self) -> float:
        return float(_lib.load().fwi_plan_workspace_bytes(self._h))

    @property
    def ckpt_interval(self) -> int:
        return int(_lib.load().fwi_plan_ckpt_interval(self._h))

    @property
    def pair_rows(self) -> int:
        """Rows per fused core item of the step pairs (0: single steps only)."""
        return int(_lib.load().fwi_plan_pair_rows(self._h))

    @property
    def adj_pair_rows(self) -> int:
        """Rows per core item of the adjoint step pairs (0: single adjoint steps only)."""
        return int(_lib.load().fwi_plan_adj_pair_rows(self._h))

    @property
    def kept_steps(self) -> int:
        """Steps whose imaging terms the forward keeps (not recomputed)."""
        return int(_lib.load().fwi_plan_kept_steps(self._h))

    @property
    def surface_rows(self) -> int:
        """End row of the FP64 near-surface band of the forward (0: none)."""
        return int(_lib.load().fwi_plan_surface_rows(self._h))

    @property
    def rows_per_item(self) -> int:
        return int(_lib.load().fwi_plan_rows_per_item(self._h))

    @property
    def item_order(self) -> int:
        return int(_lib.load().fwi_plan_item_order(self._h))

    @property
    def adj_tiling(self):
        """(rows per item, item order, grouped) of the adjoint kernels."""
        r, o, g = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.load().fwi_plan_adj_tiling(self._h, ctypes.byref(r), ctypes.byref(o), ctypes.byref(g))
        return r.value, o.value, bool(g.value)

    def graph_stats(self):
        """(graph captures, graph replays) so far."""
        c, r = ctypes.c_longlong(), ctypes.c_longlong()
        _lib.load().fwi_plan_graph_stats(self._h, ctypes.byref(c), ctypes.byref(r))
        return c.value, r.value

    def graph_nodes(self):
        """(kernel nodes of the largest captured graph, of all graphs of the last capture)."""
        a, b = ctypes.c_longlong(), ctypes.c_longlong()
        _lib.load().fwi_plan_graph_nodes(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def cells(self):
        """(sum of nx_s*nz over shots, C-PML cells of all windows)."""
        s = ctypes.c_longlong()
        c = _lib.load().fwi_plan_cells(self._h, ctypes.byref(s))
        return int(c), int(s.value)

    def shot_misfits(self) -> np.ndarray:
        """Per-shot FP64 misfits of the last misfit or gradient call (they sum, in shot order, to the total)."""
        out = np.zeros(len(self.shots), np.float64)
        _lib.check(_lib.load().fwi_plan_shot_misfits(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                      None), "fwi_plan_shot_misfits")
        return out

    def set_timing(self, stride: int):
        _lib.load().fwi_plan_set_timing(self._h, stride)

    def kernel_launches(self, kind: int) -> int:
        """Launches of a timing kind since the last set_timing (counted while timing is on)."""
        return int(_lib.load().fwi_plan_kernel_launches(self._h, kind))

    def core_cells(self) -> int:
        """Cells the step pairs advance two steps per launch (0 without step pairs)."""
        return int(_lib.load().fwi_plan_core_cells(self._h))

    def adj_core_cells(self) -> int:
        """Cells the adjoint step pairs advance two steps per launch (0 without adjoint step pairs)."""
        return int(_lib.load().fwi_plan_adj_core_cells(self._h))

    def kernel_time(self, kind: int):
        """(total ms, launches timed) of kind 0 forward step, 1 forward step storing, 2 adjoint step, 3 / 4 fused
        core step pair without / with store."""
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        _lib.check(_lib.load().fwi_plan_kernel_time(self._h, kind, ctypes.byref(ms), ctypes.byref(n)),
                   "fwi_plan_kernel_time")
        return ms.value, int(n.value)

    def set_wavelet(self, w: np.ndarray):
        self._wav = np.ascontiguousarray(w, np.float32)
        _lib.check(_lib.load().fwi_plan_set_wavelet(self._h, self._wav.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   "fwi_plan_set_wavelet")

    def set_cpml(self, ab):
        fp = ctypes.POINTER(ctypes.c_float)
        self._ta, self._tb = (np.ascontiguousarray(x, np.float32) for x in ab)
        _lib.check(_lib.load().fwi_plan_set_cpml(self._h, self._ta.ctypes.data_as(fp), self._tb.ctypes.data_as(fp)),
                   "fwi_plan_set_cpml")

    def split_gathers(self, g: torch.Tensor):
        """Per-shot [nt, nrec] views of a survey gather array."""
        return [g[self.gather_offsets[i]:self.gather_offsets[i + 1]].view(self.nt, s.nrec)
                for i, s in enumerate(self.shots)]

    def _model(self, model):
        vh, vp, vn = (t.contiguous() for t in model)
        for t in (vh, vp, vn):
            assert t.is_cuda and t.dtype == torch.float32 and t.shape == (self.nz, self.nx)
        return vh, vp, vn

    def _persistent(self, name, shape, dtype, device):
        """A plan-owned output buffer (same address on every call, so graph replays apply)."""
        buf = self._bufs.get(name)
        if buf is None or buf.device != device:
            buf = torch.empty(shape, dtype=dtype, device=device)
            self._bufs[name] = buf
        return buf

    def forward(self, model, stream=None, out: torch.Tensor | None = None) -> torch.Tensor:
        """Gathers of every shot; `out` (float32, gather_size) is filled in place when given."""
        vh, vp, vn = self._model(model)
        if out is None:
            out = torch.empty(self.gather_size, dtype=torch.float32, device=vh.device)
        assert out.is_contiguous() and out.dtype == torch.float32 and out.numel() == self.gather_size
        _lib.check(_lib.load().fwi_forward(self._h, vh.data_ptr(), vp.data_ptr(), vn.data_ptr(), out.data_ptr(),
                                           _stream_ptr(stream)), "fwi_forward")
        return out

    def misfit(self, model, obs: torch.Tensor, stream=None) -> float:
        vh, vp, vn = self._model(model)
        obs = obs.contiguous()
        m = ctypes.c_double()
        _lib.check(_lib.load().fwi_misfit(self._h, vh.data_ptr(), vp.data_ptr(), vn.data_ptr(), obs.data_ptr(),
                                          ctypes.byref(m), _stream_ptr(stream)), "fwi_misfit")
        return m.value

    def gradient(self, model, obs: torch.Tensor, want_syn: bool = False, stream=None, out: torch.Tensor | None = None):
        """Returns (misfit, grad[3, nz, nx] float64 for (v_hor, v_po, v_nmo), syn or None).  The gradient is
        written into `out` when given (and returned), else into a plan-owned buffer that is returned cloned;
        either way the library sees the same buffers from call to call and replays its captured graphs."""
        vh, vp, vn = self._model(model)
        obs = obs.contiguous()
        grad = out if out is not None else self._persistent("grad", (3, self.nz, self.nx), torch.float64, vh.device)
        assert grad.is_contiguous() and grad.dtype == torch.float64 and grad.numel() == 3 * self.nz * self.nx
        syn = self._persistent("syn", (self.gather_size,), torch.float32, vh.device) if want_syn else None
        m = ctypes.c_double()
        _lib.check(_lib.load().fwi_gradient(self._h, vh.data_ptr(), vp.data_ptr(), vn.data_ptr(), obs.data_ptr(),
                                            grad.data_ptr(), ctypes.byref(m), syn.data_ptr() if syn is not None else None,
                                            _stream_ptr(stream)), "fwi_gradient")
        return m.value, (grad if out is not None else grad.clone()), (syn.clone() if syn is not None else None)


    def gradient_src(self, model, adj_src: torch.Tensor, want_syn: bool = False, stream=None):
        """Gradient of <adj_src, gathers(model)> with respect to (v_hor, v_po, v_nmo): (grad float64, syn)."""
        vh, vp, vn = self._model(model)
        adj_src = adj_src.contiguous().float()
        assert adj_src.numel() == self.gather_size
        grad = self._persistent("grad", (3, self.nz, self.nx), torch.float64, vh.device)
        syn = self._persistent("syn", (self.gather_size,), torch.float32, vh.device) if want_syn else None
        _lib.check(_lib.load().fwi_gradient_src(self._h, vh.data_ptr(), vp.data_ptr(), vn.data_ptr(),
                                                adj_src.data_ptr(), grad.data_ptr(),
                                                syn.data_ptr() if syn is not None else None, _stream_ptr(stream)),
                   "fwi_gradient_src")
        return grad.clone(), (syn.clone() if syn is not None else None)


def chain_rule(vp: torch.Tensor, eps: torch.Tensor, delta: torch.Tensor, grad: torch.Tensor, stream=None):
    """dJ/d(vp, eps, delta) [3, nz, nx] float64 from dJ/d(v_hor, v_po, v_nmo) (fwi_chain_rule)."""
    vp, eps, delta = (t.contiguous().float() for t in (vp, eps, delta))
    grad = grad.contiguous()
    out = torch.empty_like(grad)
    _lib.check(_lib.load().fwi_chain_rule(vp.data_ptr(), eps.data_ptr(), delta.data_ptr(), grad.data_ptr(),
                                          out.data_ptr(), vp.numel(), _stream_ptr(stream)), "fwi_chain_rule")
    return out
