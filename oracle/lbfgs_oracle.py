"""CPU restatement of the reference optimizer (src/main.c:539-800) in FP64 -- TEST INFRASTRUCTURE ONLY.

Used by tests/ as the checker of fwi_amd's device optimizer (fwi_amd/csrc/lbfgs.hip); nothing in the
product imports it.  Blocks are (v_hor, v_po, v_nmo) with gradients in the same order; updates touch rows
iz > 15 (include/parallelcalls.h:6) and end with v_hor >= v_nmo (parallelcalls.cu:282, 400).  The
reference's dot products are host float sums with a /2.0 that cancels (main.c:576-676); here they are FP64.
Pinned against the dense BFGS inverse-Hessian recursion (Nocedal & Wright, Alg. 7.4 / eq. 6.17) in
tests/test_host_logic.py.
"""
from __future__ import annotations

import numpy as np

LAYER = 15


def _update(base, direction, alpha, scale=(1.0, 1.0, 1.0)):
    """m_k = base_k - alpha_k * (dir_k * scale_k) on rows > LAYER (FP64, rounded to float), v_hor >= v_nmo."""
    out = [b.copy() for b in base]
    rows = slice(LAYER + 1, None)
    for k in range(3):
        val = base[k][rows].astype(np.float64) - alpha[k] * (direction[k][rows] * scale[k])
        out[k][rows] = val.astype(np.float32)
    h, n = out[0][rows], out[2][rows]
    out[0][rows] = np.where(h < n, n, h)
    return out


def two_loop(s_list, y_list, g):
    """L-BFGS two-loop recursion of main.c:574-682 for one block; pairs oldest first; H0 = gamma I with
    gamma = s_k.y_k / y_k.y_k of the newest pair."""
    q = np.asarray(g, np.float64).copy()
    alphas, rhos = [], []
    for sl, yl in zip(reversed(s_list), reversed(y_list)):
        rho = 1.0 / np.sum(yl * sl)
        a = rho * np.sum(sl * q)
        q -= a * yl
        alphas.append(a)
        rhos.append(rho)
    alphas.reverse()
    rhos.reverse()
    gamma = np.sum(s_list[-1] * y_list[-1]) / np.sum(y_list[-1] * y_list[-1])
    r = gamma * q
    for l, (sl, yl) in enumerate(zip(s_list, y_list)):
        beta = rhos[l] * np.sum(yl * r)
        r += (alphas[l] - beta) * sl
    return r


class LbfgsOracle:
    def __init__(self, m: int = 5):
        self.m = m
        self.s, self.y = [], []
        self.back_m = None
        self.back_g = None
        self.r = None

    def sd_step(self, model, grad, alpha):
        self.back_m = [x.copy() for x in model]
        self.back_g = [np.asarray(g, np.float64).copy() for g in grad]
        norms = [np.sqrt(np.sum(np.asarray(g, np.float64) ** 2)) for g in grad]
        scale = [1.0 / nn if nn > 0 else 0.0 for nn in norms]
        return _update(model, [np.asarray(g, np.float64) for g in grad], alpha, scale)

    def direction(self, model, grad):
        g = [np.asarray(x, np.float64) for x in grad]
        s = [model[k].astype(np.float64) - self.back_m[k].astype(np.float64) for k in range(3)]
        y = [g[k] - self.back_g[k] for k in range(3)]
        self.s.append(s)
        self.y.append(y)
        if len(self.s) > self.m:  # FIFO: the reference shifts slots 1..m-1 down (main.c:791-800)
            self.s.pop(0)
            self.y.pop(0)
        r = [two_loop([p[k] for p in self.s], [p[k] for p in self.y], g[k]) for k in range(3)]
        self.r = r
        self.back_m = [x.copy() for x in model]
        self.back_g = [x.copy() for x in g]
        return r

    def trial(self, alpha):
        return _update(self.back_m, self.r, alpha)


def dense_bfgs_direction(pairs, g):
    """H g with H from the BFGS inverse update over `pairs` (oldest first), H0 = gamma I, gamma from the
    newest pair: H <- (I - rho s y^T) H (I - rho y s^T) + rho s s^T."""
    s_new, y_new = pairs[-1]
    n = g.size
    H = np.eye(n) * (s_new @ y_new) / (y_new @ y_new)
    for s, y in pairs:
        rho = 1.0 / (y @ s)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    return H @ g
