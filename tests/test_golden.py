"""The committed golden fixtures (tests/golden/golden_small64.npz, written by tests/golden/make_fixtures.py from
the oracle): the oracle must reproduce them bit for bit, and they must show the precision relations the
path promises (CPU only)."""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from helpers import rel_l2  # noqa: E402

FIXTURE = os.path.join(HERE, "golden", "golden_small64.npz")


def _fixture():
    return np.load(FIXTURE, allow_pickle=False)


def test_oracle_reproduces_golden_fixtures():
    import make_fixtures

    fresh, _ = make_fixtures.compute()
    gold = _fixture()
    assert sorted(fresh) == sorted(gold.files)
    for k in gold.files:
        a, b = np.asarray(fresh[k]), gold[k]
        assert a.dtype == b.dtype and a.shape == b.shape, k
        assert np.array_equal(a, b), k


def test_golden_fixture_precision_relations():
    g = _fixture()
    # FP32 increment form and the reference's mixed arithmetic against FP64
    assert rel_l2(g["gather_f32i"], g["gather_f64"]) < 1e-5
    assert rel_l2(g["gather_mix"], g["gather_f64"]) < 1e-5
    for k in range(3):
        assert rel_l2(g["grad_exact_f32i"][k], g["grad_exact_f64"][k]) < 1e-4, k
    assert abs(g["misfit_exact_f32i"] - g["misfit_exact_f64"]) < 1e-4 * g["misfit_exact_f64"]
    # the steepest-descent step leaves rows iz <= 15 untouched and keeps v_hor >= v_nmo
    init = np.stack([g["sq_init_vh"], g["sq_init_vp"], g["sq_init_vn"]])
    assert np.array_equal(g["sd_model"][:, :16], init[:, :16])
    assert np.all(g["sd_model"][0] >= g["sd_model"][2])
    assert not np.array_equal(g["sd_model"], init)
