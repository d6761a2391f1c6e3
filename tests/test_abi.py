"""The C-ABI library loads and exports every entry point declared in include/*.h (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("fwi_serialcalls.h", "fwi.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**([A-Za-z_][A-Za-z0-9_]*)\s*\(", src,
                             flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_reference_launchers():
    ref = {"pre_variables_h", "CPML_x_h", "CPML_z_h", "PSI_F_h", "ZETA_F_h", "propagator_F_h", "res_h",
           "SecondDerivate_P_h", "launchPropagation", "launchBackPropagation", "NORMAL_h", "UpdateModel_h",
           "compute_sy_h", "compute_qi_h", "compute_ri_h", "update_r_h", "update_LBFGS_h"}
    assert ref <= _declared()


def test_library_exports_every_declared_symbol():
    from fwi_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfwi.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding's signature table covers the same set
    assert _declared() <= set(_lib.SYMBOLS)


def test_binding_loads_without_gpu():
    from fwi_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfwi.so not built")
    L = _lib.load()
    assert L.fwi_last_error() == 0


def test_cpml_table_library_matches_oracle():
    import numpy as np

    import oracle_lib as O
    from fwi_amd import models as M
    from fwi_amd.survey import cpml_table

    p = M.CpmlParams(cpml=20, dh=10.0, dt=8e-4, f=8.0)
    a, b = cpml_table(p)
    ta, tb = O.cpml_table(p)
    np.testing.assert_array_equal(a, ta.astype(np.float32))
    np.testing.assert_array_equal(b, tb.astype(np.float32))


@pytest.mark.parametrize("kw,msg", [({"ckpt_format": 1, "arith": 0}, "increment form"),
                                    ({"ckpt_format": 3}, "ckpt_format"), ({"grad_group": 5}, "grad_group"),
                                    ({"arith": 7}, "arith")])
def test_plan_rejects_bad_options_before_touching_the_gpu(kw, msg):
    """Option errors come back from fwi_plan_create with a message, before any device call."""
    import numpy as np

    from fwi_amd import _lib
    from fwi_amd import models as M
    from fwi_amd.survey import Plan

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfwi.so not built")
    nx, nz, cpml = 64, 48, 10
    shots = [M.reference_shot(nx, cpml, 30, 6)]
    tab = (np.zeros(cpml, np.float32), np.ones(cpml, np.float32))
    with pytest.raises(RuntimeError, match=msg):
        Plan(nx, nz, 50, 10.0, 1e-3, cpml, shots, np.zeros(50, np.float32), tab, **kw)


def test_entry_points_reject_null_arguments_before_touching_the_gpu():
    """The status-returning entry points refuse a null plan, optimizer or array with a message that names the
    argument, before any device call (the legacy launchers keep the reference's unchecked semantics)."""
    import ctypes

    from fwi_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfwi.so not built")
    L = _lib.load()
    msg = lambda: L.fwi_error_message().decode()
    assert L.fwi_forward(None, None, None, None, None, None) == 1
    assert msg() == "fwi_forward: plan is null"
    assert L.fwi_gradient(None, None, None, None, None, None, None, None, None) == 1
    assert msg() == "fwi_gradient: plan is null"
    assert L.fwi_plan_set_wavelet(None, None) == 1 and msg() == "fwi_plan_set_wavelet: plan is null"
    assert L.fwi_chain_rule(None, None, None, None, None, 10, None) == 1 and msg() == "fwi_chain_rule: vp is null"
    assert L.fwi_chain_rule(None, None, None, None, None, 0, None) == 0  # nothing to do
    alpha = (ctypes.c_double * 3)(1, 1, 1)
    assert L.fwi_lbfgs_trial(None, None, None, None, alpha, None) == 1
    assert msg() == "fwi_lbfgs_trial: optimizer is null"
    assert L.fwi_plan_create(None, None, None) == 1 and msg() == "fwi_plan_create: plan is null"


def test_array_checks_fail_closed():
    """An array the runtime cannot classify is refused, not passed on to a kernel: without a GPU every array
    query fails, so a non-null host array is refused with a message naming it (on a GPU machine the same
    pageable array is classified as unregistered host memory and refused as well).  Nothing is launched."""
    import numpy as np

    from fwi_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfwi.so not built")
    L = _lib.load()
    a = np.zeros(16, np.float32)
    g = np.zeros(48, np.float64)
    o = np.zeros(48, np.float64)
    p = a.ctypes.data
    assert L.fwi_chain_rule(p, p, p, g.ctypes.data, o.ctypes.data, 16, None) == 1
    m = L.fwi_error_message().decode()
    assert m.startswith("fwi_chain_rule: vp is "), m
    assert ("cannot classify" in m) or ("host memory the GPU cannot address" in m), m
    assert not o.any()
