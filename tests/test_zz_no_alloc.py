"""GPU: fwi_dot3_noalloc, the entry point that allocates nothing (its scratch is a static device array), against
fwi_dot3 through the C ABI: two calls in a row, the second with a smaller input, equal bit for bit (the same block
partition), and a size beyond the stated maximum or a pageable output refused before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dot3_noalloc_matches_dot3():
    import torch

    from fwi_amd import _lib

    L = _lib.load()
    rng = np.random.default_rng(5)
    out = torch.empty(3, dtype=torch.float64, device="cuda")
    for n in (640017, 1000, 0):  # a smaller input after a larger one (the scratch keeps the last call's partials)
        a = torch.from_numpy(rng.standard_normal(3 * n)).cuda() if n else torch.zeros(3, dtype=torch.float64,
                                                                                      device="cuda")
        b = torch.from_numpy(rng.standard_normal(3 * n)).cuda() if n else torch.zeros(3, dtype=torch.float64,
                                                                                      device="cuda")
        ref = (ctypes.c_double * 3)()
        assert L.fwi_dot3(a.data_ptr(), b.data_ptr(), n, ref, None) == 0
        out.fill_(float("nan"))
        assert L.fwi_dot3_noalloc(a.data_ptr(), b.data_ptr(), n, out.data_ptr(), None) == 0
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        print(f"n={n}: {got}")
        assert np.array_equal(got, np.array(list(ref))), n
        if n:
            A, B = a.cpu().numpy().reshape(3, n), b.cpu().numpy().reshape(3, n)
            np.testing.assert_allclose(got, (A * B).sum(axis=1), rtol=1e-12)
        else:
            assert not got.any()
    assert L.fwi_dot3_noalloc(a.data_ptr(), b.data_ptr(), (1 << 40) + 1, out.data_ptr(), None) == 1
    assert "2^40" in L.fwi_error_message().decode()
    host = np.zeros(3)
    assert L.fwi_dot3_noalloc(a.data_ptr(), b.data_ptr(), 3, host.ctypes.data, None) == 1
    assert "out" in L.fwi_error_message().decode()
