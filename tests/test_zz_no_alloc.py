"""The allocation-free fit entry point (soap_fit_ccf_host_noalloc): through the C
ABI, two calls in a row -- the second with fewer CCFs and fewer velocities, so
that stale staging from the first call would show -- give results equal bit for
bit to soap_fit_ccf_host_err's on the same inputs, with and without per-point
errors; a call beyond the staging limit is refused before any GPU work; and
repeated calls leave the device's free memory unchanged."""
import numpy as np
import pytest
import torch  # noqa: F401  (before the library initialises HIP)

from soap_gpu_amd.engine import Engine, SoapError

pytestmark = pytest.mark.gpu


def _ccfs(rng, n_ccf, v):
    s = rng.uniform(2500, 4000, (n_ccf, 1))
    x0 = rng.uniform(-1500, 1500, (n_ccf, 1))
    y = 1.0 - rng.uniform(0.3, 0.6, (n_ccf, 1)) * np.exp(-0.5 * ((v[None, :] - x0) / s) ** 2)
    y += 0.02 * np.exp(-0.5 * ((v[None, :] - x0 - 2000) / 3000) ** 2) + 1e-4 * rng.standard_normal(y.shape)
    return y


def _same(a, b):
    for k in ("par", "model", "depth", "bis"):
        assert np.array_equal(a[k], b[k]), k


def test_noalloc_fit_equals_allocating_fit():
    eng = Engine(0)
    try:
        rng = np.random.default_rng(2026)
        v1 = np.arange(-20000, 20001, 250.0)  # 161 velocities
        v2 = np.arange(-15000, 15001, 500.0)  # 61 velocities
        y1, y2 = _ccfs(rng, 300, v1), _ccfs(rng, 40, v2)
        e1 = 1e-3 * rng.uniform(0.5, 2.0, y1.shape)
        e2 = 1e-3 * rng.uniform(0.5, 2.0, y2.shape)
        for err1, err2 in ((None, None), (e1, e2)):
            a1 = eng.fit_ccf_host_noalloc(v1, y1, err1)
            a2 = eng.fit_ccf_host_noalloc(v2, y2, err2)
            r1, r2 = eng.fit_ccf_host(v1, y1, err1), eng.fit_ccf_host(v2, y2, err2)
            _same(a1, r1)
            _same(a2, r2)
            if err1 is not None:  # the solver outcome per CCF as well
                assert np.array_equal(a1["status"], r1["status"]) and np.array_equal(a2["status"], r2["status"])
            assert np.all(np.isfinite(a1["par"])) and np.all(np.isfinite(a2["par"]))
    finally:
        eng.close()


def test_noalloc_fit_limit_and_no_device_allocation():
    import torch

    eng = Engine(0)
    try:
        rng = np.random.default_rng(5)
        v = np.arange(-20000, 20001, 250.0)
        y = _ccfs(rng, 64, v)
        eng.fit_ccf_host_noalloc(v, y)
        eng.fit_ccf_host_noalloc(v, y[:7], np.full((7, v.size), 2e-3))
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info(0)[0]
        for k in range(6):
            eng.fit_ccf_host_noalloc(v, y[: 64 - 9 * k])
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info(0)[0] == free0
        # 161 points: about 509.5 staging doubles per CCF (CCF, model, 5 parameters,
        # 2 x 91 bisector values, a status word), so 5000 CCFs exceed 2^21
        big = np.ones((5000, v.size))
        par_before = eng.fit_ccf_host_noalloc(v, y[:3])["par"]
        with pytest.raises(SoapError, match="limit"):
            eng.fit_ccf_host_noalloc(v, big)
        # the refused call ran nothing: the staging still gives the same results
        np.testing.assert_array_equal(eng.fit_ccf_host_noalloc(v, y[:3])["par"], par_before)
    finally:
        eng.close()
