"""The bisector injection on the GPU (soap_bis_inject through the C ABI) against
the oracle's restatement of CB_injection_faculae.py:191-244: quiet form with and
without samples that pass their neighbours (the GPU then sorts them), the active
form with all four branches, and a spectrum of config-1 length.  Spectra are held
to 1e-12 norm-wise (the arithmetic is the same term by term; only the device
square root and division may differ in the last bit)."""
import numpy as np
import pytest

import oracle as O
from tests.pipeline_ref import normwise

pytestmark = pytest.mark.gpu


def _spectrum(n, seed, lo=5000.0, hi=5030.0):
    rng = np.random.default_rng(seed)
    w = np.exp(np.linspace(np.log(lo), np.log(hi), n))
    f = np.ones(n)
    for c0 in rng.uniform(lo + 1, hi - 1, max(4, int((hi - lo) * 2))):
        f -= rng.uniform(0.1, 0.9) * np.exp(-0.5 * ((w - c0) / rng.uniform(0.03, 0.08)) ** 2)
    return w, np.clip(f, 0.02, 1.0)


@pytest.mark.parametrize("up,scale", [([300.0, -450.0, 80.0], 1.0), ([2.0e4, -3.0e4, 9.0e3], 1.5)])
def test_quiet_vs_oracle(engine, up, scale):
    w, f = _spectrum(6000, 3)
    ref = O.bis_inject(w, f, up, scale)
    got = engine.bis_inject(w, f, up, scale)
    assert normwise(got, ref) <= 1e-12
    assert np.array_equal(got == 0.0, ref == 0.0)


def test_active_vs_oracle(engine):
    w, f = _spectrum(8000, 5)
    f = f.copy()
    f[100] = 0.8
    kw = dict(low=[-300.0, 100.0, 40.0], dep_cut=0.8, vel_cut=120.0, velo_shift=-50.0)
    ref = O.bis_inject(w, f, [500.0, -200.0], 1.2, **kw)
    got = engine.bis_inject(w, f, [500.0, -200.0], 1.2, **kw)
    assert normwise(got, ref) <= 1e-12


def test_config1_length(engine):
    """732,585 pixels over 3800-7000 A: a quiet cube row's worth of work."""
    w, f = _spectrum(732585, 11, 3800.0, 7000.0)
    up = [150.0, -420.0, 120.0]
    ref = O.bis_inject(w, f, up, 1.0)
    got = engine.bis_inject(w, f, up, 1.0)
    assert normwise(got, ref) <= 1e-12


def test_rejects_bad_arguments(engine):
    from soap_gpu_amd.engine import SoapError
    w, f = _spectrum(100, 1)
    with pytest.raises(SoapError):
        engine.bis_inject(w, f, np.zeros(17), 1.0)
    with pytest.raises(SoapError):
        engine.bis_inject(w[:50], f, [1.0], 1.0)
