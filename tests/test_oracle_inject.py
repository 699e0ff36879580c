"""Pins of the oracle's bisector injection (oracle.bis_inject, the seed
preprocessing of CB_injection_faculae.py:191-244; the reference module runs file
I/O at import and is never imported): the resampling equals scipy's linear
interp1d, which the reference calls, on the moved samples; a constant velocity on
a linear grid is an exact Doppler resample of a linear spectrum; zero velocity
is the identity inside the open range; the active form's branches follow the
reference's if / elif order."""
import numpy as np
import pytest
from scipy.interpolate import interp1d

import oracle as O

C = 299792458.0


def _spectrum(n=3000, seed=3):
    rng = np.random.default_rng(seed)
    w = np.linspace(5000.0, 5030.0, n)
    f = np.ones(n)
    for c0 in rng.uniform(5001, 5029, 40):
        f -= rng.uniform(0.1, 0.9) * np.exp(-0.5 * ((w - c0) / rng.uniform(0.03, 0.08)) ** 2)
    return w, np.clip(f, 0.02, 1.0)


def _moved_quiet(w, f, up, scale):
    rv = np.poly1d(up)(f) * scale
    beta = rv / C
    return w + w * (np.sqrt((1 + beta) / (1 - beta)) - 1)


@pytest.mark.parametrize("up,scale", [([300.0, -450.0, 80.0], 1.0), ([2.0e4, -3.0e4, 9.0e3], 1.5)])
def test_quiet_equals_scipy_interp1d(up, scale):
    """The second case moves pixels past their neighbours (the samples need sorting)."""
    w, f = _spectrum()
    moved = _moved_quiet(w, f, up, scale)
    if scale > 1:
        assert np.any(np.diff(moved) < 0)
    ref = np.zeros_like(w)
    idx = (w < moved.max()) & (w > moved.min())
    ref[idx] = interp1d(moved, f)(w[idx])
    got = O.bis_inject(w, f, up, scale)
    np.testing.assert_array_equal(got, ref)


def test_constant_velocity_is_an_exact_resample():
    """v constant: moved = w (1 + fb) is linear in the pixel index, so a spectrum
    linear in the index resamples exactly to a + b (w / (1 + fb) - w0) / dw."""
    n = 2001
    w = np.linspace(6000.0, 6010.0, n)
    dw = w[1] - w[0]
    f = 0.3 + 0.4 * np.arange(n) / (n - 1)  # in (0, 1)
    v = 850.0
    fb = np.sqrt((1 + v / C) / (1 - v / C)) - 1
    got = O.bis_inject(w, f, [v], 1.0)
    moved = w * (1 + fb)
    inside = (w < moved.max()) & (w > moved.min())
    expect = 0.3 + 0.4 * ((w / (1 + fb) - w[0]) / dw) / (n - 1)
    np.testing.assert_allclose(got[inside], expect[inside], rtol=0, atol=1e-11)
    assert np.all(got[~inside] == 0.0) and (~inside).sum() >= 1


def test_zero_velocity_is_identity_inside():
    w, f = _spectrum(500)
    got = O.bis_inject(w, f, [0.0, 0.0], 2.0)
    np.testing.assert_array_equal(got[1:-1], f[1:-1])  # the open range excludes the end points
    assert got[0] == 0.0 and got[-1] == 0.0


def test_active_branches():
    """flux <= 0.5 uses the lower polynomial even above dep_cut (elif order);
    0.5 < flux < dep_cut the upper one; flux > dep_cut the cut velocity; flux ==
    dep_cut is not moved."""
    w, f = _spectrum(4000, seed=5)
    f = f.copy()
    f[100] = 0.8  # == dep_cut
    up, low = [500.0, -200.0], [-300.0, 100.0, 40.0]
    dep, vcut, vs, sc = 0.8, 120.0, -50.0, 1.2
    got = O.bis_inject(w, f, up, sc, low=low, dep_cut=dep, vel_cut=vcut, velo_shift=vs)
    rv = np.where(f <= 0.5, (np.poly1d(low)(f) + vs) * sc,
                  np.where((f < dep) & (f > 0.5), (np.poly1d(up)(f) + vs) * sc, (vcut + vs) * sc))
    beta = rv / C
    moved = w + w * (np.sqrt((1 + beta) / (1 - beta)) - 1)
    moved[f == dep] = w[f == dep]
    ref = np.zeros_like(w)
    idx = (w < moved.max()) & (w > moved.min())
    ref[idx] = interp1d(moved, f)(w[idx])
    np.testing.assert_array_equal(got, ref)
    # dep_cut below 0.5: pixels in (dep_cut, 0.5] still take the lower polynomial
    got2 = O.bis_inject(w, f, up, sc, low=low, dep_cut=0.3, vel_cut=vcut, velo_shift=vs)
    rv2 = np.where(f <= 0.5, (np.poly1d(low)(f) + vs) * sc, (vcut + vs) * sc)
    rv2 = np.where(f == 0.3, 0.0, rv2)
    beta2 = rv2 / C
    moved2 = w + w * (np.sqrt((1 + beta2) / (1 - beta2)) - 1)
    ref2 = np.zeros_like(w)
    idx2 = (w < moved2.max()) & (w > moved2.min())
    ref2[idx2] = interp1d(moved2, f)(w[idx2])
    np.testing.assert_array_equal(got2, ref2)
