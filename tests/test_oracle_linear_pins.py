"""Pins of the oracle's linear-grid branches (the reference's native 0.005 A
grids):

- orc_select_lines + orc_ccf (linear): against golden vectors written by the
  reference's own CCF_func.py (tests/golden/make_ccf_golden.py; CCF_func.py:5-91
  with the extended grid of run_all_SOAP.py:288-296): overlapping lines, lines
  whose holes reach the linspace extensions and their duplicated join points,
  41 and 161 velocities.
- orc_conv_raw (linear conv_inst, SOAP_resolution.cu:76-95): closed forms -- a
  flat input stays flat (the kernel is a normalised Gaussian sampled at 0.005 A),
  Gaussian widths add in quadrature, and a single spike shows the window
  m in [i - w, i + w - 2] (asymmetric, Q10) with the width taken at the output
  pixel.
"""
import math
import os

import numpy as np
import pytest

import oracle as O

C = 299792458.0
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ccf_linear_golden.npz")


def _golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _dlam(v_ms, lam):
    b = v_ms / C
    return lam * (np.sqrt((1 + b) / (1 - b)) - 1)


@pytest.mark.parametrize("case", ["a", "b"])
def test_linear_line_selection_matches_reference(case):
    g = _golden()
    sl, sw = O.select_lines(g[f"{case}_wave_t"], g[f"{case}_lines_l"], g[f"{case}_lines_w"])
    assert np.array_equal(sl, g[f"{case}_sel_l"]) and np.array_equal(sw, g[f"{case}_sel_w"])


@pytest.mark.parametrize("case", ["a", "b"])
def test_linear_ccf_matches_reference_vectors(case):
    """orc_ccf's literal linear mask equals calculate_CCF_1/2 of the reference:
    finite values to summation-order rounding, and the same non-finite values
    where the duplicated join point of the extended grid gives an edge pixel a
    zero spacing (infinite weight)."""
    g = _golden()
    wave, step, v = g[f"{case}_wave_t"], float(g[f"{case}_step"]), g[f"{case}_vgrid"]
    for s, ref in zip(g[f"{case}_spectra"], g[f"{case}_ccf"]):
        got = O.ccf(O.GRID_LINEAR, wave, step, v, g[f"{case}_sel_l"], g[f"{case}_sel_w"], s)
        fin = np.isfinite(ref)
        assert np.array_equal(np.isfinite(got), fin)
        assert np.array_equal(got[~fin], ref[~fin])
        np.testing.assert_allclose(got[fin], ref[fin], rtol=1e-13, atol=0)


def test_linear_ccf_template_entries():
    """Each template row of the reference, dotted with a unit spectrum per
    column block, equals the oracle's CCF of that block: the mask entries
    themselves agree, not only their dot products with random spectra."""
    g = _golden()
    wave, step, v = g["a_wave_t"], float(g["a_step"]), g["a_vgrid"]
    n = len(wave)
    tpl = np.zeros((len(v), n))
    tpl[g["a_tpl_row"], g["a_tpl_col"]] = g["a_tpl_val"]
    # spectra e_p summed over disjoint pixel sets with distinct powers of two:
    # each mask entry is recovered from the CCF's binary digits
    rng = np.random.default_rng(5)
    for _ in range(3):
        cols = rng.choice(n, size=40, replace=False)
        spec = np.zeros(n)
        for k, c in enumerate(cols):
            spec[c] = 2.0 ** (-3 * k)
        got = O.ccf(O.GRID_LINEAR, wave, step, v, g["a_sel_l"], g["a_sel_w"], spec)
        np.testing.assert_allclose(got, tpl @ spec, rtol=1e-14, atol=0)
    # and single pixels
    for c in rng.choice(n, size=25, replace=False):
        spec = np.zeros(n)
        spec[c] = 1.0
        got = O.ccf(O.GRID_LINEAR, wave, step, v, g["a_sel_l"], g["a_sel_w"], spec)
        np.testing.assert_array_equal(got, tpl[:, c])


def test_linear_conv_flat_stays_flat():
    """1 - A constant: o = (1 - A) * sum_d 0.005 G(d), a normalised Gaussian
    sampled at 0.005 A (sampling and truncation errors far below FP64), for
    the active form (weight 0.005) and the quiet form (weight argument)."""
    wave = 5000.0 + 0.005 * np.arange(12000)
    sig = C / 115000 / (2 * math.sqrt(2 * math.log(2)))
    w = O.conv_window(wave, sig)
    valid = np.zeros(len(wave), bool)
    valid[w + 1:len(wave) - 1 - w] = True
    o = O.conv_raw(O.GRID_LINEAR, wave, sig, np.full(len(wave), 0.3), 1)
    assert np.all(o[~valid] == 0.0)
    assert np.max(np.abs(o[valid] / 0.7 - 1)) < 1e-13
    o0 = O.conv_raw(O.GRID_LINEAR, wave, sig, np.full(len(wave), 0.3), 0, weight=0.0125)
    assert np.max(np.abs(o0[valid] / (0.7 * 2.5) - 1)) < 1e-13


def test_linear_conv_gaussian_widths_add():
    """A Gaussian bump of width s convolved by the instrument: variance
    s^2 + sigma_lambda^2 (the width varies by 0.2 % over the window)."""
    wave = 5000.0 + 0.005 * np.arange(4001)
    sig = C / 60000 / (2 * math.sqrt(2 * math.log(2)))
    lam0, s = 5010.0, 0.05
    bump = np.exp(-0.5 * ((wave - lam0) / s) ** 2)
    o = O.conv_raw(O.GRID_LINEAR, wave, sig, 1.0 - bump, 1)
    mean = np.sum(wave * o) / np.sum(o)
    var = np.sum((wave - mean) ** 2 * o) / np.sum(o)
    expect = s * s + _dlam(sig, lam0) ** 2
    assert abs(mean - lam0) < 1e-6
    assert abs(var / expect - 1) < 2e-3
    # the integral is conserved: sum_i o_i = sum_m (1 - A_m)
    assert abs(np.sum(o) / np.sum(bump) - 1) < 1e-9


def test_linear_conv_spike_window_is_asymmetric():
    """Spike in 1 - A at pixel p: o_i != 0 exactly for p - w + 2 <= i <= p + w
    (m runs over [i - w, i + w - 2], Q10), with
    o_i = 0.005 / (sigma_i sqrt(2 pi)) exp(-(p - i)^2 0.005^2 / (2 sigma_i^2)) and
    sigma_i the instrumental width at the OUTPUT pixel lambda_i.  The grid runs
    100 -> 200 A so that the window edge (50 widths at lambda_0) is ~28 widths
    at lambda_p and the edge terms are representable."""
    wave = 100.0 + 0.005 * np.arange(20000)
    sig = 30000.0
    w = O.conv_window(wave, sig)
    p = 16000
    A = np.ones(len(wave))
    A[p] = 0.0
    o = O.conv_raw(O.GRID_LINEAR, wave, sig, A, 1)
    nz = np.nonzero(o)[0]
    assert nz[0] == p - w + 2 and nz[-1] == p + w and len(nz) == 2 * w - 1
    i = np.arange(p - w + 2, p + w + 1)
    si = _dlam(sig, wave[i])
    expect = 0.005 / (si * math.sqrt(2 * math.pi)) * np.exp(-((p - i) * 0.005) ** 2 / (2 * si * si))
    np.testing.assert_allclose(o[i], expect, rtol=1e-11)
    # the width is the output pixel's: the input pixel's width would be off in the far tail
    sm = _dlam(sig, wave[p])
    wrong = 0.005 / (sm * math.sqrt(2 * math.pi)) * np.exp(-((p - i) * 0.005) ** 2 / (2 * sm * sm))
    assert np.max(np.abs(o[i] / wrong - 1)) > 1e-2
