"""Local PHOENIX handling on the GPU (soap_phoenix_interp and the phoenix_extraction
executable, in place of phoenix_extraction.py for local files) against the oracle's
restatement: vacuum-to-air conversion, the two wavelength windows and the linear
interpolation in temperature, bit for bit; the executable on a temporary
directory laid out like the reference's, with synthetic library files (the
reference ships none and nothing is downloaded)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle as O
from soap_gpu_amd.build import BIN
from tests.fits_util import write_fits

pytestmark = pytest.mark.gpu


def _library_like(n=200_000, seed=0):
    """A vacuum wavelength grid over 3000-9000 A and float32 spectra of the library's kind."""
    rng = np.random.default_rng(seed)
    w = np.sort(rng.uniform(3000.0, 9000.0, n))
    spec = {t: (rng.uniform(0.5, 1.0, n) * 1e14 * (t / 5000.0) ** 4).astype(np.float32)
            for t in (5100.0, 5200.0, 5600.0, 5700.0, 5800.0, 6000.0, 6100.0)}
    return w, spec


@pytest.mark.parametrize("temp", [5778, 5115, 6028, 5700, 5800])
def test_interp_vs_oracle_float32(engine, temp):
    w, spec = _library_like()
    t_lo, t_hi = O.phoenix_bracket(temp)
    lo, hi = spec[t_lo], spec[t_hi]
    air, flux = engine.phoenix_interp(w, lo.astype(np.float64), hi.astype(np.float64), temp, t_lo, t_hi, 3800.0,
                                      6700.0, f32=True)
    ref_air, ref_flux = O.phoenix_interp(w, lo, hi, temp, t_lo, t_hi, 3800.0, 6700.0)
    assert air.size == ref_air.size > 0
    np.testing.assert_array_equal(air, ref_air)
    np.testing.assert_array_equal(flux, ref_flux)


def test_interp_vs_oracle_float64_and_edges(engine):
    rng = np.random.default_rng(4)
    w = np.linspace(3300.0, 7200.0, 50_001)
    lo, hi = rng.uniform(1.0, 2.0, w.size), rng.uniform(1.0, 2.0, w.size)
    air, flux = engine.phoenix_interp(w, lo, hi, 5650.5, 5600.0, 5700.0, 3800.0, 6700.0, f32=False)
    ref_air, ref_flux = O.phoenix_interp(w, lo, hi, 5650.5, 5600.0, 5700.0, 3800.0, 6700.0)
    np.testing.assert_array_equal(air, ref_air)
    np.testing.assert_array_equal(flux, ref_flux)
    # nothing in the window: empty result
    air, flux = engine.phoenix_interp(w, lo, hi, 5650.5, 5600.0, 5700.0, 8000.0, 8100.0)
    assert air.size == 0 and flux.size == 0
    # a single pixel
    air, flux = engine.phoenix_interp(w[20000:20001], lo[20000:20001], hi[20000:20001], 5650.5, 5600.0, 5700.0,
                                      3800.0, 6700.0, f32=False)
    ref_air, ref_flux = O.phoenix_interp(w[20000:20001], lo[20000:20001], hi[20000:20001], 5650.5, 5600.0, 5700.0,
                                         3800.0, 6700.0)
    np.testing.assert_array_equal(air, ref_air)
    np.testing.assert_array_equal(flux, ref_flux)


def test_interp_rejects_bad_arguments(engine):
    from soap_gpu_amd.engine import SoapError
    w = np.linspace(4000.0, 5000.0, 100)
    with pytest.raises(SoapError):
        engine.phoenix_interp(w, w, w, 5700.0, 5700.0, 5700.0, 3800.0, 6700.0)  # t_low == t_high
    with pytest.raises(SoapError):
        engine.phoenix_interp(w, w[:50], w, 5700.0, 5600.0, 5700.0, 3800.0, 6700.0)


def _layout(tmp_path, with_files=True):
    src = tmp_path / "input_phoenix"
    src.mkdir()
    w, spec = _library_like(n=120_000, seed=9)
    if with_files:
        write_fits(src / "WAVE_PHOENIX-ACES-AGSS-COND-2011.fits", w)
        for t, s in spec.items():
            _, name = O.phoenix_name(t, 4.5, 0.0, 0.0)
            write_fits(str(src) + name, s)
    cfg = tmp_path / "config_conti_phoenix.cfg"
    cfg.write_text("[star]\nTstar      : 5778\nTdiff_spot : 663\nTdiff_faculae : 250\nlogg      : 4.5\n"
                   "minwave      : 3800.0\nmaxwave      : 6700.0\n\n[data_io]\n"
                   f"input_prefix: {tmp_path}\ninput_dir_source: /input_phoenix/\n")
    return src, cfg, w, spec


def test_phoenix_extraction_executable(tmp_path):
    src, cfg, w, spec = _layout(tmp_path)
    r = subprocess.run([os.path.join(BIN, "phoenix_extraction"), str(cfg)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    for temp in (5778, 5778 - 663, 5778 + 250):
        t_lo, t_hi = O.phoenix_bracket(temp)
        ref_air, ref_flux = O.phoenix_interp(w, spec[t_lo], spec[t_hi], temp, t_lo, t_hi, 3800.0, 6700.0)
        with np.load(src / f"New_Phoenix_{temp}.npz") as d:
            np.testing.assert_array_equal(d["wave"], ref_air)
            np.testing.assert_array_equal(d["flux"], ref_flux)


def test_phoenix_extraction_missing_file(tmp_path):
    src, cfg, w, spec = _layout(tmp_path)
    _, name = O.phoenix_name(5200.0, 4.5, 0.0, 0.0)
    os.remove(str(src) + name)
    r = subprocess.run([os.path.join(BIN, "phoenix_extraction"), str(cfg)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "nothing is downloaded" in r.stderr
