"""Local PHOENIX handling (phoenix_extraction.py): the numpy restatement in the oracle
pinned by what fixes it independently (the reference's inverse formula, the
textbook refractive index of air, values derived by hand from the reference's
grid and format rules, exact endpoints of the interpolation), and the library's
host-side helpers (temperature bracket, grid, file names, FITS reader) against it.
No GPU needed."""
import numpy as np
import pytest

import oracle as O
from soap_gpu_amd import engine as E
from tests.fits_util import write_fits


def test_vac2air_inverse_and_textbook_value():
    w = np.linspace(3000.0, 10000.0, 70001)
    # air2vac (N. Piskunov's fit) undoes vac2air (Birch and Downs) to 2e-12; a dropped
    # term or a wrong coefficient leaves >= 1e-6
    assert np.max(np.abs(O.phoenix_air2vac(O.phoenix_vac2air(w)) / w - 1)) < 1e-11
    # refractive index of standard air at 500 nm: 1.000279
    n = 5000.0 / O.phoenix_vac2air(np.array([5000.0]))[0]
    assert abs(n - 1.000279) < 2e-6
    assert np.all(np.diff(O.phoenix_vac2air(w)) > 0)


# (temp, bracket) from the reference's rule by hand: argmin of |grid - temp| (the
# first minimum), then (k, k + 1) if temp is above grid[k], else (k - 1, k); grid
# entry 48 is 6800 K, so just above 7000 K the bracket is (7000, 6800)
BRACKETS = [(5778, (5700, 5800)), (5115, (5100, 5200)), (6028, (6000, 6100)), (5800, (5700, 5800)),
            (2350, (2300, 2400)), (6850, (6800, 6900)), (7000, (6900, 7000)), (7050, (7000, 6800)),
            (7150, (7000, 7200)), (11600, (11400, 11600)), (2300, None), (2250, None), (11700, None)]


@pytest.mark.parametrize("temp,expect", BRACKETS)
def test_bracket_pins(temp, expect):
    got = O.phoenix_bracket(temp)
    assert (got is None and expect is None) or tuple(got) == expect


def test_library_bracket_equals_oracle():
    for t in np.arange(2000.0, 12000.0, 7.0):
        ref = O.phoenix_bracket(t)
        if ref is None:
            with pytest.raises(E.SoapError):
                E.phoenix_bracket(t)
        else:
            assert E.phoenix_bracket(t) == tuple(ref), t


def test_grid_pins_and_library():
    assert O.phoenix_grid(4.5) == (4.5, 0.0, 0.0)
    assert O.phoenix_grid(4.44, 0.3, 0.0) == (4.5, 0.5, 0.0)      # [Fe/H] > 0: alpha 0
    assert O.phoenix_grid(2.2, -1.7, 0.45) == (2.0, -1.5, 0.4)
    assert O.phoenix_grid(7.0, -5.0, -1.0) == (6.0, -4.0, -0.2)
    rng = np.random.default_rng(3)
    for lg, fe, al in zip(rng.uniform(-1, 7, 200), rng.uniform(-5, 2, 200), rng.uniform(-0.5, 1.5, 200)):
        assert E.phoenix_grid(lg, fe, al) == O.phoenix_grid(lg, fe, al)


NAMES = [((5700, 4.5, 0.0, 0.0), ("Z-0.0", "/lte05700-4.50-0.0.PHOENIX-ACES-AGSS-COND-2011-HiRes.fits")),
         ((5100, 4.0, -1.0, 0.2), ("Z-1.0", "/lte05100-4.00-1.0.Alpha=+0.20.PHOENIX-ACES-AGSS-COND-2011-HiRes.fits")),
         ((6000, 4.5, -0.5, -0.2), ("Z-0.5", "/lte06000-4.50-0.5.Alpha=-0.20.PHOENIX-ACES-AGSS-COND-2011-HiRes.fits")),
         ((5800, 5.0, 0.5, 0.0), ("Z+0.5", "/lte05800-5.00+0.5.PHOENIX-ACES-AGSS-COND-2011-HiRes.fits")),
         ((10000, 3.5, 0.0, 0.0), ("Z-0.0", "/lte010000-3.50-0.0.PHOENIX-ACES-AGSS-COND-2011-HiRes.fits"))]


@pytest.mark.parametrize("args,expect", NAMES)
def test_name_pins_and_library(args, expect):
    assert O.phoenix_name(*args) == expect
    assert E.phoenix_name(*args) == expect


def test_interp_endpoints():
    rng = np.random.default_rng(7)
    w = np.linspace(3200.0, 7300.0, 5001)
    lo = rng.uniform(1e5, 2e5, w.size).astype(np.float32)
    hi = (lo * rng.uniform(0.6, 1.9, w.size)).astype(np.float32)  # within 2x: hi - lo exact in float32
    air, f = O.phoenix_interp(w, lo, hi, 5700, 5700, 5800, 3800.0, 6700.0)
    sel = (O.phoenix_vac2air(w) >= 3800.0) & (O.phoenix_vac2air(w) <= 6700.0)
    assert air.size == sel.sum() and np.all(air >= 3800.0) and np.all(air <= 6700.0)
    np.testing.assert_array_equal(f, lo[sel].astype(np.float64))           # temp = t_low
    _, f = O.phoenix_interp(w, lo, hi, 5800, 5700, 5800, 3800.0, 6700.0)
    np.testing.assert_array_equal(f, hi[sel].astype(np.float64))           # temp = t_high
    _, f = O.phoenix_interp(w, lo, hi, 5750, 5700, 5800, 3800.0, 6700.0)
    np.testing.assert_allclose(f, 0.5 * (lo[sel].astype(np.float64) + hi[sel]), rtol=1e-15)


@pytest.mark.parametrize("dtype", [">f4", ">f8", ">i2", ">i4"])
def test_fits_reader(tmp_path, dtype):
    rng = np.random.default_rng(11)
    a = (rng.standard_normal(10007) * 1000).astype(dtype)
    p = tmp_path / "x.fits"
    write_fits(p, a, extra_cards=40)  # header longer than one 2880-byte block
    got, bp = E.fits_read(str(p))
    assert bp == {">f4": -32, ">f8": -64, ">i2": 16, ">i4": 32}[dtype]
    np.testing.assert_array_equal(got, a.astype(np.float64))


def test_fits_reader_scaling_2d_and_errors(tmp_path):
    a = np.arange(-300, 300, dtype=">i2").reshape(20, 30)
    p = tmp_path / "s.fits"
    write_fits(p, a, bscale=0.5, bzero=1000.0)
    got, bp = E.fits_read(str(p))
    np.testing.assert_array_equal(got, 1000.0 + 0.5 * a.ravel().astype(np.float64))
    with pytest.raises(E.SoapError):
        E.fits_read(str(tmp_path / "missing.fits"))
    q = tmp_path / "text.fits"
    q.write_bytes(b"not a fits file" * 300)
    with pytest.raises(E.SoapError):
        E.fits_read(str(q))
