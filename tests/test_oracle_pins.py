"""Pins of the FP64 CPU oracle (oracle/soap_oracle.c) against what fixes it
independently: the reference's data fixtures, closed forms, library routines
(numpy.interp, numpy.polyfit, scipy least squares, scipy.special.erf) and
brute-force restatements on tiny inputs.  Each test names the oracle function
and the mistake class it would catch."""
import math

import numpy as np
import pytest

import oracle as O
from soap_gpu_amd import synth

C = 299792458.0


# ------------------------------------------------------------------ a10: LB
def test_lb_params_golden(tables):
    """orc_lb_params vs the golden LB(5778) derived from the reference's
    Limb_brightening_info.npz (run_all_SOAP.py:140-145)."""
    lb = O.lb_params(5778, tables["LB_G2"], tables["LB_K0"], tables["LB_M0"], tables["Temp_array"])
    np.testing.assert_allclose(lb, [0.23161761, -0.11674743, -0.36866954, 0.25537337], atol=5e-9)


@pytest.mark.parametrize("teff", [6100, 5778, 5500, 5240, 5239, 4400, 3750, 3600])
def test_lb_params_numpy_polyfit(tables, teff):
    """orc_lb_params vs numpy.polyfit (a library, not reference code), both
    temperature branches (catches swapped template rows or reversed powers)."""
    G2, K0, M0, T = (np.array(tables[k], dtype=float) for k in ("LB_G2", "LB_K0", "LB_M0", "Temp_array"))
    mu = np.linspace(0.1, 1, 10)
    b = np.array([0, 1]) if teff >= 5240 else np.array([1, 2])
    lb = np.array([np.poly1d(np.polyfit(T[b], np.array([G2[i], K0[i], M0[i]])[b], 1))(teff) for i in range(10)])
    ref = np.polyfit(mu, lb, 3)[::-1]
    np.testing.assert_allclose(O.lb_params(teff, G2, K0, M0, T), ref, rtol=0, atol=1e-13)


# ------------------------------------------------------------------ a1: disk
@pytest.mark.parametrize("grid,count,sumlimb", [(300, 70671, 59840.92644189147), (600, 282687, 239367.515262983),
                                                (1000, 785343, 664948.4148579843)])
def test_disk_counts_and_limb_sum(grid, count, sumlimb, tables):
    """Disk cell count and sum of quadratic limb darkening (closed-form lattice
    counts; the limb sum approaches pi(1-u1/3-u2/6)(G/2)^2)."""
    m = O.disk_init(grid, 25.0, 0.0, 90.0, 696000.0, 0.29, 0.34, synth.LB_5778, tables["mu_boundary_array"])
    on = m["limb"] != -999.0
    assert on.sum() == count
    assert m["limb"][on].sum() == pytest.approx(sumlimb, rel=1e-13)
    cont = math.pi * (1 - 0.29 / 3 - 0.34 / 6) * (grid / 2) ** 2
    assert abs(m["limb"][on].sum() / cont - 1) < 3.0 / grid


def test_disk_fields_closed_form(tables):
    """Every map value against its closed form on the cell centres: mu, limb,
    LB polynomial, rigid-rotation velocity and the mu bin of the reference's
    boundary file (first boundary below mu; last bin beyond)."""
    G = 64
    prot, incl, R = 11.0, 63.0, 512345.0
    bounds = np.array(tables["mu_boundary_array"])
    m = O.disk_init(G, prot, 0.0, incl, R, 0.3, 0.2, [0.1, 0.2, 0.3, 0.4], bounds)
    iy, iz = np.divmod(np.arange(G * G), G)
    y = -1 + iy * (2.0 / G)
    z = -1 + iz * (2.0 / G)
    on = y * y + z * z <= 1
    mu = np.sqrt(1 - (y * y + z * z)[on])
    np.testing.assert_allclose(m["mu"][on], mu, rtol=0, atol=1e-15)
    np.testing.assert_allclose(m["limb"][on], 1 - 0.3 * (1 - mu) - 0.2 * (1 - mu) ** 2, atol=1e-15)
    np.testing.assert_allclose(m["lb"][on], 0.1 + 0.2 * mu + 0.3 * mu ** 2 + 0.4 * mu ** 3, atol=1e-15)
    v = y[on] * 2 * math.pi * R / (prot * 86400) * math.sin(math.radians(incl))
    np.testing.assert_allclose(m["delta"][on], v, rtol=1e-14, atol=1e-15)
    k = np.array([next((b for b in range(len(bounds)) if x > bounds[b]), len(bounds)) for x in mu])
    assert np.array_equal(m["region"][on], k)
    assert np.all(m["mu"][~on] == -999.0) and np.all(m["region"][~on] == 0)


def test_disk_differential_rotation():
    """omega(theta) = omega0 (1 + W sin^2 theta) with theta = -90 + 180 iz/G
    (SOAP_initialize.cu:347-351, quirk Q16)."""
    G = 40
    m = O.disk_init(G, 20.0, -0.18, 90.0, 696000.0, 0.29, 0.34, synth.LB_5778, [0.5])
    iy, iz = np.divmod(np.arange(G * G), G)
    y = -1 + iy * (2.0 / G)
    th = np.radians(-90 + iz * 180.0 / G)
    v = y * (360 / 20.0) * (1 - 0.18 * np.sin(th) ** 2) * math.pi / (180 * 86400) * 696000.0
    on = m["limb"] != -999
    np.testing.assert_allclose(m["delta"][on], v[on], rtol=1e-13, atol=1e-14)


# ------------------------------------------------------------------ a4: shift
def _shifted_axis(wave, d_kms):
    beta = d_kms * 1000 / C
    return wave + (-wave) * (1 - np.sqrt((1 + beta) / (1 - beta)))


@pytest.mark.parametrize("d", [2.2, -1.7, 0.41, -0.013, 7.9])
def test_shift_log_vs_numpy_interp(d):
    """orc_shift_row (log grid) is linear interpolation of the shifted spectrum
    at the grid points: numpy.interp on (shifted lambda, S).  The reference's
    lambda-difference form carries ~1e-10 relative rounding in t, so the bound
    is 5e-11 absolute (a wrong index or weight is O(1e-2))."""
    w = synth.small()
    row = w.cubes["q"][1]
    out = O.shift_row(O.GRID_LOG, w.wave, row, d, 0.7)
    ref = 0.7 * np.interp(w.wave, _shifted_axis(w.wave, d), row)
    sl = slice(20, len(w.wave) - 20)
    assert np.max(np.abs(out[sl] - ref[sl])) < 5e-11


@pytest.mark.parametrize("d", [2.2, -1.7, 0.41])
def test_shift_linear_vs_numpy_interp(d):
    """Linear grid: the reference's own expressions; equals numpy.interp."""
    wave = 5000 + 0.005 * np.arange(5000)
    row = 1 - 0.6 * np.exp(-0.5 * ((wave - 5010) / 0.04) ** 2)
    out = O.shift_row(O.GRID_LINEAR, wave, row, d, 1.0, step=0.005)
    ref = np.interp(wave, _shifted_axis(wave, d), row)
    assert np.max(np.abs(out[20:-20] - ref[20:-20])) < 1e-14


def test_shift_identity_and_linear_profile():
    """delta = 0 is a copy (Q21); a spectrum linear in lambda is reproduced
    exactly by linear interpolation at any shift."""
    w = synth.small()
    row = w.cubes["q"][0]
    assert np.array_equal(O.shift_row(O.GRID_LOG, w.wave, row, 0.0, 0.9), 0.9 * row)
    lin = 2.0 + 1e-3 * (w.wave - w.wave[0])
    for d in (1.3, -2.9):
        out = O.shift_row(O.GRID_LOG, w.wave, lin, d, 1.0)
        beta = d * 1000 / C
        sq = math.sqrt((1 + beta) / (1 - beta))
        exact = 2.0 + 1e-3 * (w.wave / sq - w.wave[0])  # the source point of pixel j is lambda_j / sq
        assert np.max(np.abs(out[30:-30] - exact[30:-30])) < 1e-12


def test_scatter_ref_equals_gather_where_written():
    """Gather mode equals the literal scatter at every pixel the scatter writes
    exactly once (Q1); gaps keep older values in ref mode."""
    wave = 5000 + 0.005 * np.arange(6000)
    row = 1 - 0.5 * np.exp(-0.5 * ((wave - 5012) / 0.05) ** 2)
    for d in (1.9, -2.4):
        g = O.shift_row(O.GRID_LINEAR, wave, row, d, 1.0, step=0.005)
        buf = np.full(len(wave), -1.0)
        O.scatter_row_ref(wave, row, d, 1.0, 0.005, buf)
        w = buf != -1.0
        assert np.array_equal(buf[w][5:-5], g[w][5:-5])
        if d > 0:
            assert (~w).sum() > 0  # gaps exist for redshifts


# ------------------------------------------------------------------ a2: quiet
def _naive_quiet(w, maps, cube, grid):
    out = np.zeros(w.n)
    for idx in range(grid * grid):
        if maps["limb"][idx] == -999.0:
            continue
        sh = _shifted_axis(w.wave, maps["delta"][idx])
        out += maps["limb"][idx] * np.interp(w.wave, sh, cube[maps["region"][idx]], left=0, right=0)
    return out


def test_quiet_star_bruteforce():
    """orc_quiet_star vs an independent per-cell numpy.interp sum (tiny grid)."""
    w = synth.small(grid=14, n=3000, n_mu=4)
    s = w.star
    maps = O.disk_init(14, s["prot"], 0.0, s["incl"], s["radius_km"], s["u1"], s["u2"], s["lb"], w.mu_bounds)
    q = O.quiet_star(O.GRID_LOG, w.wave, maps, w.cubes["q"], 14, nthreads=2)
    ref = _naive_quiet(w, maps, w.cubes["q"], 14)
    sl = slice(40, w.n - 40)
    assert np.max(np.abs(q[sl] - ref[sl])) / np.max(np.abs(ref[sl])) < 1e-11


def test_quiet_star_thread_invariant():
    """Accumulation order is the cell order for any thread count (bitwise)."""
    w = synth.small(grid=20, n=2048)
    s = w.star
    maps = O.disk_init(20, s["prot"], 0.0, s["incl"], s["radius_km"], s["u1"], s["u2"], s["lb"], w.mu_bounds)
    a = O.quiet_star(O.GRID_LOG, w.wave, maps, w.cubes["qsed"], 20, nthreads=1)
    b = O.quiet_star(O.GRID_LOG, w.wave, maps, w.cubes["qsed"], 20, nthreads=5)
    assert np.array_equal(a, b)


def test_quiet_star_symmetric_line_is_unshifted():
    """A rigidly rotating star with one symmetric line and no convective shift
    integrates to a line symmetric about its rest wavelength (RV = 0)."""
    wave = synth.log_grid(5000, 5004, n=4001)
    lines = dict(center=np.array([5002.0]), depth=np.array([0.6]), sigma_kms=np.array([4.0]))
    row = synth.line_row(wave, lines, 1.0, 0.0)
    cube = np.tile(row, (3, 1))
    maps = O.disk_init(60, synth.prot_for_vsini(3.0), 0.0, 90.0, 696000.0, 0.29, 0.34, synth.LB_5778, [0.6, 0.3])
    q = O.quiet_star(O.GRID_LOG, wave, maps, cube, 60)
    win = np.abs(wave - 5002.0) / 5002.0 * C < 40000.0  # +-40 km/s, far from the padded edges
    dip = q[win].max() - q[win]
    u = C * np.log(wave[win] / 5002.0)  # velocity axis (log-lambda is linear in Doppler shifts)
    centroid_v = np.sum(u * dip) / np.sum(dip)
    # each cell adds the same dip shifted by its velocity, so the dip centroid is
    # the limb-weighted mean velocity (non-zero only through the lattice's extra
    # limb column at y = -1; cells are y = -1 + 2 iy/G)
    on = maps["limb"] != -999.0
    v_mean = np.sum(maps["limb"][on] * C * np.arctanh(maps["delta"][on] * 1000.0 / C)) / np.sum(maps["limb"][on])
    d0 = row[win].max() - row[win]
    seed_v = np.sum(u * d0) / np.sum(d0)  # the seed line is Gaussian in lambda, not exactly in velocity
    assert abs(centroid_v - seed_v - v_mean) < 0.005  # m/s


# ------------------------------------------------------------------ a5: regions
def test_spot_at_disk_centre_area(tables):
    """A spot of radius s at disk centre covers ~pi s^2 (G/2)^2 cells."""
    G, s = 200, 0.2
    temp, bi = O.phase_maps(G, 90.0, 20, 5778, 663, 250, 0.0, [(0.0, 0.0, s, 0.0)])
    n = (bi == 0.0).sum()
    assert abs(n / (math.pi * s * s * (G / 2) ** 2) - 1) < 0.03
    assert np.all(temp[bi == 0.0] == 5778 - 663)


def test_spot_behind_disk_invisible():
    temp, bi = O.phase_maps(100, 90.0, 20, 5778, 663, 250, 0.5, [(0.0, 0.0, 0.1, 0.0)])
    assert (bi != -9999.0).sum() == 0
    assert O.region_bbox(0.0, 0.0, 0.1, 90.0, 20, 100, 0.5)[4] == 0


def test_region_membership_matches_angular_distance():
    """Independent geometry: a cell belongs to a region iff the angular distance
    between its surface point and the region centre (rotated by the phase about
    the inclined axis) is <= asin(s); the bounding box only limits the scan."""
    G, s, lat, lon, incl = 120, 0.25, 25.0, 40.0, 70.0
    for phase in (0.0, 0.1, 0.83):
        temp, bi = O.phase_maps(G, incl, 40, 5778, 663, 250, phase, [(lat, lon, s, 1.0)])
        # centre direction: latitude/longitude on the star, observer along +x,
        # rotation axis tilted by incl from the line of sight in the x-z plane
        i = math.radians(incl)
        la, lo = math.radians(lat), math.radians(lon) + 2 * math.pi * phase
        p = np.array([math.cos(la) * math.cos(lo), math.cos(la) * math.sin(lo), math.sin(la)])
        # star frame (axis = z') -> observer frame: rotate about y by (pi/2 - i)
        a = math.pi / 2 - i
        Ry = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
        cen = Ry @ p
        iy, iz = np.divmod(np.arange(G * G), G)
        y = -1 + iy * 2.0 / G
        z = -1 + iz * 2.0 / G
        on = y * y + z * z < 1
        x = np.sqrt(np.clip(1 - y * y - z * z, 0, None))
        cosang = x * cen[0] + y * cen[1] + z * cen[2]
        inside = on & (cosang >= math.sqrt(1 - s * s))
        got = bi == 1.0
        # only cells within a hair of the boundary may differ
        diff = got != inside
        assert np.all(np.abs(cosang[diff] - math.sqrt(1 - s * s)) < 1e-9)
        assert got.sum() > 0


def test_overlapping_regions_last_wins():
    """Q13: a later region overwrites an earlier one cell by cell."""
    regs = [(10.0, 0.0, 0.3, 1.0), (10.0, 0.0, 0.15, 0.0)]
    temp, bi = O.phase_maps(100, 90.0, 20, 5778, 663, 250, 0.0, regs)
    t2, b2 = O.phase_maps(100, 90.0, 20, 5778, 663, 250, 0.0, regs[1:])
    assert np.all(bi[b2 == 0.0] == 0.0)
    assert (bi == 1.0).sum() > 0


# ------------------------------------------------------------------ a6: integrate
def _planck(lam, temp):
    """loi_Planck with the constants of SOAP_integration.cu:205-212."""
    h, c, kb = 6.62606896e-34, 299792458.0, 1.380e-23
    return 2 * h * c ** 2 / (lam * 1e-10) ** 5 / (np.exp(h * c / (lam * 1e-10 * kb * temp)) - 1)


def _naive_integrate(w, maps, temp, bi, grid, solar=False):
    n = w.n
    flux = np.zeros(n)
    bconv = np.zeros(n)
    tot = np.zeros(n)
    cs, cf = w.cubes["spsed"][0], w.cubes["facsed"][0]
    t_norm = int(w.act["t_star"] + w.act["t_diff_fac"] + 0.9 - 216.8)
    for idx in np.nonzero((bi == 0.0) | (bi == 1.0))[0]:
        sh = _shifted_axis(w.wave, maps["delta"][idx])
        k, L = maps["region"][idx], maps["limb"][idx]
        it = lambda c: L * np.interp(w.wave, sh, w.cubes[c][k], left=0, right=0)
        q, qs = it("q"), it("qsed")
        if bi[idx] == 0.0:
            I, a = cs, it("sp")
        elif solar:
            I, a = _planck(w.wave, temp[idx]) / _planck(w.wave, t_norm) * cf, it("fac")
        else:
            I, a = (1 + maps["lb"][idx]) * cf, it("fac")
        flux += qs - I * q
        bconv += q - a
        tot += qs - a * I
    return flux, bconv, tot


def test_integrate_bruteforce():
    """orc_integrate_phase vs an independent numpy restatement of spot_integrate."""
    w = synth.small(grid=24, n=2500, n_mu=4)
    s, a = w.star, w.act
    maps = O.disk_init(24, s["prot"], 0.0, s["incl"], s["radius_km"], s["u1"], s["u2"], s["lb"], w.mu_bounds)
    temp, bi = O.phase_maps(24, s["incl"], a["nrho"], a["t_star"], a["t_diff_spot"], a["t_diff_fac"], 0.0,
                            w.regions[0])
    assert (bi == 0.0).sum() > 0 and (bi == 1.0).sum() > 0
    got = O.integrate_phase(O.GRID_LOG, w.wave, maps, temp, bi, w.cubes, 24, a["t_star"], a["t_diff_fac"])
    ref = _naive_integrate(w, maps, temp, bi, 24)
    sl = slice(40, w.n - 40)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g[sl] - r[sl])) / np.max(np.abs(r[sl])) < 1e-10


def test_integrate_bruteforce_solar():
    """Solar mode (SOAP_integration_solar.cu:424-433): facula intensity by the
    Planck ratio at the cell's plage temperature, against the numpy restatement;
    it differs from the (1+LB) form."""
    w = synth.small(grid=24, n=2500, n_mu=4)
    s, a = w.star, w.act
    maps = O.disk_init(24, s["prot"], 0.0, s["incl"], s["radius_km"], s["u1"], s["u2"], s["lb"], w.mu_bounds)
    temp, bi = O.phase_maps(24, s["incl"], a["nrho"], a["t_star"], a["t_diff_spot"], a["t_diff_fac"], 0.0,
                            w.regions[0])
    got = O.integrate_phase(O.GRID_LOG, w.wave, maps, temp, bi, w.cubes, 24, a["t_star"], a["t_diff_fac"],
                            solar_mode=1)
    ref = _naive_integrate(w, maps, temp, bi, 24, solar=True)
    plain = O.integrate_phase(O.GRID_LOG, w.wave, maps, temp, bi, w.cubes, 24, a["t_star"], a["t_diff_fac"])
    sl = slice(40, w.n - 40)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g[sl] - r[sl])) / np.max(np.abs(r[sl])) < 1e-10
    assert np.array_equal(got[1], plain[1])  # bconv has no intensity factor
    assert np.max(np.abs(got[0] - plain[0])) > 1e-3 * np.max(np.abs(plain[0]))


def test_integrate_spot_equals_quiet_seeds():
    """Spot and facula seeds equal to the quiet ones give bconv == 0 and flux == tot."""
    w = synth.small(grid=24, n=2048, n_mu=4)
    cubes = dict(w.cubes)
    cubes["sp"] = cubes["q"]
    cubes["fac"] = cubes["q"]
    s, a = w.star, w.act
    maps = O.disk_init(24, s["prot"], 0.0, s["incl"], s["radius_km"], s["u1"], s["u2"], s["lb"], w.mu_bounds)
    temp, bi = O.phase_maps(24, s["incl"], a["nrho"], a["t_star"], a["t_diff_spot"], a["t_diff_fac"], 0.0,
                            w.regions[0])
    flux, bconv, tot = O.integrate_phase(O.GRID_LOG, w.wave, maps, temp, bi, cubes, 24, a["t_star"],
                                         a["t_diff_fac"])
    assert np.all(bconv == 0.0)
    assert np.array_equal(flux, tot)


# ------------------------------------------------------------------ a7: resolution
def test_conv_window_matches_float_sqrt():
    """nb_pixel uses single-precision sqrtf (Q7): recompute with numpy float32."""
    wave = synth.log_grid(3800, 7000, dv_kms=0.25)
    sig = synth.sigma_inst_ms(115000)
    beta = sig / C
    dl = -wave[0] * (1 - float(np.sqrt(np.float32((1 + beta) / (1 - beta)))))
    assert O.conv_window(wave, sig) == int(O.rndup(50 * dl / (wave[1] - wave[0]), 0))


def test_resolution_gaussian_widths_add():
    """A Gaussian absorption line convolved with the instrumental Gaussian has
    sigma^2 = sigma_line^2 + sigma_inst^2 (variant 0, log grid), and the
    kernel is normalised (a flat spectrum stays flat)."""
    wave = synth.log_grid(5000, 5020, n=30001)
    sig = synth.sigma_inst_ms(60000)
    sl = 3000.0  # m/s
    lam0 = 5010.0
    line = 1 - 1e-3 * np.exp(-0.5 * ((wave - lam0) / (lam0 * sl / C)) ** 2)
    out, mm = O.lower_resolution(O.GRID_LOG, wave, sig, None, line, 0)
    # spec_norm is affine: recover the convolved profile o = (1 - out) max/(1-min)
    o = (1 - out) * mm[1] / (1 - mm[0])
    wid = O.conv_window(wave, sig)
    core = slice(wid + 1, len(wave) - wid - 2)
    x, prof = wave[core], o[core]
    dip = prof - prof.min()  # 1 - A is a small bump; its convolution too
    mean = np.sum(x * dip) / np.sum(dip)
    var = np.sum((x - mean) ** 2 * dip) / np.sum(dip)
    expect = (lam0 * sl / C) ** 2 + (lam0 * sig / C) ** 2
    assert abs(var / expect - 1) < 2e-3
    flat, mm2 = O.lower_resolution(O.GRID_LOG, wave, sig, None, np.full(len(wave), 0.25), 0)
    o2 = (1 - flat) * mm2[1] / (1 - mm2[0])
    assert np.max(np.abs(o2[core] / 0.75 - 1)) < 1e-6


def test_resolution_norm_sentinel():
    """max(o) is seeded at -60000 and ignores zeros (Q8): a spectrum whose
    convolved values are all below -60000 normalises by -60000 exactly."""
    wave = synth.log_grid(5000, 5010, n=8000)
    sig = synth.sigma_inst_ms(115000)
    spec = np.full(len(wave), 7.0e4)
    out, mm = O.lower_resolution(O.GRID_LOG, wave, sig, None, spec, 0)
    assert mm[1] == -60000.0 and mm[0] == 7.0e4
    out2, mm2 = O.lower_resolution(O.GRID_LOG, wave, sig, np.full(len(wave), 1.5), np.full(len(wave), 0.5), 1)
    assert mm2[0] == 1.0


# ------------------------------------------------------------------ a8: CCF
def test_ccf_single_line_box_times_gaussian():
    """One Gaussian line: the mask CCF is a box of width 820 m/s (relativistic)
    integrated over the line: erf closed form in velocity."""
    from scipy.special import erf
    wave = synth.log_grid(5000, 5020, n=40001)
    h = O.grid_h(wave)
    lam0, sv, depth = 5010.0, 3000.0, 0.5
    spec = 1 - depth * np.exp(-0.5 * ((wave - lam0) / (lam0 * sv / C)) ** 2)
    v = np.arange(-10000, 10001, 250.0)
    ccf = O.ccf(O.GRID_LOG, wave, h, v, np.array([lam0]), np.array([1.0]), spec)
    dv = h * C  # pixel width in velocity
    half = 410.0
    # mask integral in velocity units: (sum of mask weights) * pixel width
    vc = v
    a, b = (vc - half) / (sv * math.sqrt(2)), (vc + half) / (sv * math.sqrt(2))
    model = (2 * half - depth * sv * math.sqrt(math.pi / 2) * (erf(b) - erf(a))) / dv
    assert np.max(np.abs(ccf - model) / model) < 3e-3


def test_ccf_overwrite_later_line_wins():
    """Q17: two lines closer than the hole width; the later one overwrites the
    shared pixels.  Dense reference built independently in numpy."""
    wave = synth.log_grid(5000, 5010, n=12001)
    h = O.grid_h(wave)
    spec = np.random.default_rng(3).uniform(0.5, 1.0, len(wave))
    lines = np.array([5004.0, 5004.008])
    wts = np.array([0.9, 0.3])
    v = np.array([-500.0, 0.0, 250.0])
    got = O.ccf(O.GRID_LOG, wave, h, v, lines, wts, spec)
    for iv, vel in enumerate(v):
        mask = np.zeros(len(wave))
        for l, wt in zip(lines, wts):
            ls = l * (1 + vel / C)
            beta = 820 / C
            hw = ls * (math.sqrt((1 + beta) / (1 - beta)) - 1)
            xb, xe = math.log((ls - hw / 2) / wave[0]) / h, math.log((ls + hw / 2) / wave[0]) / h
            aa, bb = math.ceil(xb), math.ceil(xe) - 1
            lam = lambda k: wave[0] * math.exp(k * h)
            fb = (lam(aa) - (ls - hw / 2)) / (lam(aa) - lam(aa - 1))
            fa = abs(lam(bb) - (ls + hw / 2)) / (lam(bb + 1) - lam(bb))
            mask[aa:bb] = wt
            mask[aa - 1] = wt * fb
            mask[bb] = wt * fa
        assert got[iv] == pytest.approx(np.dot(mask, spec), rel=1e-9)


def test_appodisation_window():
    ap = O.appodisation(161)
    assert len(ap) == 161 and np.all(ap[:5] == 0) and np.all(ap[-5:] == 0)
    assert np.all(ap[45:116] == 1.0)
    np.testing.assert_allclose(ap[5:45], 0.5 * np.cos(2 * np.pi / 80 * np.arange(40) - np.pi) + 0.5, atol=1e-16)
    np.testing.assert_allclose(ap, ap[::-1], atol=1e-16)


# ------------------------------------------------------------------ a9: fit
def _gauss(v, c, k, x0, fwhm):
    s = fwhm / (2 * math.sqrt(2 * math.log(2)))
    return c + k * np.exp(-(v - x0) ** 2 / (2 * s * s))


def test_fit_noiseless_gaussian_exact():
    """A noiseless Gaussian CCF recovers its parameters; symmetric => span ~ 0."""
    v = np.arange(-20000, 20001, 250.0)
    y = _gauss(v, 0.98, -0.45, 123.4, 7100.0)
    r = O.fit_ccf(v, y)
    c, k, span, v0, fwhm = r["par"]
    assert v0 == pytest.approx(123.4, abs=1e-6)
    assert fwhm == pytest.approx(7100.0, abs=1e-5)
    assert c == pytest.approx(0.98, abs=1e-10) and k == pytest.approx(0.45, abs=1e-10)
    assert abs(span) < 1e-3
    np.testing.assert_allclose(r["model"], y, atol=1e-10)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_fit_matches_scipy_least_squares(seed):
    """Noisy asymmetric CCFs: the oracle's LM converges to the same
    least-squares optimum as scipy's MINPACK LM (RV/FWHM within 1e-4 m/s)."""
    from scipy.optimize import least_squares
    rng = np.random.default_rng(seed)
    v = np.arange(-20000, 20001, 250.0)
    y = _gauss(v, 1.0, -0.5, rng.uniform(-800, 800), rng.uniform(6000, 9000))
    y += 0.03 * np.exp(-((v - 2000) / 3000) ** 2) + 1e-3 * rng.standard_normal(v.size)
    r = O.fit_ccf(v, y)
    px = v - v[0]
    x0 = np.array([r["par"][0], -r["par"][1], r["par"][3] - v[0], r["par"][4]])
    sol = least_squares(lambda p: _gauss(px, *p) - y, x0 * (1 + 1e-3), method="lm", xtol=1e-15, ftol=1e-15,
                        gtol=1e-15)
    assert r["par"][3] == pytest.approx(sol.x[2] + v[0], abs=1e-4)
    assert r["par"][4] == pytest.approx(sol.x[3], abs=1e-4)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fit_weighted_matches_scipy_least_squares(seed):
    """Per-point errors (CCF_error.txt): the optimum of sum ((fit - y)/err)^2,
    as scipy's MINPACK LM finds it.  Errors that grow into the wings move the
    optimum away from the unit-error one (so dropping the weights fails), and
    scaling every error by a constant leaves it unchanged."""
    from scipy.optimize import least_squares
    rng = np.random.default_rng(10 + seed)
    v = np.arange(-20000, 20001, 250.0)
    y = _gauss(v, 1.0, -0.5, rng.uniform(-800, 800), rng.uniform(6000, 9000))
    y += 0.03 * np.exp(-((v - 2500) / 3000) ** 2) + 1e-3 * rng.standard_normal(v.size)
    err = 1e-3 * (1 + (v / 8000.0) ** 2) * rng.uniform(0.5, 2.0, v.size)
    r = O.fit_ccf(v, y, err)
    px = v - v[0]
    x0 = np.array([r["par"][0], -r["par"][1], r["par"][3] - v[0], r["par"][4]])
    sol = least_squares(lambda p: (_gauss(px, *p) - y) / err, x0 * (1 + 1e-3), method="lm", xtol=1e-15,
                        ftol=1e-15, gtol=1e-15)
    assert r["par"][3] == pytest.approx(sol.x[2] + v[0], abs=1e-4)
    assert r["par"][4] == pytest.approx(sol.x[3], abs=1e-4)
    assert r["par"][0] == pytest.approx(sol.x[0], abs=1e-9)
    r1 = O.fit_ccf(v, y)
    assert abs(r1["par"][3] - r["par"][3]) > 1.0  # the weights matter here (m/s)
    r2 = O.fit_ccf(v, y, err * 7.0)
    assert r2["par"][3] == pytest.approx(r["par"][3], abs=1e-6)
    assert np.array_equal(O.fit_ccf(v, y, np.ones(v.size))["par"], r1["par"])


def test_bisector_shape_and_depths():
    """Bisector of a skewed CCF: depth grid 1 - |k| (i+5)/100, and a line with a
    red wing gives a positive-slope bisector (span sign)."""
    v = np.arange(-20000, 20001, 250.0)
    y = _gauss(v, 1.0, -0.5, 0.0, 7000.0) - 0.02 * np.exp(-((v - 3000) / 2500) ** 2)
    r = O.fit_ccf(v, y)
    k = r["par"][1]
    np.testing.assert_allclose(r["depth"], 1 - k * (np.arange(91) + 5) / 100, atol=1e-15)
    assert np.all(np.isfinite(r["bis"]))
