"""Pins of the oracle's differentiable model (oracle/soap_oracle.c, a11 section;
no counterpart in the reference): the Gaussian-fit Hessian against finite
differences, the Newton-polished fit against scipy's least squares, and the
forward model against the quiet-star path of the reference's pipeline; the
bisector (shared by the fit and the model's BIS term) against the closed forms of
an exact Gaussian."""
import numpy as np
import pytest
from scipy.optimize import least_squares

import oracle as O
from soap_gpu_amd import synth

TRIM = 210


def _gauss(x, th):
    s = th[3] / (2 * np.sqrt(2 * np.log(2)))
    return th[0] + th[1] * np.exp(-((x - th[2]) ** 2) / (2 * s * s))


def _ccf_like(seed=0, n=81):
    rng = np.random.default_rng(seed)
    x = np.linspace(-20000, 20000, n)
    y = _gauss(x, [1.0, -0.4, 350.0, 7000.0]) + 1e-3 * rng.standard_normal(n)
    return x, y


def test_gauss_hessian_finite_differences():
    """orc_gauss_hess: gradient and exact Hessian of 0.5 sum r^2 against central
    differences of the objective and of the gradient (catches a wrong sign or a
    missing second-derivative term)."""
    x, y = _ccf_like()
    th = np.array([0.99, -0.38, 310.0, 7200.0])
    g, H, J, r2 = O.gauss_hess(x, y, th)
    obj = lambda t: 0.5 * np.sum((_gauss(x, t) - y) ** 2)
    steps = np.array([1e-6, 1e-6, 1e-3, 1e-2])
    gn = np.array([(obj(th + steps[a] * np.eye(4)[a]) - obj(th - steps[a] * np.eye(4)[a])) / (2 * steps[a])
                   for a in range(4)])
    np.testing.assert_allclose(g, gn, rtol=1e-6, atol=1e-12)
    Hn = np.array([(O.gauss_hess(x, y, th + steps[a] * np.eye(4)[a])[0] -
                    O.gauss_hess(x, y, th - steps[a] * np.eye(4)[a])[0]) / (2 * steps[a]) for a in range(4)])
    np.testing.assert_allclose(H, Hn.T, rtol=1e-5, atol=1e-9 * np.abs(H).max())
    np.testing.assert_allclose(J[:, 1], np.exp(-((x - th[2]) ** 2) / (2 * (th[3] / 2.3548200450309493) ** 2)),
                               rtol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gauss_polish_reaches_least_squares_optimum(seed):
    """orc_gauss_polish from the BIS_FIT2-tolerance fit lands on scipy's least
    squares optimum, with a vanishing gradient."""
    x, y = _ccf_like(seed)
    r = O.fit_ccf(x, y)
    th0 = np.array([r["par"][0], -r["par"][1], r["par"][3], r["par"][4]])
    th = O.gauss_polish(x, y, th0)
    g, H, _, _ = O.gauss_hess(x, y, th)
    assert np.all(np.abs(g) <= 1e-10 * np.sqrt(np.diag(H)) * np.abs(th).max())
    ref = least_squares(lambda t: _gauss(x, t) - y, th0, xtol=1e-15, ftol=1e-15, gtol=1e-15, method="lm").x
    np.testing.assert_allclose(th, ref, rtol=1e-9)


@pytest.mark.parametrize("v0,sig", [(137.0, 4000.0), (-2310.0, 3000.0), (71.0, 6000.0)])
def test_bisector_of_a_gaussian(v0, sig):
    """orc_bisector (BIS_FIT2.cpp:214-268) on an exact Gaussian CCF, centre off the
    velocity grid: the bisector is the centre at every depth to the accuracy of the
    per-interval quadratic inverse (< 3 m/s on a 500 m/s grid; the linear inverse,
    p2 = 0, is off by 6-8 m/s and a sign error in p2 by 12-16 m/s), the span
    vanishes, and translating the grid and the centre translates the bisector."""
    v = np.arange(-20000.0, 20001.0, 500.0)
    fw = sig * 2 * np.sqrt(2 * np.log(2))
    th = (1.0, -0.6, v0, fw)
    y = _gauss(v, th)
    span, depth, bis = O.bisector(v, y, th)
    assert np.abs(bis - v0).max() < 3.0
    assert abs(span) < 0.1
    np.testing.assert_allclose(depth, 1 - 0.6 * (np.arange(91) + 5) / 100, rtol=0, atol=1e-15)
    sh = 1234.5
    span2, _, bis2 = O.bisector(v + sh, y, (1.0, -0.6, v0 + sh, fw))
    np.testing.assert_allclose(bis2 - sh, bis, rtol=0, atol=1e-8)
    assert span2 == pytest.approx(span, abs=1e-8)
    # an affine change of the CCF (continuum and amplitude) leaves the bisector unchanged
    span3, _, bis3 = O.bisector(v, 3.0 + 2.5 * (y - 1.0), (3.0, -1.5, v0, fw))
    np.testing.assert_allclose(bis3, bis, rtol=0, atol=1e-8)


def test_bisector_matches_fit_span():
    """The span BIS_FIT2 reports (orc_fit_ccf par[2]) is orc_bisector at the fitted
    Gaussian (c, signed k, RV, FWHM)."""
    x, y = _ccf_like(3)
    r = O.fit_ccf(x, y)
    c, k_abs, span, rv, fw = r["par"]
    k = -k_abs if y.min() < c else k_abs
    s2, d2, b2 = O.bisector(x, y, (c, k, rv, fw))
    assert s2 == span
    np.testing.assert_array_equal(b2, r["bis"])


def _problem(grid=20, n=2400, n_mu=4, P=2, M=10, seed=5):
    w = synth.small(grid=grid, n=n, n_mu=n_mu)
    rng = np.random.default_rng(seed)
    maps = np.clip(rng.uniform(-0.2, 0.5, (P, M, M)), 0, 1)
    wave_t = w.wave[TRIM:w.n - TRIM]
    sl, sw = O.select_lines(wave_t, w.mask_l, w.mask_w)
    return w, maps, sl, sw


def test_diff_forward_without_activity_is_the_quiet_star():
    """With empty maps, F is the instrumental convolution of the reference's quiet
    star (itot, SOAP_initialize.cu:339-395) for rigid rotation vsini y and quadratic
    limb darkening; the convolution taps are the log-grid taps of the resolution
    stage."""
    w, maps, sl, sw = _problem()
    maps[:] = 0.0
    vsini, u1, u2 = 3.0, 0.29, 0.34
    r = O.diff_forward(w.wave, 20, w.mu_bounds, vsini, u1, u2, 0.3, w.cubes["q"], w.cubes["sp"], maps,
                       w.sigma_inst, TRIM, w.vgrid, sl, sw, np.zeros((2, w.n)), np.zeros(2), np.zeros(2),
                       1.0, 0.0, 0.0)
    prot = synth.prot_for_vsini(vsini)
    dm = O.disk_init(20, prot, 0.0, 90.0, synth.R_SUN_KM, u1, u2, [0, 0, 0, 0], w.mu_bounds)
    q = O.quiet_star(O.GRID_LOG, w.wave, dm, w.cubes["q"], 20)
    h = O.grid_h(w.wave)
    wid = O.conv_window(w.wave, w.sigma_inst)
    fs = O.delta_lambda(w.sigma_inst, 1.0)
    d = np.arange(-wid, wid - 1)
    u = np.expm1(d * h) / fs
    taps = np.exp(d * h) * np.expm1(h) / (fs * np.sqrt(2 * np.pi)) * np.exp(-u * u / 2)
    ref = np.zeros(w.n)
    for i in range(wid + 1, w.n - wid - 1):
        ref[i] = np.dot(taps, q[i - wid:i + wid - 1])
    for p in range(2):
        sl_ = slice(TRIM, w.n - TRIM)
        assert np.max(np.abs(r["F"][p][sl_] - ref[sl_])) <= 1e-11 * np.max(np.abs(ref))


def test_diff_forward_spot_seed_equal_to_quiet_ignores_maps():
    """kappa = 1 and S = Q: (1 - phi) Q + kappa phi S = Q for any map."""
    w, maps, sl, sw = _problem()
    args = (w.wave, 20, w.mu_bounds, 2.5, 0.3, 0.2, 1.0, w.cubes["q"], w.cubes["q"])
    tail = (w.sigma_inst, TRIM, w.vgrid, sl, sw, np.zeros((2, w.n)), np.zeros(2), np.zeros(2), 1.0, 1.0, 1.0)
    a = O.diff_forward(*args, maps, *tail)
    b = O.diff_forward(*args, np.zeros_like(maps), *tail)
    np.testing.assert_allclose(a["F"], b["F"], rtol=0, atol=1e-11 * np.abs(b["F"]).max())
    assert abs(a["loss"] - b["loss"]) <= 1e-11 * b["loss"]


def test_diff_forward_rv_follows_a_dark_spot():
    """A dark region on the approaching (y < 0) half removes blueshifted light, so
    the RV of the CCF fit increases; on the receding half it decreases."""
    w, maps, sl, sw = _problem(grid=32)
    M = maps.shape[1]
    maps[:] = 0.0
    maps[0, : M // 3, :] = 1.0  # y < 0 rows of phase 0 covered
    maps[1, -M // 3:, :] = 1.0  # y > 0 rows of phase 1 covered
    r = O.diff_forward(w.wave, 32, w.mu_bounds, 5.0, 0.29, 0.34, 0.2, w.cubes["q"], w.cubes["q"], maps,
                       w.sigma_inst, TRIM, w.vgrid, sl, sw, np.zeros((2, w.n)), np.zeros(2), np.zeros(2),
                       1.0, 0.0, 0.0)
    q0 = O.diff_forward(w.wave, 32, w.mu_bounds, 5.0, 0.29, 0.34, 0.2, w.cubes["q"], w.cubes["q"],
                        np.zeros_like(maps), w.sigma_inst, TRIM, w.vgrid, sl, sw, np.zeros((2, w.n)), np.zeros(2),
                        np.zeros(2), 1.0, 0.0, 0.0)
    rv0 = q0["theta"][0, 2]
    assert r["theta"][0, 2] > rv0 + 10 and r["theta"][1, 2] < rv0 - 10


def _fac_args(w, maps, sl, sw, vsini=2.5, u=(0.3, 0.2), kappa=0.4):
    head = (w.wave, 20, w.mu_bounds, vsini, u[0], u[1], kappa, w.cubes["q"], w.cubes["sp"], maps)
    tail = (w.sigma_inst, TRIM, w.vgrid, sl, sw, np.zeros((2, w.n)), np.zeros(2), np.zeros(2), 1.0, 1.0, 1.0)
    return head, tail


def test_diff_forward_empty_facula_maps_equal_the_spot_model():
    """Facula maps of zeros leave the model as the two-population one, bit for bit
    (psi = 0 adds zero terms only)."""
    w, maps, sl, sw = _problem()
    head, tail = _fac_args(w, maps, sl, sw)
    a = O.diff_forward(*head, *tail)
    b = O.diff_forward(*head, *tail, kappa_f=1.3, lb=[0.2, -0.1, 0.3, 0.05], cube_f=w.cubes["fac"],
                       fmaps=np.zeros_like(maps))
    np.testing.assert_array_equal(a["F"], b["F"])
    assert a["loss"] == b["loss"]


def test_diff_forward_facula_equal_to_quiet_ignores_facula_maps():
    """Fa = Q, kappa_f = 1 and LB = 0: the facula term kappa_f (1 + LB) psi Q cancels
    the -psi Q of the quiet weight for any facula map."""
    w, maps, sl, sw = _problem()
    head, tail = _fac_args(w, maps, sl, sw)
    a = O.diff_forward(*head, *tail)
    fm = np.clip(np.random.default_rng(9).uniform(-0.2, 0.6, maps.shape), 0, 1)
    b = O.diff_forward(*head, *tail, kappa_f=1.0, lb=[0, 0, 0, 0], cube_f=w.cubes["q"], fmaps=fm)
    np.testing.assert_allclose(b["F"], a["F"], rtol=0, atol=1e-12 * np.abs(a["F"]).max())


def test_diff_forward_facula_intensity_is_one_plus_lb_of_mu():
    """A disk fully covered by faculae (psi = 1, phi = 0) with Fa = Q is the quiet
    star of the reference's quiet-star path (itot, SOAP_initialize.cu:339-395) with
    the limb weight multiplied by kappa_f (1 + LB(mu)) (SOAP_integration.cu:317-323),
    LB = lb0 + lb1 mu + lb2 mu^2 + lb3 mu^3 with mu the cell's own mu; the quiet weight
    1 - psi vanishes.  A polynomial in mu of every degree up to 3 (a wrong power or
    a dropped 1 + fails), checked against an independent path: disk_init's limb
    map scaled and fed to quiet_star, then the log-grid convolution."""
    w, maps, sl, sw = _problem()
    vsini, u1, u2, kf = 3.0, 0.29, 0.34, 0.7
    lb = np.array([0.23, -0.4, 0.5, 0.25])
    zero = np.zeros_like(maps)
    r = O.diff_forward(w.wave, 20, w.mu_bounds, vsini, u1, u2, 0.3, w.cubes["q"], w.cubes["sp"], zero,
                       w.sigma_inst, TRIM, w.vgrid, sl, sw, np.zeros((2, w.n)), np.zeros(2), np.zeros(2),
                       1.0, 0.0, 0.0, kappa_f=kf, lb=lb, cube_f=w.cubes["q"], fmaps=np.ones_like(maps))
    prot = synth.prot_for_vsini(vsini)
    dm = O.disk_init(20, prot, 0.0, 90.0, synth.R_SUN_KM, u1, u2, [0, 0, 0, 0], w.mu_bounds)
    mu = dm["mu"]
    on = mu >= 0.0  # off-disk cells hold -999
    dm["limb"] = np.where(on, dm["limb"] * kf * (1 + lb[0] + lb[1] * mu + lb[2] * mu ** 2 + lb[3] * mu ** 3),
                          dm["limb"])
    q = O.quiet_star(O.GRID_LOG, w.wave, dm, w.cubes["q"], 20)
    ref = O.conv_raw(O.GRID_LOG, w.wave, w.sigma_inst, 1.0 - q, 0)
    sl_ = slice(TRIM, w.n - TRIM)
    for p in range(2):
        assert np.max(np.abs(r["F"][p][sl_] - ref[sl_])) <= 1e-11 * np.max(np.abs(ref))
