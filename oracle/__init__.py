"""ctypes front end of the FP64 CPU oracle (oracle/soap_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The engine in soap_gpu_amd/
never imports this package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

GRID_LINEAR = 0
GRID_LOG = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Region(ctypes.Structure):
    _fields_ = [("lat", ctypes.c_double), ("lon", ctypes.c_double),
                ("size", ctypes.c_double), ("type", ctypes.c_double)]


def build() -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_rndup.restype = ctypes.c_double
        L.orc_rndup.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_delta_lambda.restype = ctypes.c_double
        L.orc_delta_lambda.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_delta_lambda_cpu.restype = ctypes.c_double
        L.orc_delta_lambda_cpu.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_grid_h.restype = ctypes.c_double
        L.orc_grid_h.argtypes = [_dp, ctypes.c_long]
        L.orc_lb_params.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_disk_init.argtypes = [ctypes.c_int] + [ctypes.c_double] * 6 + [_dp, _dp, ctypes.c_int,
                                                                              _dp, _dp, _dp, _ip, _dp]
        L.orc_shift_row.argtypes = [ctypes.c_int, _dp, ctypes.c_long, ctypes.c_double, _dp,
                                    ctypes.c_double, ctypes.c_double, _dp]
        L.orc_scatter_row_ref.argtypes = [_dp, ctypes.c_long, ctypes.c_double, _dp, ctypes.c_double,
                                          ctypes.c_double, _dp]
        L.orc_quiet_star.argtypes = [ctypes.c_int, _dp, ctypes.c_long, ctypes.c_double, ctypes.c_int,
                                     _dp, _dp, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_quiet_star_ref.argtypes = [_dp, ctypes.c_long, ctypes.c_double, ctypes.c_int, _dp, _dp,
                                         _ip, _dp, ctypes.c_int, _dp]
        L.orc_region_bbox.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_int,
                                                              ctypes.c_double, _ip]
        L.orc_phase_maps.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     ctypes.POINTER(Region), ctypes.c_int, _dp, _dp]
        L.orc_integrate_phase.argtypes = ([ctypes.c_int, _dp, ctypes.c_long, ctypes.c_double, ctypes.c_int,
                                           _dp, _dp, _ip, _dp, _dp, _dp] + [_dp] * 6
                                          + [ctypes.c_int] * 4 + [_dp, _dp, _dp, ctypes.c_int])
        L.orc_conv_window.argtypes = [_dp, ctypes.c_double]
        L.orc_conv_raw.argtypes = [ctypes.c_int, _dp, ctypes.c_long, ctypes.c_double, _dp, ctypes.c_int,
                                   ctypes.c_double, _dp, ctypes.c_int]
        L.orc_lower_resolution.argtypes = [ctypes.c_int, _dp, ctypes.c_long, ctypes.c_double, _dp, _dp,
                                           ctypes.c_int, ctypes.c_double, _dp, _dp, ctypes.c_int]
        L.orc_select_lines.restype = ctypes.c_long
        L.orc_select_lines.argtypes = [_dp, ctypes.c_long, _dp, _dp, ctypes.c_long, _dp, _dp]
        L.orc_ccf.argtypes = [ctypes.c_int, _dp, ctypes.c_long, ctypes.c_double, _dp, ctypes.c_int,
                              _dp, _dp, ctypes.c_long, _dp, _dp, ctypes.c_int]
        L.orc_appodisation.argtypes = [ctypes.c_int, _dp]
        L.orc_ccf_finish.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_fit_ccf.argtypes = [_dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]
        L.orc_fit_ccf_err.argtypes = [_dp, _dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _ip]
        L.orc_gauss_hess.argtypes = [_dp, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_gauss_polish.argtypes = [_dp, _dp, ctypes.c_int, _dp]
        _d_ = ctypes.c_double
        L.orc_bisector.argtypes = [_dp, _dp, ctypes.c_int, _d_, _d_, _d_, _d_, _dp, _dp]
        L.orc_bisector.restype = _d_
        L.orc_diff_forward.argtypes = ([_dp, ctypes.c_long, ctypes.c_int, ctypes.c_int, _dp, _d_, _d_, _d_, _d_,
                                        _d_, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _d_,
                                        ctypes.c_int, _dp,
                                        ctypes.c_int, _dp, _dp, ctypes.c_long, _dp, _dp, _dp, _dp, _d_, _d_, _d_,
                                        _d_, _dp, _dp, _dp, _dp, _dp, ctypes.c_int])
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def nthreads_default() -> int:
    return int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))


# ---------------------------------------------------------------- scalars
def rndup(x, nb=0):
    return lib().orc_rndup(float(x), int(nb))


def delta_lambda(v_ms, lam):
    return lib().orc_delta_lambda(float(v_ms), float(lam))


def delta_lambda_cpu(v_ms, lam):
    return lib().orc_delta_lambda_cpu(float(v_ms), float(lam))


def grid_h(wave):
    w, wp = _d(wave)
    return lib().orc_grid_h(wp, len(w))


def lb_params(t_star, lb_g2, lb_k0, lb_m0, temps):
    g, gp = _d(lb_g2)
    k, kp = _d(lb_k0)
    m, mp = _d(lb_m0)
    t, tp = _d(temps)
    out = np.zeros(4)
    lib().orc_lb_params(int(t_star), gp, kp, mp, tp, out.ctypes.data_as(_dp))
    return out


# ------------------------------------------------------------------- a1
def disk_init(grid, prot, w_ratio, incl_deg, radius_km, u1, u2, lb4, mu_bounds):
    lb, lbp = _d(lb4)
    mb, mbp = _d(mu_bounds)
    n_mu = len(mb) + 1
    g2 = grid * grid
    mu = np.empty(g2)
    delta = np.empty(g2)
    limb = np.empty(g2)
    region = np.empty(g2, dtype=np.int32)
    lbm = np.empty(g2)
    lib().orc_disk_init(int(grid), float(prot), float(w_ratio), float(incl_deg), float(radius_km),
                        float(u1), float(u2), lbp, mbp, n_mu,
                        mu.ctypes.data_as(_dp), delta.ctypes.data_as(_dp), limb.ctypes.data_as(_dp),
                        region.ctypes.data_as(_ip), lbm.ctypes.data_as(_dp))
    return dict(mu=mu, delta=delta, limb=limb, region=region, lb=lbm)


# ------------------------------------------------------------------- a4
def shift_row(kind, wave, row, delta_kms, limb=1.0, step=0.0):
    w, wp = _d(wave)
    r, rp = _d(row)
    out = np.empty(len(w))
    lib().orc_shift_row(int(kind), wp, len(w), float(step), rp, float(delta_kms), float(limb),
                        out.ctypes.data_as(_dp))
    return out


def scatter_row_ref(wave, row, delta_kms, limb, step, buf):
    w, wp = _d(wave)
    r, rp = _d(row)
    assert buf.dtype == np.float64 and buf.flags.c_contiguous
    lib().orc_scatter_row_ref(wp, len(w), float(step), rp, float(delta_kms), float(limb),
                              buf.ctypes.data_as(_dp))
    return buf


# ------------------------------------------------------------------- a2
def quiet_star(kind, wave, maps, cube, grid, step=0.0, nthreads=None):
    w, wp = _d(wave)
    c, cp = _d(cube)
    n_mu = c.shape[0]
    de, dep = _d(maps["delta"])
    li, lip = _d(maps["limb"])
    rg, rgp = _i(maps["region"])
    out = np.empty(len(w))
    lib().orc_quiet_star(int(kind), wp, len(w), float(step), int(grid), dep, lip, rgp, cp, n_mu,
                         out.ctypes.data_as(_dp), nthreads or nthreads_default())
    return out


def quiet_star_ref(wave, maps, cube, grid, step):
    w, wp = _d(wave)
    c, cp = _d(cube)
    de, dep = _d(maps["delta"])
    li, lip = _d(maps["limb"])
    rg, rgp = _i(maps["region"])
    out = np.empty(len(w))
    lib().orc_quiet_star_ref(wp, len(w), float(step), int(grid), dep, lip, rgp, cp, c.shape[0],
                             out.ctypes.data_as(_dp))
    return out


# ------------------------------------------------------------------- a5
def region_bbox(lat, lon, size, incl_deg, nrho, grid, phase):
    b = np.zeros(5, dtype=np.int32)
    lib().orc_region_bbox(float(lat), float(lon), float(size), float(incl_deg), int(nrho), int(grid),
                          float(phase), b.ctypes.data_as(_ip))
    return b


def phase_maps(grid, incl_deg, nrho, t_star, t_diff_spot, t_diff_fac, phase, regions):
    regs = (Region * max(1, len(regions)))()
    for i, r in enumerate(regions):
        regs[i].lat, regs[i].lon, regs[i].size, regs[i].type = r
    g2 = grid * grid
    temp = np.empty(g2)
    bi = np.empty(g2)
    lib().orc_phase_maps(int(grid), float(incl_deg), int(nrho), int(t_star), int(t_diff_spot),
                         int(t_diff_fac), float(phase), regs, len(regions),
                         temp.ctypes.data_as(_dp), bi.ctypes.data_as(_dp))
    return temp, bi


# ------------------------------------------------------------------- a6
def integrate_phase(kind, wave, maps, temp, bi, cubes, grid, t_star, t_diff_fac, solar_mode=0,
                    step=0.0, nthreads=None):
    """cubes: dict with keys q, qsed, sp, spsed, fac, facsed -> [n_mu, n] arrays."""
    w, wp = _d(wave)
    n = len(w)
    keep = [_d(cubes[k]) for k in ("q", "qsed", "sp", "spsed", "fac", "facsed")]
    de, dep = _d(maps["delta"])
    li, lip = _d(maps["limb"])
    rg, rgp = _i(maps["region"])
    lbm, lbp = _d(maps["lb"])
    te, tep = _d(temp)
    bb, bbp = _d(bi)
    flux = np.empty(n)
    bconv = np.empty(n)
    tot = np.empty(n)
    lib().orc_integrate_phase(int(kind), wp, n, float(step), int(grid), dep, lip, rgp, lbp, tep, bbp,
                              *[k[1] for k in keep], keep[0][0].shape[0], int(t_star), int(t_diff_fac),
                              int(solar_mode), flux.ctypes.data_as(_dp), bconv.ctypes.data_as(_dp),
                              tot.ctypes.data_as(_dp), nthreads or nthreads_default())
    return flux, bconv, tot


# ------------------------------------------------------------------- a7
def conv_window(wave, sigma_ms):
    w, wp = _d(wave)
    return lib().orc_conv_window(wp, float(sigma_ms))


def conv_raw(kind, wave, sigma_ms, A, variant, weight=0.005, nthreads=None):
    """conv_inst alone: the unnormalised convolution of 1 - A (zero outside the valid range)."""
    w, wp = _d(wave)
    a, ap = _d(A)
    out = np.empty(len(w))
    lib().orc_conv_raw(int(kind), wp, len(w), float(sigma_ms), ap, int(variant), float(weight),
                       out.ctypes.data_as(_dp), nthreads or nthreads_default())
    return out


def lower_resolution(kind, wave, sigma_ms, quiet, spec, variant, weight=0.005, nthreads=None):
    w, wp = _d(wave)
    q, qp = _d(quiet if quiet is not None else spec)
    s, sp = _d(spec)
    out = np.empty(len(w))
    mm = np.zeros(2)
    lib().orc_lower_resolution(int(kind), wp, len(w), float(sigma_ms), qp, sp, int(variant), float(weight),
                               out.ctypes.data_as(_dp), mm.ctypes.data_as(_dp), nthreads or nthreads_default())
    return out, mm


# ------------------------------------------------------------------- a8
def select_lines(wave_t, line_l, line_w):
    w, wp = _d(wave_t)
    l, lp = _d(line_l)
    ww, wwp = _d(line_w)
    sl = np.empty(len(l))
    sw = np.empty(len(l))
    k = lib().orc_select_lines(wp, len(w), lp, wwp, len(l), sl.ctypes.data_as(_dp), sw.ctypes.data_as(_dp))
    return sl[:k].copy(), sw[:k].copy()


def ccf(kind, wave_t, step, vgrid, line_l, line_w, spec_t, nthreads=None):
    w, wp = _d(wave_t)
    v, vp = _d(vgrid)
    l, lp = _d(line_l)
    ww, wwp = _d(line_w)
    s, sp = _d(spec_t)
    out = np.empty(len(v))
    lib().orc_ccf(int(kind), wp, len(w), float(step), vp, len(v), lp, wwp, len(l), sp,
                  out.ctypes.data_as(_dp), nthreads or nthreads_default())
    return out


def appodisation(n_v):
    out = np.empty(n_v)
    lib().orc_appodisation(int(n_v), out.ctypes.data_as(_dp))
    return out


def ccf_finish(ccf_raw, appod):
    c = np.array(ccf_raw, dtype=np.float64, copy=True)
    a, ap = _d(appod)
    lib().orc_ccf_finish(len(c), ap, c.ctypes.data_as(_dp))
    return c


# ------------------------------------------------------------------- a9
def fit_ccf(v, y, err=None):
    """BIS_FIT2's fit + bisector; err: per-point errors (CCF_error.txt), None = ones."""
    vv, vp = _d(v)
    yy, yp = _d(y)
    n = len(vv)
    par = np.zeros(5)
    model = np.zeros(n)
    depth = np.zeros(91)
    bis = np.zeros(91)
    it = np.zeros(1, dtype=np.int32)
    if err is None:
        ep = None
    else:
        ee, ep = _d(err)
    st = lib().orc_fit_ccf_err(vp, yp, ep, n, par.ctypes.data_as(_dp), model.ctypes.data_as(_dp),
                               depth.ctypes.data_as(_dp), bis.ctypes.data_as(_dp), it.ctypes.data_as(_ip))
    return dict(par=par, model=model, depth=depth, bis=bis, iters=int(it[0]), status=int(st))


# ------------------------------------------------------------------- a11 (differentiable model)
def gauss_hess(x, y, theta):
    """(gradient[4], Hessian[4,4], J[n,4], sum r^2) of 0.5 sum r^2 for the BIS_FIT2 Gaussian."""
    xx, xp = _d(x)
    yy, yp = _d(y)
    th, tp = _d(theta)
    g = np.zeros(4)
    H = np.zeros((4, 4))
    J = np.zeros((len(xx), 4))
    r2 = np.zeros(1)
    lib().orc_gauss_hess(xp, yp, len(xx), tp, g.ctypes.data_as(_dp), H.ctypes.data_as(_dp),
                         J.ctypes.data_as(_dp), r2.ctypes.data_as(_dp))
    return g, H, J, float(r2[0])


def gauss_polish(x, y, theta):
    xx, xp = _d(x)
    yy, yp = _d(y)
    th = np.array(theta, dtype=np.float64, copy=True)
    lib().orc_gauss_polish(xp, yp, len(xx), th.ctypes.data_as(_dp))
    return th


def bisector(v, y, theta):
    """(span, depth[91], bis[91]) of a CCF for the Gaussian theta = (c, k signed, v0, fwhm)."""
    vv, vp = _d(v)
    yy, yp = _d(y)
    depth = np.zeros(91)
    bis = np.zeros(91)
    c, k, v0, fw = (float(x) for x in theta)
    span = lib().orc_bisector(vp, yp, len(vv), c, k, v0, fw, depth.ctypes.data_as(_dp), bis.ctypes.data_as(_dp))
    return float(span), depth, bis


def diff_forward(wave, grid, mu_bounds, vsini, u1, u2, kappa, cube_q, cube_s, maps, sigma_ms, trim, vgrid,
                 line_l, line_w, target, rv_t, fw_t, w_spec, w_rv, w_fw, bis_t=None, w_bis=0.0, nthreads=None,
                 kappa_f=0.0, lb=None, cube_f=None, fmaps=None):
    """Forward model + loss of the differentiable path; returns dict(loss, F[P,n], ccf[P,nv], theta[P,4],
    span[P]).  cube_f / fmaps (both or neither): facula seed rows and facula filling-factor maps,
    weighted by kappa_f (1 + LB(mu)) with LB the polynomial lb (constant term first)."""
    w, wp = _d(wave)
    n = len(w)
    mb, mbp = _d(mu_bounds)
    q, qp = _d(cube_q)
    s, sp = _d(cube_s)
    if (cube_f is None) != (fmaps is None):
        raise ValueError("cube_f and fmaps go together")
    f_, fp_ = (None, None) if cube_f is None else _d(cube_f)
    fm_, fmp_ = (None, None) if fmaps is None else _d(fmaps)
    lb_, lbp_ = (None, None) if lb is None else _d(lb)
    mp_, mpp = _d(maps)
    P, M = mp_.shape[0], mp_.shape[1]
    v, vp = _d(vgrid)
    ll, llp = _d(line_l)
    lw, lwp = _d(line_w)
    t, tp = _d(target)
    rt, rtp = _d(rv_t)
    ft, ftp = _d(fw_t)
    bt, btp = _d(np.zeros(P) if bis_t is None else bis_t)
    F = np.zeros((P, n))
    cc = np.zeros((P, len(v)))
    th = np.zeros((P, 4))
    span = np.zeros(P)
    loss = np.zeros(1)
    lib().orc_diff_forward(wp, n, int(grid), len(mb) + 1, mbp, float(vsini), float(u1), float(u2), float(kappa),
                           float(kappa_f), lbp_, qp, sp, fp_, P, M, mpp, fmp_, float(sigma_ms), int(trim), vp,
                           len(v), llp, lwp, len(ll), tp, rtp,
                           ftp, btp, float(w_spec), float(w_rv), float(w_fw), float(w_bis), F.ctypes.data_as(_dp),
                           cc.ctypes.data_as(_dp), th.ctypes.data_as(_dp), span.ctypes.data_as(_dp),
                           loss.ctypes.data_as(_dp), nthreads or nthreads_default())
    return dict(loss=float(loss[0]), F=F, ccf=cc, theta=th, span=span)


# ------------------------------------------------------------------- seed preprocessing
def _poly1d(c, x):
    """numpy.poly1d(c)(x): y = y * x + c_i over the coefficients, highest degree first."""
    y = np.zeros_like(x)
    for ci in np.atleast_1d(c):
        y = y * x + ci
    return y


def _delta_wavelength(v, w):
    """Relativistic Doppler offset of CB_injection_faculae.py:119-123."""
    beta = v / 299792458.
    return w * (np.sqrt((1 + beta) / (1 - beta)) - 1)


def bis_inject(wave, flux, up, scale, low=None, dep_cut=0.0, vel_cut=0.0, velo_shift=0.0):
    """Restatement of BIS_inject_spectral_quiet (low is None, CB_injection_faculae.py:227-244)
    and BIS_inject_spectral_active (:191-224), vectorised: each pixel moves by its own
    velocity, then scipy's linear interp1d on the moved samples is its documented
    algorithm for float64 data (stable argsort by x, then numpy.interp), evaluated at
    the original wavelengths in the open range of the moved ones, zero elsewhere."""
    w = np.asarray(wave, dtype=np.float64)
    f = np.asarray(flux, dtype=np.float64)
    if low is None:
        rv = _poly1d(up, f) * scale
        moved = w + _delta_wavelength(rv, w)
    else:
        a = (f < dep_cut) & (f > 0.5)  # the if / elif order of the reference
        b = (f <= 0.5) & ~a
        c = (f > dep_cut) & ~(a | b)
        rv = np.zeros_like(f)
        rv[a] = (_poly1d(up, f[a]) + velo_shift) * scale
        rv[b] = (_poly1d(low, f[b]) + velo_shift) * scale
        rv[c] = (vel_cut + velo_shift) * scale
        moved = w.copy()
        m = a | b | c
        moved[m] = w[m] + _delta_wavelength(rv[m], w[m])
    order = np.argsort(moved, kind="mergesort")
    xs, ys = moved[order], f[order]
    out = np.zeros_like(w)
    inside = (w < moved.max()) & (w > moved.min())
    out[inside] = np.interp(w[inside], xs, ys)
    return out


# ----------------------------------------------- seed preprocessing: local PHOENIX spectra
def phoenix_vac2air(wavelength):
    """vac2air of phoenix_extraction.py:13-22 (Birch and Downs 1994): lambda / n_air."""
    s = 1.e4 / wavelength
    n_air = 1 + 0.0000834254 + 0.02406147 / (130 - s**2) + 0.00015998 / (38.9 - s**2)
    return wavelength / n_air


def phoenix_air2vac(wavelength):
    """air2vac of phoenix_extraction.py:25-33 (N. Piskunov's inverse): lambda n_air."""
    s = 1.e4 / wavelength
    n_air = (1 + 0.00008336624212083 + 0.02408926869968 / (130.1065924522 - s**2)
             + 0.0001599740894897 / (38.92568793293 - s**2))
    return wavelength * n_air


def phoenix_temp_grid():
    """temp_array of phoenix_extraction.py:92-96 (entry 48 is 6800 K)."""
    return np.array([2300.0 + 100.0 * j if j < 48 else 7000.0 + 200.0 * (j - 49) for j in range(73)])


def phoenix_bracket(temp):
    """(temp_low, temp_high) of phoenix_extraction.py:136-146; None where the
    reference would index past the grid (or wrap to its last entry)."""
    g = phoenix_temp_grid()
    k = int(np.argmin(np.abs(g - temp)))
    lo, hi = (k, k + 1) if temp > g[k] else (k - 1, k)
    if lo < 0 or hi >= g.size:
        return None
    return g[lo], g[hi]


def phoenix_grid(logg, feh=0.0, alpha=0.0):
    """The nearest library grid values of phoenix_extraction.py:84-104,126-131."""
    logg_array = np.linspace(0.0, 6.0, 13)
    feh_array = np.concatenate((np.linspace(-4.0, -3.0, 2), np.linspace(-2.0, 1.0, 7))).round(decimals=2)
    alpha_array = (-0.2 + np.arange(8) * 0.2).round(decimals=2)
    lg = logg_array[np.argmin(np.abs(logg_array - logg))]
    fe = feh_array[np.argmin(np.abs(feh_array - feh))]
    al = 0.0 if fe > 0.0 else alpha_array[np.argmin(np.abs(alpha_array - alpha))]
    return float(lg), float(fe), float(al)


def phoenix_name(temp, logg, feh, alpha):
    """PHOENIX_name of phoenix_extraction.py:36-54, restated."""
    tail = '.PHOENIX-ACES-AGSS-COND-2011-HiRes.fits'
    if feh <= 0.0:
        sub = 'Z-%.1f' % abs(feh)
        if alpha == 0.0:
            name = '/lte0' + str(int(temp)) + '-%.2f-%.1f' % (logg, abs(feh)) + tail
        elif alpha > 0.0:
            name = '/lte0' + str(int(temp)) + '-%.2f-%.1f.Alpha=+%.2f' % (logg, abs(feh), alpha) + tail
        else:
            name = '/lte0' + str(int(temp)) + '-%.2f-%.1f.Alpha=%.2f' % (logg, abs(feh), alpha) + tail
    else:
        sub = 'Z+%.1f' % abs(feh)
        name = '/lte0' + str(int(temp)) + '-%.2f+%.1f' % (logg, abs(feh)) + tail
    return sub, name


def phoenix_interp(wave_vac, spec_low, spec_high, temp, t_low, t_high, wmin, wmax):
    """phoenix_extraction.py:170-185: vacuum window, vac2air, air window, and
    flux = low + (temp - t_low) / (t_high - t_low) (high - low) per pixel, with the
    spectra in their own dtype (float32 for the library: the difference is then a
    float32 difference, as in the reference's scalar loop)."""
    w = np.asarray(wave_vac, dtype=np.float64)
    keep = (w >= (wmin - 500)) & (w <= (wmax + 500))
    wv = w[keep]
    lo, hi = np.asarray(spec_low)[keep], np.asarray(spec_high)[keep]
    air = phoenix_vac2air(wv)
    sel = (air >= wmin) & (air <= wmax)
    frac = (temp - t_low) / (t_high - t_low)
    flux = lo[sel].astype(np.float64) + frac * (hi[sel] - lo[sel]).astype(np.float64)
    return air[sel], flux
