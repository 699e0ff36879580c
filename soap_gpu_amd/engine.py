"""ctypes binding of libsoapgpu (include/soap_gpu.h).

The engine is the HIP library; this module only marshals numpy arrays across
the C ABI.  Loading fails loudly when the library is missing: there is no CPU
fallback for any stage.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsoapgpu.so")

GRID_LINEAR, GRID_LOG = 0, 1
CUBE_QUIET, CUBE_QUIET_SED, CUBE_SPOT, CUBE_SPOT_SED, CUBE_FAC, CUBE_FAC_SED = range(6)
FLUX, BCONV, TOT = 0, 1, 2
CUBE_KEYS = ("q", "qsed", "sp", "spsed", "fac", "facsed")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class Star(ctypes.Structure):
    _fields_ = [("prot_days", ctypes.c_double), ("w_ratio", ctypes.c_double), ("incl_deg", ctypes.c_double),
                ("radius_km", ctypes.c_double), ("u1", ctypes.c_double), ("u2", ctypes.c_double),
                ("lb", ctypes.c_double * 4), ("grid", ctypes.c_int32)]


class Activity(ctypes.Structure):
    _fields_ = [("nrho", ctypes.c_int32), ("t_star", ctypes.c_int32), ("t_diff_spot", ctypes.c_int32),
                ("t_diff_fac", ctypes.c_int32), ("solar_mode", ctypes.c_int32)]


class Region(ctypes.Structure):
    _fields_ = [("lat_deg", ctypes.c_double), ("lon_deg", ctypes.c_double), ("size", ctypes.c_double),
                ("type", ctypes.c_double)]


# every symbol declared in include/soap_gpu.h
EXPORTS = (
    "soap_ctx_create", "soap_ctx_destroy", "soap_last_error", "soap_sync", "soap_version",
    "soap_set_wave", "soap_set_mu_bounds", "soap_set_cube", "soap_set_cube_device",
    "soap_disk_init", "soap_get_maps", "soap_set_maps",
    "soap_quiet_star", "soap_get_quiet", "soap_set_quiet",
    "soap_set_activity", "soap_integrate", "soap_get_raw", "soap_set_raw", "soap_alloc_phases", "soap_phase_maps",
    "soap_resolution", "soap_get_lower", "soap_get_norm",
    "soap_set_mask", "soap_ccf", "soap_get_ccf",
    "soap_fit", "soap_get_fit", "soap_fit_ccf_host", "soap_fit_ccf_host_err", "soap_fit_ccf_host_noalloc",
    "soap_bis_inject",
    "soap_phoenix_bracket", "soap_phoenix_grid", "soap_phoenix_name", "soap_fits_read", "soap_phoenix_interp",
    "soap_run_series",
    "soap_set_timing", "soap_get_stage_ms", "soap_stage_name", "soap_num_stages",
    "soap_get_integrate_stats",
    "soap_get_stage_work", "soap_set_fir_shape",
    "soap_diff_setup", "soap_diff_forward", "soap_diff_backward", "soap_diff_get",
    "soap_batch_create", "soap_batch_destroy", "soap_batch_last_error", "soap_batch_step", "soap_batch_sync",
    "soap_batch_stage_ms", "soap_batch_stage_name", "soap_batch_num_stages", "soap_batch_get_config",
    "soap_batch_set_config", "soap_batch_tune",
    "soap_diff_batch_create", "soap_diff_batch_destroy", "soap_diff_batch_last_error", "soap_diff_batch_bind",
    "soap_diff_batch_step", "soap_diff_batch_sync",
)

_lib = None


class SoapError(RuntimeError):
    pass


class DiffParams(ctypes.Structure):
    _fields_ = [("vsini_kms", ctypes.c_double), ("u1", ctypes.c_double), ("u2", ctypes.c_double),
                ("contrast", ctypes.c_double), ("contrast_fac", ctypes.c_double)]


class BisInjectParams(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("n_up", ctypes.c_int32), ("n_low", ctypes.c_int32),
                ("up", ctypes.c_double * 16), ("low", ctypes.c_double * 16), ("dep_cut", ctypes.c_double),
                ("vel_cut", ctypes.c_double), ("velo_shift", ctypes.c_double), ("scale", ctypes.c_double)]


class PhoenixParams(ctypes.Structure):
    _fields_ = [("temp", ctypes.c_double), ("t_low", ctypes.c_double), ("t_high", ctypes.c_double),
                ("wmin", ctypes.c_double), ("wmax", ctypes.c_double), ("f32", ctypes.c_int32)]


def load(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SoapError(f"libsoapgpu.so not built ({path}); run python -m soap_gpu_amd.build")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    sig = {
        "soap_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
        "soap_ctx_destroy": (None, [vp]),
        "soap_last_error": (ctypes.c_char_p, [vp]),
        "soap_sync": (ctypes.c_int, [vp]),
        "soap_version": (ctypes.c_char_p, []),
        "soap_set_wave": (ctypes.c_int, [vp, ctypes.c_int, _dp, ctypes.c_int64, ctypes.c_double]),
        "soap_set_mu_bounds": (ctypes.c_int, [vp, _dp, ctypes.c_int32]),
        "soap_set_cube": (ctypes.c_int, [vp, ctypes.c_int, _dp, ctypes.c_int32]),
        "soap_set_cube_device": (ctypes.c_int, [vp, ctypes.c_int, vp, ctypes.c_int32]),
        "soap_disk_init": (ctypes.c_int, [vp, ctypes.POINTER(Star)]),
        "soap_get_maps": (ctypes.c_int, [vp, _dp, _dp, _dp, _ip, _dp]),
        "soap_set_maps": (ctypes.c_int, [vp, ctypes.POINTER(Star), _dp, _dp, _ip, _dp]),
        "soap_quiet_star": (ctypes.c_int, [vp]),
        "soap_get_quiet": (ctypes.c_int, [vp, ctypes.c_int, _dp]),
        "soap_set_quiet": (ctypes.c_int, [vp, ctypes.c_int, _dp]),
        "soap_set_activity": (ctypes.c_int, [vp, ctypes.POINTER(Activity), ctypes.c_int32, _dp, _ip,
                                             ctypes.POINTER(Region)]),
        "soap_integrate": (ctypes.c_int, [vp]),
        "soap_get_raw": (ctypes.c_int, [vp, ctypes.c_int32, _dp, _dp, _dp]),
        "soap_set_raw": (ctypes.c_int, [vp, ctypes.c_int32, _dp, _dp, _dp]),
        "soap_alloc_phases": (ctypes.c_int, [vp, ctypes.c_int32]),
        "soap_phase_maps": (ctypes.c_int, [vp, ctypes.c_int32, _dp, _dp]),
        "soap_resolution": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double]),
        "soap_get_lower": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int, _dp]),
        "soap_get_norm": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int, _dp]),
        "soap_set_mask": (ctypes.c_int, [vp, _dp, _dp, ctypes.c_int32, _dp, ctypes.c_int32, ctypes.c_int32]),
        "soap_ccf": (ctypes.c_int, [vp]),
        "soap_get_ccf": (ctypes.c_int, [vp, _dp]),
        "soap_fit": (ctypes.c_int, [vp]),
        "soap_get_fit": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp]),
        "soap_fit_ccf_host": (ctypes.c_int, [vp, _dp, ctypes.c_int32, _dp, ctypes.c_int32, _dp, _dp, _dp, _dp]),
        "soap_fit_ccf_host_err": (ctypes.c_int, [vp, _dp, ctypes.c_int32, _dp, _dp, ctypes.c_int32, _dp, _dp, _dp,
                                                 _dp, _ip]),
        "soap_fit_ccf_host_noalloc": (ctypes.c_int, [vp, _dp, ctypes.c_int32, _dp, _dp, ctypes.c_int32, _dp, _dp,
                                                     _dp, _dp, _ip]),
        "soap_bis_inject": (ctypes.c_int, [vp, ctypes.POINTER(BisInjectParams), _dp, _dp, ctypes.c_int64, _dp]),
        "soap_phoenix_bracket": (ctypes.c_int, [ctypes.c_double, _dp, _dp]),
        "soap_phoenix_grid": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]),
        "soap_phoenix_name": (ctypes.c_int, [ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_char_p, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32]),
        "soap_fits_read": (ctypes.c_int, [ctypes.c_char_p, _dp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.POINTER(ctypes.c_int32)]),
        "soap_phoenix_interp": (ctypes.c_int, [vp, ctypes.POINTER(PhoenixParams), _dp, _dp, _dp, ctypes.c_int64, _dp,
                                               _dp, ctypes.POINTER(ctypes.c_int64)]),
        "soap_run_series": (ctypes.c_int, [vp, ctypes.POINTER(Star), ctypes.c_double, ctypes.c_double]),
        "soap_set_timing": (ctypes.c_int, [vp, ctypes.c_int]),
        "soap_get_stage_ms": (ctypes.c_int, [vp, _dp, ctypes.c_int32]),
        "soap_stage_name": (ctypes.c_char_p, [ctypes.c_int32]),
        "soap_num_stages": (ctypes.c_int32, []),
        "soap_get_integrate_stats": (ctypes.c_int, [vp, _dp, _dp, ctypes.POINTER(ctypes.c_int64)]),
        "soap_get_stage_work": (ctypes.c_int, [vp, ctypes.c_int32, _dp, _dp]),
        "soap_set_fir_shape": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32]),
        "soap_diff_setup": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                           _dp]),
        "soap_diff_forward": (ctypes.c_int, [vp, ctypes.POINTER(DiffParams), vp, vp, vp, vp, vp, vp, _dp, _dp,
                                             _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, _dp, _dp, _dp, _dp]),
        "soap_diff_backward": (ctypes.c_int, [vp, _dp, vp, vp, vp, vp, vp]),
        "soap_diff_get": (ctypes.c_int, [vp, ctypes.c_int32, _dp]),
        "soap_batch_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int32, _dp, _dp, ctypes.POINTER(vp)]),
        "soap_batch_destroy": (None, [vp]),
        "soap_batch_last_error": (ctypes.c_char_p, [vp]),
        "soap_batch_step": (ctypes.c_int, [vp, ctypes.c_int32]),
        "soap_batch_sync": (ctypes.c_int, [vp]),
        "soap_batch_stage_ms": (ctypes.c_int, [vp, _dp, ctypes.c_int32]),
        "soap_batch_stage_name": (ctypes.c_char_p, [ctypes.c_int32]),
        "soap_batch_num_stages": (ctypes.c_int32, []),
        "soap_batch_get_config": (ctypes.c_int, [vp, _ip, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_int64),
                                                 ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]),
        "soap_batch_set_config": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.POINTER(ctypes.c_int64), ctypes.c_int32]),
        "soap_batch_tune": (ctypes.c_int, [vp, ctypes.c_int32]),
        "soap_diff_batch_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int32, ctypes.POINTER(vp)]),
        "soap_diff_batch_destroy": (None, [vp]),
        "soap_diff_batch_last_error": (ctypes.c_char_p, [vp]),
        "soap_diff_batch_bind": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_double, vp, vp, vp, vp, vp, vp, vp,
                                                vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, vp, vp, vp, vp, vp, vp, vp]),
        "soap_diff_batch_step": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_int32]),
        "soap_diff_batch_sync": (ctypes.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _out(n, dtype=np.float64):
    a = np.empty(n, dtype=dtype)
    return a, a.ctypes.data_as(_dp if dtype == np.float64 else _ip)


def make_star(star: dict) -> Star:
    s = Star()
    s.prot_days = float(star["prot"])
    s.w_ratio = float(star.get("w_ratio", 0.0))
    s.incl_deg = float(star["incl"])
    s.radius_km = float(star["radius_km"])
    s.u1 = float(star["u1"])
    s.u2 = float(star["u2"])
    for i in range(4):
        s.lb[i] = float(star["lb"][i])
    s.grid = int(star["grid"])
    return s


class Engine:
    """One soap_ctx (one HIP stream on one device)."""

    def __init__(self, device: int = 0):
        self.L = load()
        self.ctx = ctypes.c_void_p()
        rc = self.L.soap_ctx_create(int(device), ctypes.byref(self.ctx))
        if rc != 0:
            raise SoapError(f"soap_ctx_create failed ({rc}): no usable HIP device")
        self.n = 0
        self.grid = 0
        self.n_phase = 0
        self.n_v = 0

    def close(self):
        if self.ctx:
            self.L.soap_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.soap_last_error(self.ctx)
            raise SoapError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # ---------------------------------------------------------------- inputs
    def set_wave(self, wave, kind=GRID_LOG, step=0.0):
        w, wp = _d(wave)
        self._chk(self.L.soap_set_wave(self.ctx, kind, wp, len(w), float(step)), "set_wave")
        self.n = len(w)

    def set_mu_bounds(self, bounds):
        b, bp = _d(bounds)
        self._chk(self.L.soap_set_mu_bounds(self.ctx, bp, len(b)), "set_mu_bounds")

    def set_cube(self, which, cube):
        c, cp = _d(cube)
        self._chk(self.L.soap_set_cube(self.ctx, which, cp, c.shape[0]), "set_cube")

    def set_cubes(self, cubes: dict):
        for i, k in enumerate(CUBE_KEYS):
            if k in cubes and cubes[k] is not None:
                self.set_cube(i, cubes[k])

    # ---------------------------------------------------------------- stages
    def disk_init(self, star: dict):
        self._star = make_star(star)
        self._chk(self.L.soap_disk_init(self.ctx, ctypes.byref(self._star)), "disk_init")
        self.grid = int(star["grid"])

    def get_maps(self):
        g2 = self.grid * self.grid
        mu, mup = _out(g2)
        de, dep = _out(g2)
        li, lip = _out(g2)
        rg = np.empty(g2, dtype=np.int32)
        lb, lbp = _out(g2)
        self._chk(self.L.soap_get_maps(self.ctx, mup, dep, lip, rg.ctypes.data_as(_ip), lbp), "get_maps")
        return dict(mu=mu, delta=de, limb=li, region=rg, lb=lb)

    def set_maps(self, star: dict, maps: dict):
        self._star = make_star(star)
        de, dep = _d(maps["delta"])
        li, lip = _d(maps["limb"])
        rg = np.ascontiguousarray(maps["region"], dtype=np.int32)
        lb, lbp = _d(maps["lb"])
        self._chk(self.L.soap_set_maps(self.ctx, ctypes.byref(self._star), dep, lip, rg.ctypes.data_as(_ip), lbp),
                  "set_maps")
        self.grid = int(star["grid"])

    def quiet_star(self):
        self._chk(self.L.soap_quiet_star(self.ctx), "quiet_star")

    def get_quiet(self, which):
        o, op = _out(self.n)
        self._chk(self.L.soap_get_quiet(self.ctx, which, op), "get_quiet")
        return o

    def set_quiet(self, which, q):
        a, ap = _d(q)
        self._chk(self.L.soap_set_quiet(self.ctx, which, ap), "set_quiet")

    def set_activity(self, act: dict, phases, regions_per_phase, solar_mode=0):
        A = Activity(int(act["nrho"]), int(act["t_star"]), int(act["t_diff_spot"]), int(act["t_diff_fac"]),
                     int(solar_mode))
        ph, php = _d(phases)
        counts = np.array([len(r) for r in regions_per_phase], dtype=np.int32)
        flat = [r for rs in regions_per_phase for r in rs]
        R = (Region * max(1, len(flat)))()
        for i, r in enumerate(flat):
            R[i].lat_deg, R[i].lon_deg, R[i].size, R[i].type = [float(x) for x in r]
        self._chk(self.L.soap_set_activity(self.ctx, ctypes.byref(A), len(ph), php, counts.ctypes.data_as(_ip), R),
                  "set_activity")
        self.n_phase = len(ph)

    def integrate(self):
        self._chk(self.L.soap_integrate(self.ctx), "integrate")

    def get_raw(self, phase):
        f, fp = _out(self.n)
        b, bp = _out(self.n)
        t, tp = _out(self.n)
        self._chk(self.L.soap_get_raw(self.ctx, int(phase), fp, bp, tp), "get_raw")
        return f, b, t

    def set_raw(self, phase, flux, bconv, tot):
        f, fp = _d(flux)
        b, bp = _d(bconv)
        t, tp = _d(tot)
        self._chk(self.L.soap_set_raw(self.ctx, int(phase), fp, bp, tp), "set_raw")

    def alloc_phases(self, n_phase):
        self._chk(self.L.soap_alloc_phases(self.ctx, int(n_phase)), "alloc_phases")
        self.n_phase = int(n_phase)

    def phase_maps(self, phase):
        g2 = self.grid * self.grid
        t, tp = _out(g2)
        b, bp = _out(g2)
        self._chk(self.L.soap_phase_maps(self.ctx, int(phase), tp, bp), "phase_maps")
        return t, b

    def resolution(self, sigma_ms, quiet_weight=0.005):
        self._chk(self.L.soap_resolution(self.ctx, float(sigma_ms), float(quiet_weight)), "resolution")

    def get_lower(self, phase, typ):
        o, op = _out(self.n)
        self._chk(self.L.soap_get_lower(self.ctx, int(phase), int(typ), op), "get_lower")
        return o

    def get_norm(self, phase, typ):
        o, op = _out(2)
        self._chk(self.L.soap_get_norm(self.ctx, int(phase), int(typ), op), "get_norm")
        return o

    def set_mask(self, line_l, line_w, vgrid, trim=210):
        l, lp = _d(line_l)
        w, wp = _d(line_w)
        v, vp = _d(vgrid)
        self._chk(self.L.soap_set_mask(self.ctx, lp, wp, len(l), vp, len(v), int(trim)), "set_mask")
        self.n_v = len(v)

    def ccf(self):
        self._chk(self.L.soap_ccf(self.ctx), "ccf")

    def get_ccf(self):
        o, op = _out((self.n_phase + 1) * 3 * self.n_v)
        self._chk(self.L.soap_get_ccf(self.ctx, op), "get_ccf")
        return o.reshape(self.n_phase + 1, 3, self.n_v)

    def fit(self):
        self._chk(self.L.soap_fit(self.ctx), "fit")

    def get_fit(self):
        n_ccf = (self.n_phase + 1) * 3
        par, pp = _out(n_ccf * 5)
        mod, mp = _out(n_ccf * self.n_v)
        dep, dp = _out(n_ccf * 91)
        bis, bp = _out(n_ccf * 91)
        self._chk(self.L.soap_get_fit(self.ctx, pp, mp, dp, bp), "get_fit")
        sh = (self.n_phase + 1, 3)
        return dict(par=par.reshape(*sh, 5), model=mod.reshape(*sh, self.n_v), depth=dep.reshape(*sh, 91),
                    bis=bis.reshape(*sh, 91))

    def bis_inject(self, wave, flux, up, scale, low=None, dep_cut=0.0, vel_cut=0.0, velo_shift=0.0):
        """Bisector injection of CB_injection_faculae.py:191-244 on the GPU: low=None is
        BIS_inject_spectral_quiet(wave, flux, up, scale), otherwise
        BIS_inject_spectral_active(wave, flux, up, low, dep_cut, vel_cut, velo_shift, scale).
        Coefficients in numpy.poly1d order (highest degree first), at most 16."""
        ww, wp = _d(wave)
        ff, fp = _d(flux)
        if ww.shape != ff.shape:
            raise SoapError("bis_inject: wave and flux differ in length")
        up = np.atleast_1d(np.asarray(up, dtype=np.float64))
        lo = np.atleast_1d(np.asarray([] if low is None else low, dtype=np.float64))
        if not 1 <= up.size <= 16 or lo.size > 16 or (low is not None and lo.size < 1):
            raise SoapError("bis_inject: 1..16 coefficients per polynomial")
        prm = BisInjectParams()
        prm.mode = 0 if low is None else 1
        prm.n_up, prm.n_low = up.size, lo.size
        for i, x in enumerate(up):
            prm.up[i] = float(x)
        for i, x in enumerate(lo):
            prm.low[i] = float(x)
        prm.dep_cut, prm.vel_cut, prm.velo_shift, prm.scale = (float(dep_cut), float(vel_cut), float(velo_shift),
                                                             float(scale))
        out, op = _out(ww.size)
        self._chk(self.L.soap_bis_inject(self.ctx, ctypes.byref(prm), wp, fp, ww.size, op), "bis_inject")
        return out

    def phoenix_interp(self, wave_vac, spec_low, spec_high, temp, t_low, t_high, wmin, wmax, f32=True):
        """The spectrum at temp from the two bracketing PHOENIX spectra on the GPU
        (phoenix_extraction.py:170-185): returns (wave_air, flux) of the pixels whose
        vacuum wavelength is in [wmin - 500, wmax + 500] and air wavelength in [wmin, wmax];
        f32: the spectra are float32 (the library's), their difference taken in float32."""
        ww, wp = _d(wave_vac)
        lo, lp = _d(spec_low)
        hi, hp = _d(spec_high)
        if not ww.shape == lo.shape == hi.shape:
            raise SoapError("phoenix_interp: arrays differ in length")
        prm = PhoenixParams(float(temp), float(t_low), float(t_high), float(wmin), float(wmax), int(bool(f32)))
        air, ap = _out(ww.size)
        flux, fp = _out(ww.size)
        cnt = ctypes.c_int64()
        self._chk(self.L.soap_phoenix_interp(self.ctx, ctypes.byref(prm), wp, lp, hp, ww.size, ap, fp,
                                             ctypes.byref(cnt)), "phoenix_interp")
        return air[:cnt.value].copy(), flux[:cnt.value].copy()

    def fit_ccf_host(self, v, y, err=None):
        """BIS_FIT2 on host CCFs y[n_ccf][n_v]; err: per-point errors of the same shape (None = ones).
        With errors the result also holds the solver outcome per CCF ('status', GSL codes)."""
        vv, vp = _d(v)
        yy = np.ascontiguousarray(np.atleast_2d(y), dtype=np.float64)
        n_ccf = yy.shape[0]
        par, pp = _out(n_ccf * 5)
        mod, mp = _out(n_ccf * len(vv))
        dep, dp = _out(n_ccf * 91)
        bis, bp = _out(n_ccf * 91)
        out = {}
        if err is None:
            self._chk(self.L.soap_fit_ccf_host(self.ctx, vp, len(vv), yy.ctypes.data_as(_dp), n_ccf, pp, mp, dp,
                                               bp), "fit_ccf_host")
        else:
            ee = np.ascontiguousarray(np.broadcast_to(np.atleast_2d(err), yy.shape), dtype=np.float64)
            st, sp = _out(n_ccf, np.int32)
            self._chk(self.L.soap_fit_ccf_host_err(self.ctx, vp, len(vv), yy.ctypes.data_as(_dp),
                                                   ee.ctypes.data_as(_dp), n_ccf, pp, mp, dp, bp, sp),
                      "fit_ccf_host_err")
            out["status"] = st
        out.update(par=par.reshape(n_ccf, 5), model=mod.reshape(n_ccf, -1), depth=dep.reshape(n_ccf, 91),
                   bis=bis.reshape(n_ccf, 91))
        return out

    def fit_ccf_host_noalloc(self, v, y, err=None):
        """fit_ccf_host through the allocation-free entry point (fixed per-GPU staging)."""
        vv, vp = _d(v)
        yy = np.ascontiguousarray(np.atleast_2d(y), dtype=np.float64)
        n_ccf = yy.shape[0]
        par, pp = _out(n_ccf * 5)
        mod, mp = _out(n_ccf * len(vv))
        dep, dp = _out(n_ccf * 91)
        bis, bp = _out(n_ccf * 91)
        ee = None if err is None else np.ascontiguousarray(np.broadcast_to(np.atleast_2d(err), yy.shape),
                                                           dtype=np.float64)
        st, sp = _out(n_ccf, np.int32)
        self._chk(self.L.soap_fit_ccf_host_noalloc(self.ctx, vp, len(vv), yy.ctypes.data_as(_dp),
                                                   None if ee is None else ee.ctypes.data_as(_dp), n_ccf, pp, mp,
                                                   dp, bp, sp), "fit_ccf_host_noalloc")
        return dict(par=par.reshape(n_ccf, 5), model=mod.reshape(n_ccf, -1), depth=dep.reshape(n_ccf, 91),
                    bis=bis.reshape(n_ccf, 91), status=st)

    def run_series(self, star: dict, sigma_ms, quiet_weight=0.005):
        self._star = make_star(star)
        self._chk(self.L.soap_run_series(self.ctx, ctypes.byref(self._star), float(sigma_ms), float(quiet_weight)),
                  "run_series")

    def sync(self):
        self._chk(self.L.soap_sync(self.ctx), "sync")

    # ---------------------------------------------------------------- timing
    def set_timing(self, on: bool):
        self._chk(self.L.soap_set_timing(self.ctx, 1 if on else 0), "set_timing")

    def stage_ms(self) -> dict:
        ns = self.L.soap_num_stages()
        o, op = _out(ns)
        self._chk(self.L.soap_get_stage_ms(self.ctx, op, ns), "get_stage_ms")
        return {self.L.soap_stage_name(i).decode(): float(o[i]) for i in range(ns)}

    def stage_work(self, name):
        """(bytes, flops) of one run of the named stage on the current inputs."""
        names = [self.L.soap_stage_name(i).decode() for i in range(self.L.soap_num_stages())]
        b, f = ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.soap_get_stage_work(self.ctx, names.index(name), ctypes.byref(b), ctypes.byref(f)),
                  "stage_work")
        return b.value, f.value

    def set_fir_shape(self, phases_per_block: int, pixels_per_thread: int):
        """Launch shape of the active FIR: (4, 0), (1, 3), (1, 5), (2, 3), (2, 1), (4, 1); (-1, -1) automatic."""
        self._chk(self.L.soap_set_fir_shape(self.ctx, int(phases_per_block), int(pixels_per_thread)),
                  "set_fir_shape")

    def integrate_stats(self):
        """(bytes_min, bytes_phase_sum, rows_sum) of the last integrate."""
        b = ctypes.c_double()
        bp = ctypes.c_double()
        r = ctypes.c_int64()
        self._chk(self.L.soap_get_integrate_stats(self.ctx, ctypes.byref(b), ctypes.byref(bp), ctypes.byref(r)),
                  "integrate_stats")
        return float(b.value), float(bp.value), int(r.value)


def select_mask_lines(wave, line_l, line_w, trim: int = 210):
    """Mask lines inside the trimmed wavelength range with weight > 0.1, in their
    given order (run_all_SOAP.py:285-290)."""
    wave_t = np.asarray(wave)[trim:len(wave) - trim]
    line_l = np.asarray(line_l, dtype=np.float64)
    line_w = np.asarray(line_w, dtype=np.float64)
    sel = (line_l > wave_t.min()) & (line_l < wave_t.max()) & (line_w > 0.1)
    return line_l[sel], line_w[sel]


def setup_workload(eng: Engine, w, with_mask: bool = True, trim: int = 210):
    """Upload a synth.Workload: grid, mu bounds, cubes, activity, mask."""
    if w.grid_kind == GRID_LINEAR:
        eng.set_wave(w.wave, GRID_LINEAR, step=w.meta["step"])
    else:
        eng.set_wave(w.wave, GRID_LOG)
    eng.set_mu_bounds(w.mu_bounds)
    eng.set_cubes(w.cubes)
    eng.disk_init(w.star)
    eng.set_activity(w.act, w.phases, w.regions)
    if with_mask:
        eng.set_mask(*select_mask_lines(w.wave, w.mask_l, w.mask_w, trim), w.vgrid, trim)


class Batch:
    """Many prepared Engines (log grids, activity and mask set, disk_init and
    quiet_star done) stepped together: one persistent launch per stage over all
    instances (soap_batch_* of the C ABI), optionally as a HIP graph."""

    def __init__(self, engines, sigma_ms, quiet_weight=0.005):
        self.L = load()
        self.engines = list(engines)
        n = len(self.engines)
        ctxs = (ctypes.c_void_p * n)(*[e.ctx for e in self.engines])
        sig, sp = _d(np.broadcast_to(np.asarray(sigma_ms, dtype=np.float64), (n,)))
        qw, qp = _d(np.broadcast_to(np.asarray(quiet_weight, dtype=np.float64), (n,)))
        self.b = ctypes.c_void_p()
        rc = self.L.soap_batch_create(ctxs, n, sp, qp, ctypes.byref(self.b))
        if rc != 0:
            msg = self.L.soap_batch_last_error(self.b) if self.b else b"allocation failed"
            self.close()
            raise SoapError(f"soap_batch_create failed ({rc}): {msg.decode() if msg else ''}")
        self.n_stage = self.L.soap_batch_num_stages()
        self.stage_names = [self.L.soap_batch_stage_name(i).decode() for i in range(self.n_stage)]

    def close(self):
        if self.b:
            self.L.soap_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.soap_batch_last_error(self.b)
            raise SoapError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def step(self, graph: bool = False):
        self._chk(self.L.soap_batch_step(self.b, int(bool(graph))), "batch_step")

    def sync(self):
        self._chk(self.L.soap_batch_sync(self.b), "batch_sync")

    def stage_ms(self) -> dict:
        ms = np.zeros(self.n_stage)
        self._chk(self.L.soap_batch_stage_ms(self.b, ms.ctypes.data_as(_dp), self.n_stage), "batch_stage_ms")
        return dict(zip(self.stage_names, ms.tolist()))

    def tune(self, reps: int = 2):
        self._chk(self.L.soap_batch_tune(self.b, int(reps)), "batch_tune")

    def config(self) -> dict:
        cells = ctypes.c_int32()
        g = ctypes.c_int32()
        bb = ctypes.c_int32()
        pre = ctypes.c_int32()
        grid = (ctypes.c_int64 * self.n_stage)()
        work = (ctypes.c_int64 * self.n_stage)()
        self._chk(self.L.soap_batch_get_config(self.b, ctypes.byref(cells), ctypes.byref(g), ctypes.byref(bb),
                                               ctypes.byref(pre), grid, work, self.n_stage), "batch_get_config")
        return {"taps_cells": cells.value, "fir_g": g.value, "fir_b": bb.value, "fir_pre": pre.value,
                "grid": dict(zip(self.stage_names, list(grid))), "work": dict(zip(self.stage_names, list(work)))}

    def set_config(self, taps_cells: int, fir_g: int, fir_b: int, grid: dict, fir_pre: int = 0):
        arr = (ctypes.c_int64 * self.n_stage)(*[int(grid.get(n, 0)) for n in self.stage_names])
        self._chk(self.L.soap_batch_set_config(self.b, int(taps_cells), int(fir_g), int(fir_b), int(fir_pre), arr,
                                               self.n_stage), "batch_set_config")


# ------------------------------------------------------------ local PHOENIX files (host-only, no GPU)
def phoenix_bracket(temp):
    """(t_low, t_high): the library temperatures around temp (phoenix_extraction.py:136-146)."""
    L = load()
    lo, hi = ctypes.c_double(), ctypes.c_double()
    rc = L.soap_phoenix_bracket(float(temp), ctypes.byref(lo), ctypes.byref(hi))
    if rc != 0:
        raise SoapError(f"phoenix_bracket({temp}) failed ({rc}): outside the library's temperature grid")
    return lo.value, hi.value


def phoenix_grid(logg, feh=0.0, alpha=0.0):
    """The library grid values nearest to (logg, [Fe/H], [alpha/Fe])."""
    L = load()
    a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = L.soap_phoenix_grid(float(logg), float(feh), float(alpha), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    if rc != 0:
        raise SoapError(f"phoenix_grid failed ({rc})")
    return a.value, b.value, c.value


def phoenix_name(temp, logg, feh, alpha):
    """PHOENIX_name of phoenix_extraction.py:36-54: (subdir, '/lte0...fits')."""
    L = load()
    sd, nm = ctypes.create_string_buffer(64), ctypes.create_string_buffer(256)
    rc = L.soap_phoenix_name(int(temp), float(logg), float(feh), float(alpha), sd, 64, nm, 256)
    if rc != 0:
        raise SoapError(f"phoenix_name failed ({rc})")
    return sd.value.decode(), nm.value.decode()


def fits_read(path):
    """(data as float64, BITPIX) of the primary image of a FITS file."""
    L = load()
    n, bp = ctypes.c_int64(), ctypes.c_int32()
    p = os.fsencode(path)
    rc = L.soap_fits_read(p, None, 0, ctypes.byref(n), ctypes.byref(bp))
    if rc != 0:
        raise SoapError(f"fits_read({path}) failed ({rc})")
    out = np.empty(n.value)
    rc = L.soap_fits_read(p, out.ctypes.data_as(_dp), n.value, ctypes.byref(n), ctypes.byref(bp))
    if rc != 0:
        raise SoapError(f"fits_read({path}) failed ({rc})")
    return out, bp.value
