"""Seeded synthetic inputs for the SOAP-GPU workloads.

The reference's own seed spectra (PHOENIX/RASSINE products and the solar
wavelength grid) are not part of its repository, so every workload here is
generated from a seed:

* wavelength grid: log-uniform, lambda_j = lambda_min * exp(j h);
* absorption lines: centres U(lambda_min, lambda_max), depth U(0.05, 0.9),
  Gaussian sigma U(3, 8) km/s; a seed row is prod_l (1 - d_l G_l(lambda));
* mu bins equal in mu; bin k (centre mu_k) scales depths by 1 + 0.05 (2 mu_k - 1)
  and shifts quiet lines by a convective blueshift of -300 m/s * mu_k;
* spot rows: depths x1.1, no blueshift; SED continuum x0.3;
* facula rows: quiet depths, no blueshift (suppressed); SED continuum x1.1;
* SED cubes are the normalised rows times a Planck continuum of T_star
  normalised at 5500 A.

Cube layout is [n_mu][n] float64 row-major, the layout of the reference's
seed-cube .bin files (SOAP_integration.cu:504-518).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

C_MS = 299792458.0
R_SUN_KM = 696000.0
# LB_params(5778) of run_all_SOAP.py:21-34,140-145 evaluated on the reference's
# Limb_brightening_info.npz (constant term first).
LB_5778 = (0.23161761, -0.11674743, -0.36866954, 0.25537337)


@dataclass
class Workload:
    name: str
    wave: np.ndarray
    grid_kind: int  # 1 = log-uniform
    mu_bounds: np.ndarray
    cubes: dict
    lines: dict
    mask_l: np.ndarray
    mask_w: np.ndarray
    star: dict
    act: dict
    phases: np.ndarray
    regions: list  # per phase: list of (lat, lon, size, type)
    sigma_inst: float
    vgrid: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.wave.shape[0])

    @property
    def n_mu(self) -> int:
        return int(self.mu_bounds.shape[0] + 1)


def log_grid(lmin: float, lmax: float, n: int | None = None, dv_kms: float | None = None) -> np.ndarray:
    if n is None:
        h = dv_kms * 1000.0 / C_MS
        n = int(np.ceil(np.log(lmax / lmin) / h)) + 1
    else:
        h = np.log(lmax / lmin) / (n - 1)
    return lmin * np.exp(np.arange(n, dtype=np.float64) * h)


def linear_grid(lmin: float, n: int, step: float = 0.005) -> np.ndarray:
    """lambda_j = lambda_min + j step (the reference's native seed format, 0.005 A)."""
    return lmin + step * np.arange(n, dtype=np.float64)


def is_linear(wave: np.ndarray) -> bool:
    d = np.diff(wave)
    return bool(np.all(np.abs(d - d[0]) <= 1e-9 * np.abs(wave[1:])))


def mu_bins(n_mu: int):
    bounds = 1.0 - np.arange(1, n_mu, dtype=np.float64) / n_mu  # descending, n_mu-1 values
    centres = 1.0 - (np.arange(n_mu, dtype=np.float64) + 0.5) / n_mu
    return bounds, centres


def draw_lines(rng: np.random.Generator, n_lines: int, lmin: float, lmax: float) -> dict:
    return dict(
        center=rng.uniform(lmin, lmax, n_lines),
        depth=rng.uniform(0.05, 0.9, n_lines),
        sigma_kms=rng.uniform(3.0, 8.0, n_lines),
    )


def _planck(lam_a: np.ndarray, t: float) -> np.ndarray:
    h, c, kb = 6.62607015e-34, C_MS, 1.380649e-23
    lam = lam_a * 1e-10
    return 2 * h * c * c / lam ** 5 / np.expm1(h * c / (lam * kb * t))


def sed_continuum(wave: np.ndarray, t_star: float = 5778.0) -> np.ndarray:
    return _planck(wave, t_star) / _planck(np.array([5500.0]), t_star)[0]


def line_row(wave: np.ndarray, lines: dict, depth_scale: float, shift_ms: float,
             nsig: float = 6.0) -> np.ndarray:
    """prod_l (1 - min(d_l*scale, 0.98) exp(-(lambda - lambda_l')^2 / 2 sigma_l^2)) on a
    log-uniform or a linear grid."""
    n = wave.shape[0]
    lc = lines["center"] * (1.0 + shift_ms / C_MS)
    if is_linear(wave):
        step = (wave[-1] - wave[0]) / (n - 1)
        sig_px = lc * (lines["sigma_kms"] * 1000.0 / C_MS) / step
        xc = (lc - wave[0]) / step
    else:
        h = np.log(wave[-1] / wave[0]) / (n - 1)
        sig_px = (lines["sigma_kms"] * 1000.0 / C_MS) / h
        xc = np.log(lc / wave[0]) / h
    half = int(np.ceil(nsig * sig_px.max())) + 1
    offs = np.arange(-half, half + 1)
    idx = np.floor(xc).astype(np.int64)[:, None] + offs[None, :]
    ok = (idx >= 0) & (idx < n)
    idxc = np.clip(idx, 0, n - 1)
    lam = wave[idxc]
    sig_l = lc * lines["sigma_kms"] * 1000.0 / C_MS
    g = np.exp(-0.5 * ((lam - lc[:, None]) / sig_l[:, None]) ** 2)
    d = np.minimum(lines["depth"] * depth_scale, 0.98)[:, None]
    tau = np.where(ok, np.log1p(-d * g), 0.0)
    acc = np.bincount(idxc.ravel(), weights=tau.ravel(), minlength=n)
    return np.exp(acc)


def make_cubes(wave: np.ndarray, lines: dict, n_mu: int, t_star: float = 5778.0,
               spot_contrast: float = 0.3, fac_contrast: float = 1.1, cb_ms: float = -300.0) -> dict:
    _, mu_c = mu_bins(n_mu)
    cont = sed_continuum(wave, t_star)
    q = np.empty((n_mu, wave.shape[0]))
    sp = np.empty_like(q)
    fac = np.empty_like(q)
    for k, mu in enumerate(mu_c):
        a = 1.0 + 0.05 * (2.0 * mu - 1.0)
        q[k] = line_row(wave, lines, a, cb_ms * mu)
        sp[k] = line_row(wave, lines, 1.1 * a, 0.0)
        fac[k] = line_row(wave, lines, a, 0.0)
    return dict(q=q, qsed=q * cont, sp=sp, spsed=sp * (spot_contrast * cont),
                fac=fac, facsed=fac * (fac_contrast * cont))


def prot_for_vsini(vsini_kms: float, radius_km: float = R_SUN_KM, incl_deg: float = 90.0) -> float:
    """Rotation period [days] giving vsini with the reference's velocity formula
    (SOAP_initialize.cu:348-351: v = (360/P) pi/(180*86400) R sin i)."""
    return 2.0 * np.pi * radius_km * np.sin(np.deg2rad(incl_deg)) / (vsini_kms * 86400.0)


def sigma_inst_ms(resolution: float) -> float:
    """run_all_SOAP.py:63-65"""
    fwhm = C_MS / float(resolution)
    return fwhm / (2 * np.sqrt(2 * np.log(2)))


def vel_grid(window_ms: float, step_ms: float) -> np.ndarray:
    """run_all_SOAP.py:262"""
    return np.arange(-1 * window_ms, window_ms + 1.0, step_ms)


def _star(grid, vsini, incl=90.0, u1=0.29, u2=0.34, w_ratio=0.0, radius_km=R_SUN_KM):
    return dict(prot=prot_for_vsini(vsini, radius_km, incl), w_ratio=w_ratio, incl=incl,
                radius_km=radius_km, u1=u1, u2=u2, grid=int(grid), lb=np.array(LB_5778))


def _act():
    return dict(t_star=5778, t_diff_spot=663, t_diff_fac=250, nrho=20)


def config0(seed: int = 1234, grid: int = 300) -> Workload:
    """Plain FP64, small: grid 300, 5000-5100 A log-uniform 20k points, 200 lines,
    10 mu bins, one spot (lat 30, lon 0, r 0.1), 25 phases, R=115000, 81-point CCF."""
    rng = np.random.default_rng(seed)
    wave = log_grid(5000.0, 5100.0, n=20000)
    lines = draw_lines(rng, 200, 5000.0, 5100.0)
    bounds, _ = mu_bins(10)
    cubes = make_cubes(wave, lines, 10)
    phases = np.arange(25, dtype=np.float64) / 25.0
    regs = [(30.0, 0.0, 0.1, 0.0)]
    return Workload("config0", wave, 1, bounds, cubes, lines, lines["center"].copy(), lines["depth"].copy(),
                    _star(grid, 2.0), _act(), phases, [list(regs) for _ in phases],
                    sigma_inst_ms(115000), vel_grid(20000.0, 500.0))


def config1(seed: int = 1235, n_phase: int = 100, grid: int = 1000) -> Workload:
    """Plain FP64 production shape: grid 1000, 3800-7000 A log-uniform at 0.25 km/s
    (~733k points), 8000 lines, 20 mu bins, 2 spots + 2 faculae overlapping,
    100 phases, R=115000, 4000-line mask CCF on 161 velocities."""
    rng = np.random.default_rng(seed)
    wave = log_grid(3800.0, 7000.0, dv_kms=0.25)
    lines = draw_lines(rng, 8000, 3800.0, 7000.0)
    bounds, _ = mu_bins(20)
    cubes = make_cubes(wave, lines, 20)
    regs = overlapping_regions(rng)
    phases = np.arange(n_phase, dtype=np.float64) / n_phase
    return Workload("config1", wave, 1, bounds, cubes, lines, lines["center"][:4000].copy(),
                    lines["depth"][:4000].copy(), _star(grid, 2.0), _act(), phases,
                    [list(regs) for _ in phases], sigma_inst_ms(115000), vel_grid(20000.0, 250.0))


def overlapping_regions(rng: np.random.Generator):
    """2 faculae (r 0.1-0.2) each overlapped by a spot (r 0.05-0.12); faculae first so
    the later spot wins where they overlap (SOAP_integration.cu:364-381)."""
    regs = []
    lon0 = rng.uniform(0.0, 60.0)
    for j in range(2):
        lat = rng.uniform(-35.0, 35.0)
        lon = lon0 + 160.0 * j
        rf = rng.uniform(0.1, 0.2)
        regs.append((lat, lon, rf, 1.0))
    for j in range(2):
        lat_f, lon_f, rf, _ = regs[j]
        rs = rng.uniform(0.05, 0.12)
        off = np.rad2deg(0.5 * rf)
        regs.append((lat_f + rng.uniform(-off, off), lon_f + rng.uniform(-off, off), rs, 0.0))
    return regs


def small(seed: int = 7, grid: int = 48, n: int = 4096, n_mu: int = 6, n_phase: int = 4,
          n_lines: int = 60, vsini: float = 3.0, incl: float = 70.0, w_ratio: float = 0.0,
          regions=None, lmin: float = 5000.0, lmax: float = 5060.0) -> Workload:
    """Tiny workload for brute-force parity (oracle in well under a second)."""
    rng = np.random.default_rng(seed)
    wave = log_grid(lmin, lmax, n=n)
    lines = draw_lines(rng, n_lines, lmin, lmax)
    bounds, _ = mu_bins(n_mu)
    cubes = make_cubes(wave, lines, n_mu)
    if regions is None:
        regions = [(20.0, 10.0, 0.25, 1.0), (25.0, 15.0, 0.15, 0.0), (-10.0, 200.0, 0.2, 0.0)]
    phases = np.arange(n_phase, dtype=np.float64) / n_phase
    return Workload("small", wave, 1, bounds, cubes, lines, lines["center"].copy(), lines["depth"].copy(),
                    _star(grid, vsini, incl=incl, w_ratio=w_ratio), _act(), phases,
                    [list(regions) for _ in phases], sigma_inst_ms(115000), vel_grid(20000.0, 500.0))


def linear_small(seed: int = 11, grid: int = 48, n: int = 4000, n_mu: int = 6, n_phase: int = 4,
                 n_lines: int = 50, vsini: float = 3.0, incl: float = 70.0, regions=None,
                 lmin: float = 5000.0) -> Workload:
    """Tiny workload on the reference's native linear 0.005 A grid."""
    rng = np.random.default_rng(seed)
    wave = linear_grid(lmin, n)
    lmax = wave[-1]
    # lines away from the edges: a mask hole shifted across the first trimmed pixel
    # meets the duplicated join point of the extended grid (run_all_SOAP.py:292-295)
    # and gets an infinite edge weight, as in the reference
    lines = draw_lines(rng, n_lines, lmin + 5.0, lmax - 5.0)
    bounds, _ = mu_bins(n_mu)
    cubes = make_cubes(wave, lines, n_mu)
    if regions is None:
        regions = [(20.0, 10.0, 0.25, 1.0), (25.0, 15.0, 0.15, 0.0), (-10.0, 200.0, 0.2, 0.0)]
    phases = np.arange(n_phase, dtype=np.float64) / n_phase
    return Workload("linear_small", wave, 0, bounds, cubes, lines, lines["center"].copy(), lines["depth"].copy(),
                    _star(grid, vsini, incl=incl), _act(), phases, [list(regions) for _ in phases],
                    sigma_inst_ms(115000), vel_grid(20000.0, 500.0), meta={"step": 0.005})


def linear_series(seed: int = 1239, n_phase: int = 25, grid: int = 1000, lmin: float = 5000.0,
                  lmax: float = 5500.0) -> Workload:
    """The production disk and activity (grid 1000, 20 mu bins, 2 spots + 2 faculae
    overlapping, vsini 2 km/s, R=115000, 161 velocities) on the reference's native
    linear 0.005 A seed format over lmin..lmax (100,000 pixels by default), with
    lines at the production density (8000 per 3800-7000 A) and every other line in
    the mask."""
    rng = np.random.default_rng(seed)
    n = int(round((lmax - lmin) / 0.005))
    wave = linear_grid(lmin, n)
    n_lines = int(round(8000 * (wave[-1] - lmin) / 3200.0))
    # lines away from the edges, as in linear_small (the duplicated join point)
    lines = draw_lines(rng, n_lines, lmin + 5.0, wave[-1] - 5.0)
    bounds, _ = mu_bins(20)
    cubes = make_cubes(wave, lines, 20)
    regs = overlapping_regions(rng)
    phases = np.arange(n_phase, dtype=np.float64) / n_phase
    half = n_lines // 2
    return Workload("linear_series", wave, 0, bounds, cubes, lines, lines["center"][:half].copy(),
                    lines["depth"][:half].copy(), _star(grid, 2.0), _act(), phases,
                    [list(regs) for _ in phases], sigma_inst_ms(115000), vel_grid(20000.0, 250.0),
                    meta={"step": 0.005})


def random_regions(rng: np.random.Generator, n_reg: int):
    """n_reg regions, faculae (r 0.1-0.2) first, then spots (r 0.05-0.12) each placed
    on a facula so that the two overlap (the later region wins a shared cell,
    SOAP_integration.cu:364-381)."""
    n_fac = (n_reg + 1) // 2
    regs = []
    for _ in range(n_fac):
        regs.append((rng.uniform(-60.0, 60.0), rng.uniform(0.0, 360.0), rng.uniform(0.1, 0.2), 1.0))
    for _ in range(n_reg - n_fac):
        lat_f, lon_f, rf, _t = regs[int(rng.integers(0, n_fac))]
        off = np.rad2deg(0.5 * rf)
        regs.append((lat_f + rng.uniform(-off, off), lon_f + rng.uniform(-off, off), rng.uniform(0.05, 0.12), 0.0))
    return regs


def batch_instance(rng: np.random.Generator, n: int, grid: int, n_phase: int, n_reg: int, vsini: float,
                   incl: float, name: str = "instance", n_mu: int = 10) -> Workload:
    """One instance of a batch, built like the small plain workload: log-uniform grid
    from 5000 A at the small workload's spacing (20k points per 100 A, so the
    lambda range grows with n), 2 lines per A with the usual depth / width
    draws, n_mu mu bins, R = 115000, a mask of all lines and a velocity window of
    20 km/s + 2 vsini at 0.5 km/s."""
    h = np.log(5100.0 / 5000.0) / 19999
    lmax = 5000.0 * np.exp(h * (n - 1))
    wave = 5000.0 * np.exp(np.arange(n, dtype=np.float64) * h)
    lines = draw_lines(rng, max(4, int(round(2.0 * (lmax - 5000.0)))), 5000.0, lmax)
    bounds, _ = mu_bins(n_mu)
    cubes = make_cubes(wave, lines, n_mu)
    regs = random_regions(rng, n_reg)
    phases = np.arange(n_phase, dtype=np.float64) / n_phase
    window = 20000.0 + 2000.0 * np.ceil(vsini)
    return Workload(name, wave, 1, bounds, cubes, lines, lines["center"].copy(), lines["depth"].copy(),
                    _star(grid, vsini, incl=incl), _act(), phases, [list(regs) for _ in phases],
                    sigma_inst_ms(115000), vel_grid(window, 500.0),
                    meta={"n_reg": n_reg, "vsini": vsini, "incl": incl})


def config2_params(seed: int = 1236, n_inst: int = 256):
    """Per-instance draws of the batched forward workload: (child seed, n, grid,
    phases, regions, vsini, inclination) with grid N in [200, 1000], 5k-200k lambda
    points, 10-100 phases, 1-6 overlapping regions, vsini U(1, 20) km/s,
    inclination U(10, 90) deg."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        out.append((int(rng.integers(0, 2 ** 62)), int(rng.integers(5000, 200001)), int(rng.integers(200, 1001)),
                    int(rng.integers(10, 101)), int(rng.integers(1, 7)), float(rng.uniform(1.0, 20.0)),
                    float(rng.uniform(10.0, 90.0))))
    return out


def config2_instance(params, i: int = 0) -> Workload:
    cs, n, grid, n_phase, n_reg, vsini, incl = params
    return batch_instance(np.random.default_rng(cs), n, grid, n_phase, n_reg, vsini, incl, name=f"config2[{i}]")


def config2(seed: int = 1236, n_inst: int = 256):
    """Batched forward: n_inst uneven instances (config2_params).  A generator:
    instances are built one at a time so that the host never holds all seed
    cubes."""
    for i, prm in enumerate(config2_params(seed, n_inst)):
        yield config2_instance(prm, i)


def small_batch(seed: int = 17, n_inst: int = 5):
    """Uneven tiny instances for batch parity tests (grid 24-72, 2600-6000 points,
    2-7 phases, 1-5 regions, vsini 1-20 km/s, inclination 10-90 deg)."""
    rng = np.random.default_rng(seed)
    return [batch_instance(rng, int(rng.integers(2600, 6001)), int(rng.integers(24, 73)), int(rng.integers(2, 8)),
                           int(rng.integers(1, 6)), float(rng.uniform(1.0, 20.0)), float(rng.uniform(10.0, 90.0)),
                           name=f"small_batch[{i}]", n_mu=6)
            for i in range(n_inst)]


def smooth_blob_maps(rng: np.random.Generator, n_phase: int, res: int, n_blobs: int = 6) -> np.ndarray:
    """Filling-factor maps [n_phase][res][res] over [-1, 1]^2: a sum of Gaussian blobs
    (centres in the disk, widths 0.05-0.25) clipped to [0, 1], drifting in y from
    phase to phase like rotating regions."""
    x = np.linspace(-1.0, 1.0, res)
    Y, Z = np.meshgrid(x, x, indexing="ij")
    c = rng.uniform(-0.7, 0.7, (n_blobs, 2))
    wdt = rng.uniform(0.05, 0.25, n_blobs)
    amp = rng.uniform(0.5, 1.5, n_blobs)
    out = np.empty((n_phase, res, res))
    for p in range(n_phase):
        dy = 1.6 * p / max(1, n_phase) - 0.8
        m = np.zeros((res, res))
        for b in range(n_blobs):
            yc = (c[b, 0] + dy + 1.0) % 2.0 - 1.0
            m += amp[b] * np.exp(-((Y - yc) ** 2 + (Z - c[b, 1]) ** 2) / (2 * wdt[b] ** 2))
        out[p] = np.clip(m, 0.0, 1.0)
    return out


def region_maps(regions, phases, incl_deg: float, res: int):
    """Discrete regions rendered as per-phase spot and facula filling-factor maps
    [P][res][res] (values 0 / 1 at the bilinear nodes y, z = -1 + 2 i / (res - 1)).
    A region (lat, lon, size, type) is the cap u . c >= sqrt(1 - size^2) of the
    unit sphere around its centre c (the membership test x^2 >= 1 - s^2 of
    SOAP_integration.cu:364, in the region's own frame), c at longitude
    lon + 360 phase on the star rotating about its axis, inclined by incl_deg to
    the line of sight; later regions win shared nodes (SOAP_integration.cu:364-381).
    Observer frame: x toward the observer, y along the rotation velocity of the
    disk centre (the model's delta = vsini y), z up."""
    g = np.linspace(-1.0, 1.0, res)
    Y, Z = np.meshgrid(g, g, indexing="ij")
    spot = np.zeros((len(phases), res, res))
    fac = np.zeros_like(spot)
    for p, ph in enumerate(phases):
        owner = region_owner(regions, ph, incl_deg, Y, Z)
        for j, (_, _, _, typ) in enumerate(regions):
            (spot if typ == 0 else fac)[p][owner == j] = 1.0
    return spot, fac


def region_owner(regions, phase: float, incl_deg: float, Y, Z):
    """Index of the region owning each disk point (Y, Z) at a phase (the last
    region containing it), -1 where none does or off the disk (geometry of
    region_maps)."""
    r2 = Y * Y + Z * Z
    on = r2 <= 1.0
    X = np.sqrt(np.clip(1.0 - r2, 0.0, None))
    si, ci = np.sin(np.deg2rad(incl_deg)), np.cos(np.deg2rad(incl_deg))
    owner = np.full(np.shape(Y), -1)
    for j, (lat, lon, size, _typ) in enumerate(regions):
        la, lo = np.deg2rad(lat), np.deg2rad(lon + 360.0 * phase)
        cx, cy, cz = np.cos(la) * np.cos(lo), np.cos(la) * np.sin(lo), np.sin(la)
        ox, oy, oz = cx * si + cz * ci, cy, -cx * ci + cz * si
        owner[on & (X * ox + Y * oy + Z * oz >= np.sqrt(1.0 - size * size))] = j
    return owner


def config3(seed: int = 1237, n_phase: int = 50, grid: int = 600, map_res: int = 512, n_mu: int = 10):
    """Adjoint shape: grid 600, lambda 5000-5500 A log-uniform (~100k points), 50
    phases, 8000 seeded lines (all in the mask), 512 x 512 smooth-blob spot and
    facula filling-factor maps per phase (facula blobs scaled to 0.6).  Returns
    (workload, spot maps, facula maps)."""
    rng = np.random.default_rng(seed)
    wave = log_grid(5000.0, 5500.0, n=100000)
    lines = draw_lines(rng, 8000, 5000.0, 5500.0)
    bounds, _ = mu_bins(n_mu)
    cubes = make_cubes(wave, lines, n_mu)
    maps = smooth_blob_maps(rng, n_phase, map_res)
    fmaps = 0.6 * smooth_blob_maps(rng, n_phase, map_res)
    phases = np.arange(n_phase, dtype=np.float64) / n_phase
    w = Workload("config3", wave, 1, bounds, cubes, lines, lines["center"].copy(), lines["depth"].copy(),
                 _star(grid, 2.0), _act(), phases, [[] for _ in phases], sigma_inst_ms(115000),
                 vel_grid(20000.0, 250.0), meta={"map_res": map_res})
    return w, maps, fmaps


def config4_params(seed: int = 1238, n_inst: int = 128):
    """Per-instance draws of the fitting workload: (child seed, n, grid, phases,
    regions) with grid 300-500, 30k-80k lambda points, 20-60 phases, 1-4 regions
    (the blobs of the filling-factor maps), vsini U(1, 10) km/s."""
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(0, 2 ** 62)), int(rng.integers(30000, 80001)), int(rng.integers(300, 501)),
             int(rng.integers(20, 61)), int(rng.integers(1, 5)), float(rng.uniform(1.0, 10.0)))
            for _ in range(n_inst)]


def config4_instance(params, map_res: int = 256, i: int = 0):
    """(workload, spot maps, facula maps, truth) of one fitting instance: a log-grid
    instance as in the batched workload with its 1-4 discrete regions (faculae,
    each with a spot inside, last region wins), rendered per phase as true spot and
    facula filling-factor maps (region_maps, inclination 90 deg); truth =
    dict(vsini, contrast, contrast_fac)."""
    cs, n, grid, n_phase, n_reg, vsini = params
    rng = np.random.default_rng(cs)
    w = batch_instance(rng, n, grid, n_phase, n_reg, vsini, 90.0, name=f"config4[{i}]")
    spot, fac = region_maps(w.regions[0], w.phases, 90.0, map_res)
    return w, spot, fac, dict(vsini=vsini, contrast=float(rng.uniform(0.2, 0.5)),
                              contrast_fac=float(rng.uniform(1.05, 1.3)))
