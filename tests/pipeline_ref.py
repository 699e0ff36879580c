"""Oracle pipeline over a synth.Workload: the run_all_SOAP.py flow (quiet star,
per-phase integration, resolution, mask CCF, fit, RV referencing) with every
numerical stage done by the CPU oracle.  Test infrastructure only."""
from __future__ import annotations

import numpy as np

import oracle as O

TRIM = 210  # cutoff 200 then 10 (run_all_SOAP.py:259,266,270-281)


def oracle_maps(w):
    s = w.star
    return O.disk_init(s["grid"], s["prot"], s.get("w_ratio", 0.0), s["incl"], s["radius_km"], s["u1"], s["u2"],
                       s["lb"], w.mu_bounds)


def select_mask(w, trim=TRIM):
    wave_t = w.wave[trim:len(w.wave) - trim]
    return O.select_lines(wave_t, w.mask_l, w.mask_w)


def grid_of(w):
    """(oracle grid kind, step argument): log grids carry no step; linear grids the
    0.005 A step of the reference's index formula."""
    if w.grid_kind == O.GRID_LINEAR:
        return O.GRID_LINEAR, float(w.meta["step"])
    return O.GRID_LOG, 0.0


def run_oracle(w, phases=None, nthreads=None, with_ccf=True):
    """Returns dict(maps, quiet[2], raw[P,3,n], lower[P+1,3,n], ccf[P+1,3,nv], fit...)."""
    G = w.star["grid"]
    kind, step = grid_of(w)
    maps = oracle_maps(w)
    q = O.quiet_star(kind, w.wave, maps, w.cubes["q"], G, step=step, nthreads=nthreads)
    qs = O.quiet_star(kind, w.wave, maps, w.cubes["qsed"], G, step=step, nthreads=nthreads)
    idx = list(range(len(w.phases))) if phases is None else list(phases)
    a = w.act
    raw = np.zeros((len(idx), 3, w.n))
    pm = []
    for r, p in enumerate(idx):
        temp, bi = O.phase_maps(G, w.star["incl"], a["nrho"], a["t_star"], a["t_diff_spot"], a["t_diff_fac"],
                                w.phases[p], w.regions[p])
        pm.append((temp, bi))
        raw[r] = O.integrate_phase(kind, w.wave, maps, temp, bi, w.cubes, G, a["t_star"], a["t_diff_fac"],
                                   step=step, nthreads=nthreads)
    out = dict(maps=maps, quiet=np.stack([q, qs]), raw=raw, phase_maps=pm, idx=idx)
    if not with_ccf:
        return out
    lower = np.zeros((len(idx) + 1, 3, w.n))
    mm = np.zeros((len(idx) + 1, 3, 2))
    for r in range(len(idx)):
        for t in range(3):
            quiet = q if t == 1 else qs  # SOAP_resolution.cu:414-416
            lower[r, t], mm[r, t] = O.lower_resolution(kind, w.wave, w.sigma_inst, quiet, raw[r, t], 1,
                                                       nthreads=nthreads)
    # quiet form: weight = the step_med argument of SOAP_initialize (run_all_SOAP.py:226)
    qw = step if kind == O.GRID_LINEAR else 0.005
    lq_sed, mm_sed = O.lower_resolution(kind, w.wave, w.sigma_inst, None, qs, 0, weight=qw, nthreads=nthreads)
    lq, mm_q = O.lower_resolution(kind, w.wave, w.sigma_inst, None, q, 0, weight=qw, nthreads=nthreads)
    lower[-1] = [lq_sed, lq, lq_sed]
    mm[-1] = [mm_sed, mm_q, mm_sed]
    out.update(lower=lower, mm=mm)
    out.update(ccf_stage(w, lower, nthreads))
    return out


def ccf_stage(w, lower, nthreads=None, trim=TRIM):
    wave_t = w.wave[trim:len(w.wave) - trim]
    sl, sw = select_mask(w, trim)
    kind, step = grid_of(w)
    h = O.grid_h(w.wave) if kind == O.GRID_LOG else step
    ap = O.appodisation(len(w.vgrid))
    nf = lower.shape[0]
    ccf = np.zeros((nf, 3, len(w.vgrid)))
    par = np.zeros((nf, 3, 5))
    bis = np.zeros((nf, 3, 91))
    for f in range(nf):
        for t in range(3):
            raw_ccf = O.ccf(kind, wave_t, h, w.vgrid, sl, sw, lower[f, t, trim:w.n - trim], nthreads=nthreads)
            ccf[f, t] = O.ccf_finish(raw_ccf, ap)
            r = O.fit_ccf(w.vgrid, ccf[f, t])
            par[f, t] = r["par"]
            bis[f, t] = r["bis"]
    rv = par[:, :, 3] - par[-1:, :, 3]  # vrad -= vrad[-1] (run_all_SOAP.py:396-397)
    return dict(ccf=ccf, par=par, bis=bis, rv=rv[:-1])


def normwise(a, b, lo=TRIM, hi=None):
    """max|a-b| / max|b| over the interior [lo, n-lo)."""
    n = a.shape[-1]
    hi = n - lo if hi is None else hi
    a = a[..., lo:hi]
    b = b[..., lo:hi]
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))
