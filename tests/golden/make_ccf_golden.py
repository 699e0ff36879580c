"""Writes tests/golden/ccf_linear_golden.npz: input/output vectors of the
reference's mask CCF on a linear 0.005 A grid.

Run in a checkout where the reference tree is present (it is read from
/root/reference/SOAP_code/CCF_func.py, which imports numpy only); the vectors
it writes are plain numeric arrays and are all the tests need.

What is exercised (SOAP_code/CCF_func.py:5-91 with the extended grid of
SOAP_code/run_all_SOAP.py:288-296):
  - the line selection of run_all_SOAP.py:288-290 (range and weight > 0.1);
  - calc_mask's hole edges (ceil of /0.005 index arithmetic from the extended
    grid's first point), the fractional end pixels, and later lines overwriting
    earlier ones (overlapping pairs);
  - lines near both trim edges, whose holes reach the linspace extensions and
    their duplicated join points (the extension spacing is 100/19999 A);
  - calculate_CCF_2 on random spectra, a spectrum of ones and a line spectrum.

Outputs per case: the selected lines, the CCF of each spectrum, and (for the
41-velocity case) the non-zero entries of the dense template.
"""
from __future__ import annotations

import importlib.util
import os

import numpy as np

REF = "/root/reference/SOAP_code/CCF_func.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ccf_linear_golden.npz")


def load_ccf_func():
    spec = importlib.util.spec_from_file_location("ref_ccf_func", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def make_case(rng, n_full, start, vgrid):
    """Inputs shaped as the driver forms them: full grid, trims 200 then 10."""
    step_true = 0.005
    wave_full = start + step_true * np.arange(n_full)
    step_size = round(wave_full[3] - wave_full[2], 5)
    wave_t = wave_full[200:-200][10:-10]
    lo, hi = wave_t.min(), wave_t.max()
    ls = list(rng.uniform(lo + 0.5, hi - 0.5, 36))
    ws = list(rng.uniform(0.12, 1.0, 36))
    # overlapping pairs: the second hole starts inside the first
    for _ in range(10):
        a = rng.uniform(lo + 0.5, hi - 0.5)
        ls += [a, a + rng.uniform(0.002, 0.012)]
        ws += list(rng.uniform(0.12, 1.0, 2))
    # near both trim edges: holes reach the extensions at the velocity extremes
    for _ in range(4):
        ls += [lo + rng.uniform(0.0005, 0.4), hi - rng.uniform(0.0005, 0.4)]
        ws += list(rng.uniform(0.12, 1.0, 2))
    # rejected by the selection: weight <= 0.1, outside the range
    ls += [rng.uniform(lo + 1, hi - 1), lo - 0.3, hi + 0.3]
    ws += [0.05, 0.5, 0.5]
    lines_l = np.array(ls)
    lines_w = np.array(ws)
    # run_all_SOAP.py:288-290
    keep = (lines_l > wave_t.min()) & (lines_l < wave_t.max()) & (lines_w > 0.1)
    sel_l, sel_w = lines_l[keep], lines_w[keep]
    # run_all_SOAP.py:292-295
    wave_extend_a = 100.0
    no_points = int(wave_extend_a / step_size)
    before = np.linspace(np.min(wave_t) - step_size * no_points, np.min(wave_t), no_points)
    after = np.linspace(wave_t[-1], wave_t[-1] + step_size * no_points, no_points)
    wave_ext = np.concatenate((before, wave_t, after))
    n_t = len(wave_t)
    spectra = np.stack([
        rng.uniform(0.0, 1.0, n_t),
        rng.uniform(0.0, 1.0, n_t),
        np.ones(n_t),
        1.0 - 0.6 * np.exp(-0.5 * ((wave_t[:, None] - sel_l[None, :12]) / 0.02) ** 2).sum(1).clip(0, 1.5) / 1.5,
    ])
    return dict(wave_full=wave_full, step=step_size, wave_t=wave_t, lines_l=lines_l, lines_w=lines_w, sel_l=sel_l,
                sel_w=sel_w, wave_ext=wave_ext, vgrid=vgrid, spectra=spectra)


def main():
    cf = load_ccf_func()
    rng = np.random.default_rng(20261203)
    cases = {
        "a": make_case(rng, 3420, 5000.0, np.arange(-20000, 20001.0, 1000.0)),
        "b": make_case(rng, 2420, 6211.3, np.arange(-20000, 20001.0, 250.0)),
    }
    out = {}
    for key, c in cases.items():
        tpl = cf.calculate_CCF_1(c["vgrid"], c["wave_t"], c["sel_l"], c["sel_w"], c["wave_ext"])
        ccf = np.stack([cf.calculate_CCF_2(c["vgrid"], s, tpl) for s in c["spectra"]])
        for name in ("wave_t", "step", "vgrid", "lines_l", "lines_w", "sel_l", "sel_w", "spectra"):
            out[f"{key}_{name}"] = np.asarray(c[name])
        out[f"{key}_ccf"] = ccf
        if key == "a":
            r, col = np.nonzero(tpl)
            out[f"{key}_tpl_row"] = r.astype(np.int32)
            out[f"{key}_tpl_col"] = col.astype(np.int32)
            out[f"{key}_tpl_val"] = tpl[r, col]
        print(key, "lines", len(c["sel_l"]), "v", len(c["vgrid"]), "n_t", len(c["wave_t"]),
              "finite ccf", bool(np.all(np.isfinite(ccf))))
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
