"""GPU parity at the production shapes of the benchmark workloads.

Production workload (grid 1000, 732,585-pixel log grid 3800-7000 A, 8000
lines, 20 mu bins, 2 spots + 2 faculae overlapping, 100 phases, R = 115000,
4000-line mask on 161 velocities): the whole series runs on the GPU
(soap_run_series), and the oracle checks it where it finishes in seconds:

- the disk maps of every cell and the region maps of the checked phases, bitwise;
- the quiet star (both cubes), the raw deltas and the convolved spectra of
  phases 0, 37 and 99 on pixel windows [a, b): the oracle runs on
  [a - 300, b + 300) (300 pixels exceed the largest Doppler shift, 9 pixels,
  plus the convolution half-width, 221 pixels), so its values on [a, b) are
  those of a full-length run; windows at three fixed places and around the
  pixels that set spec_norm's constants (min of A, max of the convolution);
- spec_norm's constants themselves: min(A) and max(o != 0) with the -60000 seed
  (SOAP_resolution.cu:123-232), from the oracle on the windows holding them;
- the CCFs and fits of those phases and of the quiet frame: the oracle's mask
  CCF, apodisation and BIS_FIT2 fit of the GPU's full-length lower spectra;
  the CCF to 1e-12 relative, RV / FWHM / bisector span within 1 cm/s, and the
  RV referenced to the quiet frame (run_all_SOAP.py:396-397).

Batched workload (256 uneven instances): the batch at its autotuned shapes
equals the plain per-instance path bit for bit in every CCF and fit, and in the
spectra of sampled phases, and the two smallest instances equal the oracle's
full pipeline.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from soap_gpu_amd import synth
from soap_gpu_amd.engine import BCONV, Batch, Engine, setup_workload
from tests.pipeline_ref import TRIM, ccf_stage, oracle_maps, run_oracle

pytestmark = pytest.mark.gpu

SPEC_TOL = 1e-12
RV_TOL = 0.01  # m/s
MARGIN = 300
CHECK_PHASES = (0, 37, 99)


def _log(msg):
    print(msg, flush=True)


def _norm_window(a, b):
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den)


@pytest.fixture(scope="module")
def config1_run(engine):
    w = synth.config1()
    _log(f"[config1] workload built: n={w.n} phases={len(w.phases)}")
    setup_workload(engine, w)
    engine.run_series(w.star, w.sigma_inst)
    engine.sync()
    _log("[config1] series done on the GPU")
    g = dict(maps=engine.get_maps(), quiet=np.stack([engine.get_quiet(0), engine.get_quiet(1)]),
             raw={p: np.stack(engine.get_raw(p)) for p in CHECK_PHASES},
             lower={p: np.stack([engine.get_lower(p, t) for t in range(3)]) for p in CHECK_PHASES + (-1,)},
             norm={p: np.stack([engine.get_norm(p, t) for t in range(3)]) for p in CHECK_PHASES + (-1,)},
             pmaps={p: engine.phase_maps(p) for p in CHECK_PHASES},
             ccf=engine.get_ccf(), fit=engine.get_fit())
    return w, g


def _quiet_for(t):
    return 0 if t == BCONV else 1  # SOAP_resolution.cu:414-416


def _windows(w, g):
    """Pixel windows [a, b): three fixed ones, and one around each pixel that
    sets a spec_norm constant of a checked spectrum (argmin of A = quiet - raw,
    argmax of the non-zero convolution output)."""
    n = w.n
    wins = [(int(f * n) - 2000, int(f * n) + 2000) for f in (0.15, 0.5, 0.85)]
    centres = set()
    for p in CHECK_PHASES:
        for t in range(3):
            A = g["quiet"][_quiet_for(t)] - g["raw"][p][t]
            centres.add(int(np.argmin(A)))
            mn, mx = g["norm"][p][t]
            o = (1.0 - g["lower"][p][t]) * mx / (1.0 - mn)
            nz = np.nonzero(g["lower"][p][t] != 1.0)[0]
            centres.add(int(nz[np.argmax(o[nz])]))
    for c in sorted(centres):
        a, b = max(c - 400, TRIM), min(c + 400, n - TRIM)
        if not any(a >= x and b <= y for x, y in wins):
            wins.append((a, b))
    return wins


@pytest.fixture(scope="module")
def config1_oracle(config1_run):
    """The oracle's quiet star, raw deltas and convolution on every window."""
    w, g = config1_run
    G = w.star["grid"]
    a_ = w.act
    maps = oracle_maps(w)
    pm = {p: O.phase_maps(G, w.star["incl"], a_["nrho"], a_["t_star"], a_["t_diff_spot"], a_["t_diff_fac"],
                          w.phases[p], w.regions[p]) for p in CHECK_PHASES}
    out = []
    for a, b in _windows(w, g):
        lo, hi = max(a - MARGIN, 0), min(b + MARGIN, w.n)
        sl = slice(lo, hi)
        wave = np.ascontiguousarray(w.wave[sl])
        cubes = {k: np.ascontiguousarray(v[:, sl]) for k, v in w.cubes.items()}
        q = np.stack([O.quiet_star(O.GRID_LOG, wave, maps, cubes[k], G) for k in ("q", "qsed")])
        raw, conv = {}, {}
        for p in CHECK_PHASES:
            temp, bi = pm[p]
            raw[p] = np.stack(O.integrate_phase(O.GRID_LOG, wave, maps, temp, bi, cubes, G, a_["t_star"],
                                                a_["t_diff_fac"]))
            conv[p] = np.stack([O.conv_raw(O.GRID_LOG, wave, w.sigma_inst, q[_quiet_for(t)] - raw[p][t], 1)
                                for t in range(3)])
        qconv = np.stack([O.conv_raw(O.GRID_LOG, wave, w.sigma_inst, q[k], 0, weight=0.005) for k in (0, 1)])
        out.append(dict(a=a, b=b, lo=lo, q=q, raw=raw, conv=conv, qconv=qconv))
        _log(f"[config1] oracle window [{a}, {b}) done")
    return maps, pm, out


@pytest.mark.timeout(900)
def test_config1_maps_bitwise(config1_run, config1_oracle):
    w, g = config1_run
    maps, pm, _ = config1_oracle
    for k in ("mu", "delta", "limb", "lb", "region"):
        assert np.array_equal(g["maps"][k], maps[k]), k
    for p in CHECK_PHASES:
        assert np.array_equal(g["pmaps"][p][0], pm[p][0]), p
        assert np.array_equal(g["pmaps"][p][1], pm[p][1]), p


@pytest.mark.timeout(900)
def test_config1_spectra_on_windows(config1_run, config1_oracle):
    """Quiet star, raw deltas and the unnormalised convolution (recovered from
    the GPU's spec_norm output with its own constants) on every window."""
    w, g = config1_run
    _, _, wins = config1_oracle
    worst = 0.0
    for win in wins:
        a, b, lo = win["a"], win["b"], win["lo"]
        r = slice(a - lo, b - lo)
        for k in range(2):
            e = _norm_window(g["quiet"][k][a:b], win["q"][k][r])
            assert e <= SPEC_TOL, ("quiet", k, a, e)
            worst = max(worst, e)
        for p in CHECK_PHASES:
            for t in range(3):
                e = _norm_window(g["raw"][p][t][a:b], win["raw"][p][t][r])
                assert e <= SPEC_TOL, ("raw", p, t, a, e)
                mn, mx = g["norm"][p][t]
                o = (1.0 - g["lower"][p][t][a:b]) * mx / (1.0 - mn)
                e2 = _norm_window(o, win["conv"][p][t][r])
                assert e2 <= SPEC_TOL, ("conv", p, t, a, e2)
                worst = max(worst, e, e2)
        for k, t in ((1, 0), (0, 1)):  # quiet frames: SED (flux/tot slots), normalised (bconv slot)
            mn, mx = g["norm"][-1][t]
            o = (1.0 - g["lower"][-1][t][a:b]) * mx / (1.0 - mn)
            e = _norm_window(o, win["qconv"][k][r])
            assert e <= SPEC_TOL, ("quiet conv", k, a, e)
    _log(f"[config1] worst window error {worst:.2e} over {len(wins)} windows")


@pytest.mark.timeout(900)
def test_config1_spec_norm_constants(config1_run, config1_oracle):
    """min(A) over every pixel and max of the non-zero convolution output (seed
    -60000): the GPU's constants equal the oracle's values at the pixels that set
    them, and no window holds a more extreme one."""
    w, g = config1_run
    _, _, wins = config1_oracle
    for p in CHECK_PHASES:
        for t in range(3):
            mn, mx = g["norm"][p][t]
            omin = min(float(np.min(win["q"][_quiet_for(t)] - win["raw"][p][t])) for win in wins)
            assert abs(mn - omin) <= SPEC_TOL * abs(omin), (p, t, mn, omin)
            omax = -60000.0
            for win in wins:
                c = win["conv"][p][t]
                nz = c[c != 0.0]
                if nz.size:
                    omax = max(omax, float(nz.max()))
            assert abs(mx - omax) <= SPEC_TOL * abs(omax), (p, t, mx, omax)


@pytest.mark.timeout(900)
def test_config1_ccf_fit_rv(config1_run):
    """The oracle's CCF, apodisation and fit of the GPU's full-length lower spectra."""
    w, g = config1_run
    frames = CHECK_PHASES + (-1,)
    lower = np.stack([g["lower"][p] for p in frames])
    ref = ccf_stage(w, lower)
    for i, p in enumerate(frames):
        np.testing.assert_allclose(g["ccf"][p], ref["ccf"][i], rtol=1e-12, atol=1e-13)
        par = g["fit"]["par"][p]
        assert np.max(np.abs(par[:, 3] - ref["par"][i][:, 3])) < RV_TOL, p
        assert np.max(np.abs(par[:, 4] - ref["par"][i][:, 4])) < RV_TOL, p
        assert np.max(np.abs(par[:, 2] - ref["par"][i][:, 2])) < RV_TOL, p
    rv_g = np.stack([g["fit"]["par"][p][:, 3] - g["fit"]["par"][-1][:, 3] for p in CHECK_PHASES])
    assert np.max(np.abs(rv_g - ref["rv"])) < RV_TOL


# ----------------------------------------------------------------- batched workload
N_INST = 256


def _prepare(w):
    e = Engine(0)
    setup_workload(e, w)
    e.quiet_star()
    return e


@pytest.mark.timeout(1200)
def test_config2_batch_equals_plain_and_oracle():
    prm = synth.config2_params(n_inst=N_INST)
    engines, ws_small, info = [], {}, []
    sizes = [p[1] * p[3] for p in prm]  # n x phases
    smallest = set(np.argsort(sizes)[:2].tolist())
    from concurrent.futures import ThreadPoolExecutor
    try:
        with ThreadPoolExecutor(max_workers=8) as ex:
            for c0 in range(0, N_INST, 16):
                idx = list(range(c0, min(c0 + 16, N_INST)))
                for i, w in zip(idx, ex.map(lambda i: synth.config2_instance(prm[i], i), idx)):
                    engines.append(_prepare(w))
                    info.append((w.n, len(w.phases), w.sigma_inst))
                    if i in smallest:
                        ws_small[i] = w
                if (c0 + 16) % 64 == 0:
                    _log(f"[config2] {c0 + 16}/{N_INST} instances on the device")
        # plain per-instance reference
        for e, (_, _, s) in zip(engines, info):
            e.integrate()
            e.resolution(s)
            e.ccf()
            e.fit()
        for e in engines:
            e.sync()
        sample = list(range(0, N_INST, 16))
        ref = []
        for i, e in enumerate(engines):
            P = info[i][1]
            d = dict(ccf=e.get_ccf(), fit=e.get_fit())
            if i in sample:
                d["raw"] = [np.stack(e.get_raw(p)) for p in (0, P - 1)]
                d["lower"] = [np.stack([e.get_lower(p, t) for t in range(3)]) for p in (0, P - 1, -1)]
            ref.append(d)
        _log("[config2] plain path done")
        # the smallest instances against the oracle's full pipeline
        for i, w in ws_small.items():
            o = run_oracle(w)
            r = ref[i]
            par = r["fit"]["par"]
            rv = par[:-1, :, 3] - par[-1:, :, 3]
            assert np.max(np.abs(rv - o["rv"])) < RV_TOL, i
            assert np.max(np.abs(par[:, :, 4] - o["par"][:, :, 4])) < RV_TOL, i
            assert np.max(np.abs(par[:, :, 2] - o["par"][:, :, 2])) < RV_TOL, i
            np.testing.assert_allclose(r["ccf"], o["ccf"], rtol=1e-12, atol=1e-13)
            _log(f"[config2] instance {i} (n={w.n}, P={len(w.phases)}, grid={w.star['grid']}) equals the oracle")
        b = Batch(engines, [x[2] for x in info])
        try:
            b.tune(reps=1)
            _log(f"[config2] tuned: {b.config()}")
            for graph in (False, True):
                for e, (_, _, s) in zip(engines, info):  # clobber the plain outputs: the batch must rewrite them
                    e.set_raw(0, *(np.zeros(e.n) for _ in range(3)))
                    e.resolution(s)
                    e.ccf()
                    e.fit()
                for e in engines:
                    e.sync()
                b.step(graph=graph)
                if graph:
                    b.step(graph=True)  # a replay after the capturing launch
                b.sync()
                for i, e in enumerate(engines):
                    assert np.array_equal(e.get_ccf(), ref[i]["ccf"]), (graph, i)
                    f = e.get_fit()
                    for k in ("par", "model", "depth", "bis"):
                        assert np.array_equal(f[k], ref[i]["fit"][k]), (graph, i, k)
                    if i in sample:
                        P = info[i][1]
                        for j, p in enumerate((0, P - 1)):
                            assert np.array_equal(np.stack(e.get_raw(p)), ref[i]["raw"][j]), (graph, i, p)
                        for j, p in enumerate((0, P - 1, -1)):
                            lw = np.stack([e.get_lower(p, t) for t in range(3)])
                            assert np.array_equal(lw, ref[i]["lower"][j]), (graph, i, p)
                _log(f"[config2] batched step ({'graph' if graph else 'eager'}) equals the plain path")
        finally:
            b.close()
    finally:
        for e in engines:
            e.close()
