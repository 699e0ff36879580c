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
import torch  # noqa: F401  (before the library initialises HIP)

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


# ----------------------------------------------------------------- adjoint workload
@pytest.fixture(scope="module")
def config3_model():
    import torch

    from soap_gpu_amd.diff import DiffModel
    from soap_gpu_amd.engine import select_mask_lines

    w, maps, fmaps = synth.config3()
    P, M = maps.shape[0], maps.shape[1]
    sl, sw = select_mask_lines(w.wave, w.mask_l, w.mask_w, TRIM)
    eng = Engine(0)
    model = DiffModel(eng, w.wave, w.mu_bounds, sl, sw, w.vgrid, w.star["grid"], P, M, w.sigma_inst, TRIM,
                      lb=w.star["lb"])
    dev = torch.device("cuda")
    t_ = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)  # noqa: E731
    d = dict(w=w, model=model, eng=eng, P=P, M=M, q=t_(w.cubes["q"]), s=t_(w.cubes["sp"]), f=t_(w.cubes["fac"]),
             maps=t_(maps), fmaps=t_(fmaps), zero=torch.zeros((P, w.n), dtype=torch.float64, device=dev))
    model.forward((2.3, 0.31, 0.30, 0.35, 1.15), d["q"], d["s"], d["maps"] * 0.8, d["zero"], np.zeros(P),
                  np.zeros(P), 1.0, 0.0, 0.0, f=d["f"], fmaps=d["fmaps"] * 0.8)
    F = model.get("F")
    th = model.get("theta")
    rng = np.random.default_rng(1237)
    d["T"] = t_(F + 1e-4 * np.abs(F).max() * rng.standard_normal(F.shape))
    d["rv_t"] = th[:, 2] + rng.uniform(-5.0, 5.0, P)
    d["fw_t"] = th[:, 3] + rng.uniform(-20.0, 20.0, P)
    d["bis_t"] = model.get("bis") + rng.uniform(-5.0, 5.0, P)
    d["wts"] = (1.0 / float(torch.sum(d["T"][:, TRIM:w.n - TRIM] ** 2)), 1e-3, 1e-5, 1e-3)
    d["prm"] = np.array([2.0, 0.29, 0.34, 0.3, 1.1])
    _log(f"[config3] model set up: n={w.n} grid={w.star['grid']} P={P} maps {M}x{M}")
    yield d
    eng.close()


def _c3_loss(d, prm=None, **over):
    a = dict(q=d["q"], s=d["s"], f=d["f"], maps=d["maps"], fmaps=d["fmaps"])
    a.update(over)
    w = d["wts"]
    return d["model"].forward(d["prm"] if prm is None else prm, a["q"], a["s"], a["maps"], d["T"], d["rv_t"],
                              d["fw_t"], *w[:3], bis_t=d["bis_t"], w_bis=w[3], f=a["f"], fmaps=a["fmaps"])[0]


@pytest.mark.timeout(900)
def test_config3_adjoint_vs_central_differences(config3_model):
    """At the adjoint shape (grid 600, 100k pixels, 50 phases, 512 x 512 spot and
    facula maps, 7473-line mask), every gradient of the full loss (spectra, RV,
    FWHM, bisector span) along its own direction -- the five scalars, and one
    random direction each of the quiet / spot / facula seed rows and of both maps
    -- against central differences (h = 1e-6 relative) of the GPU forward, whose
    parity with the oracle is held at small shapes in test_gpu_diff."""
    import torch

    d = config3_model
    _c3_loss(d)
    g5, gq, gs, gm, gf, gfm = d["model"].backward()
    p0 = d["prm"]
    for i, name in enumerate(("vsini", "u1", "u2", "contrast", "contrast_fac")):
        h = 1e-6 * abs(p0[i])
        e = np.eye(5)[i] * h
        fd = (_c3_loss(d, p0 + e) - _c3_loss(d, p0 - e)) / (2 * h)
        assert g5[i] == pytest.approx(fd, rel=2e-5, abs=1e-9 * np.abs(g5).max()), name
        _log(f"[config3] d/d{name}: adjoint {g5[i]:.9e} central {fd:.9e}")
    gen = torch.Generator(device="cuda").manual_seed(3)
    for name, g in (("q", gq), ("s", gs), ("f", gf), ("maps", gm), ("fmaps", gfm)):
        base = d[name]
        v = torch.randn(base.shape, dtype=torch.float64, device="cuda", generator=gen)
        h = 1e-6 * float(torch.linalg.vector_norm(base)) / float(torch.linalg.vector_norm(v))
        fd = (_c3_loss(d, **{name: base + h * v}) - _c3_loss(d, **{name: base - h * v})) / (2 * h)
        ad = float(torch.sum(g * v))
        assert ad == pytest.approx(fd, rel=2e-5, abs=1e-12), name
        _log(f"[config3] directional d/d{name}: adjoint {ad:.9e} central {fd:.9e}")


@pytest.mark.timeout(900)
def test_config3_empty_maps_give_the_convolved_quiet_star(config3_model):
    """With both maps empty the model spectrum is the instrumental convolution of
    the plain engine's quiet star (itot, SOAP_initialize.cu:297-395) at the same
    disk grid, limb darkening and vsini: F = conv(1) - o with o the plain path's
    unnormalised quiet-frame convolution (conv(1) from the oracle's taps)."""
    import torch

    d = config3_model
    w = d["w"]
    z = torch.zeros_like(d["maps"])
    d["model"].forward((2.0, 0.29, 0.34, 0.3, 1.1), d["q"], d["s"], z, d["zero"], np.zeros(d["P"]),
                       np.zeros(d["P"]), 1.0, 0.0, 0.0, f=d["f"], fmaps=z)
    F = d["model"].get("F")
    eng = Engine(0)
    try:
        eng.set_wave(w.wave)
        eng.set_mu_bounds(w.mu_bounds)
        eng.set_cubes(w.cubes)
        st = dict(w.star, lb=np.zeros(4))
        eng.disk_init(st)
        eng.quiet_star()
        eng.alloc_phases(1)
        eng.set_raw(0, *(np.zeros(w.n) for _ in range(3)))
        eng.resolution(w.sigma_inst)
        out = eng.get_lower(-1, BCONV)
        mn, mx = eng.get_norm(-1, BCONV)
    finally:
        eng.close()
    o = (1.0 - out) * mx / (1.0 - mn)
    one = O.conv_raw(O.GRID_LOG, w.wave, w.sigma_inst, np.zeros(w.n), 0)
    ref = np.where(one != 0.0, one - o, 0.0)
    sl_ = slice(TRIM, w.n - TRIM)
    for p in (0, d["P"] - 1):
        e = float(np.max(np.abs(F[p][sl_] - ref[sl_])) / np.max(np.abs(ref[sl_])))
        assert e <= 1e-11, (p, e)


# ----------------------------------------------------------------- fitting workload
@pytest.mark.timeout(900)
def test_config4_graph_equals_per_call_and_fit_descends():
    """Two instances of the fitting workload at full shape (grid 300-500,
    30k-80k pixels, 20-60 phases, 256 x 256 spot and facula maps rendered from
    their discrete regions): one graph replay of forward + adjoint equals the
    per-call path (loss bit for bit, gradients to rounding), and five L-BFGS
    iterations of the graph-captured loop lower the summed loss."""
    import torch

    from soap_gpu_amd.diff import DiffBatch
    from soap_gpu_amd.fitting import Instance, fit_graph

    prm = synth.config4_params()[:2]
    made = [synth.config4_instance(p, 256, i) for i, p in enumerate(prm)]
    insts, calls = [], []
    for w, maps, fmaps, tr in made:
        P = len(w.phases)
        start = np.clip(0.5 * maps + 0.25 * maps.mean(), 0.0, 1.0)
        fstart = np.clip(0.5 * fmaps + 0.25 * fmaps.mean(), 0.0, 1.0)
        for lst in (insts, calls):
            inst = Instance(w, start, P, 256, vsini0=tr["vsini"] * 1.2, contrast0=0.5 * (tr["contrast"] + 0.3),
                            weights=(0.0, 1e-2, 1e-4, 1e-3), fmaps0=fstart, contrast_fac0=1.15)
            rv, fw, bis = inst.observe((tr["vsini"], 0.29, 0.34, tr["contrast"], tr["contrast_fac"]), maps, fmaps)
            inst.rv_t, inst.fw_t, inst.bis_t = rv, fw, bis
            lst.append(inst)
        _log(f"[config4] instance n={w.n} grid={w.star['grid']} P={P} regions={len(w.regions[0])}")
    try:
        dev = torch.device("cuda")
        batch = DiffBatch([g.model for g in insts])
        bufs = []
        for i, g in enumerate(insts):
            t = dict(prm=torch.tensor(g.params().detach().cpu().numpy(), dtype=torch.float64, device=dev),
                     maps=g.maps().detach().contiguous(), fmaps=g.fmaps().detach().contiguous(),
                     rv=torch.tensor(g.rv_t, dtype=torch.float64, device=dev),
                     fw=torch.tensor(g.fw_t, dtype=torch.float64, device=dev),
                     bis=torch.tensor(g.bis_t, dtype=torch.float64, device=dev),
                     loss=torch.empty(1, dtype=torch.float64, device=dev),
                     g5=torch.empty(5, dtype=torch.float64, device=dev), gq=torch.empty_like(g.q),
                     gs=torch.empty_like(g.s), gf=torch.empty_like(g.q), gm=torch.empty_like(g.maps()),
                     gfm=torch.empty_like(g.maps()))
            batch.bind(i, 2.0 * float(g.theta[0]) + 5.0, t["prm"], g.q, g.s, t["maps"], g.target, t["rv"], t["fw"],
                       t["bis"], g.weights, t["loss"], t["g5"], t["gq"], t["gs"], t["gm"], f=g.f,
                       fmaps=t["fmaps"], gf=t["gf"], gfm=t["gfm"])
            bufs.append(t)
        batch.step(graph=True)
        batch.step(graph=True)  # a replay after the capturing launch
        batch.sync()
        for t, c in zip(bufs, calls):
            loss, _, _, _ = c.model.forward(c.params().detach().cpu().numpy(), c.q, c.s, c.maps().detach(),
                                            c.target, c.rv_t, c.fw_t, *c.weights[:3], bis_t=c.bis_t,
                                            w_bis=c.weights[3], f=c.f, fmaps=c.fmaps().detach())
            g5, gq, gs, gm, gf, gfm = c.model.backward()
            assert float(t["loss"][0]) == loss
            np.testing.assert_array_equal(t["g5"].cpu().numpy(), g5)
            for got, ref in ((t["gq"], gq), (t["gs"], gs), (t["gf"], gf), (t["gm"], gm), (t["gfm"], gfm)):
                torch.testing.assert_close(got, ref, rtol=1e-9, atol=1e-12 * float(ref.abs().max()))
        batch.close()
        r = fit_graph(insts, iters=5)
        _log(f"[config4] L-BFGS: loss {r['loss0']:.6e} -> best {r['best']:.6e} in {r['evals']} evaluations")
        assert r["best"] < r["loss0"]
    finally:
        for x in insts + calls:
            x.close()
