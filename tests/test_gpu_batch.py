"""Batched uneven instances (soap_batch_* of the C ABI): one persistent launch
per stage over every instance, eager and replayed as a HIP graph, at the shapes
the autotuner picks.  The batched step runs the same per-block code as the
plain path on the same inputs, so every output (raw and lower spectra, CCFs,
fit parameters, models and bisectors) must equal the plain path's bit for bit;
the plain path itself is held to the oracle here for one instance and in
test_gpu_parity for the rest of the stage set."""
import numpy as np
import pytest

from soap_gpu_amd import synth
from soap_gpu_amd.engine import Batch, Engine, SoapError, setup_workload
from tests.pipeline_ref import run_oracle
from tests.test_gpu_parity import RV_TOL

pytestmark = pytest.mark.gpu


def _prepare(w):
    e = Engine(0)
    setup_workload(e, w)
    e.quiet_star()
    return e


def _outputs(e, w):
    P = len(w.phases)
    raw = np.stack([np.stack(e.get_raw(p)) for p in range(P)])
    lower = np.stack([np.stack([e.get_lower(p, t) for t in range(3)]) for p in list(range(P)) + [-1]])
    return dict(raw=raw, lower=lower, ccf=e.get_ccf(), fit=e.get_fit())


def _same(a, b, what):
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], dict):
            _same(a[k], b[k], f"{what}.{k}")
        else:
            assert np.array_equal(a[k], b[k]), f"{what}.{k}"


def _clobber(e, w, rng):
    """Overwrite every output of an instance (raw spectra by upload, then lower
    spectra, CCF and fit from them), keeping its buffers."""
    for p in range(len(w.phases)):
        e.set_raw(p, *(rng.uniform(0.5, 1.0, w.n) for _ in range(3)))
    e.resolution(w.sigma_inst)
    e.ccf()
    e.fit()
    e.sync()


@pytest.fixture(scope="module")
def batch_case():
    ws = synth.small_batch()
    refs = []
    for w in ws:
        r = _prepare(w)
        r.integrate()
        r.resolution(w.sigma_inst)
        r.ccf()
        r.fit()
        refs.append(_outputs(r, w))
        r.close()
    engines = [_prepare(w) for w in ws]
    b = Batch(engines, [w.sigma_inst for w in ws])
    yield ws, engines, b, refs
    b.close()
    for e in engines:
        e.close()


def _check_all(ws, engines, refs, what):
    for i, (w, e, r) in enumerate(zip(ws, engines, refs)):
        _same(_outputs(e, w), r, f"{what}[{i}]")


def test_batch_eager_equals_plain(batch_case):
    ws, engines, b, refs = batch_case
    rng = np.random.default_rng(3)
    for w, e in zip(ws, engines):
        _clobber(e, w, rng)
    b.step(graph=False)
    b.sync()
    _check_all(ws, engines, refs, "eager")


def test_batch_ccf_mixed_footprints():
    """One slow and one fast rotator in a batch: the fast one's velocity window
    (+-100 km/s at 0.3 km/s pixels) gives line footprints past the CCF's 512-pixel
    LDS segment, the slow one's fit in one; both equal their plain runs."""
    rng = np.random.default_rng(41)
    ws = [synth.batch_instance(rng, 3000, 24, 3, 2, 3.0, 60.0, name="slow", n_mu=6),
          synth.batch_instance(rng, 3000, 24, 2, 2, 40.0, 80.0, name="fast", n_mu=6)]
    h = np.log(ws[1].wave[1] / ws[1].wave[0])
    assert (ws[1].vgrid[-1] - ws[1].vgrid[0]) / (299792458.0 * h) > 600
    refs = []
    for w in ws:
        r = _prepare(w)
        r.integrate()
        r.resolution(w.sigma_inst)
        r.ccf()
        r.fit()
        refs.append(_outputs(r, w))
        r.close()
    engines = [_prepare(w) for w in ws]
    b = Batch(engines, [w.sigma_inst for w in ws])
    try:
        b.step(graph=False)
        b.sync()
        _check_all(ws, engines, refs, "mixed footprints")
    finally:
        b.close()
        for e in engines:
            e.close()


def test_batch_graph_replays(batch_case):
    """Graph replays, each after eager work that overwrites every output."""
    ws, engines, b, refs = batch_case
    rng = np.random.default_rng(4)
    for rep in range(3):
        for w, e in zip(ws, engines):
            _clobber(e, w, rng)
        b.step(graph=True)
        b.sync()
        _check_all(ws, engines, refs, f"graph replay {rep}")
    b.step(graph=False)
    b.step(graph=True)
    b.sync()
    _check_all(ws, engines, refs, "eager + graph")


def test_batch_tuned_shapes(batch_case):
    ws, engines, b, refs = batch_case
    b.tune(reps=1)
    cfg = b.config()
    assert (cfg["fir_g"], cfg["fir_b"]) in ((2, 0), (4, 0), (6, 0), (1, 3), (1, 5), (2, 3), (4, -7))
    assert cfg["fir_pre"] in (0, 1) and (cfg["fir_pre"] == 0 or cfg["fir_b"] == 0)
    assert all(g % 8 == 0 and g >= 0 for g in cfg["grid"].values())
    for name in ("taps", "fir", "conv", "ccf", "fit"):
        assert cfg["work"][name] > 0, name
    rng = np.random.default_rng(5)
    shapes = [(1024, 2, 0, 0, 8), (16384, 6, 0, 0, 16), (4096, 4, 0, 0, None), (2048, 1, 3, 0, None),
              (4096, 1, 5, 0, 4), (8192, 2, 3, 0, None), (4096, 4, 0, 1, None), (2048, 2, 0, 1, 8),
              (4096, 6, 0, 1, 16), (4096, 4, -7, 0, None), (2048, 4, -7, 0, 8),
              (cfg["taps_cells"], cfg["fir_g"], cfg["fir_b"], cfg["fir_pre"], "tuned")]
    for cells, g, bb, pre, mult in shapes:
        if mult == "tuned":
            grid = cfg["grid"]
        elif mult is None:
            grid = {n: 0 for n in cfg["grid"]}
        else:
            grid = {n: mult * 8 for n in cfg["grid"]}  # persistent: far fewer blocks than items
        b.set_config(cells, g, bb, grid, fir_pre=pre)
        for w, e in zip(ws, engines):
            _clobber(e, w, rng)
        b.step(graph=True)
        b.sync()
        got = [_outputs(e, w) for w, e in zip(ws, engines)]
        if bb <= 0:  # the plain path's one-pixel FIR forms: bit for bit
            for i, r in enumerate(refs):
                _same(got[i], r, f"shapes {cells}/{g}/{bb}/{pre}/{mult}[{i}]")
        else:  # register-blocked FIR: the same sums in another order
            for i, (w, r) in enumerate(zip(ws, refs)):
                for p in range(len(w.phases)):
                    for t in range(3):
                        ref = r["raw"][p, t]
                        scale = max(np.max(np.abs(ref)), 1e-300)
                        assert np.max(np.abs(got[i]["raw"][p, t] - ref)) <= 1e-13 * scale, (cells, g, bb, i, p, t)
                np.testing.assert_allclose(got[i]["fit"]["par"][..., 3], r["fit"]["par"][..., 3], rtol=0,
                                           atol=1e-4)
    ms = b.stage_ms()
    assert set(ms) == set(cfg["grid"]) and all(v >= 0 for v in ms.values())


def test_batch_instance_vs_oracle(batch_case):
    ws, engines, b, refs = batch_case
    w = ws[0]
    o = run_oracle(w)
    par = refs[0]["fit"]["par"].reshape(len(w.phases) + 1, 3, 5)
    rv = par[:, :, 3] - par[-1:, :, 3]
    np.testing.assert_allclose(rv[:-1], o["rv"], rtol=0, atol=RV_TOL)
    np.testing.assert_allclose(par[:, :, 4], o["par"][:, :, 4], rtol=0, atol=RV_TOL)


def test_batch_rejects_changed_instance():
    ws = synth.small_batch(seed=23, n_inst=2)
    engines = [_prepare(w) for w in ws]
    b = Batch(engines, [w.sigma_inst for w in ws])
    w = ws[1]
    phases = np.arange(len(w.phases) + 3) / (len(w.phases) + 3)
    engines[1].set_activity(w.act, phases, [w.regions[0]] * len(phases))  # reallocates its spectra
    with pytest.raises(SoapError, match="changed"):
        b.step(graph=False)
    with pytest.raises(SoapError, match="changed"):
        b.step(graph=True)
    b.close()
    for e in engines:
        e.close()


def test_batch_rejects_linear_grid():
    w = synth.linear_small()
    e = Engine(0)
    setup_workload(e, w)
    e.quiet_star()
    with pytest.raises(SoapError, match="log-uniform"):
        Batch([e], [w.sigma_inst])
    e.close()
