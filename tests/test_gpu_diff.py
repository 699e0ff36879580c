"""The differentiable path on the GPU (soap_diff_*, soap_gpu_amd.diff) against the
FP64 oracle's model (orc_diff_forward), with and without the facula population:
forward parity (spectra, CCF, RV, FWHM and the bisector span), and every gradient
(vsini, u1, u2, spot and facula contrasts, quiet / spot / facula seed rows, spot
and facula filling-factor map pixels) against central finite differences of the
oracle's loss with h = 1e-6 relative, for the full loss and for the bisector-span
term alone."""
import numpy as np
import pytest
import torch

import oracle as O
from soap_gpu_amd import synth
from soap_gpu_amd.diff import DiffModel, SoapLoss
from tests.pipeline_ref import normwise

pytestmark = pytest.mark.gpu

TRIM = 210
P, M = 3, 10


LB = np.array([0.23161761, -0.11674743, -0.36866954, 0.25537337])  # LB(5778) of the reference's table


@pytest.fixture(scope="module", params=["spots", "spots+faculae"])
def problem(engine, request):
    fac = request.param == "spots+faculae"
    w = synth.small(grid=24, n=2600, n_mu=4)
    rng = np.random.default_rng(21)
    maps = np.clip(rng.uniform(-0.3, 0.7, (P, M, M)), 0.0, 1.0)
    fmaps = np.clip(rng.uniform(-0.4, 0.5, (P, M, M)), 0.0, 1.0) if fac else None
    wave_t = w.wave[TRIM:w.n - TRIM]
    sl, sw = O.select_lines(wave_t, w.mask_l, w.mask_w)
    q, s, f = w.cubes["q"], w.cubes["sp"], w.cubes["fac"] if fac else None
    params = np.array([3.0, 0.29, 0.34, 0.3] + ([1.15] if fac else []))
    z = np.zeros((P, w.n))

    def kw(prm, ff=f, fm=fmaps):
        return dict(kappa_f=prm[4], lb=LB, cube_f=ff, fmaps=fm) if fac else {}

    # targets: the model at other parameters, plus offsets on RV and FWHM
    tp = [3.2, 0.31, 0.3, 0.35] + ([1.2] if fac else [])
    tgt = O.diff_forward(w.wave, 24, w.mu_bounds, *tp[:4], q, s, maps * 0.8, w.sigma_inst, TRIM, w.vgrid, sl, sw, z,
                         np.zeros(P), np.zeros(P), 1.0, 0.0, 0.0, **kw(tp, fm=None if fmaps is None else fmaps * 0.7))
    T = tgt["F"] + 1e-4 * rng.standard_normal((P, w.n)) * np.abs(tgt["F"]).max()
    rv_t = tgt["theta"][:, 2] + rng.uniform(-20, 20, P)
    fw_t = tgt["theta"][:, 3] + rng.uniform(-50, 50, P)
    bis_t = tgt["span"] + rng.uniform(-20, 20, P)
    wts = (1.0 / np.sum(T[:, TRIM:w.n - TRIM] ** 2), 1e-3, 1e-5, 1e-3)

    def oracle_loss(prm, qq=q, ss=s, mm=maps, ff=f, fm=fmaps, weights=wts):
        return O.diff_forward(w.wave, 24, w.mu_bounds, *prm[:4], qq, ss, mm, w.sigma_inst, TRIM, w.vgrid, sl, sw, T,
                              rv_t, fw_t, *weights[:3], bis_t=bis_t, w_bis=weights[3], **kw(prm, ff, fm))

    model = DiffModel(engine, w.wave, w.mu_bounds, sl, sw, w.vgrid, 24, P, M, w.sigma_inst, TRIM,
                      lb=LB if fac else None)
    dev = torch.device("cuda")
    t_ = lambda a: None if a is None else torch.tensor(a, dtype=torch.float64, device=dev)  # noqa: E731
    return dict(w=w, fac=fac, maps=maps, fmaps=fmaps, q=q, s=s, f=f, params=params, T=T, rv_t=rv_t, fw_t=fw_t,
                bis_t=bis_t, wts=wts, oracle=oracle_loss, model=model, tq=t_(q), ts=t_(s), tf=t_(f), tm=t_(maps),
                tfm=t_(fmaps), tT=t_(T))


# the full loss, and the bisector-span term alone (its adjoint is otherwise a small
# part of the total)
WEIGHTS = {"all": None, "bis": (0.0, 0.0, 0.0, 1.0)}
# relative finite-difference steps: 1e-6 for the full loss; 1e-5 for the span term
# alone, whose value moves with the rounding of the polished fit optimum (about
# 1e-11 relative in the centre), enough to shift a 1e-6-step central difference
# of this term by ~3e-5 relative (steps from 1e-4 to 1e-5 agree to 5e-6)
HREL = {"all": 1e-6, "bis": 1e-5}


def _gpu(pb, params=None, weights=None):
    m = pb["model"]
    w = pb["wts"] if weights is None else weights
    return m.forward(pb["params"] if params is None else params, pb["tq"], pb["ts"], pb["tm"], pb["tT"],
                     pb["rv_t"], pb["fw_t"], *w[:3], bis_t=pb["bis_t"], w_bis=w[3], f=pb["tf"], fmaps=pb["tfm"])


def test_diff_forward_parity(problem):
    pb = problem
    loss, rv, fw, bis = _gpu(pb)
    ref = pb["oracle"](pb["params"])
    assert loss == pytest.approx(ref["loss"], rel=1e-10)
    F = pb["model"].get("F")
    for p in range(P):
        assert normwise(F[p], ref["F"][p]) <= 1e-12
    np.testing.assert_allclose(rv, ref["theta"][:, 2], rtol=0, atol=1e-6)
    np.testing.assert_allclose(fw, ref["theta"][:, 3], rtol=0, atol=1e-6)
    np.testing.assert_allclose(bis, ref["span"], rtol=0, atol=1e-6)
    np.testing.assert_array_equal(pb["model"].get("bis"), bis)
    np.testing.assert_allclose(pb["model"].get("ccf"), ref["ccf"], rtol=1e-11)


def _fd(f, h):
    return (f(h) - f(-h)) / (2 * h)


@pytest.mark.parametrize("terms", sorted(WEIGHTS))
def test_diff_gradients_scalars_vs_oracle_fd(problem, terms):
    pb = problem
    wts = WEIGHTS[terms] or pb["wts"]
    _gpu(pb, weights=wts)
    g5, gq, gs, gm, gf, gfm = pb["model"].backward()
    p0 = pb["params"]
    np_ = len(p0)
    if not pb["fac"]:
        assert g5[4] == 0.0
    for i, name in enumerate(("vsini", "u1", "u2", "contrast", "contrast_fac")[:np_]):
        h = HREL[terms] * abs(p0[i])
        fd = _fd(lambda e: pb["oracle"](p0 + e * np.eye(np_)[i], weights=wts)["loss"], h)
        assert g5[i] == pytest.approx(fd, rel=2e-5, abs=1e-9 * np.abs(g5).max()), name


@pytest.mark.parametrize("terms", sorted(WEIGHTS))
@pytest.mark.parametrize("which", ["q", "s", "maps", "f", "fmaps"])
def test_diff_gradients_arrays_vs_oracle_fd(problem, which, terms):
    """Directional derivatives along two random directions of the seed rows and of
    the maps (interior pixels, where the trimmed loss sees them)."""
    pb = problem
    if which in ("f", "fmaps") and not pb["fac"]:
        pytest.skip("the model without faculae has no facula inputs")
    wts = WEIGHTS[terms] or pb["wts"]
    _gpu(pb, weights=wts)
    g5, gq, gs, gm, gf, gfm = pb["model"].backward()
    g = {"q": gq, "s": gs, "maps": gm, "f": gf, "fmaps": gfm}[which].cpu().numpy()
    base = pb[which]
    key = {"q": "qq", "s": "ss", "maps": "mm", "f": "ff", "fmaps": "fm"}[which]
    rng = np.random.default_rng(5)
    for _ in range(2):
        v = rng.standard_normal(base.shape)
        h = HREL[terms] * np.linalg.norm(base) / np.linalg.norm(v)

        def fn(e):
            return pb["oracle"](pb["params"], weights=wts, **{key: base + e * v})["loss"]

        fd = _fd(fn, h)
        assert np.sum(g * v) == pytest.approx(fd, rel=2e-5, abs=1e-12), which


def test_soap_loss_autograd(problem):
    pb = problem
    params = torch.tensor(pb["params"], dtype=torch.float64, requires_grad=True)
    q = pb["tq"].clone().requires_grad_(True)
    s = pb["ts"].clone().requires_grad_(True)
    mp = pb["tm"].clone().requires_grad_(True)
    f = pb["tf"].clone().requires_grad_(True) if pb["fac"] else None
    fm = pb["tfm"].clone().requires_grad_(True) if pb["fac"] else None
    loss = SoapLoss.apply(pb["model"], params, q, s, mp, pb["tT"], pb["rv_t"], pb["fw_t"], pb["wts"], pb["bis_t"],
                          f, fm)
    (2.0 * loss).backward()
    ref = pb["oracle"](pb["params"])
    assert loss.item() == pytest.approx(ref["loss"], rel=1e-10)
    _gpu(pb)
    g5, gq, gs, gm, gf, gfm = pb["model"].backward()
    # the adjoint sums with floating-point atomics: runs agree to rounding
    np.testing.assert_allclose(params.grad.numpy(), 2 * g5[:len(pb["params"])], rtol=1e-9)
    pairs = [(q.grad, gq), (s.grad, gs), (mp.grad, gm)] + ([(f.grad, gf), (fm.grad, gfm)] if pb["fac"] else [])
    for a, b in pairs:
        torch.testing.assert_close(a, 2 * b, rtol=1e-9, atol=1e-12 * float(b.abs().max()))
