"""Parity of the linear-grid path (the reference's native 0.005 A seed format)
with the FP64 oracle: gather-form shift + interpolation for the quiet star and
the active deltas (SOAP_integration.cu:232-281, gather definition of
i(j) = max{i : ceil(p_i) <= j}), per-pixel-width conv_inst
(SOAP_resolution.cu:76-95) and the literal linear mask CCF (CCF_func.py:45-91).

Bars: spectra <= 1e-12 norm-wise on the interior, RV/FWHM/BIS within 1 cm/s.
The reference's lambda-difference interpolation weight carries ~1e-10 relative
rounding on a 0.005 A grid near 5000 A; the engine evaluates the same
expressions from the same wavelength values."""
import numpy as np
import pytest

import oracle as O
from soap_gpu_amd import synth
from soap_gpu_amd.engine import BCONV, setup_workload
from tests.pipeline_ref import normwise, oracle_maps, run_oracle
from tests.test_gpu_parity import RV_TOL, SPEC_TOL, run_engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def linear_case(engine):
    w = synth.linear_small()
    return w, run_engine(engine, w), run_oracle(w)


def test_linear_maps_and_quiet(linear_case):
    w, g, o = linear_case
    for k in ("mu", "delta", "limb", "lb"):
        assert np.array_equal(g["maps"][k], o["maps"][k]), k
    for q in range(2):
        assert normwise(g["quiet"][q], o["quiet"][q]) <= SPEC_TOL, q


def test_linear_integrate(linear_case):
    w, g, o = linear_case
    for p in range(len(w.phases)):
        for t in range(3):
            ref = o["raw"][p, t]
            if np.max(np.abs(ref)) == 0:
                assert np.max(np.abs(g["raw"][p, t])) == 0
                continue
            assert normwise(g["raw"][p, t], ref) <= SPEC_TOL, (p, t)


def test_linear_resolution_isolated(linear_case):
    w, g, o = linear_case
    qn, qs = g["quiet"]
    for p in range(len(w.phases)):
        for t in range(3):
            ref, _ = O.lower_resolution(O.GRID_LINEAR, w.wave, w.sigma_inst, qn if t == BCONV else qs,
                                        g["raw"][p, t], 1)
            assert normwise(g["lower"][p, t], ref) <= SPEC_TOL, (p, t)
    for t, q in ((0, qs), (1, qn), (2, qs)):
        ref, _ = O.lower_resolution(O.GRID_LINEAR, w.wave, w.sigma_inst, None, q, 0, weight=0.005)
        assert normwise(g["lower"][-1, t], ref) <= SPEC_TOL, t


def test_linear_end_to_end_rv(linear_case):
    w, g, o = linear_case
    fit = g["fit"]
    par = fit["par"].reshape(len(w.phases) + 1, 3, 5)
    rv = par[:, :, 3] - par[-1:, :, 3]
    np.testing.assert_allclose(rv[:-1], o["rv"], rtol=0, atol=RV_TOL)
    np.testing.assert_allclose(par[:, :, 4], o["par"][:, :, 4], rtol=0, atol=RV_TOL)
    np.testing.assert_allclose(par[:, :, 2], o["par"][:, :, 2], rtol=0, atol=RV_TOL)
    np.testing.assert_allclose(g["ccf"].reshape(o["ccf"].shape), o["ccf"], rtol=0, atol=1e-9)
    assert np.ptp(rv[:-1, 2]) > 0.05  # the active regions move the RV


def test_linear_solar_mode(engine):
    w = synth.linear_small(grid=40)
    setup_workload(engine, w, with_mask=False)
    engine.set_activity(w.act, w.phases, w.regions, solar_mode=1)
    engine.integrate()
    maps = oracle_maps(w)
    a = w.act
    for p in range(len(w.phases)):
        temp, bi = O.phase_maps(40, w.star["incl"], a["nrho"], a["t_star"], a["t_diff_spot"], a["t_diff_fac"],
                                w.phases[p], w.regions[p])
        ref = O.integrate_phase(O.GRID_LINEAR, w.wave, maps, temp, bi, w.cubes, 40, a["t_star"], a["t_diff_fac"],
                                solar_mode=1, step=0.005)
        got = engine.get_raw(p)
        for t in range(3):
            if np.max(np.abs(ref[t])) == 0:
                continue
            assert normwise(got[t], ref[t]) <= SPEC_TOL, (p, t)


def test_linear_series_equals_staged(linear_case):
    """soap_run_series on a linear grid (spec_norm and the CCF in one launch) against
    the stage-by-stage outputs of the same workload, bit for bit."""
    from soap_gpu_amd.engine import Engine
    w, g, o = linear_case
    e = Engine(0)
    try:
        setup_workload(e, w)
        e.run_series(w.star, w.sigma_inst)
        P = len(w.phases)
        lower = np.stack([np.stack([e.get_lower(p, t) for t in range(3)]) for p in list(range(P)) + [-1]])
        np.testing.assert_array_equal(lower, g["lower"])
        np.testing.assert_array_equal(e.get_ccf(), g["ccf"])
        fit = e.get_fit()
        for k in fit:
            np.testing.assert_array_equal(fit[k], g["fit"][k])
    finally:
        e.close()


def _linear_parity(engine, w):
    g, o = run_engine(engine, w), run_oracle(w)
    for q in range(2):
        assert normwise(g["quiet"][q], o["quiet"][q]) <= SPEC_TOL, q
    for p in range(len(w.phases)):
        for t in range(3):
            ref = o["raw"][p, t]
            if np.max(np.abs(ref)) == 0:
                assert np.max(np.abs(g["raw"][p, t])) == 0
                continue
            assert normwise(g["raw"][p, t], ref) <= SPEC_TOL, (p, t)
    par = g["fit"]["par"].reshape(len(w.phases) + 1, 3, 5)
    np.testing.assert_allclose(par[:-1, :, 3] - par[-1:, :, 3], o["rv"], rtol=0, atol=RV_TOL)
    np.testing.assert_allclose(par[:, :, 4], o["par"][:, :, 4], rtol=0, atol=RV_TOL)


def test_linear_grid300_shift_groups(engine):
    """A 300 x 300 disk: runs of equal shift along whole disk columns (up to ~300
    cells per bin segment) gathered as one group each, against the oracle's
    per-cell sums."""
    _linear_parity(engine, synth.linear_small(seed=12, grid=300, n=3000, n_mu=10, n_phase=2))


def test_linear_differential_rotation(engine):
    """Differential rotation: the shift varies along a column, so most groups are
    single cells and the run boundaries fall inside columns."""
    import dataclasses
    w = synth.linear_small(seed=13, grid=64, n=3000, n_phase=2)
    w = dataclasses.replace(w, star=synth._star(64, 3.0, incl=70.0, w_ratio=0.3))
    _linear_parity(engine, w)


def test_linear_fast_rotator(engine):
    """vsini 150 km/s: shifts of about 500 pixels widen the staged window past what
    the column form of the quiet star keeps in LDS, so the quiet star takes the
    per-group form under solid-body rotation too."""
    _linear_parity(engine, synth.linear_small(seed=14, grid=64, n=4000, n_phase=2, vsini=150.0))


def test_linear_production_disk(engine):
    """The production disk and activity of the linear series (grid 1000, 20 mu bins,
    2 spots + 2 faculae overlapping, vsini 2 km/s: 785,343 disk cells in 15,898
    column groups, the column form of the quiet star) over a 12 A stretch of its
    0.005 A grid, three phases, against the oracle's per-cell sums."""
    import dataclasses
    w = synth.linear_series(n_phase=25)
    wave = w.wave[:2400]
    sel = (w.lines["center"] > wave[0] + 1.0) & (w.lines["center"] < wave[-1] - 1.0)
    lines = {k: v[sel] for k, v in w.lines.items()}
    w = dataclasses.replace(w, wave=wave, cubes=synth.make_cubes(wave, lines, 20), lines=lines,
                            mask_l=lines["center"].copy(), mask_w=lines["depth"].copy(),
                            phases=w.phases[[0, 5, 12]], regions=[w.regions[i] for i in (0, 5, 12)])
    _linear_parity(engine, w)
