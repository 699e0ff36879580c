"""soap_run_series runs the quiet star on a second stream beside the active phases
(with tap scratch of its own).  Every spectrum goes through the same kernels as
the stage-by-stage calls, so all outputs (quiet star, raw and lower spectra,
min/max, CCFs, fit parameters, models, bisectors) must equal theirs bit for bit,
on repeated runs over the same buffers and with stage timing on (stages in
order)."""
import numpy as np
import pytest

import oracle as O
from soap_gpu_amd import synth
from soap_gpu_amd.engine import Engine, setup_workload
from tests.pipeline_ref import normwise

pytestmark = pytest.mark.gpu


def _outputs(e, w):
    P = len(w.phases)
    raw = np.stack([np.stack(e.get_raw(p)) for p in range(P)])
    frames = list(range(P)) + [-1]
    lower = np.stack([np.stack([e.get_lower(p, t) for t in range(3)]) for p in frames])
    norm = np.stack([np.stack([e.get_norm(p, t) for t in range(3)]) for p in frames])
    return dict(quiet=np.stack([e.get_quiet(0), e.get_quiet(1)]), raw=raw, lower=lower, norm=norm,
                ccf=e.get_ccf(), fit=e.get_fit())


def _same(a, b, what):
    for k in a:
        if isinstance(a[k], dict):
            _same(a[k], b[k], f"{what}.{k}")
        else:
            assert np.array_equal(a[k], b[k]), f"{what}.{k}"


# an odd pixel count puts every odd spectrum off the 16-byte boundary, which
# spec_norm (its own launch when staged, a part of the CCF launch in the series)
# handles with a one-pixel head; spectrum counts 114, 51 and 18 leave a partial
# group of 8 in the CCF; vsini 40 km/s widens the velocity window to +-100 km/s
# (401 velocities: 448-thread blocks, line footprints past the 512-pixel LDS
# segment of the CCF)
@pytest.mark.parametrize("n_phase,n,vsini", [(37, 6000, 6.0), (16, 6000, 6.0), (5, 6001, 6.0), (4, 5000, 40.0)])
def test_series_pipeline_equals_staged_calls(n_phase, n, vsini):
    w = synth.batch_instance(np.random.default_rng(29), n, 64, n_phase, 3, vsini, 70.0, name="series", n_mu=8)
    staged = Engine(0)
    setup_workload(staged, w)
    staged.disk_init(w.star)
    staged.quiet_star()
    staged.integrate()
    staged.resolution(w.sigma_inst)
    staged.ccf()
    staged.fit()
    ref = _outputs(staged, w)
    staged.close()
    e = Engine(0)
    setup_workload(e, w)
    for rep in range(2):  # a second series on the same buffers
        e.run_series(w.star, w.sigma_inst)
        e.sync()
        _same(_outputs(e, w), ref, f"pipelined run {rep}")
    e.set_timing(True)  # stage timing runs the stages in order
    e.run_series(w.star, w.sigma_inst)
    e.sync()
    e.set_timing(False)
    _same(_outputs(e, w), ref, "timed run")
    e.close()
    # and the normalised spectra of phase 0 against the oracle's lower_resolution
    # of the same raw spectra (the odd-length path included)
    qn, qs = ref["quiet"]
    for t in range(3):
        low, _ = O.lower_resolution(O.GRID_LOG, w.wave, w.sigma_inst, qn if t == 1 else qs, ref["raw"][0, t], 1)
        assert normwise(ref["lower"][0, t], low) <= 1e-12, t
