"""Capacity of the fixed-point tap sums (regression test).

At the largest disk grid the ABI accepts (8192, 52.7M disk cells) with one mu bin
and no rotation, every cell's weight lands on one tap: Q_j = (sum_c limb_c) S_j.
Each contribution's low word is below 2^32, so the 64-bit low-word sum of all
cells (about 2^56) cannot wrap; a 40-bit low word would have wrapped past 2^24
contributions and lost 2^-8 per wrap, about 1e-10 of this sum.  The reference
value is the exactly rounded sum (math.fsum) of the engine's own limb map."""
import math

import numpy as np
import pytest

from soap_gpu_amd import synth
from soap_gpu_amd.engine import Engine

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
def test_quiet_star_one_tap_at_grid_8192():
    G = 8192
    wave = synth.log_grid(5000.0, 5010.0, n=2048)
    rng = np.random.default_rng(8192)
    row = rng.uniform(0.2, 1.0, (1, wave.size))
    eng = Engine(0)
    try:
        eng.set_wave(wave)
        eng.set_mu_bounds(np.zeros(0))
        eng.set_cubes({"q": row, "qsed": 2.0 * row})
        star = dict(prot=1e30, w_ratio=0.0, incl=90.0, radius_km=synth.R_SUN_KM, u1=0.29, u2=0.34, grid=G,
                    lb=np.zeros(4))
        eng.disk_init(star)
        eng.quiet_star()
        q0, q1 = eng.get_quiet(0), eng.get_quiet(1)
        maps = eng.get_maps()
    finally:
        eng.close()
    on = maps["mu"] >= 0.0
    n_on = int(on.sum())
    assert n_on > 2 ** 25  # past the 40-bit low word's 2^24 contributions
    total = math.fsum(maps["limb"][on].tolist())
    for q, scale in ((q0, 1.0), (q1, 2.0)):
        np.testing.assert_allclose(q / (scale * row[0]), total, rtol=1e-14, atol=0)
