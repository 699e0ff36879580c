"""The region renderer of the synthetic fitting workload (synth.region_maps /
region_owner: discrete spots and faculae as filling-factor maps) against the
oracle's region maps (orc_phase_maps: the reference's boundary, bounding box and
per-cell inverse rotation, SOAP_integration.cu:21-202,326-390) on the same disk
cells, phases and inclinations: the same owner type at every cell, up to a few
cells on region edges at low inclination (a cap test in the observer frame against
the reference's bounding-box scan)."""
import numpy as np
import pytest

import oracle as O
from soap_gpu_amd import synth

REGIONS = [(20.0, 10.0, 0.25, 1.0), (25.0, 15.0, 0.15, 0.0), (-10.0, 200.0, 0.2, 0.0)]


@pytest.mark.parametrize("incl,max_diff", [(90.0, 0), (60.0, 0), (30.0, 4)])
def test_region_owner_matches_oracle_maps(incl, max_diff):
    G = 120
    iy, iz = np.meshgrid(np.arange(G), np.arange(G), indexing="ij")
    Y, Z = -1 + 2 * iy / G, -1 + 2 * iz / G
    types = np.array([r[3] for r in REGIONS])
    for ph in (0.0, 0.1, 0.37, 0.8):
        _, bi = O.phase_maps(G, incl, 20, 5778, 663, 250, ph, REGIONS)
        own = synth.region_owner(REGIONS, ph, incl, Y, Z).ravel()
        typ = np.where(own >= 0, types[np.maximum(own, 0)], -9999.0)
        assert (bi != -9999.0).sum() > 100  # something is on the disk
        assert int(np.sum(typ != bi)) <= max_diff, (incl, ph)


def test_region_maps_last_region_wins():
    """A spot inside a facula: its nodes are spot (1 in the spot map, 0 in the
    facula map), the rest of the facula stays facula."""
    sp, fa = synth.region_maps([(0.0, 0.0, 0.3, 1.0), (0.0, 0.0, 0.1, 0.0)], [0.0], 90.0, 101)
    assert sp[0, 50, 50] == 1.0 and fa[0, 50, 50] == 0.0
    assert sp[0, 50, 62] == 0.0 and fa[0, 50, 62] == 1.0
    assert np.all(sp + fa <= 1.0)
