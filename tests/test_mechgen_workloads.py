"""The seeded workload mechanisms: the committed files under
tests/golden/mech are what the generator writes (so the GPU tests, the bench
and the oracle read the same mechanism), every reaction is element-balanced,
and the 2,000-species mechanism has the shape of the large workload (8,000
reactions, at most two participants per side, a connected network reachable
from the initial majors)."""
import os

import numpy as np
import pytest

from conftest import MECHS


@pytest.mark.parametrize("wl", ["10sp", "53sp", "160sp", "2000sp"])
def test_committed_mechanism_is_the_generator_output(wl):
    from zerork_hip import mechgen
    m = mechgen.workload_mechanism(wl)
    with open(os.path.join(MECHS, f"synth_{wl}.mech")) as f:
        assert f.read() == m.mech_text()
    with open(os.path.join(MECHS, f"synth_{wl}.therm")) as f:
        assert f.read() == m.therm_text()


def test_large_mechanism_shape():
    from zerork_hip import mechgen
    m = mechgen.workload_mechanism("2000sp")
    assert len(m.species) == 2000 and len(m.reactions) == 8000
    comp = {s: c for s, c in zip(m.species, m.comp)}
    reach = set(m.majors)
    for r in m.reactions:
        assert 1 <= len(r.reac) <= 2 and 1 <= len(r.prod) <= 2
        tot = lambda side: {e: sum(comp[s].get(e, 0) for s in side) for e in m.elements}
        assert tot(r.reac) == tot(r.prod)
        reach.update(r.reac)
        reach.update(r.prod)
    # most species (1,746 of 2,000) take part in some reaction
    assert len(reach) >= 0.85 * len(m.species)
    kinds = [r.kind for r in m.reactions]
    assert kinds.count("falloff") == 800 and kinds.count("tb") == 2000
