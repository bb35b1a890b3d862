"""A run started from the reference's own sample deck (periodic system, electron beam from
part of the wall below, 11 electron sub-cycles per ion cycle): beam electrons every step,
ions, collisions and the accumulated-field reset on the ion steps only."""
import os

import numpy as np
import pytest

from edgpu import deck as D

pytestmark = pytest.mark.gpu

EBEAM = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "deck_periodic_ebeam")


def test_sample_ebeam_deck_runs_with_subcycles(edgpu):
    dk = D.read_deck(EBEAM)
    prob = dk.problem()
    rt = edgpu.Runtime(prob, seed=prob.seed, sort_interval=4)
    try:
        dk.apply(rt)
        rt.load_initial_particles()
        beam, ion_steps, coll = [], [], []
        for step in range(22):
            se, si = rt.step()
            beam.append(se["emit_count"][4])
            ion_steps.append(si["n_before"] > 0)
            coll.append(sum(se["collisions"][:3]))
        print("beam per step", beam[:3], "ion steps", [k for k, v in enumerate(ion_steps) if v],
              "e-n events", coll[10], coll[21], flush=True)
        assert all(b == 10 for b in beam)
        assert [k for k, v in enumerate(ion_steps) if v] == [10, 21]
        assert all(c == 0 for k, c in enumerate(coll) if k not in (10, 21))
        assert coll[10] > 0 and coll[21] > 0
        assert rt.count(0) > 0.99 * prob.n_particles
    finally:
        rt.close()
