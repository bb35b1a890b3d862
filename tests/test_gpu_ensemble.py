"""Multi-step statistical parity: 8-seed ensembles of the GPU production path (FP64 and
mixed precision) against 8-seed ensembles of the oracle in the reference's form, on the
plain FP64 reference case reduced to 32x32 cells (50 particles/cell per species, e-n MCC
and RCX, grounded walls), 60 steps.  Compared per observable: density profiles of both
species and the electron mean-energy profile across the gap (8 bins), the electron energy
distribution (10 bins), wall fluxes of both species at both walls and the collision
counts of every process; the tolerance is 4 standard errors of the difference plus a
small floor for bins that are nearly deterministic."""
import numpy as np
import pytest

import ensemble as E
from edgpu import config

pytestmark = pytest.mark.gpu

SEEDS = [123456789, 7, 11, 13, 17, 19, 23, 29]
STEPS = 60


@pytest.fixture(scope="module")
def oracle_ensemble(oracle):
    prob = config.workload(1, 8)
    runs = [E.oracle_run(oracle, prob, s, STEPS) for s in SEEDS]
    return prob, E.ensemble(runs)


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_moments_match_oracle(edgpu, oracle_ensemble, precision):
    prob, ref = oracle_ensemble
    runs = [E.gpu_run(edgpu, prob, 1000 + s, STEPS, precision) for s in range(len(SEEDS))]
    got = E.ensemble(runs)
    for k in ("dens_e", "dens_i", "energy", "flux_e", "flux_i", "coll"):
        print("%-7s gpu %s" % (k, np.array2string(got[k][0], precision=4, max_line_width=200)), flush=True)
        print("%-7s cpu %s" % ("", np.array2string(ref[k][0], precision=4, max_line_width=200)), flush=True)
    bad = E.compare(got, ref, len(SEEDS), len(SEEDS), floors={"dens_e": 2e-3, "dens_i": 2e-3, "energy": 5e-3,
                                                               "eedf": 1e-3})
    assert not bad, bad
