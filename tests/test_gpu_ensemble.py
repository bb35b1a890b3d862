"""Multi-step statistical parity: 8-seed ensembles of the GPU production path (FP64 and
mixed precision) against 8-seed ensembles of the oracle in the reference's form, on the
plain FP64 reference case reduced to 32x32 cells (50 particles/cell per species, e-n MCC
and RCX, grounded walls) over 60 steps, and at its full size (256x256 cells, 3.3 M
particles per species) over 40 steps.  Compared per observable: density profiles of both
species and the electron mean-energy profile across the gap (8 bins), the electron energy
distribution (10 bins), wall fluxes of both species at both walls and the collision
counts of every process; the tolerance is 4 standard errors of the difference plus a
small floor for bins that are nearly deterministic."""
import time
from concurrent.futures import ThreadPoolExecutor

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


FULL_STEPS = 40


@pytest.fixture(scope="module")
def oracle_ensemble_full(oracle):
    """The oracle's 8 runs at the full reference size, one thread each (the C routines
    release the interpreter while they run)."""
    prob = config.workload(1)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(len(SEEDS)) as ex:
        runs = list(ex.map(lambda s: E.oracle_run(oracle, prob, s, FULL_STEPS), SEEDS))
    print("oracle ensemble at %dx%d, %d particles per species: %.1f s" %
          (prob.gmi - 1, prob.gmj, prob.n_particles, time.perf_counter() - t0), flush=True)
    return prob, E.ensemble(runs)


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_full_size_reference_case(edgpu, oracle_ensemble_full, precision):
    """The same comparison at the reference case's own size: eight times smaller standard
    errors than at 32x32, so the same 4-sigma bound is eight times tighter."""
    prob, ref = oracle_ensemble_full
    print("case full-size ensemble, precision %d: %d particles per species, %d steps" %
          (precision, prob.n_particles, FULL_STEPS), flush=True)
    runs = [E.gpu_run(edgpu, prob, 2000 + s, FULL_STEPS, precision, sort_interval=12) for s in range(len(SEEDS))]
    got = E.ensemble(runs)
    for k in ("dens_e", "dens_i", "energy", "flux_e", "flux_i", "coll"):
        print("%-7s gpu %s" % (k, np.array2string(got[k][0], precision=5, max_line_width=200)), flush=True)
        print("%-7s cpu %s" % ("", np.array2string(ref[k][0], precision=5, max_line_width=200)), flush=True)
    bad = E.compare(got, ref, len(SEEDS), len(SEEDS), floors={"dens_e": 5e-4, "dens_i": 5e-4, "energy": 1e-3,
                                                               "eedf": 2e-4})
    assert not bad, bad
