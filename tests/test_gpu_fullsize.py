"""The headline workload at its full size (2048x1024 cells, 200 particles per cell per
species, 419 M of each), checked through properties that do not depend on the size: every
step's particle bookkeeping closes (pushed = kept + removed, appended particles counted), the
charge on the grid equals the particle count after the periodic fold and the wall doubling are
undone, and sorting (every 4 steps here) changes neither.  Both precisions."""
import numpy as np
import pytest

from edgpu import config

pytestmark = pytest.mark.gpu


def _grid_total(rt, prob, species):
    rho = rt.gather_electron_charge_density() if species == 0 else rt.gather_ion_charge_density()
    R = rho.reshape(prob.ny, prob.nx).copy()
    R[0, :] *= 0.5
    R[-1, :] *= 0.5
    return R[:, 1:prob.gmi].sum()


@pytest.mark.parametrize("precision", [0, 1])
def test_full_size_bookkeeping_and_charge(edgpu, precision):
    prob = config.workload(2)
    rt = edgpu.Runtime(prob, precision=precision, sort_interval=4, ion_sort_interval=4, capacity_factor=1.05)
    try:
        rt.initiate_electron_neutral_collisions()
        rt.initiate_ion_neutral_collisions()
        rt.load_initial_particles()
        ne, ni = rt.count(0), rt.count(1)
        assert ne == ni == prob.n_particles
        for step in range(6):
            se, si = rt.step()
            assert se["n_before"] == ne and si["n_before"] == ni
            assert se["n_after"] == ne - se["n_removed"]
            assert si["n_after"] == ni - si["n_removed"]
            ne, ni = rt.count(0), rt.count(1)
            # what was appended after the push: ionisation products (and nothing else here)
            assert ne - se["n_after"] == ni - si["n_after"] >= 0
        for species, n in ((0, ne), (1, ni)):
            total = _grid_total(rt, prob, species)
            print("precision %d species %d: %d particles, grid charge %.6f" % (precision, species, n, total),
                  flush=True)
            # FP64 weights sum to 1 to rounding; the mixed path's FP32 weights to ~1e-7 each
            assert abs(total - n) < (1e-9 if precision == 0 else 1e-6) * n
    finally:
        rt.close()


def test_combined_case_full_size(edgpu):
    """The combined case at its full size: 4096x2048 cells, 100 particles per cell per species
    (839 M of each), one grounded 256x128 inner box (so the multigrid Krylov solve to 1e-8),
    secondary yield 0.5 on both y walls, the 100 eV beam, e-n collisions and charge exchange,
    mixed precision with electrons sorted every 4 steps.  Each step's bookkeeping closes with
    the emitted electrons and the ionisation products, the box holds its potential exactly and
    the grid charge equals the particle count."""
    prob = config.workload(5)
    print("case cfg5 full size: %d particles/species, grid %dx%d" % (prob.n_particles, prob.gmi, prob.gmj),
          flush=True)
    rt = edgpu.Runtime(prob, precision=1, sort_interval=4, ion_sort_interval=8, capacity_factor=1.03,
                       add_capacity=1 << 22)
    try:
        rt.setup_wall_emission()
        rt.initiate_electron_neutral_collisions()
        rt.initiate_ion_neutral_collisions()
        rt.load_initial_particles()
        ne, ni = rt.count(0), rt.count(1)
        # the electrons placed inside the box are removed before the ions take their positions
        b = prob.inner_boxes[0]
        area = (b[1] - b[0]) * (b[3] - b[2]) / float((prob.gmi - 1) * prob.gmj)
        sigma = (prob.n_particles * area * (1.0 - area)) ** 0.5
        assert ne == ni and abs(ne - prob.n_particles * (1.0 - area)) < 6 * sigma
        nobj = prob.n_boundary_obj + len(prob.inner_boxes)
        for step in range(4):
            se, si = rt.step(krylov_rtol=1e-8, krylov_warm_start=1)
            assert se["n_before"] == ne and si["n_before"] == ni
            assert se["n_after"] == ne - se["n_removed"] and si["n_after"] == ni - si["n_removed"]
            assert sum(se["hit_count"][:nobj]) == se["n_removed"]
            ion = se["collisions"][2]
            ne, ni = rt.count(0), rt.count(1)
            assert ne == se["n_after"] + sum(se["emit_count"][:nobj]) + ion
            assert ni == si["n_after"] + ion
            print("step %d: %d electrons (%d removed, %d emitted, %d ionisations), %d ions" %
                  (step, ne, se["n_removed"], sum(se["emit_count"][:nobj]), ion, ni), flush=True)
        phi = rt.get_grid("phi").reshape(prob.ny, prob.nx)
        b = prob.inner_boxes[0]
        assert np.all(phi[b[2]:b[3] + 1, b[0]:b[1] + 1] == 0.0) and np.abs(phi).max() > 0.0
        for species, n in ((0, ne), (1, ni)):
            total = _grid_total(rt, prob, species)
            print("species %d: %d particles, grid charge %.3f" % (species, n, total), flush=True)
            assert abs(total - n) < 1e-6 * n
    finally:
        rt.close()
