"""Edge cases and error paths through the C ABI, all forced from the host: empty species, an
upload beyond the capacity, an add list too small for the step's emission, a Krylov solve
stopped before its tolerance, and calls out of order.  Each reports its status (the
reference aborts on them) and no kernel runs past a buffer."""
import numpy as np
import pytest

from edgpu import config

pytestmark = pytest.mark.gpu


def test_steps_with_empty_species(edgpu):
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, sort_interval=4)
    try:
        rt.initiate_electron_neutral_collisions()
        rt.initiate_ion_neutral_collisions()
        empty = [np.zeros(0) for _ in range(5)]
        rt.upload(0, *empty)
        rt.upload(1, *empty)
        for _ in range(3):
            se, si = rt.step()
            assert se["n_before"] == se["n_after"] == 0 and si["n_before"] == 0
        assert rt.count(0) == 0 and rt.count(1) == 0
        assert not rt.gather_electron_charge_density().any()
        assert rt.vz_bound(1) == 0.0
    finally:
        rt.close()


def test_upload_beyond_capacity(edgpu):
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, capacity_factor=1.0)  # room for exactly n_particles
    try:
        n = prob.n_particles
        arrs = [np.full(n + 1, 1.5) for _ in range(5)]
        with pytest.raises(edgpu.EdgpuError) as e:
            rt.upload(0, *arrs)
        assert e.value.code == 6
        rt.upload(0, *[a[:n] for a in arrs])  # the context stays usable
        assert rt.count(0) == n
    finally:
        rt.close()


def test_add_list_overflow_is_reported(edgpu):
    """A beam of 100 electrons per step into an add list of 16: the emission kernel stops
    writing at the capacity and the step reports it."""
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, add_capacity=16)
    try:
        rt.setup_constant_emission(1, 100.0, model=1, We_beam_eV=100.0)
        rt.load_initial_particles()
        with pytest.raises(edgpu.EdgpuError) as e:
            rt.step()
        assert e.value.code == 6
    finally:
        rt.close()


def test_krylov_stopped_before_tolerance(edgpu):
    prob = config.workload(4, 16)  # walled box with two biased inner boxes
    rt = edgpu.Runtime(prob)
    try:
        rt.load_initial_particles()
        rt.gather_ion_charge_density(host=False)
        rt.gather_electron_charge_density(host=False)
        rt.set_krylov_preconditioner(0)
        with pytest.raises(edgpu.EdgpuError) as e:
            rt.solve_potential_krylov(rtol=1e-12, max_iter=2)
        assert e.value.code == 9
        _, it, rr = rt.solve_potential_krylov(rtol=1e-8)
        assert rr <= 1e-8 and it > 2
    finally:
        rt.close()


def test_calls_out_of_order(edgpu):
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, sort_interval=0)
    try:
        with pytest.raises(edgpu.EdgpuError) as e:
            rt.sort_particles(0)  # no spare buffers without a sort interval
        assert e.value.code == 8
        with pytest.raises(edgpu.EdgpuError) as e:
            rt.cnn_record_pair()  # no network
        assert e.value.code == 8
        with pytest.raises(edgpu.EdgpuError) as e:
            rt.set_rng_state(1, 0, n_sub=1)  # one sub-cycle: n_sub is 0
        assert e.value.code == 1
    finally:
        rt.close()
