"""Edge cases and error paths through the C ABI, all forced from the host: empty species,
arrays that grow past their initial capacity, particle records in the reference's layout, an
add list too small for the step's emission, a Krylov solve
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


def test_arrays_grow_beyond_capacity(edgpu):
    """An upload beyond the capacity and appends beyond it grow the device arrays (as
    PROCESS_ADDED_ELECTRONS grows electron(:), pic2d_ElectronDynamics.f90:1458-1482): particles
    keep their values, and sorts and steps go on with the charge on the grid equal to the count."""
    prob = config.workload(1, 8)
    n = prob.n_particles
    rt = edgpu.Runtime(prob, capacity_factor=1.0)  # room for exactly n_particles
    try:
        rng = np.random.default_rng(3)
        arrs = [rng.uniform(1.0, 30.0, n + 1), rng.uniform(1.0, 30.0, n + 1)] + \
            [rng.normal(0.0, 0.05, n + 1) for _ in range(3)]
        rt.upload(0, *arrs)
        got = rt.download(0)
        assert all(np.array_equal(a, b) for a, b in zip(arrs, got[:5]))
    finally:
        rt.close()
    rt = edgpu.Runtime(prob, capacity_factor=1.0, sort_interval=2)
    try:
        rt.load_initial_particles()
        # a beam of 500 electrons per step into full arrays: the first append grows them
        rt.setup_constant_emission(3, 500.0, model=1, Te_normal_eV=1.0, Te_parallel_eV=0.1, We_beam_eV=100.0)
        for _ in range(5):
            se, _ = rt.step()
            assert rt.count(0) == se["n_after"] + sum(se["emit_count"])
        ne = rt.count(0)
        print("electrons %d in arrays sized for %d at init" % (ne, n), flush=True)
        assert ne > n
        R = rt.gather_electron_charge_density().reshape(prob.ny, prob.nx).copy()
        R[0, :] *= 0.5
        R[-1, :] *= 0.5
        assert abs(R[:, 1:prob.gmi].sum() - ne) < 1e-9 * ne
    finally:
        rt.close()


def test_particle_records_round_trip(edgpu):
    """edgpu_particles_upload_aos / _download_aos take the reference's array of TYPE particle
    (48-byte records) directly; a 56-byte record with extra fields works as well."""
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob)
    try:
        n = 10001
        rng = np.random.default_rng(5)
        rec = np.zeros(n, rt.PARTICLE)
        rec["X"], rec["Y"] = rng.uniform(0.0, 32.0, n), rng.uniform(0.0, 32.0, n)
        for f in ("VX", "VY", "VZ"):
            rec[f] = rng.normal(0.0, 0.1, n)
        rec["tag"] = rng.integers(0, 1000, n)
        rt.upload_aos(0, rec)
        soa = rt.download(0)
        for q, f in enumerate(("X", "Y", "VX", "VY", "VZ", "tag")):
            assert np.array_equal(soa[q], rec[f])
        back = rt.download_aos(0)
        assert np.array_equal(back[["X", "Y", "VX", "VY", "VZ", "tag"]], rec[["X", "Y", "VX", "VY", "VZ", "tag"]])
        wide = np.dtype([("X", "f8"), ("Y", "f8"), ("VX", "f8"), ("VY", "f8"), ("VZ", "f8"), ("tag", "i4"),
                         ("extra", "i4"), ("weight", "f8")])
        w = np.zeros(n, wide)
        for f in ("X", "Y", "VX", "VY", "VZ", "tag"):
            w[f] = rec[f]
        w["weight"] = -1.0
        rt.upload_aos(1, w)
        assert np.array_equal(rt.download(1)[0], rec["X"])
        w2 = rt.download_aos(1, wide)
        assert np.array_equal(w2["VZ"], rec["VZ"]) and np.array_equal(w2["tag"], rec["tag"])
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
