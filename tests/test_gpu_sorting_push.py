"""The sorting push: the electron push of a sort step writes the particles to the spare arrays
in the tile order of their positions before the push (edgpu_sort_with_next_push for callers
that run the phases themselves, edgpu_step when a sort is due), so the sort makes no pass over
the particles of its own.

- The particles it leaves, the removed set and the hit counts equal those of the plain push
  after a separate sort (same field, so every particle's push is the same arithmetic), bit for
  bit in both precisions, and in FP64 equal the oracle's push (orc_advance, itself equal to the
  reference's ADVANCE_ELECTRONS); the deposits agree to 1e-12.
- The order: the particles follow their tiles before the push, except where the compaction
  moved a particle from the end into a removed particle's slot.
- The fused e-n collision path behind it (candidacy in the push keyed by the new slot, the
  selection and kinematics kernels on the new arrays, the products appended): with the same
  uniform words in every slot each particle's outcome depends on the particle alone, and the
  electrons and ions it leaves equal those of the plain path.
- A pending count is dropped by a call that changes the electrons; steps with the sorting push
  keep their bookkeeping."""
import numpy as np
import pytest

from edgpu import config
from helpers import oracle_particles, rel_err, sorted_rows, synthetic_field

pytestmark = pytest.mark.gpu

TILE_W = 16


def _tile_keys(prob, x, y):
    i = np.clip(np.floor(x).astype(np.int64), 0, prob.gmi - 1)
    j = np.clip(np.floor(y).astype(np.int64), 0, prob.gmj - 1)
    ntx = (prob.gmi + TILE_W - 1) // TILE_W
    return i // TILE_W + j * ntx


def _electrons(O, prob):
    e, _ = oracle_particles(O, prob)
    if prob.inner_boxes:
        inside = np.zeros(e[0].size, bool)
        for (i0, i1, j0, j1) in prob.inner_boxes:
            inside |= (e[0] > i0) & (e[0] < i1) & (e[1] > j0) & (e[1] < j1)
        e = [np.ascontiguousarray(a[~inside]) for a in e]
    return e


def _by_tag(d):
    o = np.argsort(d[5])
    return [a[o] for a in d]


SORT_STEPS = (2, 3)  # after two plain pushes (drifted tiles), and right after a sorting push


def _run(rt, prob, e, sorting, EX, EY):
    """Four electron pushes from the same start; before pushes 2 and 3 the electrons are sorted,
    by the push itself (sorting) or by a separate sort.  Returns per push: the particles before
    it, the statistics, the removed count and the particles after it."""
    tags = np.arange(1, e[0].size + 1, dtype=np.int32)
    rt.upload(0, *e, tag=tags)
    rt.sort_particles(0)
    rt.set_grid("EX", EX)
    rt.set_grid("EY", EY)
    out = []
    for step in range(4):
        if step in SORT_STEPS:
            if sorting:
                rt.sort_with_next_push()
            else:
                rt.sort_particles(0)
        before = rt.download(0)
        st = rt.advance_electrons()
        idx, _ = rt.removed(0)
        out.append((before, st, idx.size, rt.download(0)))
    return out, rt.gather_electron_charge_density()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n,scale", [(1, 4), (2, 32), (4, 16)])
def test_sorting_push_equals_sort_then_push(edgpu, oracle, precision, n, scale):
    O = oracle
    prob = config.workload(n, scale)
    print("case sorting push cfg%d/%d precision %d: %d electrons, grid %dx%d" %
          (n, scale, precision, prob.n_particles, prob.gmi, prob.gmj), flush=True)
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=0, sort_interval=4)
    try:
        e = _electrons(O, prob)
        EX, EY = synthetic_field(prob, amp=0.08)
        sp_run, rho_s = _run(rt, prob, e, True, EX, EY)
        pl_run, rho_p = _run(rt, prob, e, False, EX, EY)
    finally:
        rt.close()
    g = O.make_geom(prob.geometry())
    B = (prob.Bx, prob.By, prob.Bz)
    nobj = prob.n_boundary_obj + len(prob.inner_boxes)
    ref = None
    for step, ((bs, ss, rs, a), (bp, sp_, rp, b)) in enumerate(zip(sp_run, pl_run)):
        # the same particles, bit for bit, whichever path sorted them
        A, Bt = _by_tag(a), _by_tag(b)
        assert np.array_equal(A[5], Bt[5]), "push %d: removed sets differ" % step
        for q in range(5):
            assert np.array_equal(A[q], Bt[q]), "push %d, component %d" % (step, q)
        assert ss["n_removed"] == sp_["n_removed"] == rs == rp
        assert list(ss["hit_count"]) == list(sp_["hit_count"])
        if precision == 0:
            # and the oracle's push of the same particles
            ref = [x.copy() for x in bs[:5]]
            rc, rem, _, hits = O.advance(g, O.make_species(True), EX, EY, B, prob.Ez, *ref, nobj)
            assert rc == 0 and int(rem.sum()) == rs
            assert list(ss["hit_count"][:nobj]) == list(hits[:nobj])
            ref = [np.ascontiguousarray(x[rem == 0]) for x in ref]
            assert np.array_equal(sorted_rows(*ref), sorted_rows(*a[:5])), "push %d against the oracle" % step
        if step in SORT_STEPS:
            # order: the tiles the particles had before the push, non-decreasing along the
            # slots except where the compaction put a particle from the end into a hole
            key_of = dict(zip(bs[5].tolist(), _tile_keys(prob, bs[0], bs[1]).tolist()))
            keys = np.array([key_of[t] for t in a[5].tolist()])
            descents = int(np.count_nonzero(np.diff(keys) < 0))
            print("push %d: %d electrons, %d removed, %d descents" % (step, a[0].size, rs, descents), flush=True)
            assert descents <= 2 * rs
    assert rel_err(rho_s, rho_p) < 1e-12


STRIDE = 11  # uniforms per slot of an injected e-n stream
# word 0 makes every electron a candidate; word 1, against the cumulative process
# probabilities of the table, gives no collision below ~35 eV, ionisation at 40-80 eV and
# elastic scattering above ~120 eV
WORDS = [1e-7, 0.065, 0.5, 0.25, 0.75, 0.1, 0.9, 0.3, 0.6, 0.45, 0.55]


@pytest.mark.parametrize("precision", [0, 1])
def test_sorting_push_collisions(edgpu, oracle, precision):
    """Candidacy in the sorting push, selection and kinematics on the arrays it wrote, the
    products appended: the same electrons and ions as the plain fused path."""
    O = oracle
    prob = config.workload(1, 8)
    print("case sorting push with e-n collisions, precision %d: %d electrons" % (precision, prob.n_particles),
          flush=True)
    rng = np.random.default_rng(41)
    e, _ = oracle_particles(O, prob)
    vv = rng.standard_normal((3, e[0].size))
    vv /= np.linalg.norm(vv, axis=0)
    speed = np.sqrt(rng.uniform(0.0, 200.0, e[0].size) / prob.energy_factor_eV)
    e[2], e[3], e[4] = [np.ascontiguousarray(vv[q] * speed) for q in range(3)]
    EX, EY = synthetic_field(prob, amp=0.05)
    results = []
    rt = edgpu.Runtime(prob, precision=precision, fuse_collisions=1, sort_interval=4)
    try:
        rt.initiate_electron_neutral_collisions()
        rt.set_uniform_stream(np.tile(np.array(WORDS), 4 * e[0].size))
        for sorting in (True, False):
            rt.upload(0, *e, tag=np.arange(1, e[0].size + 1, dtype=np.int32))
            rt.upload(1, e[0][:0], e[1][:0], e[2][:0], e[3][:0], e[4][:0])
            rt.sort_particles(0)
            rt.set_grid("EX", EX)
            rt.set_grid("EY", EY)
            rt.advance_electrons()  # drift, and products from the first push
            rt.process_added_electrons()
            rt.process_added_ions()
            if sorting:
                rt.sort_with_next_push()
            else:
                rt.sort_particles(0)
            st = rt.advance_electrons()
            rt.process_added_electrons()
            rt.process_added_ions()
            results.append((st, rt.download(0), rt.download(1)))
    finally:
        rt.close()
    (s1, e1, i1), (s2, e2, i2) = results
    print("collisions per process %s, %d electrons, %d ions" % (s1["collisions"][:4], e1[0].size, i1[0].size),
          flush=True)
    assert s1["collisions"][0] > 100 and s1["collisions"][2] > 100  # elastic and ionising
    assert list(s1["collisions"]) == list(s2["collisions"]) and s1["n_removed"] == s2["n_removed"]
    assert np.array_equal(sorted_rows(*e1), sorted_rows(*e2))
    assert np.array_equal(sorted_rows(*i1), sorted_rows(*i2))


def test_pending_count_is_dropped(edgpu, oracle):
    """An upload after edgpu_sort_with_next_push drops the count, and the next push is a plain
    one; electrons never sorted are sorted at once; only the electrons are sorted by a push."""
    O = oracle
    prob = config.workload(1, 8)
    rt = edgpu.Runtime(prob, fuse_collisions=0, sort_interval=4)
    try:
        e = _electrons(O, prob)
        tags = np.arange(1, e[0].size + 1, dtype=np.int32)
        rt.upload(0, *e, tag=tags)
        rt.sort_with_next_push()  # never sorted: sorted now
        got = rt.download(0)
        assert np.all(np.diff(_tile_keys(prob, got[0], got[1])) >= 0)
        rt.sort_with_next_push()
        half = [a[: a.size // 2] for a in e]
        rt.upload(0, *half, tag=tags[: tags.size // 2])
        EX, EY = synthetic_field(prob, amp=0.05)
        rt.set_grid("EX", EX)
        rt.set_grid("EY", EY)
        st = rt.advance_electrons()
        assert st["n_before"] == half[0].size
        with pytest.raises(edgpu.EdgpuError):
            edgpu._check(edgpu.lib().edgpu_sort_with_next_push(1), "edgpu_sort_with_next_push")
    finally:
        rt.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_steps_with_sorting_push(edgpu, oracle, precision):
    """Whole steps (collisions, exchanges, appends) with the electrons sorted by their push every
    third step: the bookkeeping closes and the grid charge equals the particle count."""
    prob = config.workload(2, 16)
    rt = edgpu.Runtime(prob, precision=precision, sort_interval=3)
    try:
        rt.initiate_electron_neutral_collisions()
        rt.initiate_ion_neutral_collisions()
        rt.load_initial_particles()
        ne, ni = rt.count(0), rt.count(1)
        for step in range(8):
            se, si = rt.step()
            assert se["n_before"] == ne and si["n_before"] == ni
            assert se["n_after"] == ne - se["n_removed"] and si["n_after"] == ni - si["n_removed"]
            ne, ni = rt.count(0), rt.count(1)
            assert ne - se["n_after"] == ni - si["n_after"] >= 0  # ionisation products
        for species, n in ((0, ne), (1, ni)):
            rho = rt.gather_electron_charge_density() if species == 0 else rt.gather_ion_charge_density()
            R = rho.reshape(prob.ny, prob.nx).copy()
            R[0, :] *= 0.5  # the wall rows' density is doubled
            R[-1, :] *= 0.5
            total = R[:, 1:prob.gmi].sum()  # column 0 is the periodic copy of column gmi
            print("precision %d species %d: %d particles, grid charge %.6f" % (precision, species, n, total),
                  flush=True)
            assert abs(total - n) < (1e-9 if precision == 0 else 1e-6) * n
    finally:
        rt.close()
