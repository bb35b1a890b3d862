"""Pinning of the oracle's ion-induced electron emission (CPU only),
PROCESS_ION_INDUCED_ELECTRON_EMISSION / INJECT_ION_IMPACT_EXTRACTED_ELECTRON
(pic2d_IonEmissions.f90:4-348): the energy window, the number of electrons (floor of the
coefficient, then one more below the remainder), the order of the draws, the wall the electrons
leave and their speed scale, and the energy bookkeeping that stops the emission once the ion's
energy above the threshold is spent."""
import numpy as np
import pytest

import oracle_ffi as OF

XMAX, YMAX = 128.0, 96.0


def _emit(oracle, setD, E, u, dirflag=4, minE=1.0, maxE=100.0, vinj=0.05, Ms=4.0, x=10.3, y=0.2):
    maxw, inj = oracle.maxwell_tables()
    v = np.sqrt(E / Ms)
    return OF.iiee_emit(setD, minE, maxE, vinj, Ms, maxw, inj, dirflag, XMAX, YMAX, x, y, 0.0, -v, 0.0, u)


def test_energy_window(oracle):
    u = np.full(32, 0.5)
    for E in (0.5, 1.0, 100.0, 150.0):
        rc, out = _emit(oracle, 1.0, E, u)
        assert rc == 0 and out.shape[0] == 0, E
    rc, out = _emit(oracle, 1.0, 50.0, u)
    assert out.shape[0] == 1


@pytest.mark.parametrize("R,n", [(0.2, 3), (0.29, 3), (0.31, 2), (0.9, 2)])
def test_count_and_draw_order(oracle, R, n):
    """setD = 2.3: two electrons from six uniforms (three each), then the seventh uniform
    decides the third, which takes uniforms eight to ten."""
    u = np.full(32, 0.5)
    u[6] = R
    rc, out = _emit(oracle, 2.3, 50.0, u)
    assert rc == 0 and out.shape[0] == n
    # every electron of a constant stream is the same
    assert np.all(out == out[0])


def test_leaves_the_wall(oracle):
    maxw, inj = oracle.maxwell_tables()
    rng = np.random.default_rng(3)
    u = rng.uniform(0, 1, 32)
    for dirflag, xy, sign in ((4, (1, 1.0e-6), +1), (2, (1, YMAX - 1.0e-6), -1)):
        rc, out = _emit(oracle, 3.0, 80.0, u, dirflag=dirflag, x=-2.0)
        assert out.shape[0] == 3
        assert np.all(out[:, 0] == 0.0)            # x clamped into the area
        assert np.all(out[:, xy[0]] == xy[1])
        assert np.all(sign * out[:, 3] >= 0.0)
        # velocities: the tables scaled by vinj, normal component from the injection table
        for q in range(3):
            vx = OF.lib().orc_maxwell_sample(OF.dp(maxw), u[3 * q]) * 0.05
            vn = OF.lib().orc_maxwell_inj_sample(OF.dp(inj), u[3 * q + 1]) * 0.05
            assert out[q, 2] == pytest.approx(vx, abs=1e-15) and abs(out[q, 3]) == pytest.approx(vn, abs=1e-15)


def test_energy_is_spent(oracle):
    """An ion 1e-6 above the threshold: the first electron is scaled to exactly that energy,
    the others are not emitted."""
    u = np.full(32, 0.7)
    rc, out = _emit(oracle, 3.0, 1.0 + 1.0e-6, u)
    assert out.shape[0] == 1
    assert (out[0, 2:] ** 2).sum() == pytest.approx(1.0e-6, rel=1e-9)


def test_ion_reflection_by_hand():
    """INJECT_REFLECTED_ION at each wall: 1e-6 off it at the clamped coordinate, the normal
    component reversed."""
    for dirflag, want in ((4, [3.5, 1.0e-6, 0.1, 0.2, 0.3]), (2, [3.5, YMAX - 1.0e-6, 0.1, 0.2, 0.3]),
                          (1, [1.0e-6, 7.0, -0.1, -0.2, 0.3]), (3, [XMAX - 1.0e-6, 7.0, -0.1, -0.2, 0.3])):
        vy = -0.2 if dirflag == 4 else 0.2 if dirflag == 2 else -0.2
        out = OF.ion_reflect(dirflag, XMAX, YMAX, 3.5, 7.0, 0.1 if dirflag in (2, 4) else 0.1, vy, 0.3)
        if dirflag in (2, 4):
            assert list(out) == [3.5, want[1], 0.1, -vy, 0.3]
        else:
            assert list(out) == [want[0], 7.0, -0.1, vy, 0.3]
    # the coordinate along the wall is clamped
    assert OF.ion_reflect(4, XMAX, YMAX, -1.0, 0.0, 0.0, -1.0, 0.0)[0] == 0.0
    assert OF.ion_reflect(1, XMAX, YMAX, 0.0, YMAX + 2.0, -1.0, 0.0, 0.0)[1] == YMAX
