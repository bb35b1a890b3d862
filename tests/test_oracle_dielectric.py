"""Pinning of the oracle's dielectric-wall path (CPU only): PREPARE_SYS_Y and
SOLVE_POISSON_FFTX_LINSYSY with dielectric walls (pic2d_Prepare_FFTX_SYSY.f90:776-897,
pic2d_ElectricFieldCalc_FFT_X.f90:100-145, 676-759), and the surface charge that wall hits and
emission leave on them (pic2d_ElectronWallCollisions.f90:168-195, pic2d_IonWallCollisions.f90:
200-219, 717-766, pic2d_SEEmission.f90:226-237).

- mode 0 against the closed form of a uniform charge between a charged dielectric and a biased
  metal wall (pins the half wall-row density, the sign of the surface charge, the gauge);
- eps = 0 against the mirror image: a dielectric wall of zero permittivity is a symmetry
  plane, so the solve equals the metal-metal solve (pinned elsewhere by a dense solve) of the
  domain mirrored about it;
- every wall combination against a dense real-space solve of the same discrete operator
  (spectral x second derivative, three-point y stencil, half rows at dielectric walls with the
  reference's permittivity term applied as an x operator, mode-0 gauge between two
  dielectrics);
- the surface charge: hand-worked weights, charge and first-moment conservation, clamping,
  the periodic exchange.
The permittivity coefficient itself, -2 pi eps n / N^2 in the wall row of mode n, is the
reference's (pic2d_Prepare_FFTX_SYSY.f90:836-837); only its eps = 0 limit and its place in the
system are pinned independently.
"""
import math

import numpy as np
import pytest

OBJ_ABOVE, OBJ_BELOW = 1, 3


def _geom(oracle, gmi, gmj, phi_above=0.0, phi_below=0.0):
    segs = [(oracle.SIDE_ABOVE, 0, gmi, OBJ_ABOVE), (oracle.SIDE_BELOW, 0, gmi, OBJ_BELOW)]
    return oracle.make_geom(dict(gmi=gmi, gmj=gmj, periodic_x=True, segments=segs,
                                 object_phi=[0.0, phi_above, 0.0, phi_below], n_boundary_obj=4, inner_boxes=[]))


def _mask(below, above):
    return (1 << OBJ_BELOW if below else 0) | (1 << OBJ_ABOVE if above else 0)


def _periodic(rows, N):
    """node 0 and node N+1 = gmi as the periodic images of nodes N and 1"""
    rows[..., 0] = rows[..., N]
    rows[..., N + 1] = rows[..., 1]
    return rows


def _solve(oracle, g, N, gmj, rho_i, below, above, eps_b, eps_a, sb, sa, N_ppc=1):
    a_eq = oracle.prepare_sys_y_walls(N, gmj, below, above, eps_b, eps_a)
    return oracle.solve_poisson_fftx_walls(g, N_ppc, np.zeros_like(rho_i), rho_i, a_eq, _mask(below, above), sb, sa,
                                           below and above)


@pytest.mark.parametrize("eps", [1.0, 4.0])
def test_mode0_closed_form(oracle, eps):
    """Uniform charge r between a dielectric wall below carrying a uniform surface charge s and
    a metal wall above at phi_a: phi(j) = A + B j - r j^2/2 with B = -s from the wall-row law
    phi_1 - phi_0 = -(r/2 + s), and phi(gmj) = phi_a.  The permittivity does not enter mode 0."""
    N, gmj = 16, 10
    gmi = N + 1
    r, s, phi_a, N_ppc = 0.37, -0.8, 0.3, 4
    f = 1.0 / N_ppc
    g = _geom(oracle, gmi, gmj, phi_above=phi_a)
    rho = np.full((gmj + 1, gmi + 1), r * N_ppc)
    sb = np.full(gmi + 1, s * N_ppc)
    got = _solve(oracle, g, N, gmj, rho.ravel(), True, False, eps, 1.0, sb, np.zeros(gmi + 1), N_ppc)
    got = got.reshape(gmj + 1, gmi + 1)
    j = np.arange(gmj + 1, dtype=float)
    B = -s
    A = phi_a - B * gmj + r * gmj ** 2 / 2
    want = A + B * j - r * j ** 2 / 2
    assert f * N_ppc == 1.0
    assert np.max(np.abs(got - want[:, None])) < 1e-11 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("seed", [1, 2])
def test_zero_permittivity_is_a_mirror_plane(oracle, seed):
    """eps = 0 below: the solve equals rows gmj..2 gmj of the metal-metal solve of the domain
    mirrored about the wall, whose mirror row holds rho_0 + 2 sigma (twice the wall row's
    charge, the row being a full row there)."""
    N, gmj = 16, 7
    gmi = N + 1
    rng = np.random.default_rng(seed)
    phi_a = 0.45
    rho = _periodic(rng.standard_normal((gmj + 1, gmi + 1)), N)
    sigma = _periodic(rng.standard_normal(gmi + 1), N)
    got = _solve(oracle, _geom(oracle, gmi, gmj, phi_above=phi_a), N, gmj, rho.ravel(), True, False, 0.0, 1.0,
                 sigma, np.zeros(gmi + 1)).reshape(gmj + 1, gmi + 1)
    gmj2 = 2 * gmj
    rho2 = np.empty((gmj2 + 1, gmi + 1))
    for jj in range(gmj2 + 1):
        rho2[jj] = rho[abs(jj - gmj)]
    rho2[gmj] = rho[0] + 2.0 * sigma
    a_eq = oracle.prepare_sys_y(N, gmj2)
    g2 = _geom(oracle, gmi, gmj2, phi_above=phi_a, phi_below=phi_a)
    mirror = oracle.solve_poisson_fftx(g2, 1, np.zeros(rho2.size), rho2.ravel(), a_eq).reshape(gmj2 + 1, gmi + 1)
    scale = np.abs(mirror).max()
    assert np.max(np.abs(got - mirror[gmj:])) < 1e-12 * scale


def _dense(N, gmj, below, above, eps_b, eps_a, y, phi_b, phi_a):
    """Dense real-space solve over the FFT columns m = 0..N-1 (node N for m = 0) and rows
    0..gmj: interior rows phi_{j-1} + Dxx phi_j - 2 phi_j + phi_{j+1} = -y_j; a metal row
    phi = wall value; a dielectric row (phi_1 - phi_0) + Dxx/2 phi_0 - K phi_0 = -y_0 with K the
    x operator of eigenvalue 2 pi eps |n| / N^2; with both walls dielectric the mode-0 part of
    row 0 is replaced by mean(phi_0) = 0."""
    m = np.arange(N)
    nabs = np.minimum(m, N - m).astype(float)
    F = np.fft.fft(np.eye(N), axis=0)
    Finv = np.fft.ifft(np.eye(N), axis=0)

    def op(eig):
        return np.real(Finv @ np.diag(eig) @ F)

    Dxx = op(-(2 * math.pi * nabs / N) ** 2)
    I = np.eye(N)
    n = (gmj + 1) * N
    A = np.zeros((n, n))
    b = np.zeros(n)
    blk = lambda j: slice(j * N, (j + 1) * N)
    for j in range(1, gmj):
        A[blk(j), blk(j - 1)] = I
        A[blk(j), blk(j)] = Dxx - 2 * I
        A[blk(j), blk(j + 1)] = I
        b[blk(j)] = -y[j]
    for j, diel, eps, nb, phi in ((0, below, eps_b, 1, phi_b), (gmj, above, eps_a, gmj - 1, phi_a)):
        if not diel:
            A[blk(j), blk(j)] = I
            b[blk(j)] = phi
            continue
        K = op(2 * math.pi * eps * nabs / N ** 2)
        A[blk(j), blk(j)] = 0.5 * Dxx - I - K
        A[blk(j), blk(nb)] = I
        b[blk(j)] = -y[j]
    if below and above:
        P0 = np.full((N, N), 1.0 / N)
        Q = I - P0
        A[blk(0)] = Q @ A[blk(0)]
        A[blk(0), blk(0)] += P0
        b[blk(0)] = Q @ b[blk(0)]
    return np.linalg.solve(A, b).reshape(gmj + 1, N)


@pytest.mark.parametrize("below,above,eps_b,eps_a", [(True, False, 4.0, 1.0), (False, True, 1.0, 2.5),
                                                      (True, True, 3.0, 1.5), (True, True, 1.0, 1.0)])
def test_dense_real_space_solve(oracle, below, above, eps_b, eps_a):
    N, gmj = 8, 6
    gmi = N + 1
    N_ppc = 2
    rng = np.random.default_rng(11 + 2 * below + above)
    phi_b, phi_a = 0.6, -0.35
    rho = _periodic(rng.standard_normal((gmj + 1, gmi + 1)), N)
    sb = _periodic(rng.standard_normal(gmi + 1), N)
    sa = _periodic(rng.standard_normal(gmi + 1), N)
    g = _geom(oracle, gmi, gmj, phi_above=phi_a, phi_below=phi_b)
    got = _solve(oracle, g, N, gmj, rho.ravel(), below, above, eps_b, eps_a, sb, sa, N_ppc)
    got = got.reshape(gmj + 1, gmi + 1)
    f = 1.0 / N_ppc
    y = f * rho.copy()
    if below:
        y[0] = 0.5 * y[0] + f * sb
    if above:
        y[gmj] = 0.5 * y[gmj] + f * sa
    cols = np.array([N] + list(range(1, N)))   # node of FFT column m
    ref = _dense(N, gmj, below, above, eps_b, eps_a, y[:, cols], phi_b, phi_a)
    assert np.max(np.abs(got[:, cols] - ref)) < 1e-12 * max(1.0, np.abs(ref).max())
    assert np.allclose(got[:, 0], got[:, N]) and np.allclose(got[:, gmi], got[:, 1])


# ---------------------------------------------------------------------------------------
# surface charge

def _records(oracle, g, obj, dirflag, x, q, count=None, fold=False):
    nx = g.ixmax - g.ixmin + 1
    sb, sa = np.zeros(nx), np.zeros(nx)
    oracle.surface_charge(g, _mask(True, True), obj, dirflag, x, q, sb, sa, count)
    if fold:
        oracle.surface_charge_fold(g, sb)
        oracle.surface_charge_fold(g, sa)
    return sb, sa


def test_surface_charge_hand_worked(oracle):
    g = _geom(oracle, 17, 9)
    # electron below at x = 3.25: -0.75 at node 3, -0.25 at node 4
    sb, sa = _records(oracle, g, [OBJ_BELOW], [4], [3.25], -1)
    assert sb[3] == -0.75 and sb[4] == -0.25 and np.count_nonzero(sb) == 2 and not sa.any()
    # ion of charge 2 above at x = 3.25: 2 - 0.5 at node 3, 2 * 0.25 at node 4
    sb, sa = _records(oracle, g, [OBJ_ABOVE], [2], [3.25], 2)
    assert sa[3] == 1.5 and sa[4] == 0.5 and not sb.any()
    # three electrons emitted at x = 10.5 below: +1.5 at nodes 10 and 11
    sb, _ = _records(oracle, g, [OBJ_BELOW], [4], [10.5], 1, count=[3])
    assert sb[10] == 1.5 and sb[11] == 1.5
    # clamping: x < 0 charges node 0, x > gmi charges node gmi (iright = min(ileft+1, iend))
    sb, _ = _records(oracle, g, [OBJ_BELOW, OBJ_BELOW], [4, 4], [-0.3, 17.4], -1)
    assert sb[0] == -1.0 and sb[17] == -1.0 and np.count_nonzero(sb) == 2
    # metal objects, side walls and other dirflags leave nothing
    sb, sa = _records(oracle, g, [0, 2, OBJ_BELOW, OBJ_ABOVE], [1, 3, 1, 3], [2.5, 2.5, 2.5, 2.5], -1)
    assert not sb.any() and not sa.any()


@pytest.mark.parametrize("q", [-1, 1, 2, 3])
def test_surface_charge_conservation_and_first_moment(oracle, q):
    """Every record adds q in total, at the centre of charge x (records inside the segment)."""
    gmi = 33
    g = _geom(oracle, gmi, 9)
    rng = np.random.default_rng(q + 10)
    n = 5000
    x = rng.uniform(0.0, gmi, n)
    df = rng.choice([2, 4], n)
    obj = np.where(df == 2, OBJ_ABOVE, OBJ_BELOW)
    sb, sa = _records(oracle, g, obj, df, x, q)
    nodes = np.arange(gmi + 1)
    for row, d in ((sb, 4), (sa, 2)):
        sel = df == d
        assert row.sum() == pytest.approx(q * sel.sum(), rel=1e-12)
        assert (nodes * row).sum() == pytest.approx(q * x[sel].sum(), rel=1e-12)


def test_surface_charge_periodic_exchange(oracle):
    """The end nodes' charge moves to their periodic images (node 0 to gmi-1, node gmi to 1):
    the total is kept and the ends are empty."""
    gmi = 17
    g = _geom(oracle, gmi, 9)
    sb, _ = _records(oracle, g, [OBJ_BELOW] * 3, [4] * 3, [0.25, 16.5, 17.0], -1, fold=True)
    want = np.zeros(gmi + 1)
    want[16] = -0.75 - 0.5     # node 0's share and x = 16.5's left share
    want[1] = -0.25 - 0.5 - 1.0    # node 1's share, x = 16.5's right share (node 17) and x = 17
    assert np.allclose(sb, want, rtol=0, atol=1e-15)
    assert sb[0] == 0.0 and sb[gmi] == 0.0
