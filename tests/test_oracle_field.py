"""Pinning of the oracle's field path (CPU only): the Numerical-Recipes real FFT against
numpy.fft, the FFT_X Poisson solve against discrete eigenfunctions and a dense
brute-force solve of the same discrete operator, the field stencil on linear
potentials and the Krylov solve against scipy's sparse direct solver."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _geom(oracle, gmi, gmj, periodic=True, phi=(0.0, 0.0, 0.0, 0.0), boxes=(), box_phi=()):
    if periodic:
        segs = [(oracle.SIDE_ABOVE, 0, gmi, 1), (oracle.SIDE_BELOW, 0, gmi, 3)]
    else:
        segs = [(oracle.SIDE_LEFT, 0, gmj, 0), (oracle.SIDE_ABOVE, 0, gmi, 1), (oracle.SIDE_RIGHT, 0, gmj, 2),
                (oracle.SIDE_BELOW, 0, gmi, 3)]
    return oracle.make_geom(dict(gmi=gmi, gmj=gmj, periodic_x=periodic, segments=segs,
                                 object_phi=list(phi) + list(box_phi), n_boundary_obj=4, inner_boxes=list(boxes)))


# ---------------------------------------------------------------------------------------
# REAL8FT / FOUR1_8 (pic2d_ElectricFieldCalc_FFT_X.f90:1568-1710)

@pytest.mark.parametrize("N", [8, 64, 256])
def test_real8ft_against_numpy(oracle, N):
    """Forward transform = sum x_j exp(+2 pi i jk/N) = conj(rfft); packing: [0] = F_0,
    [1] = F_{N/2}, then (Re, Im) pairs; the inverse returns N/2 times the input."""
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N)
    d = x.copy()
    oracle.lib().orc_real8ft(oracle.dp(d), N, 1)
    F = np.fft.rfft(x)
    assert d[0] == pytest.approx(F[0].real, abs=1e-12)
    assert d[1] == pytest.approx(F[N // 2].real, abs=1e-12)
    assert np.allclose(d[2::2], F[1:N // 2].real, atol=1e-12)
    assert np.allclose(d[3::2], -F[1:N // 2].imag, atol=1e-12)
    oracle.lib().orc_real8ft(oracle.dp(d), N, -1)
    assert np.allclose(d * 2.0 / N, x, atol=1e-13)


# ---------------------------------------------------------------------------------------
# SOLVE_POISSON_FFTX_LINSYSY with PREPARE_SYS_Y (metal walls)

def _solve(oracle, g, N, gmj, rho_i, N_ppc=1):
    a_eq = oracle.prepare_sys_y(N, gmj)
    return oracle.solve_poisson_fftx(g, N_ppc, np.zeros_like(rho_i), rho_i, a_eq)


@pytest.mark.parametrize("n,m,kind", [(0, 1, "cos"), (3, 2, "cos"), (5, 7, "sin"), (8, 3, "cos"), (16, 1, "cos")])
def test_poisson_discrete_eigenfunctions(oracle, n, m, kind):
    """phi = trig(2 pi n i/N) sin(pi m j/gmj) solves the x-spectral / y-FD operator with
    eigenvalue -(4 pi^2 n^2/N^2 + 2 - 2 cos(pi m/gmj)) exactly (grounded walls)."""
    N, gmj = 32, 16
    gmi = N + 1
    g = _geom(oracle, gmi, gmj)
    i = np.arange(gmi + 1)[None, :]
    j = np.arange(gmj + 1)[:, None]
    trig = np.cos if kind == "cos" else np.sin
    phi = trig(2 * math.pi * n * i / N) * np.sin(math.pi * m * j / gmj)
    lam = 4 * math.pi ** 2 * n ** 2 / N ** 2 + 2 - 2 * math.cos(math.pi * m / gmj)
    N_ppc = 7
    rho_i = (N_ppc * lam * phi).ravel()
    got = _solve(oracle, g, N, gmj, rho_i, N_ppc).reshape(gmj + 1, gmi + 1)
    assert np.max(np.abs(got - phi)) < 1e-12


def _dense_operator(N, gmj):
    """x spectral second derivative (exact k^2 per mode) kron identity + y three-point
    stencil on the interior rows j = 1..gmj-1 (walls eliminated)."""
    kk = (2 * math.pi * np.minimum(np.arange(N), N - np.arange(N)) / N) ** 2
    Dxx = np.real(np.fft.ifft(-kk[:, None] * np.fft.fft(np.eye(N), axis=0), axis=0))
    ny = gmj - 1
    Dyy = np.diag(-2 * np.ones(ny)) + np.diag(np.ones(ny - 1), 1) + np.diag(np.ones(ny - 1), -1)
    return np.kron(np.eye(ny), Dxx) + np.kron(Dyy, np.eye(N))


def test_poisson_dense_brute_force(oracle):
    """Random charge and biased walls on a tiny grid against a dense solve."""
    N, gmj = 8, 6
    gmi = N + 1
    phi_b, phi_a = 0.5, -0.3
    g = _geom(oracle, gmi, gmj, phi=(0.0, phi_a, 0.0, phi_b))
    rng = np.random.default_rng(5)
    rho = rng.standard_normal((gmj + 1, gmi + 1))
    rho[:, 0] = rho[:, N]          # periodic images as the fold leaves them
    rho[:, gmi] = rho[:, 1]
    got = _solve(oracle, g, N, gmj, rho.ravel()).reshape(gmj + 1, gmi + 1)
    # unknowns: columns 1..N (column N is the image of column 0), rows 1..gmj-1
    A = _dense_operator(N, gmj)
    cols = [(c % N) for c in range(1, N + 1)]
    b = -rho[1:gmj, 1:N + 1].copy()
    b[0, :] -= phi_b
    b[-1, :] -= phi_a
    # reorder columns 1..N to the FFT order 0..N-1 (column N is node 0)
    order = np.argsort(cols)
    sol = np.linalg.solve(A, b[:, order].ravel()).reshape(gmj - 1, N)
    ref = np.empty((gmj - 1, N))
    ref[:, order] = sol
    assert np.max(np.abs(got[1:gmj, 1:N + 1] - ref)) < 1e-12
    assert np.allclose(got[0, :], phi_b) and np.allclose(got[gmj, :], phi_a)
    assert np.allclose(got[:, 0], got[:, N]) and np.allclose(got[:, gmi], got[:, 1])


# ---------------------------------------------------------------------------------------
# CALCULATE_ELECTRIC_FIELD_FFTX_LINSYSY (pic2d_ElectricFieldCalc_FFT_X.f90:1324-1547)

def test_efield_linear_potential_walled(oracle):
    """phi = a + b i + c j gives EX = -2 f2 b, EY = -2 f2 c everywhere when f1 = 2 f2
    (centred differences inside, one-sided at the walls and corners)."""
    gmi, gmj = 10, 7
    g = _geom(oracle, gmi, gmj, periodic=False)
    i = np.arange(gmi + 1)[None, :]
    j = np.arange(gmj + 1)[:, None]
    b, c = 0.3, -0.7
    phi = (1.5 + b * i + c * j + 0.0 * i * j).ravel()
    f2 = 0.005
    EX, EY = oracle.efield(g, phi, 2 * f2, f2)
    assert np.allclose(EX, -2 * f2 * b, atol=1e-15)
    assert np.allclose(EY, -2 * f2 * c, atol=1e-15)


def test_efield_periodic_and_accumulation(oracle):
    N, gmj = 16, 8
    gmi = N + 1
    g = _geom(oracle, gmi, gmj)
    i = np.arange(gmi + 1)[None, :]
    j = np.arange(gmj + 1)[:, None]
    kx = 2 * math.pi / N
    phi = (np.sin(kx * i) + 0.25 * j + 0.0 * i).ravel()
    f2 = 0.01
    acc = (np.ones_like(phi), np.zeros_like(phi))
    EX, EY = oracle.efield(g, phi, 2 * f2, f2, acc)
    EX = EX.reshape(gmj + 1, gmi + 1)
    EY = EY.reshape(gmj + 1, gmi + 1)
    # centred difference of the sine, periodic images at both x edges
    want = f2 * (np.sin(kx * (i - 1)) - np.sin(kx * (i + 1)))
    assert np.allclose(EX[1:gmj, :], np.broadcast_to(want, (gmj + 1, gmi + 1))[1:gmj, :], atol=1e-15)
    assert np.allclose(EY, -2 * f2 * 0.25, atol=1e-15)
    assert np.allclose(acc[0], 1.0 + EX.ravel()) and np.allclose(acc[1], EY.ravel())


# ---------------------------------------------------------------------------------------
# Krylov path (SOLVE_POTENTIAL_WITH_PETSC operator and right-hand side)

def _scipy_reference(gmi, gmj, periodic, dirichlet, rhs):
    """Direct solve of the 5-point Laplacian with the given Dirichlet values."""
    nx = gmi + 1
    ncol = gmi - 1 if periodic else nx
    idx = {}
    for jj in range(gmj + 1):
        for ii in range(ncol):
            if (ii, jj) not in dirichlet:
                idx[(ii, jj)] = len(idx)
    A = sp.lil_matrix((len(idx), len(idx)))
    b = np.zeros(len(idx))
    for (ii, jj), r in idx.items():
        A[r, r] = -4.0
        b[r] = rhs[jj * nx + ii]
        for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            i2, j2 = ii + di, jj + dj
            if periodic:
                i2 %= ncol
            if (i2, j2) in dirichlet:
                b[r] -= dirichlet[(i2, j2)]
            else:
                A[r, idx[(i2, j2)]] += 1.0
    x = spla.spsolve(A.tocsr(), b)
    phi = np.zeros((gmj + 1) * nx)
    for (ii, jj), v in dirichlet.items():
        phi[jj * nx + ii] = v
    for (ii, jj), r in idx.items():
        phi[jj * nx + ii] = x[r]
    if periodic:
        for jj in range(gmj + 1):
            phi[jj * nx + gmi - 1] = phi[jj * nx + 0]
            phi[jj * nx + gmi] = phi[jj * nx + 1]
    return phi


@pytest.mark.parametrize("periodic", [False, True])
def test_krylov_against_scipy(oracle, periodic):
    gmi, gmj = (17 if periodic else 16), 12
    wall = (0.0, 0.2, 0.0, -0.1) if periodic else (0.1, 0.2, -0.3, -0.1)
    boxes = [(4, 7, 4, 6), (10, 12, 5, 8)]
    box_phi = (0.8, -0.6)
    g = _geom(oracle, gmi, gmj, periodic, wall, boxes, box_phi)
    rng = np.random.default_rng(6)
    nn = (gmi + 1) * (gmj + 1)
    rho_e = rng.uniform(0.5, 1.5, nn)
    rho_i = rng.uniform(0.5, 1.5, nn)
    b = np.zeros(nn)
    oracle.lib().orc_krylov_rhs(g, 3, oracle.dp(rho_e), oracle.dp(rho_i), oracle.dp(b))
    x = np.zeros(nn)
    res = np.zeros(1)
    it = oracle.lib().orc_krylov_cg(g, oracle.dp(b), oracle.dp(x), 1e-12, 5000, oracle.dp(res))
    assert it >= 0 and res[0] <= 1e-12
    # Dirichlet set from the definition: walls (first segment containing the node; the
    # bottom/top rows include the corners), then the inner boxes (inclusive node ranges)
    d = {}
    ncol = gmi - 1 if periodic else gmi + 1
    for ii in range(ncol):
        d[(ii, 0)] = wall[3]
        d[(ii, gmj)] = wall[1]
    if not periodic:
        for jj in range(1, gmj):
            d[(0, jj)] = wall[0]
            d[(gmi, jj)] = wall[2]
    for (i0, i1, j0, j1), ph in zip(boxes, box_phi):
        for ii in range(i0, i1 + 1):
            for jj in range(j0, j1 + 1):
                d[(ii, jj)] = ph
    rhs = -(rho_i - rho_e) / 3.0
    ref = _scipy_reference(gmi, gmj, periodic, d, rhs)
    assert np.max(np.abs(x - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))


def test_wall_segment_overlap_rules(oracle):
    """Where two wall segments share an end node: the Krylov right-hand side takes the
    first segment that contains it (the search EXITs, pic2d_ElectricFieldCalc_PETSc.F90:66-80),
    the FFT_X wall row the last one (every part overwrites its range,
    pic2d_ElectricFieldCalc_FFT_X.f90:100-121)."""
    N, gmj = 16, 6
    gmi = N + 1
    segs = [(oracle.SIDE_ABOVE, 0, gmi, 1), (oracle.SIDE_BELOW, 0, 8, 3), (oracle.SIDE_BELOW, 8, gmi, 4)]
    phi = [0.0, 0.25, 0.0, -0.5, 0.75]
    g = oracle.make_geom(dict(gmi=gmi, gmj=gmj, periodic_x=True, segments=segs, object_phi=phi, n_boundary_obj=5,
                              inner_boxes=[]))
    nn = (gmi + 1) * (gmj + 1)
    z = np.zeros(nn)
    b = np.zeros(nn)
    oracle.lib().orc_krylov_rhs(g, 1, oracle.dp(z), oracle.dp(z), oracle.dp(b))
    assert b[7] == -0.5 and b[8] == -0.5 and b[9] == 0.75
    sol = oracle.solve_poisson_fftx(g, 1, z, z, oracle.prepare_sys_y(N, gmj))
    assert sol[7] == pytest.approx(-0.5, abs=1e-14) and sol[8] == pytest.approx(0.75, abs=1e-14)


# ---------------------------------------------------------------------------------------
# vacuum gaps (pic2d_CurProblemValues.f90:334-390, pic2d_WallPotentials.f90:29-52)

def _gap_box(oracle, gmi, gmj, V):
    """left electrode at 0, right electrode at V, the walls above and below vacuum gaps that
    join them; the segments of the gaps run from the left wall to the right one"""
    segs = [(oracle.SIDE_LEFT, 0, gmj, 0), (oracle.SIDE_ABOVE, 0, gmi, 1), (oracle.SIDE_RIGHT, 0, gmj, 2),
            (oracle.SIDE_BELOW, 0, gmi, 3)]
    return oracle.make_geom(dict(gmi=gmi, gmj=gmj, periodic_x=False, segments=segs, object_phi=[0.0, 7.0, V, -7.0],
                                 n_boundary_obj=4, inner_boxes=[], vacuum_gaps=[1, 3]))


def test_vacuum_gap_connections(oracle):
    """Each gap's start touches the left electrode (its segment's end above, its start below)
    and its end the right one; the gap's own potential is never used."""
    g = _gap_box(oracle, 12, 9, 2.0)
    assert (g.seg_gap[1], g.seg_gap_from[1], g.seg_gap_to[1]) == (1, 0, 2)
    assert (g.seg_gap[3], g.seg_gap_from[3], g.seg_gap_to[3]) == (1, 0, 2)
    assert g.seg_gap[0] == 0 and g.seg_gap[2] == 0
    # an end that touches no other object is refused, as the reference stops there
    with pytest.raises(ValueError):
        oracle.make_geom(dict(gmi=12, gmj=9, periodic_x=False, segments=[(oracle.SIDE_BELOW, 2, 5, 0),
                                                                         (oracle.SIDE_BELOW, 5, 12, 1)],
                              object_phi=[0.0, 1.0], n_boundary_obj=2, inner_boxes=[], vacuum_gaps=[0]))


def test_vacuum_gaps_between_electrodes_closed_form(oracle):
    """No charge, electrodes at 0 (left) and V (right) joined by gaps along the walls above and
    below: the gaps carry V i/gmi, and the discrete Laplace solution is phi = V i/gmi everywhere."""
    gmi, gmj, V = 12, 9, 2.5
    g = _gap_box(oracle, gmi, gmj, V)
    n = (gmi + 1) * (gmj + 1)
    b = np.zeros(n)
    oracle.lib().orc_krylov_rhs(oracle.C.byref(g), 1, oracle.dp(np.zeros(n)), oracle.dp(np.zeros(n)), oracle.dp(b))
    x = np.zeros(n)
    res = np.zeros(1)
    it = oracle.lib().orc_krylov_cg(oracle.C.byref(g), oracle.dp(b), oracle.dp(x), 1e-13, 10000, oracle.dp(res))
    assert it >= 0
    want = np.tile(V * np.arange(gmi + 1) / gmi, gmj + 1)
    assert np.max(np.abs(x - want)) < 1e-11
    P = x.reshape(gmj + 1, gmi + 1)
    assert P[0, 0] == 0.0 and P[gmj, gmi] == V
