"""Pins of the FP64 oracle: each reference function restated in oracle/oracle.c
is checked against a library routine, a closed form, an invariant or a brute
force evaluation chosen so that a dropped term, a wrong sign or a shifted
index fails.  The reference ships no tests or golden data for this path
(its own checks are runtime Stops and monitors), so these pins are what fixes
the oracle.
"""
import math

import numpy as np
import pytest

import oracle
from channel_rl_amd import cases

rng = np.random.default_rng(2024)


def mk(nx=18, ny=17, nz=18, stretched=False, **kw):
    y = cases.tanh_y(ny, 1.5) if stretched else cases.uniform_y(ny)
    x = cases.periodic_faces(nx, 2 * math.pi)
    z = cases.periodic_faces(nz, math.pi)
    args = dict(nu=1e-2, CFL=0.5, dPdx=0.0, x_mass_cte=0, LES_model=2, Dirichlet_nu_t=1)
    args.update(kw)
    return oracle.Oracle(x, y, z, **args)


def centres(f):
    return cases.centres_with_ghosts(np.asarray(f))


def fields(o, fu, fv, fw):
    """Evaluate callables on every point of U, V, W (ghosts included)."""
    xg, yg, zg = centres(o.x), centres(o.y), centres(o.z)
    X, Y, Z = o.x, o.y, o.z
    U = fu(X[:, None, None], yg[None, :, None], zg[None, None, :]) * np.ones((o.nx, o.nyg, o.nzg))
    V = fv(xg[:, None, None], Y[None, :, None], zg[None, None, :]) * np.ones((o.nxg, o.ny, o.nzg))
    W = fw(xg[:, None, None], yg[None, :, None], Z[None, None, :]) * np.ones((o.nxg, o.nyg, o.nz))
    return (np.asfortranarray(U), np.asfortranarray(V), np.asfortranarray(W))


zero = lambda x, y, z: 0.0 * x * y * z  # noqa: E731


# ------------------------------------------------------------ library routines

@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 12, 30, 32, 64, 96, 128, 192, 384])
def test_fft_matches_numpy(n):
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = oracle.fft(a, -1)
    b = oracle.fft(a, +1)
    assert np.max(np.abs(f - np.fft.fft(a))) <= 1e-14 * n * np.max(np.abs(a))
    assert np.max(np.abs(b - np.fft.ifft(a) * n)) <= 1e-14 * n * np.max(np.abs(a))


def test_zgtsv_matches_lapack():
    la = pytest.importorskip("scipy.linalg.lapack")
    x1, info1 = oracle.zgtsv([], [2.0 + 1j], [], [3.0 - 1j])
    assert info1 == 0 and abs(x1[0] - (3.0 - 1j) / (2.0 + 1j)) <= 1e-15
    for n, pivot in [(2, False), (7, False), (40, False), (40, True), (97, True)]:
        d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        dl = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
        du = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
        if not pivot:
            d = d + 4.0 * np.sign(d.real)
        b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        x, info = oracle.zgtsv(dl, d, du, b)
        r = la.zgtsv(dl.copy(), d.copy(), du.copy(), b.copy())
        assert info == 0 and r[4] == 0
        assert np.max(np.abs(x - r[3])) <= 1e-13 * np.max(np.abs(r[3]))
        # and it solves the system
        A = np.diag(d) + np.diag(dl, -1) + np.diag(du, 1)
        assert np.max(np.abs(A @ x - b)) <= 1e-12 * np.max(np.abs(b)) * max(1, np.max(np.abs(A)))


# ------------------------------------------------------------ set-up constants

def test_modified_wavenumbers_closed_form():
    o = mk(34, 9, 22)
    p = o.poisson()
    g = o.grid()
    for kk, n, d in ((p["kxx"], 32, g["dxmin"]), (p["kzz"], 20, g["dzmin"])):
        k = np.arange(n)
        ref = 2.0 * (np.cos(2 * np.pi * k / n) - 1.0) / d ** 2
        assert np.allclose(kk, ref, rtol=1e-13, atol=1e-12 / d ** 2)
        assert np.allclose(kk[1:], kk[1:][::-1], rtol=1e-12)


@pytest.mark.parametrize("stretched", [False, True])
def test_dyy_rows(stretched):
    o = mk(10, 21, 10, stretched=stretched)
    p, g = o.poisson(), o.grid()
    d, l, u = p["d"], p["l"], p["u"]
    n = len(d)
    # Neumann rows and interior rows sum to zero (constants are in the null space)
    rows = d.copy()
    rows[:-1] += u
    rows[1:] += l
    assert np.max(np.abs(rows)) <= 1e-12 * np.max(np.abs(d))
    A = np.diag(d) + np.diag(l, -1) + np.diag(u, 1)
    yg = g["yg"][1:-1]
    if not stretched:
        # exact second derivative of a quadratic on a uniform grid (interior rows)
        assert np.allclose((A @ yg ** 2)[1:-1], 2.0, rtol=1e-10)
    # left null vector w_j = y_j - y_{j-1} (finite-volume conservation)
    w = np.diff(o.y)
    assert np.max(np.abs(w @ A)) <= 1e-12 * np.max(np.abs(A)) * np.max(w) * n


@pytest.mark.parametrize("stretched", [False, True])
def test_interpolation_weights(stretched):
    o = mk(10, 21, 10, stretched=stretched)
    g = o.grid()
    assert np.allclose(g["w0"] + g["w1"], 1.0, rtol=0, atol=1e-15)
    yg = g["yg"]
    f = 3.0 * yg - 1.0
    yi = g["w0"] * f[:-1] + g["w1"] * f[1:]
    assert np.allclose(yi, 3.0 * o.y - 1.0, rtol=0, atol=1e-13)


def test_rk3_order_conditions():
    o = mk()
    t, c = o.rk()
    # Butcher form of the low-storage Wray scheme: stages at 0, 8/15, 2/3
    A = np.array([[0, 0, 0], [c[0, 0], 0, 0], [c[1, 0], c[1, 1], 0]])
    b = c[2]
    cc = A.sum(axis=1)
    assert np.allclose(cc, [0, 8 / 15, 2 / 3])
    assert np.allclose(t, [8 / 15, 2 / 3, 1.0])
    assert math.isclose(b.sum(), 1.0)
    assert math.isclose(b @ cc, 0.5)
    assert math.isclose(b @ cc ** 2, 1 / 3)
    assert math.isclose(b @ A @ cc, 1 / 6)


# ------------------------------------------------------------ boundary conditions

def test_boundary_condition_images():
    o = mk(12, 9, 14)
    raw = [np.asfortranarray(rng.standard_normal(o.shape(w))) for w in (0, 1, 2)]
    o.set_state(*raw)
    o.apply_boundary_conditions()
    U, V, W = o.state()
    nx, ny, nz, nxg, nyg, nzg = o.nx, o.ny, o.nz, o.nxg, o.nyg, o.nzg
    # x: faces 1 <- nx-1, nx <- 2 ; centres 1 <- nxg-2, nxg-1 <- 2, nxg <- 3 (0-based below)
    assert np.array_equal(U[0, 1:-1, 1:-3], U[nx - 2, 1:-1, 1:-3])
    assert np.array_equal(U[nx - 1, 1:-1, 1:-3], U[1, 1:-1, 1:-3])
    for F in (V, W):
        assert np.array_equal(F[0], F[nxg - 3]) and np.array_equal(F[nxg - 2], F[1]) and np.array_equal(F[nxg - 1], F[2])
    # z: U, V centres; W faces
    for F in (U, V):
        assert np.array_equal(F[:, :, 0], F[:, :, nzg - 3]) and np.array_equal(F[:, :, nzg - 2], F[:, :, 1])
        assert np.array_equal(F[:, :, nzg - 1], F[:, :, 2])
    assert np.array_equal(W[:, :, 0], W[:, :, nz - 2]) and np.array_equal(W[:, :, nz - 1], W[:, :, 1])
    # walls: V = 0 on faces, U and W mirrored in the ghost rows
    assert not V[:, 0].any() and not V[:, ny - 1].any()
    assert np.array_equal(U[:, 0], -U[:, 1]) and np.array_equal(U[:, nyg - 1], -U[:, nyg - 2])
    assert np.array_equal(W[:, 0], -W[:, 1]) and np.array_equal(W[:, nyg - 1], -W[:, nyg - 2])
    # untouched interior
    assert np.array_equal(U[1:nx - 1, 1:nyg - 1, 1:nzg - 3], raw[0][1:nx - 1, 1:nyg - 1, 1:nzg - 3])


# ------------------------------------------------------------ test filter

def filter_matrix_1d(n, w3, ends):
    M = np.zeros((n, n))
    for i in range(1, n - 1):
        M[i, i - 1:i + 2] = w3
    if ends:
        M[0, :2] = [2 / 3, 1 / 3]
        M[-1, -2:] = [1 / 3, 2 / 3]
    return M


def test_filter_brute_force():
    a = rng.standard_normal((7, 6, 8))
    out = oracle.filter_xzy(a)
    s = np.array([1 / 6, 2 / 3, 1 / 6])
    Fx, Fz = filter_matrix_1d(7, s, False), filter_matrix_1d(8, s, False)
    Fy = filter_matrix_1d(6, s, True)
    ref = np.einsum("ai,bj,ck,ijk->abc", Fx, Fy, Fz, a)[1:-1, :, 1:-1]
    assert np.max(np.abs(out - ref)) <= 1e-14 * np.max(np.abs(a))


def test_filter_closed_forms():
    n1, n2, n3 = 9, 8, 10
    x = np.arange(n1)[:, None, None] * 0.3
    y = np.arange(n2)[None, :, None] * 0.5
    z = np.arange(n3)[None, None, :] * 0.2
    one = np.ones((n1, n2, n3))
    assert np.allclose(oracle.filter_xzy(2.5 * one), 2.5, rtol=1e-15)
    lin = (1 + 2 * x + 3 * z - y) * one
    assert np.allclose(oracle.filter_xzy(lin)[:, 1:-1], lin[1:-1, 1:-1, 1:-1], rtol=1e-14)
    quad = (y * y) * one
    want = (y * y + 0.25 / 3)[:, 1:-1, :]
    assert np.allclose(oracle.filter_xzy(quad)[:, 1:-1, :], want, rtol=1e-13)


# ------------------------------------------------------------ right-hand sides

def rhs_with(o, U, V, W, nut=None, dPdx=0.0):
    o.set(oracle.NUT, np.zeros(o.shape(oracle.NUT)) if nut is None else nut)
    o.set_scalars(dPdx=dPdx)
    return o.compute_rhs(U, V, W)


def iu(o, a):
    return a[1:o.nx - 1, 1:o.nyg - 1, 1:o.nzg - 1]


def iv(o, a):
    return a[1:o.nxg - 1, 1:o.ny - 1, 1:o.nzg - 1]


def iw(o, a):
    return a[1:o.nxg - 1, 1:o.nyg - 1, 1:o.nz - 1]


def test_rhs_uniform_flow():
    o = mk(stretched=True)
    U, V, W = fields(o, lambda x, y, z: 1.7 + 0 * x, lambda x, y, z: -0.4 + 0 * x, lambda x, y, z: 0.9 + 0 * x)
    nut = np.asfortranarray(rng.uniform(0, 1e-3, o.shape(oracle.NUT)))
    ru, rv, rw = rhs_with(o, U, V, W, nut, dPdx=0.31)
    # exact up to the rounding of the y-weighted face values on the stretched grid
    assert np.max(np.abs(iu(o, ru) - 0.31)) <= 1e-12
    assert np.max(np.abs(iv(o, rv))) <= 1e-12 and np.max(np.abs(iw(o, rw))) <= 1e-12


def test_rhs_x_mode_convection_closed_form():
    o = mk(34, 9, 10, nu=0.0)
    k, A = 3.0, 0.7
    dx = o.x[1] - o.x[0]
    U, V, W = fields(o, lambda x, y, z: A * np.sin(k * x), zero, zero)
    ru, rv, rw = rhs_with(o, U, V, W)
    xf = o.x[1:o.nx - 1][:, None, None]
    ref = -(A * A) * np.cos(k * dx / 2) ** 2 * np.sin(k * dx) / dx * np.sin(2 * k * xf)
    assert np.max(np.abs(iu(o, ru) - ref)) <= 1e-12
    assert np.max(np.abs(iv(o, rv))) == 0 and np.max(np.abs(iw(o, rw))) == 0


def test_rhs_x_mode_diffusion_modified_wavenumber():
    o = mk(34, 9, 10, nu=0.03)
    k, A = 4.0, 1.3
    dx = o.x[1] - o.x[0]
    xg = centres(o.x)
    U, V, W = fields(o, zero, zero, lambda x, y, z: A * np.sin(k * x) + 0 * z)
    ru, rv, rw = rhs_with(o, U, V, W)
    ref = -0.03 * 4 / dx ** 2 * np.sin(k * dx / 2) ** 2 * A * np.sin(k * xg[1:o.nxg - 1])[:, None, None]
    assert np.max(np.abs(iw(o, rw) - ref)) <= 1e-12


def test_rhs_y_diffusion_of_quadratic():
    o = mk(10, 17, 10, nu=0.02)
    U, V, W = fields(o, lambda x, y, z: y * y + 0 * x, zero, zero)
    ru, _, _ = rhs_with(o, U, V, W, dPdx=0.5)
    assert np.allclose(iu(o, ru), 2 * 0.02 + 0.5, rtol=1e-11)


@pytest.mark.parametrize("stretched", [False, True])
def test_rhs_variable_viscosity_linear(stretched):
    # U = y, nu_t = a + b y at centres: d/dy(nu_t dU/dy) = b exactly on any grid
    o = mk(10, 17, 10, stretched=stretched, nu=0.0)
    U, V, W = fields(o, lambda x, y, z: y + 0 * x, zero, zero)
    yg = centres(o.y)
    nut = np.asfortranarray(np.broadcast_to((0.2 + 0.05 * yg)[None, :, None], o.shape(oracle.NUT)))
    ru, _, _ = rhs_with(o, U, V, W, nut)
    assert np.allclose(iu(o, ru), 0.05, rtol=1e-10)


@pytest.mark.parametrize("stretched", [False, True])
def test_rhs_uv_flux_weights(stretched):
    # U = y (linear), V = c: -d(uv)/dy = -c exactly with the y-weights
    o = mk(10, 17, 10, stretched=stretched, nu=0.0)
    U, V, W = fields(o, lambda x, y, z: y + 0 * x, lambda x, y, z: 0.6 + 0 * x, zero)
    ru, rv, _ = rhs_with(o, U, V, W)
    assert np.allclose(iu(o, ru), -0.6, rtol=1e-11)
    assert np.max(np.abs(iv(o, rv))) <= 1e-12  # d(uv)/dx = 0, d(v^2)/dy = 0


def analytic_fields():
    """smooth periodic velocity and its continuous RHS (nu constant, nu_t = 0)"""
    a, b = 1.0, 2.0

    def u(x, y, z): return np.sin(x) * np.cos(b * z) * (1 + 0.3 * y)
    def v(x, y, z): return 0.5 * np.cos(x) * np.sin(b * z) * np.sin(y)
    def w(x, y, z): return 0.4 * np.sin(2 * x) * np.cos(b * z) + 0.2 * y
    h = 1e-5

    def d(f, i, x, y, z):
        e = [0, 0, 0]; e[i] = h
        return (f(x + e[0], y + e[1], z + e[2]) - f(x - e[0], y - e[1], z - e[2])) / (2 * h)

    def lap(f, x, y, z):
        s = 0
        for i in range(3):
            e = [0, 0, 0]; e[i] = 1e-4
            s = s + (f(x + e[0], y + e[1], z + e[2]) - 2 * f(x, y, z) + f(x - e[0], y - e[1], z - e[2])) / 1e-8
        return s

    def rhs(c, x, y, z, nu):
        fs = (u, v, w)
        f = fs[c]
        conv = sum(d(lambda X, Y, Z: f(X, Y, Z) * fs[j](X, Y, Z), j, x, y, z) for j in range(3))
        return -conv + nu * lap(f, x, y, z)
    return u, v, w, rhs, a


def test_rhs_second_order_convergence():
    errs = []
    for n in (18, 34):
        o = mk(n, n - 1, n, nu=0.05)
        u, v, w, rhs, _ = analytic_fields()
        U, V, W = fields(o, u, v, w)
        ru, rv, rw = rhs_with(o, U, V, W)
        xg, yg, zg = centres(o.x), centres(o.y), centres(o.z)
        # compare away from the first/last interior rows (no wall condition applied here)
        s = slice(3, -3)
        eu = iu(o, ru)[:, s, :] - rhs(0, o.x[1:o.nx - 1, None, None], yg[1:o.nyg - 1][None, s, None],
                                      zg[1:o.nzg - 1][None, None, :], 0.05)
        ev = iv(o, rv)[:, s, :] - rhs(1, xg[1:o.nxg - 1, None, None], o.y[1:o.ny - 1][None, s, None],
                                      zg[1:o.nzg - 1][None, None, :], 0.05)
        ew = iw(o, rw)[:, s, :] - rhs(2, xg[1:o.nxg - 1, None, None], yg[1:o.nyg - 1][None, s, None],
                                      o.z[1:o.nz - 1][None, None, :], 0.05)
        errs.append([np.max(np.abs(e)) for e in (eu, ev, ew)])
    ratios = np.array(errs[0]) / np.array(errs[1])
    assert np.all(ratios > 3.3), (errs, ratios)


# ------------------------------------------------------------ strain, DSM

def test_strain_of_linear_shear():
    o = mk(12, 11, 12)
    U, V, W = fields(o, lambda x, y, z: 2.0 * y + 0 * x, zero, lambda x, y, z: -1.5 * y + 0 * x)
    Sij, S = o.compute_Sij(U, V, W)
    s = (slice(1, o.nxg - 1), slice(1, o.nyg - 1), slice(1, o.nzg - 1))
    assert np.allclose(Sij[s + (3,)], 1.0, rtol=1e-12)   # S12 = a/2
    assert np.allclose(Sij[s + (5,)], -0.75, rtol=1e-12)  # S23 = b/2
    for q in (0, 1, 2, 4):
        assert np.max(np.abs(Sij[s + (q,)])) <= 1e-12
    assert np.allclose(S[s], math.sqrt(2.0 ** 2 + 1.5 ** 2), rtol=1e-12)


def test_dsm_contraction_matches_full_tensors():
    c, f = cases.small_case(12, 11, 12, seed=9)
    o = oracle.Oracle(c.x, c.y, c.z, **c.kwargs())
    o.set_state(*f)
    o.apply_boundary_conditions()
    U, V, W = o.state()
    L, M, LM, MM = o.dsm_parts(U, V, W)
    idx = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]

    def full(T):
        F = np.zeros(T.shape[:3] + (3, 3))
        for q, (a, b) in enumerate(idx):
            F[..., a, b] = T[..., q]
            F[..., b, a] = T[..., q]
        return F
    Lf, Mf = full(L), full(M)
    for j in range(1, o.nyg - 1):
        sl = (slice(1, o.nxg - 1), j, slice(1, o.nzg - 1))
        assert math.isclose(np.einsum("ikab,ikab->", Lf[sl], Mf[sl]), LM[j], rel_tol=1e-10, abs_tol=1e-14)
        assert math.isclose(np.einsum("ikab,ikab->", Mf[sl], Mf[sl]), MM[j], rel_tol=1e-10)


def test_dsm_linear_shear():
    o = mk(14, 13, 14)
    g = 3.0
    U, V, W = fields(o, lambda x, y, z: g * y + 0 * x, zero, zero)
    L, M, LM, MM = o.dsm_parts(U, V, W)
    dy = o.y[1] - o.y[0]
    s = (slice(1, o.nxg - 1), slice(3, o.nyg - 3), slice(1, o.nzg - 1))
    assert np.allclose(L[s + (0,)], g * g * dy * dy / 3, rtol=1e-9)
    assert np.allclose(M[s + (3,)], 1.5 * g * g, rtol=1e-10)
    for q in (1, 2, 4, 5):
        assert np.max(np.abs(L[s + (q,)])) <= 1e-10
    for q in (0, 1, 2, 4, 5):
        assert np.max(np.abs(M[s + (q,)])) <= 1e-10
    nut = o.compute_eddy_viscosity(U, V, W)
    assert np.max(np.abs(nut[:, 3:o.nyg - 3, :])) <= 1e-12


def test_dsm_galilean_invariance():
    c, f = cases.small_case(12, 11, 12, seed=4)
    o = oracle.Oracle(c.x, c.y, c.z, **c.kwargs())
    o.set_state(*f)
    o.apply_boundary_conditions()
    U, V, W = o.state()
    n0 = o.compute_eddy_viscosity(U, V, W)
    n1 = o.compute_eddy_viscosity(U + 3.7, V, W - 1.1)
    assert np.max(np.abs(n1 - n0)) <= 1e-9 * np.max(np.abs(n0))


def test_dsm_constant_field_is_zero():
    o = mk(10, 9, 10)
    U, V, W = fields(o, lambda x, y, z: 1.0 + 0 * x, zero, zero)
    nut = o.compute_eddy_viscosity(U, V, W)
    assert np.all(nut == 0.0)  # <M:M> = 0 -> 0/0 clipped to 0


# ------------------------------------------------------------ AMD model

def _amd_expected(A, dxyz):
    """AMD of the linear field u_i = A_ik x_k on a uniform grid, from the
    model's tensor definition: -C sum_ij S_ij sum_k (D_k A_ik)(D_k A_jk) /
    sum A_ml^2, clipped at 0."""
    S = 0.5 * (A + A.T)
    G = A * np.asarray(dxyz)[None, :]          # undivided differences dh_k u_i
    N = np.einsum("ij,ik,jk->", S, G, G)
    return max(-0.3 * N / np.sum(A * A), 0.0), N


def test_amd_linear_field_closed_form():
    o = mk(14, 13, 10, LES_model=3)
    dxyz = (o.x[1] - o.x[0], o.y[1] - o.y[0], o.z[1] - o.z[0])
    assert len({round(d, 12) for d in dxyz}) == 3  # distinct spacings catch swapped directions
    A = np.array([[0.9, 0.35, -0.25], [0.5, -1.7, 0.45], [-0.65, 0.8, 0.6]])
    nu_e, N = _amd_expected(A, dxyz)
    if N > 0:
        A = -A
        nu_e, N = _amd_expected(A, dxyz)
    assert nu_e > 0
    U, V, W = fields(o, lambda x, y, z: A[0, 0] * x + A[0, 1] * y + A[0, 2] * z,
                     lambda x, y, z: A[1, 0] * x + A[1, 1] * y + A[1, 2] * z,
                     lambda x, y, z: A[2, 0] * x + A[2, 1] * y + A[2, 2] * z)
    nut = o.compute_eddy_viscosity(U, V, W)
    inner = nut[1:-2, 1:-1, 1:-2]
    assert np.max(np.abs(inner - nu_e)) <= 1e-10 * nu_e
    # Dirichlet ghost rows mirror the first interior row
    assert np.allclose(nut[1:-2, 0, 1:-2], -nut[1:-2, 1, 1:-2])


def test_amd_shear_and_constant_fields_are_zero():
    o = mk(10, 9, 10, LES_model=3)
    # constant field: 0/0, clipped to 0
    U, V, W = fields(o, lambda x, y, z: 1.0 + 0 * x, zero, zero)
    assert np.all(o.compute_eddy_viscosity(U, V, W) == 0.0)
    # plane shear U = y: no normal strain, and dh_y U dh_y V S12 = 0 -> N = 0
    U, V, W = fields(o, lambda x, y, z: y + 0 * x, zero, zero)
    assert np.all(o.compute_eddy_viscosity(U, V, W) == 0.0)


def test_amd_galilean_invariance():
    o = mk(12, 11, 12, stretched=True, LES_model=3)
    rng = np.random.default_rng(3)
    U, V, W = (np.asfortranarray(rng.standard_normal(s)) for s in ((o.nx, o.nyg, o.nzg), (o.nxg, o.ny, o.nzg),
                                                                   (o.nxg, o.nyg, o.nz)))
    a = o.compute_eddy_viscosity(U, V, W)
    b = o.compute_eddy_viscosity(U + 0.75, V - 0.5, W + 0.25)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.abs(a).max()
    assert np.abs(a).max() > 0


def test_amd_wall_offset_ghosts():
    o = mk(10, 9, 10, LES_model=3, Dirichlet_nu_t=2, dPdx=0.49)
    rng = np.random.default_rng(4)
    U, V, W = (np.asfortranarray(rng.standard_normal(s)) for s in ((o.nx, o.nyg, o.nzg), (o.nxg, o.ny, o.nzg),
                                                                   (o.nxg, o.nyg, o.nz)))
    nut = o.compute_eddy_viscosity(U, V, W)
    off = 0.38 * 0.7 * (o.y[1] - o.y[0])
    assert np.allclose(nut[:, 0, :], -nut[:, 1, :] + off)
    assert np.allclose(nut[:, -1, :], -nut[:, -2, :] + off)


# ------------------------------------------------------------ true pressure

def test_pressure_makes_the_momentum_rhs_divergence_free():
    # the pressure solves Laplacian(p) = div(rhs) with dP/dy = bc at the
    # walls: the corrected fluxes rhs - grad p are discretely divergence free
    # at every unique centre (wall faces: rhs_v = bc = dP/dy, so zero flux)
    from helpers import projected_state
    case, field = cases.small_case(18, 17, 18, seed=3)
    U, V, W = projected_state(case, field)
    o = oracle.Oracle(case.x, case.y, case.z, **case.kwargs())
    o.set_state(U, V, W)
    rng = np.random.default_rng(5)
    o.set(oracle.NUT, np.asfortranarray(rng.uniform(0.0, 2e-3, o.shape(oracle.NUT))))
    o.set_scalars(dPdx=0.3)
    P = o.compute_pressure()
    RU, RV, RW = o.compute_rhs(U, V, W)
    nx, ny, nz, nxg, nyg, nzg = o.nx, o.ny, o.nz, o.nxg, o.nyg, o.nzg
    RU[0], RU[nx - 1] = RU[nx - 2], RU[1]
    RW[:, :, 0], RW[:, :, nz - 1] = RW[:, :, nz - 2], RW[:, :, 1]
    g = o.grid()
    xg, yg, zg = g["xg"], g["yg"], g["zg"]
    x, y, z = o.x, o.y, o.z
    cu = RU - (P[1:, :, :] - P[:-1, :, :]) / (xg[1:] - xg[:-1])[:, None, None]
    cv = RV - (P[:, 1:, :] - P[:, :-1, :]) / (yg[1:] - yg[:-1])[None, :, None]
    cv[:, 0, :] = 0.0
    cv[:, ny - 1, :] = 0.0
    cw = RW - (P[:, :, 1:] - P[:, :, :-1]) / (zg[1:] - zg[:-1])[None, None, :]
    I, J, K = slice(1, nxg - 2), slice(1, nyg - 1), slice(1, nzg - 2)   # unique centres, 0-based
    div = ((cu[1:nxg - 2] - cu[0:nxg - 3])[:, J, K] / (x[1:nxg - 2] - x[0:nxg - 3])[:, None, None] +
           (cv[I, 1:nyg - 1] - cv[I, 0:nyg - 2])[:, :, K] / (y[1:] - y[:-1])[None, :, None] +
           (cw[I, J, 1:nzg - 2] - cw[I, J, 0:nzg - 3]) / (z[1:nzg - 2] - z[0:nzg - 3])[None, None, :])
    raw = ((RU[1:nxg - 2] - RU[0:nxg - 3])[:, J, K] / (x[1:nxg - 2] - x[0:nxg - 3])[:, None, None])
    assert np.abs(raw).max() > 0
    assert np.abs(div).max() <= 1e-9 * np.abs(raw).max(), (np.abs(div).max(), np.abs(raw).max())
    # the wall ghost rows carry the wall data: P(2) - P(1) = dy bc with the
    # same bc on every image of a cell
    assert np.allclose(P[0], P[nxg - 3]) and np.allclose(P[:, :, nzg - 2], P[:, :, 1])


# ------------------------------------------------------------ statistics

def test_wall_shear_of_poiseuille_flow():
    # U = A y (2 - y) evaluated at every centre (ghost rows included): the
    # centred difference across each wall is exact for a quadratic, so
    # tau = 2 A nu at both walls, u_tau = sqrt(2 A nu), Re_tau = u_tau / nu
    A, nu = 1.7, 0.01
    o = mk(10, 13, 10, nu=nu)
    U, V, W = fields(o, lambda x, y, z: A * y * (2 - y) + 0 * x, zero, zero)
    o.set_state(U, V, W)
    o.set(oracle.NUT, np.zeros(o.shape(oracle.NUT), order="F"))
    w = o.wall_shear()
    assert abs(w["tau_bottom"] - 2 * A * nu) < 1e-13 and abs(w["tau_top"] - 2 * A * nu) < 1e-13
    assert abs(w["u_tau"] - np.sqrt(2 * A * nu)) < 1e-13
    assert abs(w["Re_tau"] - np.sqrt(2 * A * nu) / nu) < 1e-10
    # an eddy viscosity constant in the wall rows adds in proportion
    o.set(oracle.NUT, np.full(o.shape(oracle.NUT), 0.5 * nu, order="F"))
    assert abs(o.wall_shear()["tau_bottom"] - 2 * A * 1.5 * nu) < 1e-13


def test_plane_statistics_of_known_fields():
    o = mk(10, 9, 10)
    U, V, W = fields(o, lambda x, y, z: 2.0 + np.sin(x) + 0 * y, lambda x, y, z: 0.5 + 0 * x * y * z,
                     lambda x, y, z: np.cos(2 * z) + 0 * x * y)
    o.set_state(U, V, W)
    st = o.plane_stats()
    # sin over the unique x faces and cos(2z) over the unique z faces average to 0
    assert np.allclose(st[0], 2.0) and np.allclose(st[1], 0.5) and np.allclose(st[2], 0.0, atol=1e-14)
    assert np.allclose(st[3], 4.0 + 0.5) and np.allclose(st[4], 0.25) and np.allclose(st[5], 0.5)


# ------------------------------------------------------------ RL agents

def test_rl_states_of_a_linear_profile():
    from oracle import rl
    s_, nu, dref = 3.0, 1e-3, 0.01
    o = mk(18, 13, 10, nu=nu)
    U, V, W = fields(o, lambda x, y, z: s_ * np.where(y < 1.0, y, 2.0 - y) + 0 * x, zero, zero)
    nut = np.zeros((o.nxg, o.nyg, o.nzg))
    yg = np.concatenate([[0.0], centres(o.y)])
    draws = np.array([[[0.2, 0.3], [0.7, 0.9]], [[0.4, 0.5], [0.6, 0.1]]])
    st, rw = rl.observe(U, nut, yg, nu, 2, dref, draws)
    ut = np.sqrt(nu * s_)
    for w in range(2):
        for a in range(2):
            yind = 3 if draws[w, a, 0] < 0.5 else 4
            yr = draws[w, a, 1]
            l = yr * yg[yind] + (1 - yr) * yg[yind + 1]   # distance from the wall (symmetric grid)
            assert abs(st[w, a, 3] - l) < 1e-13 and abs(st[w, a, 4] - ut) < 1e-13
            assert abs(st[w, a, 2] - s_ * l) < 1e-12            # linear interpolation of a linear profile
            assert abs(st[w, a, 0] - s_ * l / ut) < 1e-10
            assert abs(st[w, a, 1] - (s_ * l / ut - s_ * l / ut * np.log(l * ut / nu))) < 1e-9
            assert abs(rw[w, a] - (100 - 5000 * abs(ut - 0.1) / 0.1)) < 1e-9


def test_rl_actions_interpolate_between_agents():
    from oracle import rl
    nx, nzg = 14, 11
    alpha = np.full((nx, 2, nzg), 0.5)
    f = np.array([[1.0, 2.0, 4.0], [0.5, 1.5, 1.0]])
    out = rl.act(alpha, 3, f)
    nxa = (nx - 2) // 3
    for w in range(2):
        for a in range(3):
            # at each agent's face the action itself; halfway to the next, the mean
            assert np.allclose(out[a * nxa + 1, w, :], 0.5 * f[w, a], rtol=1e-15)
        assert np.allclose(out[0, w], out[nx - 2, w]) and np.allclose(out[nx - 1, w], out[1, w])
    # equal factors: uniform alpha scaled
    out = rl.act(alpha, 3, np.full((2, 3), 1.2))
    assert np.allclose(out, 0.6, rtol=1e-15)


# ------------------------------------------------------------ wall model

def test_wall_model_identity():
    o = mk(12, 11, 12, nu=1e-3)
    U, V, W = fields(o, lambda x, y, z: np.where(y < 1, y, 2 - y) * 4.0 + 0 * x, zero, zero)
    alpha = np.zeros(o.shape(oracle.ALPHA), order="F")
    alpha[:, 0, :] = 0.8
    alpha[:, 1, :] = 0.6
    o.set(oracle.ALPHA, alpha)
    nut = o.compute_wall_model(U, np.zeros(o.shape(oracle.NUT), order="F"))
    yg = centres(o.y)
    dyw = yg[1] - yg[0]
    dub = (U[:, 1, :] - U[:, 0, :]) / dyw
    dut = (U[:, o.nyg - 2, :] - U[:, o.nyg - 1, :]) / dyw
    assert np.allclose((1e-3 + nut[1:o.nxg - 1, 0, :]) * dub[0, 0], 0.8, rtol=1e-13)
    assert np.allclose((1e-3 + nut[1:o.nxg - 1, 1, :]) * dub[0, 0], 0.8, rtol=1e-13)
    assert np.allclose((1e-3 + nut[1:o.nxg - 1, o.nyg - 1, :]) * dut[0, 0], 0.6, rtol=1e-13)
    assert np.all(nut[:, 2:o.nyg - 2, :] == 0.0)


@pytest.mark.parametrize("yplus", [3.0, 8.0, 30.0, 300.0, 5000.0])
def test_loglaw_inverse(yplus):
    nu, h = 1e-4, 0.05
    ut = yplus * nu / h
    up = yplus if yplus < 11.0 else math.log(yplus) / 0.41 + 5.2
    got = oracle.loglaw_utau(up * ut, h, nu)
    assert math.isclose(got, ut, rel_tol=1e-12)


# ------------------------------------------------------------ Poisson / projection

@pytest.mark.parametrize("stretched", [False, True])
def test_poisson_manufactured_solution(stretched):
    o = mk(18, 17, 14, stretched=stretched)
    p, g = o.poisson(), o.grid()
    nxp, nzp, n = o.nx - 2, o.nz - 2, o.ny - 1
    ptrue = rng.standard_normal((nxp, n, nzp))
    dx, dz = g["dxmin"], g["dzmin"]
    lap = (np.roll(ptrue, -1, 0) - 2 * ptrue + np.roll(ptrue, 1, 0)) / dx ** 2
    lap += (np.roll(ptrue, -1, 2) - 2 * ptrue + np.roll(ptrue, 1, 2)) / dz ** 2
    A = np.diag(p["d"]) + np.diag(p["l"], -1) + np.diag(p["u"], 1)
    lap += np.einsum("jm,imk->ijk", A, ptrue)
    rhs = np.zeros(o.shape(oracle.P_), order="F")
    rhs[1:o.nxg - 2, 1:o.nyg - 1, 1:o.nzg - 2] = lap
    sol = o.solve_poisson_equation(rhs)
    got = sol[1:o.nxg - 2, 1:o.nyg - 1, 1:o.nzg - 2]
    # the 00-mode row scaling pins the plane mean of the first row to zero
    want = ptrue - ptrue[:, 0, :].mean()
    assert np.max(np.abs(got - want)) <= 1e-10 * np.max(np.abs(ptrue))
    assert abs(got[:, 0, :].mean()) <= 1e-12
    # periodic images
    assert np.array_equal(sol[o.nxg - 2, 1:o.nyg - 1, 1:o.nzg - 1], sol[1, 1:o.nyg - 1, 1:o.nzg - 1])
    assert np.array_equal(sol[1:o.nxg - 1, 1:o.nyg - 1, o.nzg - 2], sol[1:o.nxg - 1, 1:o.nyg - 1, 1])


@pytest.mark.parametrize("stretched", [False, True])
def test_projection_divergence_free_and_idempotent(stretched):
    c, f = cases.small_case(18, 17, 18, seed=12, stretched=stretched)
    o = oracle.Oracle(c.x, c.y, c.z, **c.kwargs())
    o.set_state(*f)
    o.apply_boundary_conditions()
    _, before = o.check_divergence()
    o.compute_projection_step()
    o.apply_boundary_conditions()
    _, after = o.check_divergence()
    assert before > 1.0 and after <= 1e-13 * before * 100
    s1 = o.state()
    o.compute_projection_step()
    o.apply_boundary_conditions()
    for a, b in zip(o.state(), s1):
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


# ------------------------------------------------------------ mass flow, dt

def test_mass_flow_of_uniform_flow_and_correction():
    o = mk(12, 11, 12, stretched=True)
    U, V, W = fields(o, lambda x, y, z: 2.5 + 0 * x, zero, zero)
    assert math.isclose(o.compute_mean_mass_flow_U(U), 2.5, rel_tol=1e-14)
    c, f = cases.small_case(12, 11, 12, seed=2)
    o = oracle.Oracle(c.x, c.y, c.z, **c.kwargs())
    U = np.asfortranarray(f[0])
    q0 = o.compute_mean_mass_flow_U(U)
    o.set_scalars(Qflow_x_0=q0)
    d = o.compute_dPx_for_constant_mass_flow(U)
    U2 = U.copy(order="F")
    U2[1:-1, 1:-1, 1:-1] += d
    # after the shift the reference functional (rows 3 and nyg-2) equals Qflow_x_0
    assert abs(o.compute_dPx_for_constant_mass_flow(U2)) <= 1e-13 * abs(q0)
    sel = U2[1:o.nx - 1, [2, o.nyg - 3], 1:o.nzg - 2]
    assert math.isclose(0.5 * sel.sum(axis=1).mean(), q0, rel_tol=1e-13)


def test_dt_rule():
    o = mk(12, 11, 12, stretched=True, nu=1e-3, CFL=0.5)
    g = o.grid()
    U, V, W = fields(o, lambda x, y, z: 2.0 + 0 * x, lambda x, y, z: 0.0 * x, lambda x, y, z: -4.0 + 0 * x)
    o.set_state(U, V, W)
    dxs = np.diff(g["xg"])[1:o.nxg - 1]
    dzs = np.diff(g["zg"])[1:o.nzg - 1]
    conv = min(0.5 * dxs.max() / 2.0, 0.5 * dzs.max() / 4.0)  # V = 0 gives +inf in y
    vis = 0.5 * min(g["dxmin"], g["dymin"], g["dzmin"]) ** 2 / 1e-3
    assert o.compute_dt() == min(conv, vis, 1e-3)
    o2 = mk(12, 11, 12, nu=1e-3, CFL=-0.004)
    o2.set_state(U, V, W)
    assert o2.compute_dt() == 0.004


# ------------------------------------------------------------ RK3

def test_rk3_amplification_of_advected_diffused_mode():
    nu, U0, k, A = 0.02, 0.8, 2.0, 0.5
    o = mk(34, 21, 10, nu=nu, CFL=-2e-3, LES_model=0)
    U, V, W = fields(o, lambda x, y, z: U0 + 0 * x, zero, lambda x, y, z: A * np.sin(k * x) + 0 * y + 0 * z)
    o.set_state(U, V, W)
    o.apply_boundary_conditions()
    # wall stress equal to the viscous one so the wall-model eddy viscosity is zero
    Ub = o.get(0)
    yg = centres(o.y)
    alpha = np.zeros(o.shape(oracle.ALPHA), order="F")
    alpha[:, 0, :] = nu * (Ub[:, 1, :] - Ub[:, 0, :]) / (yg[1] - yg[0])
    alpha[:, 1, :] = nu * (Ub[:, o.nyg - 2, :] - Ub[:, o.nyg - 1, :]) / (yg[1] - yg[0])
    o.set(oracle.ALPHA, alpha)
    dt = o.compute_dt()
    o.compute_time_step_RK3()
    W1 = o.get(2)
    dx = o.x[1] - o.x[0]
    lam = -1j * U0 * np.sin(k * dx) / dx - nu * 4 / dx ** 2 * np.sin(k * dx / 2) ** 2
    z = lam * dt
    R = 1 + z + z * z / 2 + z ** 3 / 6
    xg = centres(o.x)[1:o.nxg - 1]
    ref = (A * R * np.exp(1j * k * xg)).imag
    mid = slice(6, o.nyg - 6)  # rows the walls cannot reach in one step
    got = W1[1:o.nxg - 1, mid, 1:o.nz - 1]
    assert np.max(np.abs(got - ref[:, None, None])) <= 1e-12


# ------------------------------------------------------------ threading

def _threaded_outputs():
    """DSM and AMD eddy viscosities, right-hand sides, two RK3 steps with the
    wall model and mass-flow correction, and the pressure, of a small
    stretched channel."""
    out = {}
    for les in (2, 3):
        case, field = cases.small_case(26, 17, 22, seed=5, LES_model=les)
        o = oracle.Oracle(case.x, case.y, case.z, **case.kwargs())
        o.set_state(*field)
        o.apply_boundary_conditions()
        U, V, W = o.state()
        out[f"nut{les}"] = o.compute_eddy_viscosity(U, V, W)
        for q, r in enumerate(o.compute_rhs(U, V, W)):
            out[f"rhs{les}{q}"] = r
        o.set_scalars(dt=0.0, dPdx=case.dPdx, Qflow_x_0=0.5)
        o.set_wall_model(1, h_wm=0.1)
        for _ in range(2):
            o.compute_dt()
            o.compute_time_step_RK3()
        for q, a in enumerate(o.state()):
            out[f"state{les}{q}"] = a
        out[f"p{les}"] = o.compute_pressure()
    return out


def test_threaded_oracle_matches_one_thread(tmp_path):
    """The OpenMP loops write disjoint outputs by the serial operations, so
    the oracle on all threads equals the oracle on one thread bit for bit."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    path = str(tmp_path / "one_thread.npz")
    env = dict(os.environ, OMP_NUM_THREADS="1", PYTHONPATH=root + os.pathsep + os.path.join(root, "tests"))
    subprocess.check_call([sys.executable, os.path.abspath(__file__), path], env=env, timeout=300)
    one = np.load(path)
    many = _threaded_outputs()
    assert sorted(one.files) == sorted(many)
    for k in one.files:
        assert np.array_equal(one[k], many[k], equal_nan=True), k


if __name__ == "__main__":
    import sys
    assert oracle.num_threads() == 1
    np.savez(sys.argv[1], **_threaded_outputs())
