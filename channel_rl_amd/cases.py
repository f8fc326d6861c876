"""Channel cases: grids, physical parameters and seeded initial fields.

Grids follow the reference's face convention (code/input_output.f90:108-124):
periodic directions have faces x(1..nx) with x(1)=0, x(nx-1)=L and x(nx) the
image of x(2); the wall-normal faces run from y(1)=0 to y(ny)=2.  The
stretched wall-normal grid, the log-law and laminar initial fields and the
counter-based random numbers have no reference counterpart (the reference
draws unseeded rand(), code/input_output.f90:127-162).
"""
import math

import numpy as np

M64 = (1 << 64) - 1


def periodic_faces(n, L):
    """n faces, n-2 unique cells over [0, L]."""
    dx = L / (n - 2)
    return np.arange(n, dtype=np.float64) * dx


def uniform_y(ny, Ly=2.0):
    return np.linspace(0.0, Ly, ny)


def tanh_y(ny, gamma=1.5):
    """y_j = 1 + tanh(gamma (2j/(ny-1) - 1)) / tanh(gamma), j = 0..ny-1."""
    j = np.arange(ny, dtype=np.float64)
    y = 1.0 + np.tanh(gamma * (2.0 * j / (ny - 1) - 1.0)) / math.tanh(gamma)
    y[0], y[-1] = 0.0, 2.0
    return y


def uniform_noise(seed, stream, shape):
    """Counter-based uniform numbers in [-1, 1): splitmix64 of (seed, stream,
    Fortran linear index).  Identical on every host for a given seed."""
    n = int(np.prod(shape))
    with np.errstate(over="ignore"):
        idx = np.arange(n, dtype=np.uint64)
        z = (np.uint64(seed & M64) * np.uint64(0x9E3779B97F4A7C15)
             + np.uint64((stream * 0xD1B54A32D192ED03) & M64) + idx * np.uint64(0xBF58476D1CE4E5B9))
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (2.0 * u - 1.0).reshape(shape, order="F")


def bulk_velocity_plus(re_tau):
    """U_b/u_tau from the correlation Re_tau = 0.09 Re_b^0.88, Re_b = 2 delta U_b / nu."""
    re_b = (re_tau / 0.09) ** (1.0 / 0.88)
    return re_b / (2.0 * re_tau)


def centres_with_ghosts(f):
    m = 0.5 * (f[:-1] + f[1:])
    g = np.empty(len(m) + 2)
    g[1:-1] = m
    g[0] = m[0] - 2.0 * (m[0] - f[0])
    g[-1] = m[-1] + 2.0 * (f[-1] - m[-1])
    return g


def loglaw_plus(yplus, kappa=0.41, B=5.2):
    yp = np.maximum(yplus, 1e-300)
    return np.where(yp < 11.0, yp, np.log(yp) / kappa + B)


class Case:
    """Parameters of one channel plus its raw (not yet projected) initial field."""

    def __init__(self, name, nx, ny, nz, Lx, Lz, y, nu, dPdx, CFL, x_mass_cte=1, LES_model=2,
                 Dirichlet_nu_t=1, wall="equilibrium", h_wm=0.0, steps=10):
        self.name = name
        self.nx, self.ny, self.nz = nx, ny, nz
        self.x = periodic_faces(nx, Lx)
        self.z = periodic_faces(nz, Lz)
        self.y = np.asarray(y, dtype=np.float64)
        self.nu, self.dPdx, self.CFL = nu, dPdx, CFL
        self.x_mass_cte, self.LES_model, self.Dirichlet_nu_t = x_mass_cte, LES_model, Dirichlet_nu_t
        self.wall, self.h_wm, self.steps = wall, h_wm, steps
        self.nxg, self.nyg, self.nzg = nx + 1, ny + 1, nz + 1
        self.wm_nu_limit = 0.0  # the wall model's viscosity limit (0: the reference's formula)

    @property
    def cells(self):
        return (self.nx - 2) * (self.ny - 1) * (self.nz - 2)

    def kwargs(self):
        return dict(nu=self.nu, CFL=self.CFL, dPdx=self.dPdx, x_mass_cte=self.x_mass_cte,
                    LES_model=self.LES_model, Dirichlet_nu_t=self.Dirichlet_nu_t)

    def field_shapes(self):
        return ((self.nx, self.nyg, self.nzg), (self.nxg, self.ny, self.nzg), (self.nxg, self.nyg, self.nz))

    def initial_field(self, profile, ub, amp, seed, local=False):
        """Mean profile (callable of wall distance) on U plus uniform noise on
        every component; arrays in the reference's shapes.  The noise is
        +-amp*ub everywhere, or with local=True +-amp times the mean profile
        at each value's own wall distance (a fluctuation of amp relative to
        the local mean, vanishing at the walls like the mean itself)."""
        yg = centres_with_ghosts(self.y)
        su, sv, sw = self.field_shapes()
        dist = np.minimum(yg, 2.0 - yg)
        mean = profile(dist)
        U = np.zeros(su, order="F")
        U[:, :, :] = mean[None, :, None]
        if local:
            dv = np.minimum(self.y, 2.0 - self.y)
            au, av, aw = amp * mean[None, :, None], amp * profile(dv)[None, :, None], amp * mean[None, :, None]
        else:
            au = av = aw = amp * ub
        U += au * uniform_noise(seed, 1, su)
        V = av * uniform_noise(seed, 2, sv)
        W = aw * uniform_noise(seed, 3, sw)
        return np.asfortranarray(U), np.asfortranarray(V), np.asfortranarray(W)


def config(name, seed=1234):
    """The benchmark configurations: returns (Case, (U, V, W) raw initial field).

    small      : 32x33x32 (faces 34 33 34), 2pi x 2 x pi, Re_tau 180, tanh y (1.5),
                 laminar parabola + 5% noise, DSM, equilibrium wall model at the
                 first off-wall cell, CFL 0.5, constant mass flow.
    wmles      : 192x97x192 (faces 194 97 194), Re_tau 2000 in the reference's
                 units (nu, dPdx of input_parameters_2000), log-law + 10% noise
                 (relative to the local mean, as in `large` and the ensemble),
                 DSM, equilibrium wall model at h_wm = 0.1.
    large      : 384x193x384 (faces 386 193 386), 4pi x 2 x 2pi, Re_tau 5200.
    """
    if name == "small":
        re_tau = 180.0
        c = Case(name, 34, 33, 34, 2 * math.pi, math.pi, tanh_y(33, 1.5), nu=1.0 / re_tau, dPdx=1.0, CFL=0.5,
                 h_wm=0.0, steps=100)
        ub = bulk_velocity_plus(re_tau)
        c.re_tau = re_tau
        f = c.initial_field(lambda d: 1.5 * ub * d * (2.0 - d) / 1.0, ub, 0.05, seed)
        return c, f
    if name == "wmles":
        nu, dpdx = 2.061855670103093e-05, 0.001705857682121
        utau = math.sqrt(dpdx)
        re_tau = utau / nu
        c = Case(name, 194, 97, 194, 2 * math.pi, math.pi, tanh_y(97, 1.5), nu=nu, dPdx=dpdx, CFL=0.5,
                 h_wm=0.1, steps=2000)
        ub = bulk_velocity_plus(re_tau) * utau
        c.re_tau = re_tau
        f = c.initial_field(lambda d: utau * loglaw_plus(d * utau / nu), ub, 0.10, seed, local=True)
        return c, f
    if name == "large":
        re_tau = 5200.0
        nu = 1.0 / re_tau
        c = Case(name, 386, 193, 386, 4 * math.pi, 2 * math.pi, tanh_y(193, 1.5), nu=nu, dPdx=1.0, CFL=0.5,
                 h_wm=0.1, steps=500)
        ub = bulk_velocity_plus(re_tau)
        c.re_tau = re_tau
        f = c.initial_field(lambda d: loglaw_plus(d / nu), ub, 0.10, seed, local=True)
        return c, f
    raise KeyError(name)


ENSEMBLE_GRIDS = ((34, 33, 34), (66, 49, 66), (98, 65, 98))


def ensemble(n=64, seed=42, field_seed=1234):
    """The batched ensemble: n channels, 2pi x 2 x pi, grids (faces) drawn
    uniformly from ENSEMBLE_GRIDS and Re_tau log-uniform in [180, 5200] with
    the seeded generator `seed`; nu = 1/Re_tau (u_tau = 1 units), constant
    bulk flow (mass-flow mode 2: the reference's near-wall functional lets the
    bulk of these channels run away within a few hundred steps), the wall
    model's viscosity held below a quarter of the explicit diffusion bound
    (wm_nu_limit: where the resolved wall gradient of the finer low-Re_tau
    grids nearly vanishes, the reference's formula asks for a viscosity the
    explicit step cannot take), DSM, wall model at h_wm = 0.1, initial field
    log law + 10%
    noise relative to the local mean (noise seed field_seed + channel index):
    noise of 10% of the bulk velocity flips the sign of the wall gradient in
    the viscous sublayer of the low-Re_tau fine grids, where the wall model's
    viscosity alpha/(dU/dy) - nu then turns negative and the run diverges
    within a few steps.  Returns [(Case, field)]."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        nx, ny, nz = ENSEMBLE_GRIDS[int(rng.integers(len(ENSEMBLE_GRIDS)))]
        re_tau = float(np.exp(rng.uniform(np.log(180.0), np.log(5200.0))))
        nu = 1.0 / re_tau
        c = Case(f"ensemble{q}", nx, ny, nz, 2 * math.pi, math.pi, tanh_y(ny, 1.5), nu=nu, dPdx=1.0, CFL=0.5,
                 x_mass_cte=2, h_wm=0.1, steps=1000)
        c.re_tau = re_tau
        c.wm_nu_limit = 0.25
        ub = bulk_velocity_plus(re_tau)
        f = c.initial_field(lambda d, nu=nu: loglaw_plus(d / nu), ub, 0.10, field_seed + q, local=True)
        out.append((c, f))
    return out


def small_case(nx, ny, nz, seed=7, stretched=True, LES_model=2, x_mass_cte=1, CFL=0.5, amp=0.1):
    """A test-sized channel with the small configuration's physics."""
    re_tau = 180.0
    y = tanh_y(ny, 1.5) if stretched else uniform_y(ny)
    c = Case(f"test{nx}x{ny}x{nz}", nx, ny, nz, 2 * math.pi, math.pi, y, nu=1.0 / re_tau, dPdx=1.0, CFL=CFL,
             x_mass_cte=x_mass_cte, LES_model=LES_model, h_wm=0.0, steps=10)
    ub = bulk_velocity_plus(re_tau)
    f = c.initial_field(lambda d: 1.5 * ub * d * (2.0 - d), ub, amp, seed)
    return c, f
