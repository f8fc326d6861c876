"""GPU parity: every C-ABI entry point against the FP64 oracle on the same
seeded inputs.

Tolerances (FP64): stencil stages whose arithmetic is the reference's own
operation order (boundary conditions, wall model, dt) must agree to 1e-14
relative and are expected bit-for-bit; the RHS multiplies by reciprocal
spacings where the reference divides (a few ulp per term), also 1e-14
relative to the field's maximum; stages with reordered sums or a
different algorithm (DSM plane sums, FFT/Thomas Poisson, mass flow) to
1e-12 relative; whole RK3 steps to 1e-11 relative after 1 step and 1e-9
after 10 steps (round-off growth through the dynamic model).
"""
import numpy as np
import pytest

from channel_rl_amd import FP64, Channel, cases
from channel_rl_amd import NU_T, P, ALPHA_X
from helpers import interior, oracle_for, projected_state, rel

pytestmark = pytest.mark.gpu

# periodic lengths nx-2, nz-2: 8, 16, 32 (radix 4 and 2), 20 and 30 (radix 5,
# 3 and 2), so every FFT pass the kernels have runs against the oracle
SIZES = [(10, 9, 10), (18, 17, 18), (34, 33, 34), (22, 13, 32)]


def make_pair(nx, ny, nz, seed=5, **kw):
    case, field = cases.small_case(nx, ny, nz, seed=seed, **kw)
    U, V, W = projected_state(case, field)
    o = oracle_for(case)
    o.set_state(U, V, W)
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    g.set_state(U, V, W)
    return case, o, g, (U, V, W)


def random_nut(case, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.uniform(0.0, 2e-3, (case.nxg, case.nyg, case.nzg)))


@pytest.mark.parametrize("n", SIZES)
def test_boundary_conditions_bitwise(gpu, n):
    case, o, g, _ = make_pair(*n)
    rng = np.random.default_rng(1)
    raw = [np.asfortranarray(rng.standard_normal(s)) for s in case.field_shapes()]
    o.set_state(*raw)
    g.set_state(*raw)
    o.apply_boundary_conditions()
    g.apply_boundary_conditions()
    for a, b in zip(g.state(), o.state()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("n", SIZES)
def test_rhs_matches_oracle(gpu, n):
    case, o, g, (U, V, W) = make_pair(*n)
    nut = random_nut(case, 3)
    o.set(NU_T, nut)
    g.set_field(NU_T, nut)
    o.set_scalars(dPdx=0.37)
    g.set_scalars(dPdx=0.37)
    ru, rv, rw = o.compute_rhs(U, V, W)
    gu, gv, gw = g.compute_rhs()
    for q, (gr, orr) in enumerate(((gu, ru), (gv, rv), (gw, rw))):
        ref = interior(q, orr, case)
        assert gr.shape == ref.shape
        assert rel(gr, ref) <= 1e-14, f"rhs component {q}: rel err {rel(gr, ref)}"


@pytest.mark.parametrize("n", SIZES)
def test_eddy_viscosity_matches_oracle(gpu, n):
    case, o, g, (U, V, W) = make_pair(*n)
    ref = o.compute_eddy_viscosity(U, V, W)
    g.compute_eddy_viscosity()
    got = g.get_field(NU_T)
    assert np.all(np.isfinite(got))
    assert rel(got, ref) <= 1e-12, rel(got, ref)


@pytest.mark.parametrize("n", SIZES)
def test_wall_model_matches_oracle(gpu, n):
    case, o, g, (U, V, W) = make_pair(*n)
    nut = random_nut(case, 4)
    alpha = o.equilibrium_alpha(U)
    g.set_field(ALPHA_X, alpha)
    g.set_field(NU_T, nut)
    ref = o.compute_wall_model(U, nut)
    g.compute_wall_model()
    got = g.get_field(NU_T)
    assert rel(got, ref) <= 1e-14, rel(got, ref)


@pytest.mark.parametrize("n", SIZES)
def test_equilibrium_wall_stress(gpu, n):
    case, o, g, (U, V, W) = make_pair(*n)
    for h in (0.0, 0.1):
        ref = o.equilibrium_alpha(U, h_wm=h)
        g.compute_equilibrium_wall_stress(h_wm=h)
        got = g.get_field(ALPHA_X)
        assert rel(got, ref) <= 1e-13, (h, rel(got, ref))


@pytest.mark.parametrize("n", SIZES + [(10, 4, 10), (10, 5, 10), (10, 50, 12), (10, 80, 10), (12, 98, 10),
                                       (10, 130, 12), (10, 193, 10), (10, 194, 10)])
def test_projection_matches_oracle(gpu, n):
    """The y solve of ny - 1 rows: up to 192 two-ended in registers (3 and 4
    rows, the smallest systems; 49 and 79, odd, where the bottom half is a
    row shorter; 97 and 129, the 64- and 96-row halves, partly filled; 192,
    the 96-row halves full; the small sizes), more streamed through memory
    before the last 32 (193 rows: 161 streamed, in groups of 8)."""
    case, field = cases.small_case(*n, seed=11)
    o = oracle_for(case)
    o.set_state(*field)
    o.apply_boundary_conditions()
    U, V, W = o.state()
    g = Channel(case.x, case.y, case.z, **case.kwargs())
    g.set_state(U, V, W)
    o.compute_projection_step()
    o.apply_boundary_conditions()
    g.compute_projection_step()
    for a, b in zip(g.state(), o.state()):
        assert rel(a, b) <= 1e-12, rel(a, b)
    p_ref = o.get(P)[1:-1, 1:-1, 1:-1]
    p_got = g.get_field(P)[1:-1, 1:-1, 1:-1]
    assert rel(p_got, p_ref) <= 1e-11, rel(p_got, p_ref)
    md, ma = g.check_divergence()
    scale = np.max(np.abs(field[0])) / (case.x[1] - case.x[0])
    assert ma <= 1e-12 * scale, (ma, scale)


@pytest.mark.parametrize("n", SIZES)
def test_dt_exact(gpu, n):
    case, o, g, _ = make_pair(*n, CFL=0.5)
    assert g.compute_dt() == o.compute_dt()


@pytest.mark.parametrize("n", SIZES)
def test_mass_flow_matches_oracle(gpu, n):
    case, o, g, (U, V, W) = make_pair(*n)
    q_ref = o.compute_mean_mass_flow_U(U)
    q_got = g.compute_mean_mass_flow_U()
    assert abs(q_got - q_ref) <= 1e-12 * abs(q_ref)  # sums of ~1e4 terms in another order
    o.set_scalars(Qflow_x_0=q_ref)
    g.set_scalars(Qflow_x_0=q_ref)
    d_ref = o.compute_dPx_for_constant_mass_flow(U)
    d_got = g.compute_dPx_for_constant_mass_flow()
    assert abs(d_got - d_ref) <= 1e-12 * abs(q_ref)


def run_pair(nx, ny, nz, steps, precision=FP64, seed=5, **kw):
    case, o, g, (U, V, W) = make_pair(nx, ny, nz, seed=seed, **kw)
    if precision != FP64:
        g = Channel(case.x, case.y, case.z, precision=precision, **case.kwargs())
        g.set_state(U, V, W)
    q0 = o.compute_mean_mass_flow_U(U)
    o.set_scalars(dt=0.0, t=0.0, dPdx=case.dPdx, Qflow_x_0=q0)
    g.set_scalars(dt=0.0, t=0.0, dPdx=case.dPdx, Qflow_x_0=q0)
    g.set_wall_model(1, h_wm=case.h_wm)
    o.set_wall_model(1, h_wm=case.h_wm)
    for _ in range(steps):
        o.compute_dt()
        o.compute_time_step_RK3()
        g.compute_dt(sync=False)
        g.compute_time_step_RK3()
    return case, o, g


def test_rk3_one_step(gpu):
    case, o, g = run_pair(18, 17, 18, 1)
    for a, b in zip(g.state(), o.state()):
        assert rel(a, b) <= 1e-11, rel(a, b)
    assert abs(g.scalars()["t"] - o.scalars()["t"]) <= 1e-15


def test_rk3_ten_steps_small_config(gpu):
    # 32x33x32 cells (faces 34 33 34): the small configuration, 10 steps
    case, o, g = run_pair(34, 33, 34, 10)
    for a, b in zip(g.state(), o.state()):
        assert rel(a, b) <= 1e-9, rel(a, b)
    _, ma = g.check_divergence()
    assert ma <= 1e-10, ma


@pytest.mark.parametrize("precision", [FP64, "fp32"])
def test_rhs_tile_shapes_agree_bitwise(gpu, precision):
    # every tile shape of the fused RHS + RK kernel (tile, z chunk) does the
    # same arithmetic per cell: one RK3 step from the same state must agree
    # bit for bit; the autotuner leaves the state alone and picks one of the
    # shapes
    from channel_rl_amd import FP32_BULK, load_library, KERNEL_RHS
    prec = FP64 if precision == FP64 else FP32_BULK
    n = load_library().chan_tile_shape_count(KERNEL_RHS)
    assert n >= 2
    case, field = cases.small_case(34, 33, 34, seed=11)
    U, V, W = projected_state(case, field)
    out = []
    for shape in range(n):
        g = Channel(case.x, case.y, case.z, precision=prec, **case.kwargs())
        g.set_state(U, V, W)
        g.set_scalars(dt=0.0, t=0.0, dPdx=case.dPdx, Qflow_x_0=g.compute_mean_mass_flow_U())
        g.set_wall_model(1, h_wm=case.h_wm)
        g.set_tile_shape(shape)
        assert g.tile_shape() == shape
        g.compute_dt(sync=False)
        g.compute_time_step_RK3()
        out.append(g.state())
    for shape in range(1, n):
        for a, b in zip(out[shape], out[0]):
            assert np.array_equal(a, b), f"shape {shape}"
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    g.set_state(U, V, W)
    before = g.state()
    t = g.autotune(reps=2)["rhs"]
    assert t.shape == (n,) and np.all(t > 0)
    assert 0 <= g.tile_shape() < n and t[g.tile_shape()] == t.min()
    for a, b in zip(g.state(), before):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [FP64, "fp32"])
def test_dsm_tile_shapes_match_oracle(gpu, precision):
    """Every tile shape of the DSM pass A (30 or 14 outputs per row, in
    blocks of 128, 256 or 512 threads) gives the oracle's eddy viscosity (FP64: 1e-12; FP32 bulk: 1e-3 of its
    maximum, the FP32 rounding of the inputs through the differences of
    filtered products in L_ij and M_ij: 1.5e-4 measured), on grids whose rows
    fill the tiles unevenly; the autotuner
    picks the fastest DSM shape as well and leaves the state alone."""
    from channel_rl_amd import FP32_BULK, KERNEL_DSM, load_library
    prec = FP64 if precision == FP64 else FP32_BULK
    tol = 1e-12 if prec == FP64 else 1e-3
    n = load_library().chan_tile_shape_count(KERNEL_DSM)
    assert n == 4
    for dims in ((34, 33, 34), (22, 13, 32)):
        case, o, g, (U, V, W) = make_pair(*dims)
        ref = o.compute_eddy_viscosity(U, V, W)
        for shape in range(n):
            g = Channel(case.x, case.y, case.z, precision=prec, **case.kwargs())
            g.set_state(U, V, W)
            g.set_tile_shape(shape, KERNEL_DSM)
            assert g.tile_shape(KERNEL_DSM) == shape
            g.compute_eddy_viscosity()
            got = g.get_field(NU_T)
            assert rel(got, ref) <= tol, (dims, shape, rel(got, ref))
        before = g.state()
        t = g.autotune(reps=2)["dsm"]
        assert t.shape == (n,) and np.all(t > 0) and t[g.tile_shape(KERNEL_DSM)] == t.min()
        for a, b in zip(g.state(), before):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("n,dirichlet", [((10, 9, 10), 1), ((18, 17, 18), 0), ((34, 33, 34), 2)])
def test_amd_matches_oracle(gpu, n, dirichlet):
    case, field = cases.small_case(*n, seed=6, LES_model=3)
    case.Dirichlet_nu_t = dirichlet
    U, V, W = projected_state(case, field)
    o = oracle_for(case)
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    g.set_state(U, V, W)
    o.set_scalars(dPdx=0.36)
    g.set_scalars(dPdx=0.36)
    ref = o.compute_eddy_viscosity(U, V, W)
    g.compute_eddy_viscosity()
    got = g.get_field(NU_T)
    assert np.all(np.isfinite(got)) and np.abs(ref).max() > 0
    assert rel(got, ref) <= 1e-12, rel(got, ref)


def test_autotune_without_the_dynamic_model(gpu):
    """A channel with the AMD model (LES_model 3) has no DSM pass and no |S|
    buffer: the autotuner times the RHS shapes only, reports 0 for the DSM
    shapes and leaves the state and the step as they were."""
    from channel_rl_amd import KERNEL_DSM, load_library
    case, o, g, (U, V, W) = make_pair(18, 17, 18, LES_model=3)
    before = g.state()
    t = g.autotune(reps=1)
    assert np.all(t["rhs"] > 0)
    assert t["dsm"].shape == (load_library().chan_tile_shape_count(KERNEL_DSM),) and np.all(t["dsm"] == 0)
    for a, b in zip(g.state(), before):
        assert np.array_equal(a, b)


def test_rk3_with_amd(gpu):
    case, o, g = run_pair(18, 17, 18, 3, LES_model=3)
    for a, b in zip(g.state(), o.state()):
        assert rel(a, b) <= 1e-10, rel(a, b)


@pytest.mark.parametrize("n", SIZES)
def test_true_pressure_matches_oracle(gpu, n):
    case, o, g, (U, V, W) = make_pair(*n)
    nut = random_nut(case, 4)
    o.set(NU_T, nut)
    g.set_field(NU_T, nut)
    o.set_scalars(dPdx=0.3)
    g.set_scalars(dPdx=0.3)
    ref = o.compute_pressure()
    got = g.compute_pressure()
    assert np.all(np.isfinite(got))
    assert rel(got, ref) <= 1e-11, rel(got, ref)


@pytest.mark.parametrize("n", SIZES)
def test_check_divergence_matches_oracle(gpu, n):
    """code/projection.f90:223-245: the signed max of the divergence (the
    reference starts it at 0) and the max of its magnitude, each a max over
    the same per-cell expression: bit for bit on unprojected random fields,
    whose divergence is of order one everywhere."""
    case, o, g, _ = make_pair(*n)
    rng = np.random.default_rng(21)
    raw = [np.asfortranarray(rng.standard_normal(s)) for s in case.field_shapes()]
    o.set_state(*raw)
    g.set_state(*raw)
    ref = o.check_divergence()
    assert ref[0] > 0.0
    assert g.check_divergence() == ref
    # dV/dy = -1 everywhere (V = -y on the faces, U = W = 0): the signed max
    # stays at its start value 0, the magnitude is 1
    zu, zw = np.zeros(case.field_shapes()[0], order="F"), np.zeros(case.field_shapes()[2], order="F")
    V = np.asfortranarray(np.broadcast_to(-case.y[None, :, None], case.field_shapes()[1]))
    o.set_state(zu, V, zw)
    g.set_state(zu, V, zw)
    md, ma = g.check_divergence()
    assert (md, ma) == o.check_divergence()
    assert md == 0.0 and abs(ma - 1.0) <= 1e-13


@pytest.mark.parametrize("n", SIZES + [(12, 98, 10), (10, 130, 12), (10, 193, 10), (10, 194, 10)])
def test_solve_poisson_equation_matches_oracle(gpu, n):
    """code/projection.f90:74-128 on a random right-hand side (the 00 mode
    incompatible, as the solve does not require compatibility: its D(2)*1.5
    row makes the system regular): the solution at every centre the oracle
    writes, 1e-11 relative (FFT and Thomas against a recursive FFT and
    zgtsv-equivalent elimination in another summation order)."""
    case, o, g, _ = make_pair(*n)
    rng = np.random.default_rng(31)
    rhs = np.asfortranarray(rng.standard_normal((case.nxg, case.nyg, case.nzg)))
    ref = o.solve_poisson_equation(rhs)
    g.set_field(P, rhs)
    g.solve_poisson_equation()
    got = g.get_field(P)
    sl = (slice(1, case.nxg - 1), slice(1, case.nyg - 1), slice(1, case.nzg - 1))
    assert np.all(np.isfinite(got[sl]))
    assert rel(got[sl], ref[sl]) <= 1e-11, rel(got[sl], ref[sl])


def test_rk3_hundred_steps_config1(gpu):
    """The small configuration as specified (32x33x32 cells, faces 34 33 34,
    Re_tau 180, laminar parabola + 5% seeded noise, DSM, equilibrium wall
    model at the first off-wall cell, CFL 0.5, constant mass flow) for 100
    RK3 steps against the oracle: fields within 1e-12 relative (round-off of
    the reordered DSM plane sums and FFT/Thomas grows through the dynamic
    model; the growth is printed at 10, 50 and 100 steps), time and dt
    within 1e-12, and the divergence at round-off."""
    case, field = cases.config("small", seed=1234)
    U, V, W = projected_state(case, field)
    o = oracle_for(case)
    o.set_state(U, V, W)
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    g.set_state(U, V, W)
    q0 = o.compute_mean_mass_flow_U(U)
    for x in (o, g):
        x.set_scalars(dt=0.0, t=0.0, dPdx=case.dPdx, Qflow_x_0=q0)
        x.set_wall_model(1, h_wm=case.h_wm)
    errs = {}
    for s in range(1, 101):
        o.compute_dt()
        o.compute_time_step_RK3()
        g.compute_dt(sync=False)
        g.compute_time_step_RK3()
        if s in (10, 50, 100):
            errs[s] = max(rel(a, b) for a, b in zip(g.state(), o.state()))
    print(f"config 1 field error after 10/50/100 steps: {errs[10]:.2e} {errs[50]:.2e} {errs[100]:.2e}")
    assert errs[100] <= 1e-12, errs
    so, sg = o.scalars(), g.scalars()
    assert abs(sg["t"] - so["t"]) <= 1e-12 * so["t"] and abs(sg["dt"] - so["dt"]) <= 1e-12 * so["dt"]
    _, ma = g.check_divergence()
    assert ma <= 1e-10, ma


def test_rk3_constant_bulk_flow_matches_oracle(gpu):
    """Mass-flow mode 2 (constant bulk flow, no reference counterpart): five
    steps against the oracle at 1e-11, and the bulk mean of
    code/mass_flow.f90:24-62 held at Qflow_x_0 to round-off."""
    case, o, g = run_pair(34, 33, 34, 5, x_mass_cte=2)
    for a, b in zip(g.state(), o.state()):
        assert rel(a, b) <= 1e-11, rel(a, b)
    q0 = g.scalars()["Qflow_x_0"]
    assert abs(g.compute_mean_mass_flow_U() - q0) <= 1e-13 * abs(q0)
    assert abs(o.compute_mean_mass_flow_U(o.state()[0]) - q0) <= 1e-13 * abs(q0)


def test_wall_viscosity_limit(gpu):
    """chan_set_wall_viscosity_limit (no reference counterpart): with limit L
    each wall face's total viscosity alpha/(dU/dy) is held in
    [0, L (yg(2)-yg(1))^2 / dt] before the x average of
    code/wallmodel.f90:37; L = 0 is the reference's formula (the oracle's)."""
    case, o, g, (U, V, W) = make_pair(18, 17, 18)
    rng = np.random.default_rng(8)
    U = np.array(U, order="F")
    # reversed, vanishing and ordinary wall gradients at both walls
    U[:, 1, :] *= rng.choice([-1.0, 1e-6, 1.0], size=(case.nx, case.nzg))
    U[:, 0, :] = -U[:, 1, :]
    U[:, -2, :] *= rng.choice([-1.0, 1e-6, 1.0], size=(case.nx, case.nzg))
    U[:, -1, :] = -U[:, -2, :]
    alpha = np.asfortranarray(rng.uniform(0.5, 2.0, (case.nx, 2, case.nzg)))
    nut = random_nut(case, 5)
    g.set_state(U, V, W)
    g.set_field(ALPHA_X, alpha)
    dt, lim = 1e-3, 0.25
    g.set_scalars(dt=dt)
    g.set_field(NU_T, nut)
    g.compute_wall_model()
    o.set(ALPHA_X, alpha)
    assert rel(g.get_field(NU_T), o.compute_wall_model(U, nut)) <= 1e-14   # limit 0: the reference
    g.set_wall_viscosity_limit(lim)
    g.set_field(NU_T, nut)
    g.compute_wall_model()
    got = g.get_field(NU_T)
    yg = o.grid()["yg"]
    dyw = yg[1] - yg[0]
    numax = lim * dyw * dyw / dt
    tot = np.empty((case.nx, 2, case.nzg))
    tot[:, 0, :] = alpha[:, 0, :] / ((U[:, 1, :] - U[:, 0, :]) / dyw)
    tot[:, 1, :] = alpha[:, 1, :] / ((U[:, -2, :] - U[:, -1, :]) / dyw)
    tot = np.where(tot >= 0.0, np.minimum(tot, numax), 0.0) - case.nu
    cen = 0.5 * (tot[:-1] + tot[1:])                 # x centres 2..nxg-1
    # centres 1..nxg from their canonical cells: 1 <- nxg-2, nxg-1 <- 2, nxg <- 3
    ref = np.concatenate([cen[-2:-1], cen[:-1], cen[:2]])
    for row, wall in ((0, 0), (1, 0), (-2, 1), (-1, 1)):
        assert np.max(np.abs(got[:, row, :] - ref[:, wall, :])) <= 1e-14 * numax
    assert np.array_equal(got[:, 2:-2, :], g.get_field(NU_T)[:, 2:-2, :])


@pytest.mark.parametrize("precision", [FP64, "fp32"])
def test_run_steps_folds_compute_dt_bitwise(gpu, precision):
    """run_steps with the dynamic model folds compute_dt into the first DSM
    pass of the step (per-column minima of |U| and |V|, per-plane maxima of
    dz/|W|, a last block of the ratio kernel): three steps give the same dt
    and fields, bit for bit, as compute_dt + compute_time_step_RK3, for every
    DSM tile shape.  The fields are shifted so that the convective limit
    binds (a small CFL and viscosity keep the viscous limit and the cap
    away), once through U and once through W."""
    from channel_rl_amd import FP32_BULK, KERNEL_DSM, WALL_EQUILIBRIUM, load_library
    prec = FP64 if precision == FP64 else FP32_BULK
    for u0, w0 in ((3.0, 50.0), (50.0, 3.0)):
        case, field = cases.small_case(22, 13, 32, seed=9)
        case.nu, case.CFL = 1e-4, 1e-3
        rng = np.random.default_rng(4)
        U = np.asfortranarray(u0 * (1.0 + 0.05 * rng.uniform(-1, 1, field[0].shape)))
        V = np.asfortranarray(0.1 * rng.uniform(-1, 1, field[1].shape))
        W = np.asfortranarray(w0 * (1.0 + 0.05 * rng.uniform(-1, 1, field[2].shape)))
        for shape in range(load_library().chan_tile_shape_count(KERNEL_DSM)):
            chans = []
            for _ in range(2):
                g = Channel(case.x, case.y, case.z, precision=prec, **case.kwargs())
                g.set_state(U, V, W)
                g.apply_boundary_conditions()
                g.compute_projection_step()
                g.set_scalars(dt=0.0, t=0.0, dPdx=0.0, Qflow_x_0=0.0)
                g.set_wall_model(WALL_EQUILIBRIUM, h_wm=0.1)
                g.set_tile_shape(shape, KERNEL_DSM)
                chans.append(g)
            a, b = chans
            for step in range(3):
                dt_a = a.compute_dt()
                a.compute_time_step_RK3()
                b.run_steps(1)
                dt_b = b.scalars()["dt"]
                assert dt_a == dt_b, (u0, w0, shape, step, dt_a, dt_b)
                assert dt_a < 0.5 * case.CFL * (case.y[1] - case.y[0]) ** 2 / case.nu  # convective limit binds
            for x, y in zip(a.state(), b.state()):
                assert np.array_equal(x, y), (u0, w0, shape)
            del a, b, chans


@pytest.mark.parametrize("precision", [FP64, "fp32"])
def test_deferred_dsm_pass_b_bitwise(gpu, precision, monkeypatch):
    """The step's first two substeps leave the dynamic model's pass B
    (code/subgrid.f90:512-535 with the wall rows of code/wallmodel.f90:39-42)
    to the RHS kernel: it reads |S| with per-row coefficients from the ratio
    launch, and a small kernel completes |S|'s wall rows and periodic images.
    Three steps give the same fields, nu_t and P, bit for bit, as with pass B
    materialising nu_t in every substep (CHANLES_NO_DEFERRED_NUT=1), for the
    equilibrium law, for a fixed alpha_x and for the learned wall model
    (an untrained network, checked against the law), on a grid whose tiles do
    not divide the box and for every RHS tile shape."""
    from channel_rl_amd import FP32_BULK, KERNEL_RHS, WALL_EQUILIBRIUM, WALL_MLP, WallNetwork, load_library
    prec = FP64 if precision == FP64 else FP32_BULK
    case, field = cases.small_case(22, 13, 34, seed=11)
    nshape = load_library().chan_tile_shape_count(KERNEL_RHS)
    net = WallNetwork(seed=3)
    for wall in ("equilibrium", "fixed", "network"):
        for shape in range(0, nshape, 3 if wall != "network" else 5):
            chans = []
            for deferred in (False, True):
                if deferred:
                    monkeypatch.delenv("CHANLES_NO_DEFERRED_NUT", raising=False)
                else:
                    monkeypatch.setenv("CHANLES_NO_DEFERRED_NUT", "1")
                g = Channel(case.x, case.y, case.z, precision=prec, **case.kwargs())
                g.set_state(*field)
                g.apply_boundary_conditions()
                g.compute_projection_step()
                g.set_scalars(dt=0.0, t=0.0, dPdx=0.0, Qflow_x_0=0.0)
                if wall == "equilibrium":
                    g.set_wall_model(WALL_EQUILIBRIUM, h_wm=0.1)
                elif wall == "network":
                    g.set_wall_network(net, max_rel_err=0.05)
                    g.set_wall_model(WALL_MLP, h_wm=0.1)
                else:
                    g.set_field(ALPHA_X, np.full(g.shape(ALPHA_X), 2e-4))
                g.set_tile_shape(shape, KERNEL_RHS)
                chans.append(g)
            a, b = chans
            a.run_steps(3)
            b.run_steps(3)
            for x, y in zip(a.state(), b.state()):
                assert np.array_equal(x, y), (wall, shape)
            for f in (NU_T, P, ALPHA_X):
                assert np.array_equal(a.get_field(f), b.get_field(f)), (wall, shape, f)
            assert a.scalars() == b.scalars()
            a.close()
            b.close()
            del a, b, chans


@pytest.mark.parametrize("precision", [FP64, "fp32"])
def test_amd_step_folds_compute_dt_and_wall_stress_bitwise(gpu, precision):
    """With the AMD model (LES_model 3), run_steps folds compute_dt into the
    first substep's AMD kernel (per-column minima of |U| and |V|, per-plane
    maxima of dz/|W|, the terms of the image column and plane included) and
    a last block of the wall rows' launch, which also computes the
    equilibrium wall stress before the wall rows: three steps give the same
    dt and fields, bit for bit, as compute_dt + compute_time_step_RK3, with
    the convective limit binding once through U and once through W (the
    folded wall stress itself is checked against the oracle by
    test_rk3_with_amd)."""
    from channel_rl_amd import FP32_BULK, WALL_EQUILIBRIUM
    prec = FP64 if precision == FP64 else FP32_BULK
    for u0, w0 in ((3.0, 50.0), (50.0, 3.0)):
        case, field = cases.small_case(22, 13, 32, seed=9, LES_model=3)
        case.nu, case.CFL = 1e-4, 1e-3
        rng = np.random.default_rng(4)
        U = np.asfortranarray(u0 * (1.0 + 0.05 * rng.uniform(-1, 1, field[0].shape)))
        V = np.asfortranarray(0.1 * rng.uniform(-1, 1, field[1].shape))
        W = np.asfortranarray(w0 * (1.0 + 0.05 * rng.uniform(-1, 1, field[2].shape)))
        chans = []
        for _ in range(2):
            g = Channel(case.x, case.y, case.z, precision=prec, **case.kwargs())
            g.set_state(U, V, W)
            g.apply_boundary_conditions()
            g.compute_projection_step()
            g.set_scalars(dt=0.0, t=0.0, dPdx=0.0, Qflow_x_0=0.0)
            g.set_wall_model(WALL_EQUILIBRIUM, h_wm=0.1)
            chans.append(g)
        a, b = chans
        for step in range(3):
            dt_a = a.compute_dt()
            a.compute_time_step_RK3()
            b.run_steps(1)
            dt_b = b.scalars()["dt"]
            assert dt_a == dt_b, (u0, w0, step, dt_a, dt_b)
            assert dt_a < 0.5 * case.CFL * (case.y[1] - case.y[0]) ** 2 / case.nu  # convective limit binds
        for x, y in zip(a.state(), b.state()):
            assert np.array_equal(x, y), (u0, w0)
        assert np.array_equal(a.get_field(NU_T), b.get_field(NU_T))
        assert np.array_equal(a.get_field(ALPHA_X), b.get_field(ALPHA_X))
        del a, b, chans


@pytest.mark.parametrize("les", [2, 3])
def test_captured_and_launched_steps_agree_bitwise(gpu, les, monkeypatch):
    """run_steps replays a captured graph of the step; with
    CHANLES_NO_GRAPH=1 the same launches go to the stream one by one, and
    profile_steps times them one by one: all three give the same fields, dt
    and nu_t, bit for bit, with every fold of the step in use (compute_dt,
    the wall stress and, for the dynamic model, the deferred pass B and the
    mass-flow shift)."""
    from channel_rl_amd import WALL_EQUILIBRIUM
    case, field = cases.small_case(22, 13, 34, seed=12, LES_model=les)
    chans = []
    for graph in (True, False, True):
        if graph:
            monkeypatch.delenv("CHANLES_NO_GRAPH", raising=False)
        else:
            monkeypatch.setenv("CHANLES_NO_GRAPH", "1")
        g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
        g.set_state(*field)
        g.apply_boundary_conditions()
        g.compute_projection_step()
        g.initialize_mass_flow()
        g.set_scalars(dt=0.0, t=0.0, dPdx=case.dPdx, Qflow_x_0=g.scalars()["Qflow_x_0"])
        g.set_wall_model(WALL_EQUILIBRIUM, h_wm=0.1)
        chans.append(g)
    a, b, c = chans
    a.run_steps(4)
    b.run_steps(4)
    c.profile_steps(4)
    for other in (b, c):
        for x, y in zip(a.state(), other.state()):
            assert np.array_equal(x, y), les
        assert np.array_equal(a.get_field(NU_T), other.get_field(NU_T))
        assert a.scalars() == other.scalars()
