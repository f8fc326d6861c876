"""GPU checks at the benchmark sizes.

The 192x97x192 FP64 channel (faces 194 97 194, Re_tau 2000, dynamic
Smagorinsky, wall model at h_wm = 0.1) and the 384x193x384 FP32-bulk channel
(faces 386 193 386, Re_tau 5200).  At these sizes every kernel runs many
tiles, several rounds of blocks, the z-chunk lengths chosen for the grid and
the tuned RHS tile, which the small parity cases do not reach.

Against the oracle (one stage of the same state; the oracle takes seconds
for one stage at 192x97x192): the DSM eddy viscosity, the right-hand sides
and the projection, at the small cases' tolerances.  Size-independent
properties of whole RK3 steps: the field stays finite and divergence-free to
round-off, the constant-mass-flow correction holds its quantity at the
initial bulk flow, the step is deterministic (two
channels from the same start agree bit for bit) and the projection is
idempotent (projecting a projected field changes it by round-off only).
"""
import numpy as np
import pytest

import oracle
from channel_rl_amd import FP32_BULK, FP64, NU_T, WALL_EQUILIBRIUM, Channel, cases
from channel_rl_amd.ensemble import make_channel
from helpers import interior, oracle_for, projected_state, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wmles():
    case, field = cases.config("wmles")
    return case, field, projected_state(case, field)


def div_scale(case, U):
    return np.max(np.abs(U)) / np.min(np.diff(case.x))


def held_flow(U):
    """The quantity the constant-mass-flow correction holds at Qflow_x_0
    after every step (code/mass_flow.f90:120-165, one domain): the mean over
    x faces 2..nx-1 and z centres 2..nzg-2 of the average of U on rows 3 and
    nyg-2."""
    nx, nyg, nzg = U.shape
    return float(np.mean(0.5 * (U[1:nx - 1, 2, 1:nzg - 2] + U[1:nx - 1, nyg - 3, 1:nzg - 2])))


def test_fullsize_stages_match_oracle(gpu, wmles):
    case, _, (U, V, W) = wmles
    o = oracle_for(case)
    o.set_state(U, V, W)
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    g.set_state(U, V, W)
    ref = o.compute_eddy_viscosity(U, V, W)
    g.compute_eddy_viscosity()
    got = g.get_field(NU_T)
    assert np.all(np.isfinite(got))
    assert rel(got, ref) <= 1e-12, rel(got, ref)
    # the right-hand sides with the same eddy viscosity in both (the GPU's)
    o.set(oracle.NUT, got)
    g.set_scalars(dPdx=case.dPdx)
    o.set_scalars(dPdx=case.dPdx)
    ref_r = o.compute_rhs(U, V, W)
    got_r = g.compute_rhs()
    for q in range(3):
        ref_q = interior(q, ref_r[q], case)
        assert rel(got_r[q], ref_q) <= 1e-14, (q, rel(got_r[q], ref_q))


@pytest.mark.parametrize("dirichlet", [1, 2])
def test_fullsize_amd_matches_oracle(gpu, wmles, dirichlet):
    """The AMD eddy viscosity (LES model 3, its LDS-tiled z march with
    several chunks per column and the ghost rows of two wall conditions) on
    the 192x97x192 state against the oracle, 1e-12 as on the small grids."""
    import copy
    case, _, (U, V, W) = wmles
    case = copy.copy(case)
    case.LES_model, case.Dirichlet_nu_t = 3, dirichlet
    o = oracle_for(case)
    o.set_scalars(dPdx=0.37)
    ref = o.compute_eddy_viscosity(U, V, W)
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    g.set_state(U, V, W)
    g.set_scalars(dt=0.0, t=0.0, dPdx=0.37, Qflow_x_0=0.0)
    g.compute_eddy_viscosity()
    got = g.get_field(NU_T)
    assert np.all(np.isfinite(got))
    assert rel(got, ref) <= 1e-12, rel(got, ref)


def test_fullsize_projection_matches_oracle_and_is_idempotent(gpu, wmles):
    case, field, _ = wmles
    o = oracle_for(case)
    o.set_state(*field)
    o.apply_boundary_conditions()
    U, V, W = o.state()
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    g.set_state(U, V, W)
    o.compute_projection_step()
    o.apply_boundary_conditions()
    g.compute_projection_step()
    once = g.state()
    for a, b in zip(once, o.state()):
        assert rel(a, b) <= 1e-12, rel(a, b)
    _, maxdiv = g.check_divergence()
    assert maxdiv <= 1e-12 * div_scale(case, U), (maxdiv, div_scale(case, U))
    g.compute_projection_step()
    for a, b in zip(g.state(), once):
        assert rel(a, b) <= 1e-13, rel(a, b)


def test_fullsize_fp64_steps_keep_invariants_and_are_deterministic(gpu, wmles):
    case, field, _ = wmles
    chans = [make_channel(case, field, precision=FP64, wall=WALL_EQUILIBRIUM) for _ in range(2)]
    for ch in chans:
        ch.run_steps(3)
        ch.synchronize()
    a, b = (ch.state() for ch in chans)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    U = a[0]
    assert all(np.all(np.isfinite(x)) for x in a)
    _, maxdiv = chans[0].check_divergence()
    assert maxdiv <= 1e-12 * div_scale(case, U), (maxdiv, div_scale(case, U))
    q0 = chans[0].scalars()["Qflow_x_0"]
    q = held_flow(U)
    assert abs(q - q0) <= 1e-12 * abs(q0), (q, q0)
    for ch in chans:
        ch.close()


def test_fullsize_fp32_bulk_step_keeps_invariants(gpu):
    case, field = cases.config("large")
    ch = make_channel(case, field, precision=FP32_BULK, wall=WALL_EQUILIBRIUM)
    ch.run_steps(2)
    ch.synchronize()
    U, V, W = ch.state()
    assert all(np.all(np.isfinite(x)) for x in (U, V, W))
    _, maxdiv = ch.check_divergence()
    # the FP64 solve's correction is stored in the FP32 fields: divergence at
    # FP32 round-off of the velocity gradients
    assert maxdiv <= 1e-5 * div_scale(case, U), (maxdiv, div_scale(case, U))
    q0 = ch.scalars()["Qflow_x_0"]
    q = held_flow(U)
    assert abs(q - q0) <= 1e-6 * abs(q0), (q, q0)
    ch.close()


@pytest.mark.parametrize("les", [2, 3])
def test_step_folds_bitwise_at_full_size(gpu, wmles, les, monkeypatch):
    """At 192x97x192 FP64 (every kernel over many tiles and rounds of
    blocks): run_steps, with compute_dt in the first eddy-viscosity launches
    and the dynamic model's pass B deferred to the RHS, against compute_dt +
    compute_time_step_RK3 with pass B materialised in every substep
    (CHANLES_NO_DEFERRED_NUT=1): two steps agree bit for bit (the wall stress
    and the mass-flow shift ride in the same launches on both sides)."""
    case, field, (U, V, W) = wmles
    chans = []
    for folded in (True, False):
        if folded:
            monkeypatch.delenv("CHANLES_NO_DEFERRED_NUT", raising=False)
        else:
            monkeypatch.setenv("CHANLES_NO_DEFERRED_NUT", "1")
        kw = case.kwargs()
        kw["LES_model"] = les
        g = Channel(case.x, case.y, case.z, precision=FP64, **kw)
        g.set_state(U, V, W)
        g.initialize_mass_flow()
        g.set_scalars(dt=0.0, t=0.0, dPdx=case.dPdx, Qflow_x_0=g.scalars()["Qflow_x_0"])
        g.set_wall_model(WALL_EQUILIBRIUM, h_wm=case.h_wm)
        chans.append(g)
    a, b = chans
    a.run_steps(2)
    for _ in range(2):
        b.compute_dt()
        b.compute_time_step_RK3()
    for x, y in zip(a.state(), b.state()):
        assert np.array_equal(x, y), les
    assert np.array_equal(a.get_field(NU_T), b.get_field(NU_T))
    assert a.scalars() == b.scalars()
    a.close()
    b.close()
