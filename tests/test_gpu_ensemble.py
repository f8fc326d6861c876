"""The batched ensemble: members advanced by one graph launch per step must
end bit for bit where the same channels end when stepped one at a time
(every member runs the same kernels in the same order on its own data)."""
import numpy as np
import pytest

from channel_rl_amd import FP32_BULK, FP64, WALL_EQUILIBRIUM, Channel, Ensemble, cases
from channel_rl_amd.ensemble import make_channel

pytestmark = pytest.mark.gpu

MEMBERS = [((18, 17, 18), FP64, 1), ((34, 33, 34), FP32_BULK, 2), ((26, 25, 26), FP64, 3)]


def _members():
    out = []
    for (nx, ny, nz), prec, seed in MEMBERS:
        case, field = cases.small_case(nx, ny, nz, seed=seed)
        out.append((case, field, prec))
    return out


def _state(ch):
    return list(ch.state()) + [np.array([ch.scalars()[k] for k in ("dt", "t", "dPdx", "Qflow_x_0")])]


def test_ensemble_matches_members_stepped_alone(gpu):
    mem = _members()
    alone = [make_channel(c, f, precision=p) for c, f, p in mem]
    batch = [make_channel(c, f, precision=p) for c, f, p in mem]
    ens = Ensemble(batch)
    assert len(ens) == 3
    for ch in alone:
        ch.run_steps(4)
    ens.run_steps(4)
    ens.synchronize()
    for a, b in zip(alone, batch):
        for x, y in zip(_state(a), _state(b)):
            assert np.array_equal(x, y)


def test_ensemble_recaptures_after_member_changes(gpu):
    mem = _members()
    alone = [make_channel(c, f, precision=p) for c, f, p in mem]
    batch = [make_channel(c, f, precision=p) for c, f, p in mem]
    ens = Ensemble(batch)
    ens.run_steps(2)
    for ch in alone:
        ch.run_steps(2)
    # step one member on its own (its rotation moves), change another's wall model
    batch[0].run_steps(1)
    alone[0].run_steps(1)
    for ch in (batch[2], alone[2]):
        ch.set_wall_model(WALL_EQUILIBRIUM, h_wm=0.1)
    ens.run_steps(3)
    for ch in alone:
        ch.run_steps(3)
    ens.synchronize()
    for a, b in zip(alone, batch):
        for x, y in zip(_state(a), _state(b)):
            assert np.array_equal(x, y)


def test_ensemble_rejects_bad_members(gpu):
    case, field = cases.small_case(18, 17, 18, seed=1)
    ch = make_channel(case, field)
    with pytest.raises(Exception):
        Ensemble([ch, ch])
    with pytest.raises(Exception):
        Ensemble([])


def test_config5_batched_ensemble_matches_members_alone(gpu, wall_network):
    """The ensemble configuration as specified: 64 channels drawn by seed 42
    (faces 34 33 34 / 66 49 66 / 98 65 98, Re_tau log-uniform in
    [180, 5200]), FP32 bulk with FP64 Poisson, the learned wall model in the
    loop at h_wm = 0.1: 10 batched steps end bit for bit where the same
    channels end when each is stepped alone."""
    from channel_rl_amd import WALL_MLP
    net = wall_network[0]
    mem = cases.ensemble(64)
    assert len({(c.nx, c.ny, c.nz) for c, _ in mem}) == 3
    alone = [make_channel(c, f, precision=FP32_BULK, wall=WALL_MLP, net=net) for c, f in mem]
    batch = [make_channel(c, f, precision=FP32_BULK, wall=WALL_MLP, net=net) for c, f in mem]
    ens = Ensemble(batch)
    for ch in alone:
        ch.run_steps(10)
    ens.run_steps(10)
    ens.synchronize()
    for q, (a, b) in enumerate(zip(alone, batch)):
        for x, y in zip(_state(a), _state(b)):
            assert np.all(np.isfinite(x))
            assert np.array_equal(x, y), f"member {q}"
        assert a.wall_model_stats() == b.wall_model_stats()
    err = np.mean([ch.wall_model_stats()[0] for ch in batch])
    print(f"ensemble mean wall-stress deviation from the equilibrium law after 10 steps: {err:.4f}")
    ens.close()


def test_config5_one_step_per_grid_against_oracle(gpu):
    """One batched step of the first member of each grid size of the
    ensemble draw (its Re_tau and seeded log-law field; FP64, dynamic
    Smagorinsky, equilibrium wall model at h_wm = 0.1, constant mass flow)
    against the oracle: 1e-11 relative, as a single channel's step."""
    import oracle
    from helpers import rel
    first = {}
    for c, f in cases.ensemble(64):
        first.setdefault((c.nx, c.ny, c.nz), (c, f))
    assert len(first) == 3
    chans, oracles = [], []
    for c, f in first.values():
        g = make_channel(c, f, precision=FP64, wall=WALL_EQUILIBRIUM)
        o = oracle.Oracle(c.x, c.y, c.z, **c.kwargs())
        o.set_state(*g.state())
        s = g.scalars()
        o.set_scalars(dt=s["dt"], t=s["t"], dPdx=s["dPdx"], Qflow_x_0=s["Qflow_x_0"])
        o.set_wall_model(1, h_wm=c.h_wm)
        chans.append(g)
        oracles.append(o)
    ens = Ensemble(chans)
    ens.run_steps(1)
    ens.synchronize()
    for (c, _), g, o in zip(first.values(), chans, oracles):
        o.compute_dt()
        o.compute_time_step_RK3()
        for a, b in zip(g.state(), o.state()):
            assert rel(a, b) <= 1e-11, ((c.nx, c.ny, c.nz), rel(a, b))
        assert abs(g.scalars()["t"] - o.scalars()["t"]) <= 1e-15
    ens.close()


def _grouped_members():
    """Four FP64 channels of one shape (different Re_tau), three FP32-bulk
    channels of another, one AMD channel: two groups and one lone member."""
    out = []
    for q in range(4):
        case, field = cases.small_case(18, 17, 18, seed=10 + q)
        case.nu = case.nu / (1.0 + q)
        out.append((case, field, FP64))
    for q in range(3):
        case, field = cases.small_case(26, 17, 18, seed=20 + q)
        out.append((case, field, FP32_BULK))
    case, field = cases.small_case(18, 17, 18, seed=30, LES_model=3)
    out.append((case, field, FP64))
    return out


def test_grouped_ensemble_matches_members_alone(gpu):
    """Grouped launches (one launch per stage for all members of a shape):
    after 7 steps, bit for bit the members stepped alone and the same
    ensemble without grouping (every member its own kernels)."""
    mem = _grouped_members()
    alone = [make_channel(c, f, precision=p, wall=WALL_EQUILIBRIUM) for c, f, p in mem]
    grouped = [make_channel(c, f, precision=p, wall=WALL_EQUILIBRIUM) for c, f, p in mem]
    single = [make_channel(c, f, precision=p, wall=WALL_EQUILIBRIUM) for c, f, p in mem]
    for ch in alone + grouped + single:
        ch.set_wall_model(WALL_EQUILIBRIUM, h_wm=0.1)
    eg, es = Ensemble(grouped), Ensemble(single)
    es.set_grouping(False)
    assert eg.group_count() == 3 and es.group_count() == len(mem)
    for ch in alone:
        ch.run_steps(7)
    eg.run_steps(7)
    es.run_steps(7)
    eg.synchronize()
    es.synchronize()
    for q, (a, b, c) in enumerate(zip(alone, grouped, single)):
        for x, y, z in zip(_state(a), _state(b), _state(c)):
            assert np.all(np.isfinite(x))
            assert np.array_equal(x, y), f"member {q} grouped"
            assert np.array_equal(x, z), f"member {q} ungrouped"
    eg.close()
    es.close()


def test_group_tuning_leaves_the_state_and_picks_valid_shapes(gpu):
    """tune_groups times each grid's grouped step under every DSM and RHS
    tile shape and restores the members' fields, scalars and wall stress
    after each trial: afterwards the members step exactly as untuned copies
    with the picked shapes set."""
    from channel_rl_amd import ALPHA_X, KERNEL_DSM, KERNEL_RHS, load_library
    from channel_rl_amd.ensemble import tune_groups, tune_members
    L = load_library()
    case, field = cases.small_case(18, 17, 18, seed=21)
    tuned = [make_channel(case, field) for _ in range(2)]
    ref = [make_channel(case, field) for _ in range(2)]
    before = [_state(ch) + [ch.get_field(ALPHA_X)] for ch in tuned]
    picks = tune_groups(tuned, tune_members(tuned, reps=1), steps=2)
    (r, d), = picks.values()
    assert 0 <= r < L.chan_tile_shape_count(KERNEL_RHS) and 0 <= d < L.chan_tile_shape_count(KERNEL_DSM)
    for ch, b in zip(tuned, before):
        assert ch.tile_shape(KERNEL_RHS) == r and ch.tile_shape(KERNEL_DSM) == d
        for x, y in zip(_state(ch) + [ch.get_field(ALPHA_X)], b):
            assert np.array_equal(x, y)
    for ch in ref:
        ch.set_tile_shape(r, KERNEL_RHS)
        ch.set_tile_shape(d, KERNEL_DSM)
    e1, e2 = Ensemble(tuned), Ensemble(ref)
    e1.run_steps(3)
    e2.run_steps(3)
    e1.synchronize()
    e2.synchronize()
    for a, b in zip(tuned, ref):
        for x, y in zip(_state(a), _state(b)):
            assert np.array_equal(x, y)
    e1.close()
    e2.close()
