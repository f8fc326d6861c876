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
