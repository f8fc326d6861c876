"""The RL agents' observations and actions on the GPU against the oracle's
restatement of code/smarties_stat.f90 (oracle/rl.py), with the same
random draws."""
import numpy as np
import pytest

from oracle import rl
from channel_rl_amd import ALPHA_X, FP64, NU_T, Channel, cases
from helpers import oracle_for, projected_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nagents", [((18, 17, 18), 4), ((34, 33, 34), 8)])
def test_rl_observe_matches_oracle(gpu, n, nagents):
    case, field = cases.small_case(*n, seed=8)
    U, V, W = projected_state(case, field)
    rng = np.random.default_rng(2)
    nut = np.asfortranarray(rng.uniform(0.0, 1e-3, (case.nxg, case.nyg, case.nzg)))
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    g.set_state(U, V, W)
    g.set_field(NU_T, nut)
    draws = rng.uniform(0.0, 1.0, (2, nagents, 2))
    st, rw = g.rl_observe(nagents, 0.9, draws)
    yg = np.concatenate([[0.0], oracle_for(case).grid()["yg"]])
    st_r, rw_r = rl.observe(U, nut, yg, case.nu, nagents, 0.9, draws)
    # a wall cell with a negative gradient gives NaN in both (sqrt of a
    # negative stress), as in the reference
    assert np.array_equal(np.isnan(st), np.isnan(st_r)) and np.array_equal(np.isnan(rw), np.isnan(rw_r))
    ok = ~np.isnan(st_r)
    assert np.max(np.abs(st[ok] - st_r[ok])) <= 1e-13 * np.abs(st_r[ok]).max()
    ok = ~np.isnan(rw_r)
    assert np.max(np.abs(rw[ok] - rw_r[ok])) <= 1e-11 * np.abs(rw_r[ok]).max()


def test_rl_act_matches_oracle(gpu):
    case, field = cases.small_case(26, 17, 18, seed=9)
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    rng = np.random.default_rng(3)
    alpha = np.asfortranarray(rng.uniform(0.5, 1.5, (case.nx, 2, case.nzg)))
    g.set_field(ALPHA_X, alpha)
    f = rng.uniform(0.8, 1.2, (2, 6))
    g.rl_act(6, f)
    assert np.array_equal(g.get_field(ALPHA_X), rl.act(alpha, 6, f))


def test_rl_rejects_bad_agent_counts(gpu):
    case, field = cases.small_case(18, 17, 18, seed=9)
    g = Channel(case.x, case.y, case.z, precision=FP64, **case.kwargs())
    with pytest.raises(Exception):
        g.rl_observe(5, 1.0, np.zeros((2, 5, 2)))   # 5 does not divide nx-2 = 16
