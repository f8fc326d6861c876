"""CPU oracle of the RL agents' observations and actions (test
infrastructure only): a plain restatement of code/smarties_stat.f90:24-258
with the random draws passed in.

observe(U, nu_t, yg, nu, nagents, dPdx_ref, draws) -> (states, rewards)
act(alpha, nagents, factors) -> new alpha
Arrays use the reference's shapes (U (nx,nyg,nzg), nu_t (nxg,nyg,nzg),
alpha (nx,2,nzg)); yg is 1-based (element 0 unused).
"""
import numpy as np


def _sqrt(x):
    """sqrt with NaN for a negative argument, as the compiled reference gives."""
    return float(np.sqrt(x)) if x >= 0 else float("nan")


def _log(x):
    return float(np.log(x)) if x > 0 else (float("-inf") if x == 0 else float("nan"))


def observe(U, nut, yg, nu, nagents, dPdx_ref, draws):
    nx, nyg, _ = U.shape
    nxa = (nx - 2) // nagents
    u = lambda i, j, k: U[i - 1, j - 1, k - 1]  # noqa: E731
    n = lambda i, j, k: nut[i - 1, j - 1, k - 1]  # noqa: E731
    states = np.zeros((2, nagents, 6))
    rewards = np.zeros((2, nagents))
    draws = np.asarray(draws, dtype=np.float64).reshape(2, nagents, 2)
    for wall in range(2):
        for a in range(nagents):
            ib = a * nxa + 2
            yind = 3 if draws[wall, a, 0] < 0.5 else 4
            yr = draws[wall, a, 1]
            if wall == 0:
                # code/smarties_stat.f90:61-108
                ut = _sqrt((nu + 0.25 * (n(ib + 1, 2, 2) + n(ib, 2, 2) + n(ib + 1, 1, 2) + n(ib, 1, 2))) *
                               (u(ib, 2, 2) - u(ib, 1, 2)) / (yg[2] - yg[1]))
                l = yr * yg[yind] + (1.0 - yr) * yg[yind + 1]
                jh, jl = yind, yind + 1
                lin = lambda i, k: yr * u(i, jh, k) + ((1.0 - yr) * u(i, jl, k))  # noqa: E731
                Uloc = 0.2 * lin(ib, 2) + 0.2 * lin(ib - 1, 2) + 0.2 * lin(ib + 1, 2) + 0.2 * lin(ib, 1) + \
                    0.2 * lin(ib, 3)
                dU = 0.2 * (u(ib, jl, 2) - u(ib, jh, 2) + u(ib - 1, jl, 2) - u(ib - 1, jh, 2) + u(ib + 1, jl, 2) -
                            u(ib + 1, jh, 2) + u(ib, jl, 1) - u(ib, jh, 1) + u(ib, jl, 3) - u(ib, jh, 3)) / \
                    (yg[jl] - yg[jh])
                tw = (nu + 0.25 * (n(ib, 1, 2) + n(ib + 1, 1, 2) + n(ib, 2, 2) + n(ib + 1, 2, 2))) * \
                    (u(ib, 2, 2) - u(ib, 1, 2)) / (yg[2] - yg[1])
            else:
                # code/smarties_stat.f90:120-159, 171-178 (u_tau divides by
                # the bottom spacing, as the reference does)
                ut = _sqrt((nu + 0.25 * (n(ib + 1, nyg, 2) + n(ib, nyg, 2) + n(ib + 1, nyg - 1, 2) +
                                             n(ib, nyg - 1, 2))) *
                               (u(ib, nyg - 1, 2) - u(ib, nyg, 2)) / (yg[2] - yg[1]))
                jh, jl = nyg - yind + 1, nyg - yind
                l = 2.0 - (yr * yg[jh] + (1.0 - yr) * yg[jl])
                lin = lambda i, k: yr * u(i, jh, k) + ((1.0 - yr) * u(i, jl, k))  # noqa: E731
                Uloc = 0.2 * lin(ib, 2) + 0.2 * lin(ib - 1, 2) + 0.2 * lin(ib + 1, 2) + 0.2 * lin(ib, 1) + \
                    0.2 * lin(ib, 3)
                dU = 0.2 * (u(ib, jl, 2) - u(ib, jh, 2) + u(ib - 1, jl, 2) - u(ib - 1, jh, 2) + u(ib + 1, jl, 2) -
                            u(ib + 1, jh, 2) + u(ib, jl, 1) - u(ib, jh, 1) + u(ib, jl, 3) - u(ib, jh, 3)) / \
                    (yg[jh] - yg[jl])
                tw = (nu + 0.25 * (n(ib, nyg - 1, 2) + n(ib + 1, nyg - 1, 2) + n(ib, nyg, 2) + n(ib + 1, nyg, 2))) * \
                    (u(ib, nyg - 1, 2) - u(ib, nyg, 2)) / (yg[nyg] - yg[nyg - 1])
            states[wall, a] = [(dU * l) / ut, Uloc / ut - (dU * l) / ut * _log(l * ut / nu), Uloc, l, ut,
                               dPdx_ref]
            sq = _sqrt(dPdx_ref)
            rewards[wall, a] = 100.0 - 5000.0 * abs(_sqrt(tw) - sq) / sq
    return states, rewards


def act(alpha, nagents, factors):
    """code/smarties_stat.f90:190,213,231-258."""
    alpha = np.array(alpha, dtype=np.float64, order="F")
    nx, _, nzg = alpha.shape
    nxa = (nx - 2) // nagents
    nzp = nzg - 3
    factors = np.asarray(factors, dtype=np.float64).reshape(2, nagents)
    acts = np.array([[alpha[a * nxa + 2 - 1, w, 1] * factors[w, a] for a in range(nagents)] for w in range(2)])
    for j in range(1, nxa + 1):
        for a in range(nagents):
            an = a + 1 if a + 1 < nagents else 0
            for w in range(2):
                alpha[a * nxa + j + 1 - 1, w, :] = (acts[w, a] * (nxa - j + 1)) / nxa + (acts[w, an] * (j - 1)) / nxa
    alpha[0] = alpha[nx - 2]
    alpha[nx - 1] = alpha[1]
    for iz in range(1, nzp + 1):
        alpha[:, :, iz] = (alpha[:, :, 1] * (nzp - iz + 1)) / nzp + (alpha[:, :, nzg - 1] * (iz - 1)) / nzp
    return alpha
