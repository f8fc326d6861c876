"""Training data and training of the learned wall-stress closure.

The network (`WallNetwork`) learns the wall stress from data the solver
generates: a small channel (32x33x32 cells, faces 34 33 34, Re_tau 180,
constant mass flow, dynamic Smagorinsky, equilibrium wall model at the first
off-wall cell) is advanced from its seeded initial field, and every few steps
the flow at every wall cell is sampled at several matching heights.  Each
sample is also evaluated with the viscosity divided by a log-uniform factor
in [1, nu_factor_max], which moves the same velocity profiles to higher
Reynolds numbers (the features and targets are functions of h u/nu, h^2
du/dn/nu, h sqrt(tau)/nu and h^3 dpdx/nu^2 only), so the network covers the
matching-height Reynolds numbers of Re_tau 180..5200 channels.  Targets are
the equilibrium log law's stress.
"""
import numpy as np

from . import ALPHA_X, Channel, WallNetwork, WALL_EQUILIBRIUM, cases

HEIGHTS = (0.0, 0.05, 0.1, 0.15, 0.2, 0.3, 0.4, 0.5)


def solver_samples(n_samples, seed=7, develop_steps=20, every=4, heights=HEIGHTS, nu_factor_max=30.0,
                   kappa=0.41, B=5.2, precision=0):
    """(X, Y) with n_samples rows drawn from runs of the small channel.

    Each round draws a viscosity factor per height, takes the equilibrium
    stress of the current flow at that height and viscosity, advances one step
    and samples again with that stress as the lagged wall stress (the wall
    model sees the previous evaluation's stress in the same way), then
    advances the remaining steps of the round."""
    case, field = cases.config("small", seed=seed)
    ch = Channel(case.x, case.y, case.z, precision=precision, **case.kwargs())
    ch.set_state(*field)
    ch.apply_boundary_conditions()
    ch.compute_projection_step()
    ch.initialize_mass_flow()
    q0 = ch.scalars()["Qflow_x_0"]
    ch.set_scalars(dt=0.0, t=0.0, dPdx=case.dPdx, Qflow_x_0=q0)
    ch.set_wall_model(WALL_EQUILIBRIUM, h_wm=0.0, kappa=kappa, B=B)
    ch.run_steps(develop_steps)
    rng = np.random.default_rng(seed)
    Xs, Ys, n = [], [], 0
    shape = (ch.nx, 2, ch.nzg)
    while n < n_samples:
        nus = [case.nu / float(np.exp(rng.uniform(0.0, np.log(nu_factor_max)))) for _ in heights]
        prev = [ch.wall_samples(h_wm=h, nu=nu, kappa=kappa, B=B, stress=True)[2] for h, nu in zip(heights, nus)]
        ch.run_steps(1)
        for h, nu, tp in zip(heights, nus, prev):
            ch.set_field(ALPHA_X, np.asfortranarray(tp.reshape(shape, order="F")))
            X, Y = ch.wall_samples(h_wm=h, nu=nu, kappa=kappa, B=B)
            Xs.append(X)
            Ys.append(Y)
            n += len(X)
        ch.run_steps(max(every - 1, 0))
    X = np.concatenate(Xs)[:n_samples]
    Y = np.concatenate(Ys)[:n_samples]
    ch.close()
    return X, Y


# Boxes, in the features' own asinh-scaled units, over which the three
# features other than the local Reynolds number are drawn during training:
# f1 (wall-normal gradient, either sign), f2 (lagged wall stress, >= 0), f3
# (driving pressure gradient, either sign; in a constant-mass-flow start-up
# the correction drives it far beyond a developed flow's value).  Every wall
# cell of the benchmark channels, their transients included, lies inside.
NUISANCE_BOX = ((-12.0, 12.0), (0.0, 10.0), (-25.0, 25.0))


def decorrelate(X, rng, box=NUISANCE_BOX):
    """A copy of X with f1..f3 replaced by independent uniform draws over
    `box`.  The equilibrium target is a function of the local Reynolds number
    f0 alone; in the solver's samples the other features follow it (the
    lagged stress f2 is the previous step's equilibrium stress, so a network
    can read the target off it), and a network that keys on those relations
    errs wherever a flow breaks them (a start-up transient, another grid or
    Reynolds number).  Drawing them afresh, independent of the target, makes
    the data itself say that they carry no information."""
    Xe = np.array(X, dtype=np.float32, copy=True)
    for q, (lo, hi) in enumerate(box):
        Xe[:, q + 1] = rng.uniform(lo, hi, len(Xe)).astype(np.float32)
    return Xe


def train(X, Y, epochs=20, batch=1024, lr=1e-3, lr_final=1e-5, seed=0, device=0, box=NUISANCE_BOX):
    """A WallNetwork fitted to (X, Y): standardisation from the data (with
    f1..f3 decorrelated), then Adam with the learning rate lr in the first
    epoch, decaying on a cosine to lr_final in the last (the optimiser state
    carries over between epochs); returns (network, per-epoch loss).  Every
    epoch draws f1..f3 afresh (`decorrelate`); box=None trains on X as
    sampled."""
    net = WallNetwork(seed=seed, device=device)
    rng = np.random.default_rng(seed)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    net.fit_normalization(X if box is None else decorrelate(X, rng, box), Y)
    losses = []
    for e in range(epochs):
        c = 0.5 * (1.0 + np.cos(np.pi * e / max(epochs - 1, 1)))
        Xe = X if box is None else decorrelate(X, rng, box)
        losses.append(net.train(Xe, Y, epochs=1, batch=batch, lr=lr_final + (lr - lr_final) * c, seed=seed + e)[0])
    return net, np.array(losses)


def stress_error(net, X, Y):
    """Relative error of the network's stress along u_par against the targets:
    |tau_net - tau_eq| / tau_eq per sample (tau recovered from o0 =
    log(1 + h^2 tau / nu^2))."""
    P = net.forward(X).astype(np.float64)
    t_net = np.expm1(P[:, 0])
    t_eq = np.expm1(Y[:, 0].astype(np.float64))
    ok = t_eq > 0
    return np.abs(t_net[ok] - t_eq[ok]) / t_eq[ok]
