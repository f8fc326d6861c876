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


def train(X, Y, epochs=20, batch=1024, lr=1e-3, lr_final=1e-5, seed=0, device=0, input_noise=1.0):
    """A WallNetwork fitted to (X, Y): standardisation from the data, then
    Adam with the learning rate lr in the first epoch, decaying on a cosine to
    lr_final in the last (the optimiser state carries over between epochs);
    returns (network, per-epoch loss).  input_noise > 0 adds, every epoch,
    fresh Gaussian noise of that many standard deviations to the features
    f1..f3 (wall-normal gradient, lagged stress, pressure gradient): the
    equilibrium target is a function of the local Reynolds number f0, and the
    noise keeps the network from keying on how the other features happen to
    correlate with it in the training flow."""
    net = WallNetwork(seed=seed, device=device)
    net.fit_normalization(X, Y)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    sd = X.std(0)
    rng = np.random.default_rng(seed)
    losses = []
    for e in range(epochs):
        c = 0.5 * (1.0 + np.cos(np.pi * e / max(epochs - 1, 1)))
        Xe = X
        if input_noise > 0.0:
            Xe = X.copy()
            Xe[:, 1:] += (input_noise * sd[1:] * rng.standard_normal((len(X), 3))).astype(np.float32)
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
