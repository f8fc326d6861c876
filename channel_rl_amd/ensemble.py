"""The batched ensemble: many channels of uneven size and Re_tau advanced
together by one graph launch per step (`Ensemble`), and the measurement of
its throughput against the same channels stepped one at a time."""
import time

import numpy as np

from . import ALPHA_X, FP64, FP32_BULK, WALL_EQUILIBRIUM, WALL_MLP, Channel, Ensemble, cases


def make_channel(case, field, precision=FP64, wall=WALL_EQUILIBRIUM, net=None, max_rel_err=0.05):
    """A channel at its projected initial field with its wall model set."""
    ch = Channel(case.x, case.y, case.z, precision=precision, **case.kwargs())
    ch.set_state(*field)
    ch.apply_boundary_conditions()
    ch.compute_projection_step()
    ch.initialize_mass_flow()
    ch.set_scalars(dt=0.0, t=0.0, dPdx=case.dPdx, Qflow_x_0=ch.scalars()["Qflow_x_0"])
    if net is not None:
        ch.set_wall_network(net, max_rel_err)
    ch.set_wall_model(wall, h_wm=case.h_wm)
    if getattr(case, "wm_nu_limit", 0.0) > 0.0:
        ch.set_wall_viscosity_limit(case.wm_nu_limit)
    return ch


def tune_members(channels, reps=3):
    """The autotuner on one channel of each shape and precision, its tile
    shapes then set on the others of that kind (channels of one shape keep
    one set of shapes, so they still form one group).  Returns {(nx, ny, nz,
    precision): (rhs shape, dsm shape)}."""
    from . import KERNEL_DSM, KERNEL_RHS
    picked = {}
    for ch in channels:
        key = (ch.nx, ch.ny, ch.nz, ch.precision)
        if key not in picked:
            ch.autotune(reps=reps)
            picked[key] = (ch.tile_shape(KERNEL_RHS), ch.tile_shape(KERNEL_DSM))
        else:
            ch.set_tile_shape(picked[key][0], KERNEL_RHS)
            ch.set_tile_shape(picked[key][1], KERNEL_DSM)
    return picked


def tune_groups(channels, picked, steps=12):
    """The tile shapes of each grid's members refined on their grouped launch:
    a group runs many times the blocks of one channel, so the shapes the
    autotuner picks for one channel (tune_members, `picked`) need not be the
    group's best.  For every grid with two or more members, the grouped step
    is timed over `steps` steps for each DSM shape (the RHS shape fixed), then
    for each RHS shape (the best DSM shape fixed); the fastest pair is set on
    the members.  Every member's fields, scalars and wall stress are restored
    after each trial, so the tuning leaves the state as it was.  Returns
    {(nx, ny, nz, precision): (rhs shape, dsm shape)} with the refined picks."""
    from . import KERNEL_DSM, KERNEL_RHS, load_library
    L = load_library()
    nr, nd = L.chan_tile_shape_count(KERNEL_RHS), L.chan_tile_shape_count(KERNEL_DSM)
    out = dict(picked)
    for key in picked:
        members = [ch for ch in channels if (ch.nx, ch.ny, ch.nz, ch.precision) == key]
        if len(members) < 2:
            continue
        saved = [(ch.state(), ch.scalars(), ch.get_field(ALPHA_X)) for ch in members]
        # one ensemble for all the trials: a change of shape makes it capture
        # its step again
        ens = Ensemble(members)

        def trial(r, d):
            for ch in members:
                ch.set_tile_shape(r, KERNEL_RHS)
                ch.set_tile_shape(d, KERNEL_DSM)
            try:
                ens.run_steps(2)
                ens.synchronize()
                t0 = time.perf_counter()
                ens.run_steps(steps)
                ens.synchronize()
                return time.perf_counter() - t0
            finally:
                for ch, (st, sc, a) in zip(members, saved):
                    ch.set_state(*st)
                    ch.set_scalars(**sc)
                    ch.set_field(ALPHA_X, a)

        try:
            r0, d0 = picked[key]
            td = [trial(r0, d) for d in range(nd)]
            d1 = int(np.argmin(td))
            tr = [trial(r, d1) for r in range(nr)]
            r1 = int(np.argmin(tr))
        finally:
            ens.close()
        for ch in members:
            ch.set_tile_shape(r1, KERNEL_RHS)
            ch.set_tile_shape(d1, KERNEL_DSM)
        out[key] = (r1, d1)
    return out


def time_sequential(channels, steps):
    """Seconds for `steps` steps of every channel, one channel after the other
    (each with its own captured step)."""
    for ch in channels:
        ch.run_steps(1)
    for ch in channels:
        ch.synchronize()
    t0 = time.perf_counter()
    for ch in channels:
        ch.run_steps(steps)
    for ch in channels:
        ch.synchronize()
    return time.perf_counter() - t0


def time_batched(ens, steps):
    ens.run_steps(1)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.run_steps(steps)
    ens.synchronize()
    return time.perf_counter() - t0


def wall_error(channels):
    """Ensemble mean of the learned wall model's deviation from the
    equilibrium stress and of the fallback fraction (last evaluation)."""
    st = np.array([ch.wall_model_stats() for ch in channels])
    return float(st[:, 0].mean()), float(st[:, 1].mean())


def main(argv=None):
    """Run and time the grouped ensemble of the 64-channel draw (or of its
    members with a given number of x faces), FP32 bulk with the learned wall
    model (trained on 2x10^5 samples, 4 epochs), for profiling:

        python -m channel_rl_amd.ensemble --faces 98 --steps 30
    """
    import argparse
    import sys

    from . import wallnet
    ap = argparse.ArgumentParser()
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--faces", type=int, default=0, help="only the members with this many x faces (0: all)")
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--shapes-file", default="",
                    help="tile shapes per grid: read from this JSON file when it exists (no tuning), "
                         "else tuned and written there")
    a = ap.parse_args(argv)
    X, Y = wallnet.solver_samples(200_000, seed=7)
    net, _ = wallnet.train(X, Y, epochs=4)
    mem = [(c, f) for c, f in cases.ensemble(a.members) if a.faces in (0, c.nx)]
    if not mem:
        print(f"no member with {a.faces} x faces", file=sys.stderr)
        return 2
    chans = [make_channel(c, f, precision=FP32_BULK, wall=WALL_MLP, net=net) for c, f in mem]
    import json
    import os
    from . import KERNEL_DSM, KERNEL_RHS
    if a.shapes_file and os.path.exists(a.shapes_file):
        with open(a.shapes_file) as fh:
            picks = {tuple(map(int, k.split())): v for k, v in json.load(fh).items()}
        for ch in chans:
            r, d = picks[(ch.nx, ch.ny, ch.nz, ch.precision)]
            ch.set_tile_shape(r, KERNEL_RHS)
            ch.set_tile_shape(d, KERNEL_DSM)
    else:
        picks = tune_groups(chans, tune_members(chans))
        if a.shapes_file:
            with open(a.shapes_file, "w") as fh:
                json.dump({" ".join(map(str, k)): list(v) for k, v in picks.items()}, fh)
    ens = Ensemble(chans)
    ens.run_steps(10)
    ens.synchronize()
    dt = time_batched(ens, a.steps)
    print(f"{len(mem)} members in {ens.group_count()} groups: {1e3 * dt / a.steps:.3f} ms per step, "
          f"{sum(c.cells for c, _ in mem) * a.steps / dt / 1e6:.1f} Mcells/s", flush=True)
    ens.close()
    for c in chans:
        c.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
