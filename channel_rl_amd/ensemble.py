"""The batched ensemble: many channels of uneven size and Re_tau advanced
together by one graph launch per step (`Ensemble`), and the measurement of
its throughput against the same channels stepped one at a time."""
import time

import numpy as np

from . import FP64, FP32_BULK, WALL_EQUILIBRIUM, WALL_MLP, Channel, Ensemble, cases


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
