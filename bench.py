#!/usr/bin/env python3
"""Channel LES throughput on one MI355X: cell-RK3-steps per second.

One step is compute_dt + compute_time_step_RK3 (three substeps of dynamic
Smagorinsky, equilibrium wall model, RHS + RK combine, FFT/tridiagonal
pressure projection, then the constant-mass-flow correction) over the whole
channel resident in HBM.  Default workload: 192x97x192 (faces 194 97 194),
Re_tau 2000, FP64, seeded log-law + 10% initial field (synthetic).

Prints one JSON line: value = unique cells x steps / second / 1e6 (Mcells/s),
plus the roofline of the dominant kernel (HIP-event durations measured in this
run on the library's stream) and the CPU oracle timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure
# MI355X vector peaks (vendor figures): FP64 78.6 TFLOP/s, FP32 157.3 TFLOP/s
VALU_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
# algorithmic flops per unique cell and launch, counted from the implemented
# formulas (an FMA counts 2): the DSM pass A is centre velocities and |S| S_ij
# (102: the off-diagonal strains 20 each), the separable 3x3x3 test filter of
# 15 inputs and 3 staggered velocities (18 x 15 = 270), hat(S) and 4|hat(S)|
# (85), L_ij, M_ij and the two contractions (62)
ALG_FLOPS = {"k_dsm_fused": 519}

# Algorithmic HBM values per unique cell per launch (FP64: x8 bytes; FP32
# bulk: bulk values x4, pressure values x8), the compulsory traffic of each
# kernel: every input read once, every output written once.
#   k_rhs_rk : reads U,V,W,nu_t (+Uo,F1 in substeps 2,3), writes U* (+F1 in
#              substep 1): 10, 13, 13 -> 12 on average per launch
#   k_dsm_fused: reads U,V,W, writes |S| (pass A of the dynamic model; the
#              filters, L_ij, M_ij and the plane sums stay on chip)
#   k_dsm_nut: reads |S|, writes nu_t (pass B)
#   k_fftx_fwd: reads U,V,W (divergence), writes p-hat (complex, half spectrum = 1 double/cell)
#   k_fftz: reads and writes p-hat (1 complex per 2 cells each way: 2 doubles/cell)
#   k_tridiag: reads p-hat once, writes it once (the rows of a system stay in
#              registers between the sweeps): 2
#   k_fftx_inv: reads p-hat (1), writes p (1)
#   k_project: reads p, U,V,W, writes U,V,W: 7
# A whole step of the dynamic model (mass-flow mode 1) moves 107 values per
# cell (856 B FP64, 548 B FP32 bulk): compute_dt rides in the first pass A,
# pass B runs in the last substep only (the RHS converts |S| in the others)
# and the mass-flow shift in the last projection.  step_alg_gbs sums these
# per launch over the launches the profiled steps make.
ALG_VALUES = {
    "k_rhs_rk": (12, 0),
    "k_dsm_fused": (4, 0),
    "k_dsm_nut": (2, 0),
    "k_fftx_fwd": (3, 1),
    "k_fftz": (0, 2),
    "k_tridiag": (0, 2),
    "k_fftx_inv": (0, 2),
    "k_project": (6, 1),
    # the AMD eddy viscosity (LES model 3): reads U,V,W, writes nu_t
    "k_amd": (4, 0),
    # compute_dt on its own (reads U,V,W) and the mass-flow shift on its own
    # (reads and writes U), where the step does not fold them
    "k_dt_partial": (3, 0),
    "k_shift_u": (2, 0),
}


def parse():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="wmles", choices=["small", "wmles", "large"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--les", default="dsm", choices=["dsm", "amd"],
                    help="subgrid model: dynamic Smagorinsky (the workload's) or AMD (LES model 3)")
    ap.add_argument("--profile-steps", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="at most this many oracle steps for the CPU baseline (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=5.0,
                    help="stop each CPU baseline run after the step that passes this many seconds")
    ap.add_argument("--cpu-runs", type=int, default=3, help="runs of the CPU baseline sample")
    ap.add_argument("--ensemble", type=int, default=64, help="channels of the batched ensemble (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=1000,
                    help="steps of the batched ensemble (its wall-stress error is read after them)")
    ap.add_argument("--ensemble-timed", type=int, default=100,
                    help="steps timed for each form of the ensemble")
    ap.add_argument("--tile-shapes", default="",
                    help="RHS,DSM tile shape indices to use instead of the autotuner's picks (counter runs that "
                         "must launch the same instances as a traced run)")
    ap.add_argument("--profile-compare", type=int, default=20000,
                    help="steps of the FP32-bulk vs FP64 mean-profile comparison (0: skip); the second "
                         "half is averaged")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _fluct_diff(a, b):
    """rms of the difference of two channels' U fluctuations (about their
    plane means) relative to the rms of the first's: ~1.4 once two runs are
    decorrelated, 0 when identical."""
    import channel_rl_amd as cr
    Ua, Ub = a.get_field(cr.U), b.get_field(cr.U)
    fa = Ua - Ua.mean(axis=(0, 2), keepdims=True)
    fb = Ub - Ub.mean(axis=(0, 2), keepdims=True)
    return float(np.sqrt(np.mean((fa - fb) ** 2)) / np.sqrt(np.mean(fa ** 2)))


def compare_fp32_profiles(case, field, args, cr):
    """The FP32 bulk against FP64 over a long run of this channel: both from
    the same initial field, the first half discarded, plane statistics
    averaged every 50 steps over the second half; how far the two
    trajectories have separated when the averaging starts (fluctuation rms
    of the difference over that of one run: ~1.4 for independent fields),
    the relative differences of the mean U+ profile, <u'v'>+ and u_tau, and
    ms/step of each precision."""
    from channel_rl_amd import statistics as S
    from channel_rl_amd.ensemble import make_channel
    n = args.profile_compare
    half = n // 2
    chans = {p: make_channel(case, field, precision=p, wall=cr.WALL_EQUILIBRIUM) for p in (cr.FP64, cr.FP32_BULK)}
    ms = {}
    for p, c in chans.items():
        t0 = time.perf_counter()
        c.run_steps(half)
        c.synchronize()
        ms[p] = 1e3 * (time.perf_counter() - t0) / half
    sep = _fluct_diff(chans[cr.FP64], chans[cr.FP32_BULK])
    res = {}
    for p, c in chans.items():
        a = S.ProfileAverager(c)
        a.run(n - half, 50)
        res[p] = a.result()
    sep_end = _fluct_diff(chans[cr.FP64], chans[cr.FP32_BULK])
    for c in chans.values():
        c.close()
    d = S.profile_difference(res[cr.FP64], res[cr.FP32_BULK])
    out = {"steps": n, "averaged_steps": n - half, "samples": int(res[cr.FP64]["samples"]),
           "fluct_separation_at_average_start": round(sep, 3), "fluct_separation_at_end": round(sep_end, 3),
           "Uplus_rel_diff": d["Uplus"], "uvplus_rel_diff": d["uvplus"], "u_tau_rel_diff": d["u_tau"],
           "Re_tau_fp64": round(float(res[cr.FP64]["Re_tau"]), 2),
           "Re_tau_fp32": round(float(res[cr.FP32_BULK]["Re_tau"]), 2),
           "fp32_ms_per_step": round(ms[cr.FP32_BULK], 4), "fp64_ms_per_step": round(ms[cr.FP64], 4)}
    log(f"[bench] FP32 bulk vs FP64, {n} steps (second half averaged, separation {sep:.2f} -> {sep_end:.2f}): "
        f"U+ {d['Uplus']:.2e}, <u'v'>+ {d['uvplus']:.2e}, u_tau {d['u_tau']:.2e}")
    return out


def measure_batched(ch, case, args, cr, cases):
    """The learned wall model and the batched ensemble: train the network on
    data of the small channel (10^6 samples, 20 epochs), run this channel with
    it, then the ensemble's 64 channels four ways, each timed over the same
    steps: plain (FP64, equilibrium wall model, each channel's own captured
    step, one channel after the other: the single-channel form), plain with
    the batched form's settings (FP32 bulk + FP64 Poisson, the network in the
    loop, one after the other), batched with every member's own kernels (one
    graph of per-member branches), and batched with grouped launches (one
    launch per stage for all channels of a shape).  The grouped ensemble then
    runs to the configured step count and reports its wall-stress error."""
    from channel_rl_amd import ensemble as E
    from channel_rl_amd import wallnet
    t0 = time.perf_counter()
    X, Y = wallnet.solver_samples(1_000_000, seed=7)
    t1 = time.perf_counter()
    net, losses = wallnet.train(X, Y, epochs=20)
    t2 = time.perf_counter()
    Xv, Yv = wallnet.solver_samples(40_000, seed=8, develop_steps=80)
    val = float(np.mean(wallnet.stress_error(net, Xv, Yv)))
    log(f"[bench] wall network: samples {t1 - t0:.1f} s, training {t2 - t1:.1f} s, loss {losses[-1]:.3g}, "
        f"validation error {val:.4f}")
    # this channel with the network as its wall model
    ch.set_wall_network(net, 0.05)
    ch.set_wall_model(cr.WALL_MLP, h_wm=case.h_wm)
    ch.run_steps(5)
    ch.synchronize()
    n = 50
    s0 = time.perf_counter()
    ch.run_steps(n)
    ch.synchronize()
    ms_mlp = 1e3 * (time.perf_counter() - s0) / n
    err1, fb1 = ch.wall_model_stats()
    log(f"[bench] {case.name} with the network: {ms_mlp:.3f} ms/step, error {err1:.4f}, fallback {fb1:.4f}")
    mem = cases.ensemble(args.ensemble)
    ncell = sum(c.cells for c, _ in mem)
    k = args.ensemble_timed
    times, tuned = {}, {}
    for name, prec, wall in (("plain_fp64_equilibrium", cr.FP64, cr.WALL_EQUILIBRIUM),
                             ("plain_fp32_network", cr.FP32_BULK, cr.WALL_MLP)):
        chans = [E.make_channel(c, f, precision=prec, wall=wall, net=net) for c, f in mem]
        tuned[name] = E.tune_members(chans)
        times[name] = E.time_sequential(chans, k)
        for c in chans:
            c.close()
        log(f"[bench] ensemble {name}: {len(mem) * k / times[name]:.1f} channel-steps/s")
    chans = [E.make_channel(c, f, precision=cr.FP32_BULK, wall=cr.WALL_MLP, net=net) for c, f in mem]
    tuned["batched"] = E.tune_groups(chans, E.tune_members(chans))
    ens = cr.Ensemble(chans)
    ens.set_grouping(False)
    times["batched_member_kernels"] = E.time_batched(ens, k)
    log(f"[bench] ensemble batched_member_kernels: {len(mem) * k / times['batched_member_kernels']:.1f} "
        "channel-steps/s")
    ens.set_grouping(True)
    groups = ens.group_count()
    times["batched_grouped"] = E.time_batched(ens, k)
    log(f"[bench] ensemble batched_grouped ({groups} groups): {len(mem) * k / times['batched_grouped']:.1f} "
        "channel-steps/s")
    done = 2 * (k + 1)
    if args.ensemble_steps > done:
        ens.run_steps(args.ensemble_steps - done)
        ens.synchronize()
        done = args.ensemble_steps
    werr, wfb = E.wall_error(chans)
    finite = all(np.all(np.isfinite(c.get_field(cr.U))) for c in chans)
    ens.close()
    for c in chans:
        c.close()
    cps = {name: round(len(mem) * k / t, 1) for name, t in times.items()}
    tg = times["batched_grouped"]
    out = {
        "workload": f"{len(mem)} channels, faces drawn from 34 33 34 / 66 49 66 / 98 65 98, Re_tau 180..5200, "
                    f"{ncell} cells, wall model at h_wm = 0.1",
        "timed_steps": k,
        "channel_steps_per_s": cps,
        "groups": groups,
        "tile_shapes": {name: {f"{k[0]} {k[1]} {k[2]}": {"rhs": v[0], "dsm": v[1]} for k, v in t.items()}
                        for name, t in tuned.items()},
        # the batched + learned-wall-model form against the plain single-channel form
        "speedup": round(times["plain_fp64_equilibrium"] / tg, 3),
        # grouped batching alone, same precision and wall model on both sides
        "batching_speedup_same_settings": round(times["plain_fp32_network"] / tg, 3),
        "grouping_speedup_over_member_kernels": round(times["batched_member_kernels"] / tg, 3),
        "batched_mcells_per_s": round(ncell * k / tg / 1e6, 1),
        "batched_alg_hbm_frac": round(ncell * 584 * k / tg / (HBM_PEAK_GBS * 1e9), 4),
        "steps_run": done,
        "all_finite": bool(finite),
        "ensemble_wall_stress_error": round(werr, 4),
        "ensemble_fallback_fraction": round(wfb, 4),
        "network": {"samples": len(X), "epochs": 20, "train_s": round(t2 - t1, 2), "final_loss": float(losses[-1]),
                    "validation_error": round(val, 4)},
        "single_channel_mlp_ms_per_step": round(ms_mlp, 4),
        "single_channel_mlp_error": round(err1, 4),
        "single_channel_mlp_fallback": round(fb1, 4),
    }
    log(f"[bench] ensemble after {done} steps: wall error {werr:.4f}, fallback {wfb:.4f}, finite {finite}")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, ch, q0, args):
    """The oracle (the reference's algorithm restated in C, OpenMP over its
    loops with disjoint outputs) on this host's cores, FP64, on a bounded
    sample of the same workload: args.cpu_runs runs of whole RK3 steps
    (compute_dt + compute_time_step_RK3) from the GPU's current state, each
    until args.cpu_seconds have passed (at most args.cpu_steps steps); the
    mean and spread over the runs, and ms per step of each stage."""
    import oracle
    flags = oracle.use_fast_build()
    U, V, W = ch.state()
    nthr = oracle.num_threads()
    log(f"[bench] cpu baseline: oracle ({flags}) on {nthr} OpenMP thread(s), nproc {os.cpu_count()}, "
        f"{cpu_model()}")
    rates, steps_all = [], 0
    st = np.zeros(len(oracle.STAGE_NAMES))
    for r in range(args.cpu_runs):
        o = oracle.Oracle(case.x, case.y, case.z, **case.kwargs())
        o.set_state(U, V, W)
        o.set_scalars(dt=0.0, t=0.0, dPdx=0.0, Qflow_x_0=q0)
        o.set_wall_model(1, h_wm=case.h_wm)
        o.compute_dt()
        o.compute_time_step_RK3()   # warm-up step, untimed
        st0 = oracle.stage_times()
        c0 = time.perf_counter()
        done = 0
        while done < args.cpu_steps:
            o.compute_dt()
            o.compute_time_step_RK3()
            done += 1
            if time.perf_counter() - c0 >= args.cpu_seconds:
                break
        sec = time.perf_counter() - c0
        st += oracle.stage_times() - st0
        rates.append(case.cells * done / sec / 1e6)
        steps_all += done
        log(f"[bench] cpu run {r + 1}: {done} step(s) in {sec:.1f} s, {rates[-1]:.4f} Mcells/s")
        del o
    stages = {name: round(float(1e3 * t / steps_all), 2) for name, t in zip(oracle.STAGE_NAMES, st) if t > 0}
    return {"value": round(float(np.mean(rates)), 4), "unit": "Mcells/s", "cores": nthr, "kind": "port",
            "runs": [round(x, 4) for x in rates], "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "flags": flags, "ms_per_step_by_stage": stages,
            "sample": f"{args.cpu_runs} run(s), {steps_all} RK3 step(s) in all, of the same {case.nx}x{case.ny}x"
                      f"{case.nz} channel from the GPU run's state (after one untimed step each), FP64 C oracle "
                      f"(OpenMP over loops with disjoint outputs), {nthr} thread(s), mean over the runs"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: one channel runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding the channel across GPUs is not built)", file=sys.stderr)
        return 2
    import channel_rl_amd as cr
    from channel_rl_amd import cases

    prec = cr.FP64 if args.precision == "fp64" else cr.FP32_BULK
    case, field = cases.config(args.config, seed=1234)
    if args.les == "amd":
        case.LES_model = 3
    ch = cr.Channel(case.x, case.y, case.z, precision=prec, **case.kwargs())
    # initial field: boundary conditions, projection to div = 0, initial bulk flow
    ch.set_state(*field)
    ch.apply_boundary_conditions()
    ch.compute_projection_step()
    ch.initialize_mass_flow()
    q0 = ch.scalars()["Qflow_x_0"]
    ch.set_scalars(dt=0.0, t=0.0, dPdx=case.dPdx, Qflow_x_0=q0)
    ch.set_wall_model(cr.WALL_EQUILIBRIUM, h_wm=case.h_wm)
    cells = case.cells
    if args.tile_shapes:
        rs, ds = (int(v) for v in args.tile_shapes.split(","))
        ch.set_tile_shape(rs, cr.KERNEL_RHS)
        if case.LES_model == 2:
            ch.set_tile_shape(ds, cr.KERNEL_DSM)
        shapes = {"rhs": ch.tile_shape(cr.KERNEL_RHS), "dsm": ch.tile_shape(cr.KERNEL_DSM), "set": "--tile-shapes"}
        log(f"[bench] tile shapes set: RHS {shapes['rhs']}, DSM {shapes['dsm']}")
    else:
        tune = ch.autotune(reps=5)
        shapes = {"rhs": ch.tile_shape(cr.KERNEL_RHS), "dsm": ch.tile_shape(cr.KERNEL_DSM),
                  "rhs_us": [round(1e3 * t, 1) for t in tune["rhs"]], "dsm_us": [round(1e3 * t, 1) for t in tune["dsm"]]}
        log(f"[bench] autotuned tile shapes: RHS {shapes['rhs']} ({' '.join(map(str, shapes['rhs_us']))} us), "
            f"DSM {shapes['dsm']} ({' '.join(map(str, shapes['dsm_us']))} us)")
    log(f"[bench] {case.name} faces {case.nx} {case.ny} {case.nz}, {cells} cells, {args.precision}")

    ch.run_steps(args.warmup)
    ch.synchronize()
    log(f"[bench] warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    ch.run_steps(args.steps)
    ch.synchronize()
    t1 = time.perf_counter()
    sec = t1 - t0
    ms_step = 1e3 * sec / args.steps
    value = cells * args.steps / sec / 1e6
    log(f"[bench] timed {args.steps} steps: {ms_step:.3f} ms/step, {value:.1f} Mcells/s")
    _, maxdiv = ch.check_divergence()

    # per-kernel HIP-event durations on the library's stream
    prof = ch.profile_steps(args.profile_steps)
    tot = sum(v[0] for v in prof.values())
    ranked = sorted(prof.items(), key=lambda kv: -kv[1][0])
    for name, (ms, n) in ranked[:12]:
        log(f"[bench]   {name:22s} {ms / args.profile_steps:8.3f} ms/step  {1e3 * ms / n:9.2f} us/launch  "
            f"{100 * ms / tot:5.1f}%")
    dom = next(n for n, _ in ranked if n in ALG_VALUES)
    ms_tot, launches = prof[dom]
    avg_s = ms_tot / launches / 1e3
    bulk_b = 8 if prec == cr.FP64 else 4
    vb, vp = ALG_VALUES[dom]
    alg_bytes = cells * (vb * bulk_b + vp * 8)
    achieved = alg_bytes / avg_s / 1e9
    # HBM bytes per launch of the same kernel from the two rocprofv3 counter
    # passes over this workload, scaled by the copy kernel (null if absent)
    from channel_rl_amd import rocprof_summary
    tag = f"{case.nx - 2}x{case.ny}x{case.nz - 2}_{args.precision}" + ("_amd" if args.les == "amd" else "")
    traffic_all = rocprof_summary.load_traffic(os.path.join(ROOT, "profiles", f"hbm_traffic_channel_step_{tag}.md"))
    traffic = traffic_all.get(dom)
    # every stage of the step: HIP-event time per launch, algorithmic bytes
    # (where counted) and the HBM bytes of the counter passes, each as a
    # rate and as a fraction of the HBM peak
    stages = {}
    step_alg_bytes = 0
    for name, (ms, n) in ranked:
        if name in ALG_VALUES:
            b0, b1 = ALG_VALUES[name]
            step_alg_bytes += cells * (b0 * bulk_b + b1 * 8) * n / args.profile_steps
        us = 1e3 * ms / n
        st = {"us_per_launch": round(us, 2), "launches_per_step": round(n / args.profile_steps, 2)}
        if name in ALG_VALUES:
            b0, b1 = ALG_VALUES[name]
            ab = cells * (b0 * bulk_b + b1 * 8)
            st["alg_bytes"] = ab
            st["alg_hbm_frac"] = round(ab / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
        # the profiles name the tridiagonal kernels by their form (k_tridiag_tw: rows in registers)
        tb = traffic_all.get(name, traffic_all.get(name + "_tw"))
        if tb is not None:
            st["hbm_bytes"] = int(tb)
            st["hbm_frac"] = round(tb / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
        stages[name] = st
    copy_gbs, copy_ms = ch.stream_copy(1 << 30, reps=10)
    log(f"[bench] stream copy 1 GiB: {copy_gbs:.0f} GB/s (read + write), {copy_ms:.3f} ms")
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None if traffic is None else int(traffic),
                "avg_launch_us": round(avg_s * 1e6, 2), "alg_bytes_per_launch": alg_bytes,
                "step_ms_profiled": round(tot / args.profile_steps, 3),
                "copy_kernel_gbs": round(copy_gbs, 1),
                "step_alg_gbs": round(step_alg_bytes / (ms_step * 1e-3) / 1e9, 1),
                "step_alg_bytes_per_cell": round(step_alg_bytes / cells, 1)}
    if dom in ALG_FLOPS:
        # the same kernel against the vector ALU peak of its arithmetic type
        # (no MFMA on this path): its bound is issue and latency, not HBM
        pk = VALU_PEAK_TFLOPS["fp64" if prec == cr.FP64 else "fp32"]
        tf = ALG_FLOPS[dom] * cells / avg_s / 1e12
        roofline["compute"] = {"unit": "TFLOP/s", "alg_flop_per_cell": ALG_FLOPS[dom], "achieved": round(tf, 2),
                               "peak": pk, "frac": round(tf / pk, 4)}

    fp32 = None
    if args.profile_compare > 0:
        fp32 = compare_fp32_profiles(case, field, args, cr)

    batched = None
    if args.ensemble > 0:
        batched = measure_batched(ch, case, args, cr, cases)

    cpu = None
    if args.cpu_steps > 0:
        cpu = cpu_baseline(case, ch, q0, args)
        log(f"[bench] cpu: {cpu['value']} Mcells/s ({cpu['runs']})")

    out = {
        "metric": "channel cell-RK3-steps/s (Mcells/s)",
        "value": round(value, 2),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == cr.FP64 else "f32-bulk/f64-poisson",
        "data": "synthetic",
        "config": {"workload": f"channel LES RK3 step, {case.nx - 2}x{case.ny}x{case.nz - 2} "
                               f"(faces {case.nx} {case.ny} {case.nz}), Re_tau {case.re_tau:.0f}, "
                               f"{'AMD' if args.les == 'amd' else 'dynamic Smagorinsky'}, equilibrium wall model at h_wm = {case.h_wm}, constant mass flow, "
                               f"{args.precision} bulk",
                   "cells": cells, "parallelism": "single GPU", "max_abs_div": maxdiv},
        "tile_shapes": shapes,
        "roofline": roofline,
        "stages": stages,
        "cpu_baseline": cpu,
        "batched": batched,
        "fp32_bulk": fp32,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
