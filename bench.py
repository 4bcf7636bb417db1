"""Benchmark: ns/day and pair-interactions/s of the 30,000-atom SPME water
box in mixed precision on one MI355X, with the accuracy against FP64, the
pair kernel's roofline and the FP64 CPU oracle timed on the host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c2|batch]

One step = one MD step of the whole system (VV stage 1 with SHAKE, list
rebuild every 20 steps, SPME + real-space Ewald/LJ forces, VV stage 2
with SHAKE), all on the device.  Prints one JSON line.  The default
workload is the 30k-atom box; c2 is the 192k-atom box and batch the 256
uneven replica boxes (one context, every kernel covering all replicas).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "ns/day & pair-interactions/s, 30k-atom SPME water; mixed vs FP64 @ stated err"
FP32_PEAK = 157.3e12          # MI355X dense FP32 vector, FLOP/s
HBM_PEAK = 8.0e12             # MI355X HBM3E, B/s
FLOP_IN, FLOP_OUT = 64, 8     # per in-cutoff pair / per listed pair beyond rc

WORKLOADS = {
    "c1": ("tip3p_10k_mixed: 10,000 rigid TIP3P waters (30,000 atoms), 67.0 A cubic box, seed 1; "
           "SPME 64^3 order 4, alpha 0.32; cutoff 12 A, LJ switch from 10 A, skin 2 A, list every "
           "20 steps; SHAKE; dt 2 fs; FP32 pair/spline math, int64 fixed-point forces"),
    "c2": ("tip3p_64k_mixed: 64,000 rigid TIP3P waters (192,000 atoms), 124.1 A cubic box, seed 2; "
           "SPME 128^3 order 4, alpha 0.32; cutoff 12 A, switch 10 A, skin 2 A; SHAKE; dt 2 fs; "
           "mixed"),
    "batch": ("tip3p_replicas_256_mixed: 256 independent TIP3P boxes of 216-512 waters (0.0334 "
              "molecules/A^3), 280-400 K, seed 3; cutoff 9 A, switch 7.5 A, alpha 0.35, SPME mesh "
              "~1 A rounded up to a multiple of 4, order 4; SHAKE; dt 2 fs; mixed; one context"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def sync_fn():
    """hipDeviceSynchronize of the HIP runtime libmdgpu.so links (the one
    runtime in the process; every GPU call of the benchmark goes through
    the library)."""
    import ctypes
    import mdbas_amd as M
    M.load()
    rt = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                  "libamdhip64.so.7"))

    def sync():
        if rt.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    return sync


def copy_calibration(mib: int):
    """One device-to-device copy of mib MiB through the HIP runtime (a copy
    kernel that reads and writes exactly that many bytes): the known size
    against which the FETCH_SIZE / WRITE_SIZE counters of a profiled run are
    scaled."""
    import ctypes
    rt = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                  "libamdhip64.so.7"))
    nbytes = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if rt.hipMalloc(ctypes.byref(a), nbytes) or rt.hipMalloc(ctypes.byref(b), nbytes):
        raise RuntimeError("hipMalloc failed")
    try:
        if rt.hipMemset(a, 1, nbytes) or rt.hipDeviceSynchronize():
            raise RuntimeError("hipMemset failed")
        if rt.hipMemcpy(b, a, nbytes, 3) or rt.hipDeviceSynchronize():   # device to device
            raise RuntimeError("hipMemcpy failed")
    finally:
        rt.hipFree(a)
        rt.hipFree(b)
    return {"copy_calibration_mib": mib, "bytes_read": mib << 20, "bytes_written": mib << 20}


def cpu_baseline(boxes, budget_s=25.0, max_steps=10):
    """The FP64 oracle (a from-scratch port of the reference's path) on one
    host core: per box one list build, then MD steps (VV stages and
    energies); seconds per step = step + build/update, summed over boxes."""
    import oracle as O
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    t_all, steps_done, descr = 0.0, [], []
    per_box = budget_s / len(boxes)
    for wb in boxes:
        cfg = dict(wb.params)
        o = O.OracleSystem(wb, cfg)
        x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
        vx, vy, vz = wb.vx.copy(), wb.vy.copy(), wb.vz.copy()
        t0 = time.perf_counter()
        o.build_list(x, y, z, build=True)
        t_list = time.perf_counter() - t0
        f, e = o.energy(x, y, z)
        steps, t_step, tb = 0, 0.0, time.perf_counter()
        while True:
            t0 = time.perf_counter()
            o.vv(1, x, y, z, vx, vy, vz, f)
            f, e = o.energy(x, y, z)
            o.vv(2, x, y, z, vx, vy, vz, f)
            t_step += time.perf_counter() - t0
            steps += 1
            if time.perf_counter() - tb > per_box * 0.6 or steps >= max_steps:
                break
        t_all += t_step / steps + t_list / cfg["update"]
        steps_done.append(steps)
        descr.append((t_list, t_step / steps))
    return t_all, steps_done, descr


def pair_traffic(name):
    """HBM bytes per pair-kernel launch of the 30k-atom box, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated against a
    copy of known size); None for other workloads."""
    p = os.path.join(HERE, "profiles", "pair_kernel_traffic.json")
    if name != "c1" or not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)["bytes_per_launch"]


def spme_traffic(name):
    """HBM bytes of the mixed SPME kernels per step of the 30k-atom box, from
    the same committed FETCH_SIZE / WRITE_SIZE runs; None for other
    workloads."""
    p = os.path.join(HERE, "profiles", "spme_traffic.json")
    if name != "c1" or not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)["bytes_per_step"]


DRIFT_STEPS = 1000           # fixed drift window, independent of --steps
CPU_STEPS = {"c1": 100, "c2": 20}   # oracle MD steps timed on one core (c1: 5 list rebuilds)


def drift_per_water(recs, dt_ps, n_waters):
    """Slope of a linear fit to E_tot(t), kcal/mol/ns per water."""
    e = recs["tot"] / 418.4
    t_ns = np.arange(len(e)) * dt_ps * 1e-3
    return float(np.polyfit(t_ns, e, 1)[0]) / n_waters


def oracle_md(wb, nsteps):
    """The FP64 oracle's MD loop (oc_md_run: the first list build, the
    step-0 energy, then nsteps steps with a list rebuild every 20) on one
    pinned host core; returns (seconds, E_tot per step)."""
    import oracle as O
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    err, fo, et, ep, ek, sec = o.md_run(nsteps, *s)
    if err:
        raise RuntimeError(f"oracle MD run failed: {err}")
    return sec, et


def oracle_forces(wb):
    import oracle as O
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    o.build_list(x, y, z)
    return o.energy(x, y, z)


def drift_and_fp64(args, M, wb, cfg, name, ctx, sync):
    """Energy drift over a fixed window of DRIFT_STEPS steps after the
    warm-up, mixed and FP64 from the same state (the FP64 run is also timed:
    ns/day of the plain FP64 method on the GPU), then both device paths
    from the initial state for the CPU leg's steps, to set beside the
    oracle's trajectory.  Closes ctx."""
    dt, W = cfg["timeStep"], args.warmup
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    ctx.initial_energy()
    ctx.run(W, use_graph=True)
    rd = M.records_to_dict(ctx.run(DRIFT_STEPS, use_graph=True))
    drift_mixed = drift_per_water(rd, dt, wb.n_waters)
    c64 = M.Context(wb, precision=M.MDG_FP64)
    c64.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    c64.initial_energy()
    c64.run(W, use_graph=True)
    sync()
    t0 = time.perf_counter()
    r64 = M.records_to_dict(c64.run(DRIFT_STEPS, use_graph=True))
    sync()
    wall64 = time.perf_counter() - t0
    drift_64 = drift_per_water(r64, dt, wb.n_waters)
    fp64_timing = {"ns_per_day": dt * 1e-3 * DRIFT_STEPS * 86400.0 / wall64,
                   "ms_per_step": wall64 / DRIFT_STEPS * 1e3, "steps": DRIFT_STEPS,
                   "launch": "default shapes (not autotuned)"}
    log(f"FP64 on the GPU: {fp64_timing}; drift mixed {drift_mixed:.3f} fp64 {drift_64:.3f}")

    # both device paths from the initial state for the CPU leg's steps, to
    # set beside the oracle's trajectory
    ncpu = CPU_STEPS[name]
    dev_tot = {}
    for prec, c in ((M.MDG_MIXED, ctx), (M.MDG_FP64, c64)):
        c.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        r0 = c.initial_energy()
        rr = M.records_to_dict(c.run(ncpu, use_graph=True))
        dev_tot[prec] = np.concatenate([[r0.tot], rr["tot"]])
    c64.close()
    ctx.close()

    return drift_mixed, drift_64, fp64_timing, dev_tot


def single_box(args, M, configs, name):
    cfg = configs.CONFIGS[name]
    wb = configs.build(cfg)
    N, K, W = wb.n, args.steps, args.warmup
    dt = cfg["timeStep"]
    log(f"workload {cfg['name']}: {N} atoms, K={K}, W={W}")
    sync = sync_fn()

    # accuracy at the initial coordinates: mixed vs the device FP64 path,
    # and both against the FP64 oracle
    forces, pots = {}, {}
    for prec in (M.MDG_FP64, M.MDG_MIXED):
        c = M.Context(wb, precision=prec)
        c.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        pots[prec] = c.initial_energy().pot
        forces[prec] = np.stack(c.download()[6:9])
        c.close()
    f64, fmx = forces[M.MDG_FP64], forces[M.MDG_MIXED]
    eps_rms = rel(fmx, f64)
    eps_max = float((np.linalg.norm(fmx - f64, axis=0) / np.linalg.norm(f64, axis=0)).max())
    log(f"force error mixed vs FP64: rms {eps_rms:.3e} max {eps_max:.3e}")
    acc_oracle = None
    if not args.no_cpu:
        fo, eo = oracle_forces(wb)
        acc_oracle = {
            "force_rms_rel_err_mixed_vs_oracle": rel(fmx, fo),
            "force_max_rel_err_mixed_vs_oracle": float(
                (np.linalg.norm(fmx - fo, axis=0) / np.linalg.norm(fo, axis=0)).max()),
            "force_rms_rel_err_fp64_vs_oracle": rel(f64, fo),
            "pot_rel_err_mixed_vs_oracle": abs(pots[M.MDG_MIXED] / eo["pot"] - 1.0),
            "pot_rel_err_fp64_vs_oracle": abs(pots[M.MDG_FP64] / eo["pot"] - 1.0)}
        log(f"against the oracle: {acc_oracle}")

    ctx = M.Context(wb, precision=M.MDG_MIXED)
    if args.prune_skin is not None:
        ctx.set_prune(args.prune_skin)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    ctx.initial_energy()
    if args.launch:
        ctx.set_launch(*[int(v) for v in args.launch.split(",")])
    tuning = None if args.no_tune or args.launch else ctx.autotune(60)   # before the timed region
    log(f"autotuned launch shapes: {tuning}")
    if W:
        ctx.run(W, use_graph=True)
    sync()
    p0 = ctx.prune_stats()["pruned_launches"]
    t0 = time.perf_counter()
    recs = ctx.run(K, use_graph=True)                  # returns after a stream sync
    sync()
    wall = time.perf_counter() - t0
    R = M.records_to_dict(recs)
    ms_step = wall / K * 1e3
    ns_day = dt * 1e-3 * K * 86400.0 / wall
    pin = float(R["nInCut"].mean())
    pis = float(R["nInCut"].sum()) / wall
    stats = ctx.list_stats()
    pst = ctx.prune_stats()
    prune = dict(skin_A=pst["skin"], pruned_steps=(pst["pruned_launches"] - p0) / K,
                 pruned_entries=int(pst["pruned_entries"]), full_entries=int(stats["cluster_pairs"]))
    log(f"pruned list: {prune}")
    log(f"timed {K} steps: {wall:.3f} s, {ms_step:.4f} ms/step, {ns_day:.1f} ns/day")

    # kernel durations: up to 200 more steps as plain launches on one
    # stream, with HIP events around the pair kernel, SPME, the integrator,
    # the list rebuild and the whole step
    ctx.set_profiling(True)
    Kp = min(K, 200)
    recs2 = M.records_to_dict(ctx.run(Kp, use_graph=0))
    ph = ctx.phase_times()
    ctx.set_profiling(False)
    log(f"phases (ms): {ph}")
    plist = float(stats["atom_pairs"])
    pin2 = float(recs2["nInCut"].mean())
    flop = FLOP_IN * pin2 + FLOP_OUT * (plist - pin2)
    achieved = flop / (ph["pair_ms"] * 1e-3)
    G = cfg["grid"] ** 3
    spme_bytes = 56 * N + 32 * G

    drift_mixed = drift_64 = fp64_timing = None
    dev_tot = {}
    if args.profile_run:
        ctx.close()
    else:
        drift_mixed, drift_64, fp64_timing, dev_tot = drift_and_fp64(args, M, wb, cfg, name, ctx,
                                                                     sync)

    cpu = None
    traj_vs_oracle = None
    ncpu = CPU_STEPS[name]
    if not args.no_cpu and not args.profile_run:
        log(f"CPU baseline (FP64 oracle, 1 core, {ncpu} steps)...")
        sec, et = oracle_md(wb, ncpu)
        scale = abs(et[0])
        traj_vs_oracle = {
            "steps": ncpu,
            "max_rel_dEtot_fp64": float(np.abs(dev_tot[M.MDG_FP64] - et).max() / scale),
            "max_rel_dEtot_mixed": float(np.abs(dev_tot[M.MDG_MIXED] - et).max() / scale),
            # a 0.2 ps window: the fitted slope is mostly the energy's own
            # fluctuation, so it is given for the three runs side by side
            "window_slope_kcal_mol_ns_per_water": {
                "oracle": drift_per_water({"tot": et[1:]}, dt, wb.n_waters),
                "fp64": drift_per_water({"tot": dev_tot[M.MDG_FP64][1:]}, dt, wb.n_waters),
                "mixed": drift_per_water({"tot": dev_tot[M.MDG_MIXED][1:]}, dt, wb.n_waters)}}
        log(f"trajectory vs oracle: {traj_vs_oracle}")
        nreb = ncpu // cfg["update"]
        cpu = dict(value=dt * 1e-3 * ncpu * 86400.0 / sec, unit="ns/day", cores=1,
                   kind="port", pair_interactions_per_s=pin * ncpu / sec,
                   s_per_step=sec / ncpu,
                   sample=(f"FP64 oracle MD loop on 1 pinned host core ({os.uname().machine}, "
                           f"nproc {os.cpu_count()}): {ncpu} steps of the same {N}-atom box "
                           f"with {nreb} list rebuilds, plus the first list build and the step-0 "
                           f"energy, all inside the timed {sec:.1f} s"))
        log(f"cpu: {cpu}")

    return {
        "metric": METRIC if name == "c1" else METRIC.replace("30k-atom", f"{N // 1000}k-atom"),
        "value": ns_day,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[name], "atoms": N, "dt_fs": dt * 1e3,
                   "precision": "mixed"},
        "pair_interactions_per_s": pis,
        "pair_interactions_per_step": pin,
        "listed_pairs": plist,
        "cluster_pairs": int(stats["cluster_pairs"]),
        "pruned_list": prune,
        "pair_interactions_per_s_pair_kernel": pin2 / (ph["pair_ms"] * 1e-3),
        "fp64_gpu": fp64_timing,
        "accuracy": {"force_rms_rel_err_vs_fp64": eps_rms, "force_max_rel_err_vs_fp64": eps_max,
                     "vs_oracle": acc_oracle,
                     "energy_drift_kcal_mol_ns_per_water_mixed": drift_mixed,
                     "energy_drift_kcal_mol_ns_per_water_fp64": drift_64,
                     "drift_window_steps": DRIFT_STEPS,
                     "trajectory_vs_oracle": traj_vs_oracle},
        "phases_ms": ph,
        "tuning": tuning,
        "roofline": {"kernel": "k_pair_mixed", "bound": "valu",
                     "achieved": achieved / 1e12, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK, "traffic": pair_traffic(name),
                     "flop_per_launch": flop, "launch_ms": ph["pair_ms"]},
        "roofline_spme": {"bound": "hbm", "achieved": spme_bytes / (ph["spme_ms"] * 1e-3) / 1e9,
                          "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "frac": spme_bytes / (ph["spme_ms"] * 1e-3) / HBM_PEAK,
                          "bytes_per_step": spme_bytes, "traffic": spme_traffic(name)},
        "rebuild_ms_amortised": ph["rebuild_ms"] / cfg["update"],
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (ns_day / cpu["value"]) if cpu else None,
    }


def batch(args, M, configs):
    cfgs = configs.replica_configs(256)
    boxes = [configs.build(c) for c in cfgs]
    R, K, W = len(boxes), args.steps, args.warmup
    N = sum(b.n for b in boxes)
    log(f"workload batch: {R} replicas, {N} atoms, K={K}, W={W}")
    sync = sync_fn()
    b = M.BatchContext(boxes, precision=M.MDG_MIXED)
    if args.prune_skin is not None:
        b.set_prune(args.prune_skin)
    b.upload_boxes()
    b.initial_energy()
    if args.launch:
        b.set_launch(*[int(v) for v in args.launch.split(",")])
    tuning = None if args.no_tune or args.launch else b.autotune(60)
    log(f"autotuned launch shapes: {tuning}")
    if W:
        b.run(W, use_graph=True)
    sync()
    p0 = b.prune_stats()["pruned_launches"]
    t0 = time.perf_counter()
    recs = b.run(K, use_graph=True)
    sync()
    wall = time.perf_counter() - t0
    pst = b.prune_stats()
    prune = dict(skin_A=pst["skin"], pruned_steps=(pst["pruned_launches"] - p0) / K,
                 pruned_entries=int(pst["pruned_entries"]),
                 full_entries=int(b.list_stats()["cluster_pairs"]))
    log(f"pruned list: {prune}")
    D = M.batch_records(recs, R)
    dt = cfgs[0]["timeStep"]
    ns_day = dt * 1e-3 * K * 86400.0 / wall          # every replica advances this much
    pis = float(D["nInCut"].sum()) / wall
    log(f"timed {K} steps: {wall:.3f} s, {wall / K * 1e3:.4f} ms/step, {ns_day:.1f} ns/day per "
        f"replica, {R * ns_day:.0f} replica-ns/day")
    b.set_profiling(True)
    b.run(min(K, 100), use_graph=0)
    ph = b.phase_times()
    b.set_profiling(False)
    b.close()
    cpu = None
    if not args.no_cpu:
        sample = [boxes[k] for k in (0, R // 3, 2 * R // 3, R - 1)]
        t_sum, steps, descr = cpu_baseline(sample, budget_s=20.0, max_steps=5)
        t_all = t_sum * R / len(sample)                # all replicas, one after another
        cpu = dict(value=R * dt * 1e-3 * 86400.0 / t_all, unit="replica-ns/day", cores=1,
                   kind="port", s_per_step_all_replicas=t_all,
                   sample=(f"FP64 oracle on 1 host core: replicas 0, {R // 3}, {2 * R // 3}, "
                           f"{R - 1} ({steps} MD steps each after one list build), scaled to all "
                           f"{R} run one after another"))
        log(f"cpu: {cpu}")
    return {
        "metric": "replica-ns/day & pair-interactions/s, 256 uneven SPME water replicas, mixed",
        "value": R * ns_day,
        "unit": "replica-ns/day",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS["batch"], "replicas": R, "atoms": N, "dt_fs": dt * 1e3,
                   "precision": "mixed"},
        "ns_per_day_per_replica": ns_day,
        "pruned_list": prune,
        "pair_interactions_per_s": pis,
        "phases_ms": ph,
        "tuning": tuning,
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (R * ns_day / cpu["value"]) if cpu else None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c1")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--launch", type=str, default=None,
                    help="fixed launch shapes 'spread,pairblocks,spmefirst,spmeprio' instead of "
                         "autotuning")
    ap.add_argument("--prune-skin", type=float, default=None,
                    help="pruned-list distance beyond the cutoff, A (0: full list every step; "
                         "default: the library's choice for the box)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes (skip the autotuner)")
    ap.add_argument("--profile-run", action="store_true",
                    help="stop after the timed and profiled passes (no drift, FP64 or CPU "
                         "legs): the short run a profiler wraps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device once (the known-size run that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        print(json.dumps(copy_calibration(args.copy_calibration)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the sharded "
              "multi-GPU design is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    import mdbas_amd as M
    from mdbas_amd import configs
    if args.workload == "batch":
        out = batch(args, M, configs)
    else:
        out = single_box(args, M, configs, args.workload)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
