#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the force-and-integrate step on one MI355X.

Workload: 250,000 beads uniform in a 140 A cube (1.2 A minimum separation),
springs < 8 A with lognormal k (mean 0.5 kJ/mol/A^2), Amber 12-6 + Coulomb
with eps(r) = r at 12 A cutoffs, viscous Euler dt = 1 fs, neighbour list
rebuilt every 10 steps (and on drift).  Inputs are generated on the host from a
seed, uploaded once; the timed region is K full steps with everything resident
in HBM.

Printed JSON (one line): throughput, the roofline of the dominant kernel (HIP
events on the engine's stream over the timed region), the FP64-oracle force
error at the final state, and the oracle's CPU time on a bounded sample.  The
headline workload also times the spring network's construction from the 8 A
cutoff on the GPU (network_construction), with the reference's own Topology
build of a smaller network of the same density beside it.

Two helper modes serve the HBM-traffic measurement of the kernels
(rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE):
  --copy-calibration    a device-to-device copy of a known size, the yardstick
                        that converts the counters' KiB into bytes;
  --pmc-summary F C W D reads the four counter runs (FETCH bench/copy dirs,
                        WRITE bench/copy dirs) and writes
                        profiles/pmc_traffic.json, which later bench lines cite;
  --pmc-valu-summary D C  reads one run of the eight SQ_INSTS_VALU_* type
                        counters (dir D, made by command C) and writes
                        profiles/pmc_valu.json (executed FP32/FP64 FLOP).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 250k beads (fwd & adjoint), % roofline, force err vs FP64"
WORKLOADS = {
    1: "{n}-bead spring network, Amber 12-6 + Coulomb eps(r)=r, 12 A cutoffs, springs < 8 A, "
       "rebuild every 10 steps",
    2: "{n} CG beads, IMPALA curved double membrane (R 200 A, offsets 30 A), 4 rigid bodies, "
       "springs < 10 A, linear steric 4 A",
    3: "512 independent CA networks ({n} beads in all, 200-20,000 each, springs < 9-14 A), "
       "linear steric 4 A + Coulomb 16 A, one launch sequence replayed from a HIP graph, autotuned launch configuration",
    4: "adjoint of {T} steps ({ck}) of a {n}-bead network, Amber 12-6 + Coulomb eps(r)=r, "
       "12 A cutoffs, springs < 10 A; loss = MSD to a target displaced along 3 smooth modes (5 A)",
}
HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, GB/s
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector/matrix, TFLOP/s
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64, TFLOP/s
PAIR_FLOP = 40              # algorithmic flops per unique LJ+Coulomb pair
ADJ_PAIR_FLOP = 140         # adjoint: flops per unique LJ+Coulomb pair (both sides, H w and dL/dq)


COPY_BYTES = 1 << 30        # copy-calibration size


def host_cpu():
    """The host's CPU model and the threads this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_threads():
    """OpenMP threads for the CPU baselines: OMP_NUM_THREADS when set (the
    host's share), else every CPU this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))


def cpu_reference_baseline(seed, n_full, n_sample=50000, steps=2):
    """The reference's own step (SpringNetwork::computeStep compiled from its
    sources, -O3 OpenMP) on all host threads, over a bounded sample of the
    headline workload: the same generator, density and force field at
    n_sample beads, one untimed step (initRun, first neighbour build), then
    `steps` timed steps.  The reference has a constant dielectric only, so it
    runs the Coulomb term with eps(r) = eps; the pair count and loop structure
    are those of the eps(r) = eps r workload."""
    from biospring_amd import workloads
    from oracle import pyoracle as po
    thr = cpu_threads()
    po.ref_lib().ref_set_threads(thr)
    s = workloads.generate(1, seed=seed, n=n_sample)
    t = time.perf_counter()
    r = po.Reference(s)
    t_create = time.perf_counter() - t
    r.step(1)
    t = time.perf_counter()
    r.step(steps)
    dt = time.perf_counter() - t
    r.close()
    log(f"reference CPU baseline: {n_sample} beads, {steps} steps in {dt:.1f}s on {thr} threads "
        f"(network set-up {t_create:.1f}s)")
    out = {"value": n_sample * steps / dt, "unit": "particle-steps/s", "cores": thr, "kind": "reference",
           "sample": f"{steps} timed steps (after 1 untimed) of the reference's computeStep on a {n_sample}-bead "
                     f"network of the same workload (density, springs < 8 A, Amber 12-6 + Coulomb at 12 A; "
                     f"constant dielectric, the reference's only model), OpenMP on {thr} threads; the full "
                     f"workload has {n_full} beads", "seconds_per_step": dt / steps}
    out.update(host_cpu())
    return out


def hip_sync():
    ctypes.CDLL("libamdhip64.so").hipDeviceSynchronize()


def copy_calibration(reps=5):
    """hipMemcpy device-to-device of COPY_BYTES (a blit kernel) `reps` times."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(COPY_BYTES)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(COPY_BYTES)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(COPY_BYTES))
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(COPY_BYTES), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_bytes": COPY_BYTES, "reps": reps}))


def _counter_means(d, counter):
    """Mean counter value per kernel name over every dispatch in the
    rocprofv3 CSV output under directory d."""
    import csv
    import glob
    acc = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r.get("Counter_Name") != counter:
                    continue
                k = r["Kernel_Name"]
                acc.setdefault(k, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def pmc_summary(fetch_dir, fetch_copy_dir, write_dir, write_copy_dir):
    def pick(m, key):
        c = [v for k, v in m.items() if key in k]
        return c[0] if c else None
    out = {}
    cf = pick(_counter_means(fetch_copy_dir, "FETCH_SIZE"), "copy")
    cw = pick(_counter_means(write_copy_dir, "WRITE_SIZE"), "copy")
    sf = COPY_BYTES / (cf * 1024.0)
    sw = COPY_BYTES / (cw * 1024.0)
    fm = _counter_means(fetch_dir, "FETCH_SIZE")
    wm = _counter_means(write_dir, "WRITE_SIZE")
    for name, keys in (("pair", ("k_pair_q<1, 2, 0, false>", "k_pair<1, 2, 0, false>")),
                       ("spring_integrate", ("k_spring<16, true, false",)), ("cluster_list", ("k_cluster_list",))):
        f = next((v for v in (pick(fm, k) for k in keys) if v is not None), None)
        w = next((v for v in (pick(wm, k) for k in keys) if v is not None), None)
        if f is None or w is None:
            continue
        out[name] = {"fetch_bytes": f * 1024.0 * sf, "write_bytes": w * 1024.0 * sw,
                     "bytes": f * 1024.0 * sf + w * 1024.0 * sw}
    out["calibration"] = {"copy_bytes": COPY_BYTES, "copy_fetch_kib": cf, "copy_write_kib": cw,
                          "fetch_scale": sf, "write_scale": sw}
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


VALU_COUNTERS = ("SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_FMA_F32",
                 "SQ_INSTS_VALU_TRANS_F32", "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64",
                 "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_TRANS_F64")


def pmc_valu_summary(valu_dir, command):
    """profiles/pmc_valu.json from one rocprofv3 --pmc run of the eight
    SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_{F32,F64} counters: executed FLOP per
    launch = 64 lanes x (ADD + MUL + 2 FMA + TRANS) wave instructions."""
    out = {"command": command, "flop_rule": "64 lanes x (ADD + MUL + 2 FMA + TRANS) wave instructions, per type"}
    means = {c: _counter_means(valu_dir, c) for c in VALU_COUNTERS}
    for name, keys in (("pair", ("k_pair_q<1, 2, 0, false>", "k_pair<1, 2, 0, false>")),
                       ("spring_integrate", ("k_spring<16, true, false",))):
        row = {}
        for c in VALU_COUNTERS:
            cand = [v for k, v in means[c].items() if any(key in k for key in keys)]
            if cand:
                row[c] = cand[0]
        if len(row) != len(VALU_COUNTERS):
            continue
        flop = lambda t: 64.0 * (row[f"SQ_INSTS_VALU_ADD_{t}"] + row[f"SQ_INSTS_VALU_MUL_{t}"] +
                                 2.0 * row[f"SQ_INSTS_VALU_FMA_{t}"] + row[f"SQ_INSTS_VALU_TRANS_{t}"])
        out[name] = {"counters_per_dispatch": row, "executed_fp32_flop": flop("F32"),
                     "executed_fp64_flop": flop("F64")}
    path = os.path.join(ROOT, "profiles", "pmc_valu.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", type=int, default=1, choices=(1, 2, 3, 4),
                    help="1: 250k-bead LJ + Coulomb network (headline); 2: 50k CG beads, IMPALA curved double "
                         "membrane, 4 rigid bodies of 2,000 beads; 3: 512 independent CA networks of "
                         "200-20,000 beads stepped together from a HIP graph; 4: adjoint gradients of a "
                         "200-step trajectory loss on a 100k-bead network (--steps is the trajectory length)")
    ap.add_argument("--checkpoint", type=int, default=20,
                    help="adjoint: checkpoint interval (20, as config 4 states; 0: the engine's choice, the whole "
                         "trajectory stored when it fits in half the free HBM)")
    ap.add_argument("--no-adjoint", action="store_true", help="workload 1: skip the adjoint leg (config 4)")
    ap.add_argument("--systems", type=int, default=512, help="workload 3: number of systems")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch configuration first (always done for workload 3)")
    ap.add_argument("--n", type=int, default=0, help="bead count (default: the workload's own)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-construction", action="store_true",
                    help="workload 1: skip timing the spring-network construction")
    ap.add_argument("--no-trajectory", action="store_true",
                    help="workload 1: skip timing the run with trajectory writers")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH", "FETCH_COPY", "WRITE", "WRITE_COPY"))
    ap.add_argument("--pmc-valu-summary", nargs=2, metavar=("VALU_DIR", "COMMAND"))
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.pmc_summary:
        return pmc_summary(*args.pmc_summary)
    if args.pmc_valu_summary:
        return pmc_valu_summary(*args.pmc_valu_summary)
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.workload == 4:
        return bench_adjoint(args)
    from biospring_amd import _abi, workloads
    from biospring_amd.network import SpringNetwork

    t = time.perf_counter()
    bodies, off, parts = [], None, None
    if args.workload == 3:
        s, off, parts = workloads.batch(args.systems, seed=args.seed)
        s.params.graph_steps = 16
    else:
        s = workloads.generate(args.workload, seed=args.seed, n=args.n)
    log(f"generated n={s.n} springs={s.m} in {time.perf_counter() - t:.1f}s digest={s.digest()}")
    if args.workload == 2:
        s.params = workloads.impala_params(s, "double")
        bodies = workloads.rigid_groups(s, 4, min(2000, s.n // 8))
    net = SpringNetwork.from_system(s)
    for g in bodies:
        net.addRigidBody(g)
    if off is not None:
        net.setSystems(off)
    t = time.perf_counter()
    net.setup()
    log(f"setup {time.perf_counter() - t:.1f}s")
    tuned = None
    if args.autotune or args.workload == 3:
        t = time.perf_counter()
        tuned = net.autotune(10)
        log(f"autotune {tuned} in {time.perf_counter() - t:.1f}s")
    net.run(args.warmup)
    pairs_unique = net.countPairs(_abi.NS_ELECTROSTATIC if args.workload == 3 else _abi.NS_STERIC)
    cnt0 = net.counters()

    # kernel timing (HIP events) from a separate profiled run of the same
    # length; the timed run below uses the launch configuration as set
    # (events are not recorded inside a replayed graph)
    net.profile(True)
    net.run(min(args.steps, 50))
    prof = net.getProfile()
    net.profile(False)
    cnt0 = net.counters()
    hip_sync()
    t0 = time.perf_counter()
    net.run(args.steps)
    hip_sync()
    t1 = time.perf_counter()
    cnt1 = net.counters()
    elapsed = t1 - t0
    n = s.n
    value = n * args.steps / elapsed
    edges = cnt1["edges"]

    spring_ms = prof["spring"]["ms"] / max(prof["spring"]["launches"], 1)
    pair_ms = prof["pair"]["ms"] / max(prof["pair"]["launches"], 1)
    rebuild_ms = prof["rebuild"]["ms"] / max(prof["rebuild"]["launches"], 1)
    # {j, k, L0} per directed edge; {j, L0} when every spring has one
    # stiffness (the engine then reads no per-edge k)
    k_uniform = bool(s.m > 0 and np.all(s.k.view(np.uint32) == s.k.view(np.uint32)[0]))
    b_alg = 4 * (n + 1) + (8 if k_uniform else 12) * edges + 64 * n
    f_alg = PAIR_FLOP * pairs_unique
    spring_gbs = b_alg / (spring_ms * 1e-3) / 1e9
    pair_tflops = f_alg / (pair_ms * 1e-3) / 1e12
    kernels = {
        "spring_integrate": {"avg_ms": spring_ms, "bound": "hbm", "achieved": spring_gbs, "peak": HBM_PEAK_GBS,
                             "unit": "GB/s", "frac": spring_gbs / HBM_PEAK_GBS, "alg_bytes": b_alg,
                             "uniform_stiffness": k_uniform},
        "pair": {"avg_ms": pair_ms, "bound": "valu", "achieved": pair_tflops, "peak": FP32_PEAK_TFLOPS,
                 "unit": "TFLOP/s", "frac": pair_tflops / FP32_PEAK_TFLOPS, "alg_flops": f_alg,
                 "note": "compute-bound FP32/FP64 vector ALU work priced against the FP32 peak"},
        "rebuild_chain_avg_ms_per_step": rebuild_ms,
    }
    dom = "pair" if pair_ms >= spring_ms else "spring_integrate"
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tfile) and args.workload == 1 and n == 250000:
        # HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE runs
        with open(tfile) as fh:
            tt = json.load(fh)
        for name in ("pair", "spring_integrate"):
            if name in tt:
                kernels[name]["traffic_bytes"] = tt[name]["bytes"]
        traffic = tt.get(dom, {}).get("bytes")
    vfile = os.path.join(ROOT, "profiles", "pmc_valu.json")
    if os.path.exists(vfile) and args.workload == 1 and n == 250000:
        # executed VALU FLOPs per launch from the committed SQ_INSTS_VALU_* run,
        # over this run's average launch time
        with open(vfile) as fh:
            vv = json.load(fh)
        for name in ("pair", "spring_integrate"):
            if name in vv:
                ms = kernels[name]["avg_ms"]
                e32, e64 = vv[name]["executed_fp32_flop"], vv[name]["executed_fp64_flop"]
                kernels[name]["executed_valu"] = {"fp32_tflops": e32 / (ms * 1e-3) / 1e12,
                                                  "fp64_tflops": e64 / (ms * 1e-3) / 1e12,
                                                  "fp32_peak": FP32_PEAK_TFLOPS, "fp64_peak": FP64_PEAK_TFLOPS}
    d = kernels[dom]
    roofline = {"bound": d["bound"], "achieved": d["achieved"], "peak": d["peak"], "unit": d["unit"],
                "frac": d["frac"], "traffic": traffic, "kernel": dom}

    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload].format(n=n),
                   "n_beads": n, "rigid_bodies": [len(g) for g in bodies],
                   "systems": 1 if off is None else len(off) - 1,
                   "graph_steps": int(s.params.graph_steps) if tuned is None else tuned[1],
                   "spring_lanes": None if tuned is None else tuned[0],
                   "pair_layout": int(s.params.pair_layout) if tuned is None else tuned[2], "springs": s.m, "directed_edges": edges, "unique_pairs": pairs_unique,
                   "seed": args.seed, "digest": s.digest(),
                   "rebuilds_in_timed_region": cnt1["rebuilds"] - cnt0["rebuilds"],
                   "list_capacity": cnt1["list_cap"]},
        "roofline": roofline,
        "kernels": kernels,
        "adjoint": None,
    }
    if args.workload == 1 and not args.no_adjoint:
        # the adjoint half of the metric: config 4 as stated (100k beads, 200
        # steps, checkpoints every 20 steps with recomputation)
        out["adjoint"] = adjoint_leg(seed=args.seed, n=0, T=200, K=args.checkpoint or 20, warmup=5,
                                     fp64=not args.no_fp64, cpu=not args.no_cpu)

    if args.workload == 1 and not args.no_construction:
        out["network_construction"] = bench_construction(s, cpu=not args.no_cpu)
    if args.workload == 1 and not args.no_trajectory:
        out["trajectory_output"] = bench_trajectory(s, args.steps, out["ms_per_step"], every=max(1, args.steps // 4))

    x, v = net.getState()
    # Conservative force: the viscous term -gamma v is an exact product in both
    # codes and is compared separately in the tests.
    pc = s.params.copy()
    pc.viscosity_enable = 0
    net.setParams(pc)
    F, E = net.computeForces()
    if not args.no_fp64 or not args.no_cpu:
        from oracle import pyoracle as po
        if off is None:
            checks = [(s, slice(0, n), bodies)]
        else:  # a sample of the systems, each with its own oracle
            sizes = np.diff(off)
            pick = sorted(set(np.argsort(sizes)[-3:].tolist() + [0, len(sizes) // 2, len(sizes) - 1]))
            checks = [(parts[k], slice(off[k], off[k + 1]), []) for k in pick]
        if not args.no_fp64:
            t = time.perf_counter()
            per_max, glob_num, glob_den, erel = 0.0, 0.0, 0.0, 0.0
            Es = None if off is None else net.getSystemEnergies()
            for sy, sl, bb in checks:
                o = po.Oracle(sy, params=pc)
                o.set_state(x[sl], v[sl])
                for g in bb:
                    o.add_rigid_body(g)
                Fd, Ed = o.compute_forces(po.F64)
                dyn = (sy.flags & 1) == 0
                Fg = F[sl]
                dF = np.linalg.norm(Fg[dyn] - Fd[dyn], axis=1)
                nr = np.linalg.norm(Fd[dyn], axis=1)
                per_max = max(per_max, float((dF / np.maximum(nr, 1e-6 * nr.max())).max()))
                glob_num += float(np.sum((Fg[dyn] - Fd[dyn]) ** 2))
                glob_den += float(np.sum(Fd[dyn] ** 2))
                Eg = E if off is None else Es[int(np.searchsorted(off, sl.start))]
                erel = max(erel, max(abs(Eg[k] - Ed[k]) / max(abs(Ed[k]), 1e-3) for k in range(6)))
                o.close()
            if args.workload == 1:
                out["trajectory_err_vs_fp64"] = trajectory_check(args.seed)
            out["force_err_vs_fp64"] = {"per_particle_max": per_max, "global_rel_l2": (glob_num / glob_den) ** 0.5,
                                        "energy_rel_max": erel, "forces": "all terms but the viscous one",
                                        "state": f"after {args.warmup + args.steps + min(args.steps, 50)} steps",
                                        "systems_checked": len(checks)}
            log(f"fp64 check {time.perf_counter() - t:.1f}s")
        if not args.no_cpu and args.workload == 1 and po.ref_available():
            out["cpu_baseline"] = cpu_reference_baseline(args.seed, n)
        if not args.no_cpu:
            # bounded sample: the oracle steps the same system (or successive
            # systems of the batch) from the current state
            done, dt = 0, 0.0
            for sy, sl, bb in (checks if off is None else
                               [(parts[k], slice(off[k], off[k + 1]), []) for k in range(len(parts))]):
                o = po.Oracle(sy, params=s.params)
                o.set_state(x[sl], v[sl])
                for g in bb:
                    o.add_rigid_body(g)
                steps = args.cpu_steps if off is None else 2
                t = time.perf_counter()
                o.step(po.F32, steps)
                dt += time.perf_counter() - t
                done += sy.n * steps
                o.close()
                if dt > 12.0:
                    break
            port = {"value": done / dt, "unit": "particle-steps/s",
                                   "cores": int(po.oracle_lib().orc_num_threads()), "kind": "port",
                                   "sample": (f"{args.cpu_steps} full step(s) of the same {n}-bead workload"
                                              if off is None else
                                              f"2 steps of each system in turn, {done // 2} beads in all")
                                             + ", float-faithful oracle, OpenMP"}
            port.update(host_cpu())
            if "cpu_baseline" in out:
                out["cpu_baseline"]["port"] = port
            else:
                out["cpu_baseline"] = port
    print(json.dumps(out), flush=True)


def bench_construction(s, cutoff=8.0, reps=3, cpu=True, cpu_sample=5000):
    """Spring network from a cutoff (Topology::add_springs_from_cutoff) over
    the workload's own positions: the GPU build (wall time of
    bs_springs_from_cutoff, which leaves the table on the device, and HIP-event
    times of its two passes), checked spring for spring against the
    generator's host build; beside it the reference's own Topology on a
    bounded sample of the same density."""
    from biospring_amd import workloads
    from biospring_amd.network import SpringNetwork
    net = SpringNetwork(s.params)
    net.setParticles(*s.particle_arrays())
    net.springsFromCutoff(cutoff)  # first call: code objects, allocations
    net.profile(True)
    walls = []
    for _ in range(reps):
        net.setSprings([], [], [], [])
        hip_sync()
        t = time.perf_counter()
        m = net.springsFromCutoff(cutoff)
        walls.append(time.perf_counter() - t)
    prof = net.getProfile()
    net.profile(False)
    a, b, l0, _ = net.getSprings()
    same = bool(np.array_equal(a, s.id1) and np.array_equal(b, s.id2)
                and np.array_equal(l0.view(np.uint32), s.l0.view(np.uint32)))
    net.close()
    n = s.n
    count_ms = prof["topo_count"]["ms"] / max(prof["topo_count"]["launches"], 1)
    fill_ms = prof["topo_fill"]["ms"] / max(prof["topo_fill"]["launches"], 1)
    # count pass: positions in, counts out; fill pass: positions and row
    # offsets in, (j, l0) out per spring
    b_alg = 40 * n + 8 * m
    kern_ms = count_ms + fill_ms
    wall = float(np.median(walls))
    out = {"cutoff": cutoff, "springs": m, "identical_to_host_build": same,
           "wall_ms": wall * 1e3, "springs_per_s": m / wall,
           "kernel_ms": {"grid_and_count": count_ms, "fill_and_sort": fill_ms},
           "roofline": {"bound": "hbm", "achieved": b_alg / (kern_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": b_alg / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "alg_bytes": b_alg,
                        "note": "27-cell candidate gathers are served from L2 and not counted"},
           "wall_note": "wall time of the call: cell grid, both passes and the table written on the device, where "
                        "bs_setup reads it; the host copy is downloaded only when asked for (bs_get_springs)"}
    if cpu:
        from oracle import pyoracle as po
        if po.ref_available():
            smp = workloads.generate(1, seed=2, n=cpu_sample)
            t = time.perf_counter()
            r = po.ref_springs_from_cutoff(smp.x, smp.flags, cutoff)
            dt = time.perf_counter() - t
            out["cpu_baseline"] = {"value": len(r[0]) / dt, "unit": "springs/s", "cores": 1, "kind": "reference",
                                   "sample": f"Topology::add_springs_from_cutoff on a {cpu_sample}-bead network of "
                                             f"the same density ({len(r[0])} springs), compiled reference"}
        else:
            t = time.perf_counter()
            r = po.springs_from_cutoff(s.x, s.flags, cutoff)
            dt = time.perf_counter() - t
            out["cpu_baseline"] = {"value": len(r[0]) / dt, "unit": "springs/s", "cores": 1, "kind": "port",
                                   "sample": f"the oracle's cell-grid build on the same {n} beads"}
    log(f"construction: {m} springs in {wall * 1e3:.1f} ms wall, kernels {kern_ms:.2f} ms, same={same}")
    return out


def bench_trajectory(s, steps, ms_plain, every=50):
    """The same run with an XTC and a CSV writer taking a frame every `every`
    iterations (files in a temporary directory), timed until the last frame
    is on disk; beside the run without writers."""
    import tempfile
    from biospring_amd import _abi
    from biospring_amd.network import SpringNetwork
    net = SpringNetwork.from_system(s)
    net.setup()
    net.run(10)
    with tempfile.TemporaryDirectory() as d:
        net.addTrajectory(_abi.TRAJ_XTC, os.path.join(d, "t.xtc"), every)
        net.addTrajectory(_abi.TRAJ_CSV, os.path.join(d, "t.csv"), every)
        hip_sync()
        t = time.perf_counter()
        net.run(steps)
        net.flushTrajectories()
        dt = time.perf_counter() - t
        net.closeTrajectories()
        xtc_bytes = os.path.getsize(os.path.join(d, "t.xtc"))
        with open(os.path.join(d, "t.csv")) as fh:
            csv_rows = sum(1 for _ in fh) - 1
    net.close()
    frames = len(range(10 + (-10) % every, 10 + steps, every))
    ms = dt / steps * 1e3
    log(f"trajectory: {frames} frames in {steps} steps, {ms:.3f} ms/step vs {ms_plain:.3f}")
    return {"writers": ["xtc", "csv"], "every": every, "steps": steps, "frames": frames,
            "csv_rows_written": csv_rows,
            "ms_per_step": ms, "ms_per_step_without": ms_plain, "overhead": ms / ms_plain - 1.0,
            "xtc_bytes_per_frame": xtc_bytes / max(frames, 1)}


def trajectory_check(seed, n=1000, marks=(1, 10, 100, 1000)):
    """Position deviation from the FP64 oracle's trajectory after 1, 10, 100
    and 1000 steps of a workload-1 network of n beads: the engine, the engine
    with compensated positions, and the reference's own float path (the
    float-faithful oracle) as the yardstick."""
    from biospring_amd import workloads
    from biospring_amd.network import SpringNetwork
    from oracle import pyoracle as po

    t = time.perf_counter()
    s = workloads.generate(1, seed=seed, n=n)
    runs = {}
    for name, comp in (("gpu", 0), ("gpu_compensated", 1)):
        p = s.params.copy()
        p.compensated_positions = comp
        net = SpringNetwork.from_system(s, params=p)
        done, xs = 0, []
        for m in marks:
            net.run(m - done)
            done = m
            xs.append(net.getState()[0].astype(np.float64))
        runs[name] = xs
        net.close()
    for name, mode in (("fp64", po.F64), ("reference_float_path", po.F32)):
        o = po.Oracle(s)
        done, xs = 0, []
        for m in marks:
            o.step(mode, m - done)
            done = m
            xs.append(o.state(mode)[0].astype(np.float64))
        runs[name] = xs
        o.close()
    out = {"n_beads": n, "steps": list(marks)}
    for name in ("gpu", "gpu_compensated", "reference_float_path"):
        d = [np.linalg.norm(a - b, axis=1) for a, b in zip(runs[name], runs["fp64"])]
        out[name] = {"rms_A": [float(np.sqrt(np.mean(v * v))) for v in d], "max_A": [float(v.max()) for v in d]}
    log(f"trajectory check {time.perf_counter() - t:.1f}s: max dev at {marks[-1]} steps gpu "
        f"{out['gpu']['max_A'][-1]:.2e}, compensated {out['gpu_compensated']['max_A'][-1]:.2e}, "
        f"reference float path {out['reference_float_path']['max_A'][-1]:.2e} A")
    return out


def adjoint_leg(seed=1, n=0, T=200, K=20, warmup=5, fp64=True, cpu=True, check_n=1000, fd=True):
    """The checkpointed adjoint of config 4 (forward, recomputation of every
    segment and the FP64 reverse sweep), timed against the plain forward of
    the same T steps.  Accuracy: the gradient of a check_n-bead network of the
    same kind against the FP64 oracle's, and, on 1,000 beads with cutoffs
    beyond the box, against central differences (h = 1e-3) of the FP64 loss.
    CPU baseline: the FP64 oracle's forward + adjoint on a 10k-bead network of
    the same kind."""
    from biospring_amd import _abi, workloads
    from biospring_amd.network import SpringNetwork

    t = time.perf_counter()
    s = workloads.generate(4, seed=seed, n=n)
    y = workloads.fitting_target(s, seed=seed)
    n = s.n
    log(f"adjoint: generated n={n} springs={s.m} in {time.perf_counter() - t:.1f}s digest={s.digest()}")
    net = SpringNetwork.from_system(s)
    net.setup()
    x0, v0 = net.getState()
    pairs_unique = net.countPairs(_abi.NS_STERIC)
    net.adjoint(min(max(warmup, 1), T), y, K)
    net.setState(x0, v0)
    net.profile(True)
    net.adjoint(T, y, K)
    prof = net.getProfile()
    net.profile(False)
    edges = net.counters()["edges"]

    def timed(fn):
        net.setState(x0, v0)
        hip_sync()
        t0 = time.perf_counter()
        r = fn()
        hip_sync()
        return time.perf_counter() - t0, r

    t_fwd, _ = timed(lambda: net.run(T))
    t_adj, (loss, dk, dl, dq) = timed(lambda: net.adjoint(T, y, K))
    # the same in two calls: the recorded forward, then the backward sweep
    # (segment recomputation, reverse steps, gradients to the host) from the
    # MSD loss's seed formed on the host, as a caller with its own loss does
    t_rec, _ = timed(lambda: net.adjointForward(T, K))
    xT, _ = net.getState()
    g_seed = 2.0 * (xT.astype(np.float64) - y.astype(np.float64)) / n
    hip_sync()
    t0 = time.perf_counter()
    dk2, dl2, dq2 = net.adjointBackward(g_seed)
    hip_sync()
    t_bwd = time.perf_counter() - t0
    two_call_rel = max(float(np.abs(u - v).max() / max(np.abs(v).max(), 1e-300))
                       for u, v in ((dk2, dk), (dl2, dl), (dq2, dq)))
    net.close()
    log(f"adjoint: forward {t_fwd * 1e3:.1f} ms; bs_adjoint {t_adj * 1e3:.1f} ms ({t_adj / t_fwd:.2f}x); "
        f"recorded forward {t_rec * 1e3:.1f} ms + backward {t_bwd * 1e3:.1f} ms ({t_bwd / t_fwd:.2f}x); "
        f"two-call gradients vs one call {two_call_rel:.1e}")

    ap_ms = prof["adj_pair"]["ms"] / max(prof["adj_pair"]["launches"], 1)
    as_ms = prof["adj_spring"]["ms"] / max(prof["adj_spring"]["launches"], 1)
    f_alg = ADJ_PAIR_FLOP * pairs_unique
    # springs with a dynamic end: one read-modify-write of their FP32 dk, dL0
    # sums per step (16 B); {j, k, L0} per edge, {j, L0} with one stiffness
    owned = int(s.m)
    k_uniform = bool(s.m > 0 and np.all(s.k.view(np.uint32) == s.k.view(np.uint32)[0]))
    b_alg = 4 * (n + 1) + (8 if k_uniform else 12) * edges + 16 * owned + 280 * n
    kernels = {
        "adj_pair": {"avg_ms": ap_ms, "bound": "valu", "achieved": f_alg / (ap_ms * 1e-3) / 1e12,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "alg_flops": f_alg,
                     "note": "FP32 terms with FP64 sums, priced against the FP32 peak (pairs at the initial state)"},
        "adj_spring": {"avg_ms": as_ms, "bound": "hbm", "achieved": b_alg / (as_ms * 1e-3) / 1e9,
                       "peak": HBM_PEAK_GBS, "unit": "GB/s", "alg_bytes": b_alg},
    }
    for k in kernels.values():
        k["frac"] = k["achieved"] / k["peak"]
    kernels["forward_pair_avg_ms"] = prof["pair"]["ms"] / max(prof["pair"]["launches"], 1)
    kernels["forward_spring_avg_ms"] = prof["spring"]["ms"] / max(prof["spring"]["launches"], 1)
    kernels["rebuild_avg_ms"] = prof["rebuild"]["ms"] / max(prof["rebuild"]["launches"], 1)
    dom = "adj_pair" if ap_ms >= as_ms else "adj_spring"
    d = kernels[dom]
    roofline = {"bound": d["bound"], "achieved": d["achieved"], "peak": d["peak"], "unit": d["unit"],
                "frac": d["frac"], "traffic": None, "kernel": dom}
    out = {
        "workload": WORKLOADS[4].format(n=n, T=T, ck=(f"checkpoints every {K} steps, segments recomputed"
                                                      if K > 0 else
                                                      "the engine's checkpoint choice: the whole trajectory "
                                                      "kept in HBM when it fits in half the free memory")),
        "n_beads": n, "steps": T, "checkpoint_every": K, "springs": s.m, "directed_edges": edges,
        "unique_pairs_initial": pairs_unique, "seed": seed, "digest": s.digest(),
        "value": n * T / t_adj, "unit": "particle-steps/s (forward + adjoint)",
        "forward_ms": t_fwd * 1e3,
        "adjoint_ms": t_adj * 1e3, "adjoint_call_over_forward": t_adj / t_fwd,
        "recorded_forward_ms": t_rec * 1e3, "backward_ms": t_bwd * 1e3,
        "adjoint_over_forward": t_bwd / t_fwd,
        "adjoint_over_forward_note": ("backward sweep (recomputation of the checkpointed segments, the reverse "
                                      "steps and the gradients' download) over the plain forward of the same "
                                      "steps; adjoint_call_over_forward is the whole bs_adjoint call, its own "
                                      "recorded forward included"),
        "two_call_gradients_rel_diff": two_call_rel,
        "forward_particle_steps_per_s": n * T / t_fwd,
        "loss": loss, "roofline": roofline, "kernels": kernels,
    }
    if fp64 or cpu:
        from oracle import pyoracle as po
    if fp64:
        t = time.perf_counter()
        sm = workloads.generate(4, seed=seed, n=check_n)
        ym = workloads.fitting_target(sm, seed=seed)
        g = SpringNetwork.from_system(sm).adjoint(T, ym, K)
        o = po.Oracle(sm)
        r = o.adjoint(T, ym.astype(np.float64))
        o.close()
        rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
        out["grad_err_vs_fp64"] = {"n_beads": sm.n, "steps": T, "checkpoint_every": K,
                                   "loss_rel": abs(g[0] - r[0]) / r[0],
                                   "dk": rel(g[1], r[1]), "dl0": rel(g[2], r[2]), "dq": rel(g[3], r[3]),
                                   "measure": "max |g_gpu - g_fp64| / max |g_fp64|"}
        if fd:
            # Central differences (h = 1e-3) of the FP64 loss for the parameter
            # of largest gradient of each kind.  With 12 A cutoffs, pairs cross
            # them during 200 steps and the loss jumps there (the forces are
            # cut, not smoothed), which a frozen-pair-set gradient does not
            # see; the differences are therefore taken on 1,000 beads with
            # cutoffs of 60 A, beyond the box, where no pair crosses and the
            # loss is smooth.
            sf = workloads.generate(4, seed=seed, n=1000)
            yf = workloads.fitting_target(sf, seed=seed)
            pf = sf.params.copy()
            pf.steric_cutoff = pf.electrostatic_cutoff = 60.0
            yf64 = yf.astype(np.float64)
            gf = SpringNetwork.from_system(sf, params=pf).adjoint(T, yf, K)
            fdr = {}
            for name, gi, base in (("k", 1, sf.k), ("l0", 2, sf.l0), ("q", 3, sf.charge)):
                i = int(np.argmax(np.abs(gf[gi])))
                lo = []
                for sgn in (1.0, -1.0):
                    a = base.astype(np.float64).copy()
                    a[i] += sgn * 1e-3
                    oo = po.Oracle(sf, params=pf)
                    oo.set_arrays(**{name: a})
                    lo.append(oo.loss(T, yf64))
                    oo.close()
                c = (lo[0] - lo[1]) / 2e-3
                fdr[name] = {"index": i, "central_difference": c, "gpu": float(gf[gi][i]),
                             "gpu_rel_err": abs(gf[gi][i] - c) / abs(c)}
            out["grad_err_vs_fp64"]["central_difference_h1e-3"] = {
                "n_beads": sf.n, "cutoffs_A": 60.0, "steps": T, "params": fdr}
        log(f"adjoint: fp64 gradient check {time.perf_counter() - t:.1f}s")
    if cpu:
        sc = workloads.generate(4, seed=seed, n=10000)
        yc = workloads.fitting_target(sc, seed=seed)
        o = po.Oracle(sc)
        steps = 2
        t = time.perf_counter()
        o.adjoint(steps, yc.astype(np.float64))
        dt = time.perf_counter() - t
        o.close()
        out["cpu_baseline"] = {"value": sc.n * steps / dt, "unit": "particle-steps/s (forward + adjoint)",
                               "cores": int(po.oracle_lib().orc_num_threads()), "kind": "port",
                               "sample": f"FP64 oracle forward + adjoint of {steps} steps of a {sc.n}-bead network "
                                         f"of the same kind, OpenMP (the reference has no adjoint)"}
    return out


def bench_adjoint(args):
    """Workload 4 alone: the adjoint leg as the bench line."""
    leg = adjoint_leg(seed=args.seed, n=args.n, T=args.steps, K=args.checkpoint, warmup=args.warmup,
                      fp64=not args.no_fp64, cpu=not args.no_cpu, check_n=2000)
    out = {
        "metric": METRIC, "value": leg["value"], "unit": "particle-steps/s", "n_gpus": 1, "steps": leg["steps"],
        "warmup": args.warmup, "ms_per_step": leg["adjoint_ms"] / leg["steps"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32 trajectory, fp64 adjoint", "data": "synthetic",
        "config": {k: leg[k] for k in ("workload", "n_beads", "springs", "directed_edges", "unique_pairs_initial",
                                       "checkpoint_every", "seed", "digest")},
        "roofline": leg["roofline"],
        "kernels": leg["kernels"],
        "adjoint": {k: leg[k] for k in ("forward_ms", "adjoint_ms", "adjoint_call_over_forward", "recorded_forward_ms",
                                        "backward_ms", "adjoint_over_forward", "loss")},
    }
    for k in ("grad_err_vs_fp64", "cpu_baseline"):
        if k in leg:
            out[k] = leg[k]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
