"""Benchmark of the MBIR_4D ICD/NHICD hot path on one MI355X.

Workload (one GPU): 4D interlaced-view MBIR, finest multi-resolution stage of
a 256 x 256 x 8-slice Cahn-Hilliard microstructure over 16 time frames, 1024
interlaced views (K = 8 sub-frames) over one half turn, 256 detector bins,
Poisson I0 = 1e4, zingers in 0.1% of the bins, per-detector ring offsets
N(0, 0.01); generalized-Huber data term, ring correction and variance
estimation on.  A step is one full ICD iteration of that stage (HICD + NHICD
colour phases, offset solves after every phase, VSC selection, variance
update, convergence check), inputs resident in HBM.

Prints one JSON line.  ``value`` is voxel updates per second over the timed
steps (wall clock, all kernels included); the roofline block covers the
voxel-update kernel alone (HIP events on the library's stream).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "time-to-converged MAP (1e-4 rel) & voxel updates/s, learned vs plain init"


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles",
                               "pmc_traffic_issue_k_icd_mw_interlaced4d_256x256x8x16.md")


def measured_traffic():
    """HBM bytes per voxel-kernel launch of the full workload, from the
    committed FETCH_SIZE / WRITE_SIZE counter summary (a counter run cannot
    happen inside the timed run itself)."""
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_scan(small: bool):
    from timbir_amd import datagen
    if small:
        return datagen.interlaced4d(N=64, nz=4, n_frames=4, n_views=256, K=8, ch_steps=16,
                                    frame_steps=2)
    return datagen.interlaced4d()


def cpu_baseline(scan, stage_kw, threads: int, budget_s: float = 20.0):
    """The oracle (CPU restatement, reference schedule) on the same stage:
    voxel updates per second over whole ICD iterations, bounded in time."""
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    from oracle import pyorc
    o = pyorc.Problem(scan, opt=pyorc.options(schedule=pyorc.SCHED_REFERENCE, num_threads=threads),
                      **stage_kw)
    o.init_error_sino()
    t0 = time.perf_counter()
    it = 0
    while True:
        it += 1
        o.icd_iteration(it)
        el = time.perf_counter() - t0
        if el > budget_s or it >= 3:
            break
    upd = o.update_count
    return dict(value=upd / el, unit="voxel updates/s", cores=threads, kind="port",
                sample=f"{it} full ICD iteration(s) of the same stage, reference schedule, "
                       f"{upd} voxel updates in {el:.1f} s")


def time_to_converged(scan, finest_cap: int = 0, tol: float = 1e-4, init: str = "zero"):
    """Plain initialisation (init "zero", the reference's, or "fbp": stage 0
    from the filtered back-projection) under reconstruct()'s multires
    schedule on the GPU, every stage stopped by the reference's own rule (convergence
    ratio below convg_thresh after iteration 1, XT_ICD_update.c:1066-1104 and
    :1188): `seconds` is the wall time from the start of the first stage to
    that stop on the finest stage, with the MAP cost evaluated every finest
    iteration.  With finest_cap > 0 the finest stage then runs on to
    finest_cap iterations so that the cost trajectory shows how far that stop
    is from the MAP point: `cost_gap_at_stop` = (C_stop - C_cap) / |C_cap|."""
    import timbir_amd as tb
    from timbir_amd.batch import multires_schedule
    sched = multires_schedule(scan["proj_rows"], scan["proj_cols"])
    t0 = time.perf_counter()
    prev = None
    var_est = 1.0
    trace = []
    stages = []
    stop_t = stop_it = None
    for s, (mxy, mz) in enumerate(sched):
        ts = time.perf_counter()
        st = tb.Stage(scan, mult_xy=mxy, mult_z=mz, stage_idx=s,
                      prev_mult_z=sched[s - 1][1] if s else 0, var_est=var_est)
        if prev is not None:
            # the handoff of reconstruct(): upsampled from the previous stage in HBM
            st.init_from_stage(prev)
            st.set_offset(prev.offset())
            prev.close()
        elif init == "fbp":
            st.init_fbp()
        st.init_error_sino()
        finest = s == len(sched) - 1
        rule_it = None
        cap = max(finest_cap, 1000 if finest_cap <= 0 else finest_cap) if finest else 1000
        for it in range(1, cap + 1):
            conv, _ = st.icd_iteration(it)
            if finest:
                c = st.cost()[0]
                st.synchronize()
                trace.append((time.perf_counter() - t0, c))
            if conv and it > 1 and rule_it is None:
                rule_it = it
                if finest:
                    stop_t, stop_it = trace[-1][0], it
                    if finest_cap <= 0:
                        break
                else:
                    break
        st.synchronize()
        ks = st.stats()
        stages.append(dict(mult=[mxy, mz], rule_stop_iteration=rule_it,
                           iterations_run=it, seconds=time.perf_counter() - ts,
                           var_est=st.var_est, voxel_kernel_ms=ks["voxel_ms"],
                           other_kernels_ms=ks["other_ms"], updates=ks["updates"],
                           voxel_launches=ks["voxel_launches"],
                           lds_launches=ks["lds_launches"],
                           global_launches=ks["global_launches"]))
        if not finest:
            prev = st
            var_est = st.var_est
        else:
            st.close()
    marks = [i for i in (1, 2, 5, 10, 20, 50, 100, 200, 400, 800) if i <= len(trace)]
    out = dict(seconds=stop_t, finest_rule_stop_iteration=stop_it, stages=stages,
               finest_iterations_run=len(trace), cost_at_stop=trace[stop_it - 1][1] if stop_it
               else None, finest_cost_at={str(i): trace[i - 1][1] for i in marks},
               schedule=[list(x) for x in sched],
               init="zero (plain)" if init == "zero" else "filtered back-projection at stage 0",
               rule="convergence ratio < convg_thresh (reference rule)")
    if finest_cap > 0:
        ccap = trace[-1][1]
        out["c_cap"] = ccap
        out["cost_gap_at_stop"] = ((trace[stop_it - 1][1] - ccap) / abs(ccap)) if stop_it else None
        k = next((i for i, (t, c) in enumerate(trace) if abs(c - ccap) <= tol * abs(ccap)), None)
        out["seconds_to_within_tol_of_c_cap"] = trace[k][0] if k is not None else None
        out["tol"] = tol
    return out


def copy_calibration(nbytes: int = 1 << 30, reps: int = 3):
    """A plain device-to-device copy of nbytes, reps times: a counter run of
    this command tells how FETCH_SIZE / WRITE_SIZE read for known traffic."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--small", action="store_true", help="tiny workload for a quick check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the time-to-converged run (plain init, reference stop rule)")
    ap.add_argument("--converge", action="store_true",
                    help="run the finest stage on past the stop rule to --finest-cap "
                         "iterations, to show the remaining cost gap")
    ap.add_argument("--finest-cap", type=int, default=400,
                    help="finest-stage iterations run by --converge")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of known size (the scale for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--deploy4d", action="store_true",
                    help="also time the multires schedule to the stop rule on the deployment-"
                         "size scan (512x512x32 x 16 frames, 2048 views; about a minute to "
                         "generate)")
    ap.add_argument("--precision3d", action="store_true",
                    help="also time the multires schedule to the stop rule on the single-"
                         "frame precision-study scan (512x512x32, 1440 views)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1 "
                                   "(a z-slab split over GPUs is designed, not built)"}))
        sys.exit(2)

    import timbir_amd as tb
    tb.set_options(verbose=0, write_files=0)
    t0 = time.perf_counter()
    scan = make_scan(args.small)
    log(f"scan generated in {time.perf_counter() - t0:.1f} s")
    # the finest stage of reconstruct()'s schedule for this scan: (1, 1) after
    # (4, 4) and (2, 2)
    stage_kw = dict(mult_xy=1, mult_z=1, stage_idx=2, prev_mult_z=2)
    st = tb.Stage(scan, **stage_kw)
    info = st.info
    st.init_error_sino()
    it = 0
    for _ in range(args.warmup):
        it += 1
        st.icd_iteration(it)
    st.synchronize()
    st.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it += 1
        st.icd_iteration(it)
    st.synchronize()
    el = time.perf_counter() - t0
    s = st.stats()
    cost = st.cost()[0]
    value = s["updates"] / el
    vox_s = s["voxel_ms"] / 1e3
    achieved = s["alg_bytes"] / vox_s / 1e9 if vox_s > 0 else 0.0
    traffic = None if args.small else measured_traffic()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Cahn-Hilliard phantom, Poisson counts, zingers, rings)",
        "config": {
            "workload": "4D interlaced-view MBIR, finest stage: 256x256x8 slices x 16 frames, "
                        "1024 views K=8, 256 bins, I0=1e4, zingers 0.1%, rings N(0,0.01)"
                        if not args.small else "small check workload",
            "N_x": info["N_x"], "N_z": info["N_z"], "N_time": info["N_time"],
            "N_p": info["N_p"], "N_r": info["N_r"], "views_per_frame":
                info["N_p"] // info["N_time"],
        },
        "voxel_kernel": {"updates": s["updates"], "zero_skips": s["zero_skips"],
                         "launches": s["voxel_launches"], "time_ms": s["voxel_ms"],
                         "updates_per_s": s["updates"] / vox_s if vox_s > 0 else None,
                         "entries": s["entries"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes_per_launch": s["alg_bytes"] / max(s["voxel_launches"], 1),
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None},
        "cost_after": cost,
    }
    if not args.no_converge and not args.small:
        log("time to converged ...")
        line["time_to_converged"] = time_to_converged(
            scan, finest_cap=args.finest_cap if args.converge else 0)
        log("time to converged from the filtered back-projection ...")
        line["time_to_converged_fbp_init"] = time_to_converged(
            scan, finest_cap=args.finest_cap if args.converge else 0, init="fbp")
    if args.deploy4d:
        from timbir_amd import datagen
        log("deployment-size scan ...")
        big = datagen.deploy4d()
        log("time to converged, deployment size ...")
        line["time_to_converged_deploy4d"] = time_to_converged(big)
        del big
    if args.precision3d:
        from timbir_amd import datagen
        log("time to converged, precision-study scan ...")
        line["time_to_converged_precision3d"] = time_to_converged(datagen.precision3d())
    if not args.no_cpu:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(scan, stage_kw, args.cpu_threads)
    st.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
