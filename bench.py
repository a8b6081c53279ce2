"""Benchmark of the MBIR_4D ICD/NHICD hot path on one MI355X.

Workload (one GPU): the deployment scan, 4D interlaced-view MBIR of a
512 x 512 x 32-slice Cahn-Hilliard microstructure over 16 time frames, 2048
interlaced views (K = 8 sub-frames) over one half turn, 512 detector bins,
Poisson I0 = 1e4, zingers in 0.1% of the bins, per-detector ring offsets
N(0, 0.01); generalized-Huber data term, ring correction and variance
estimation on (datagen.deploy4d).

A step is one full ICD iteration of its finest multi-resolution stage (HICD +
NHICD colour phases, offset solves after every phase, VSC selection, variance
update, convergence check), inputs resident in HBM.  ``value`` is voxel
updates per second over the timed steps (wall clock, all kernels included);
the roofline block covers the voxel-update kernel alone (HIP events on the
library's stream).

``time_to_converged`` times starts of timbir_amd.converge (by default zero
under reconstruct()'s schedule, the filtered back-projection at the finest
stage, and the learned initialiser) to the workload's "1e-4 relative cost
change", and reports how far each is then from C*, the lowest MAP cost of a
long run cached in tests/golden/converged_deploy4d.npz; ``--until map``
runs on to |C - C*| <= 1e-4 |C*| with the object within 0.5% RMSE of x*.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "time-to-converged MAP (1e-4 rel) & voxel updates/s, learned vs plain init"
WORKLOAD = ("4D interlaced-view MBIR, deployment scan, finest stage: 512x512x32 slices x 16 "
            "frames, 2048 views K=8, 512 bins, I0=1e4, zingers 0.1%, rings N(0,0.01)")
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_k_icd_mwg_deploy4d_512x512x32x16.md")
GOLDEN = os.path.join(ROOT, "tests", "golden", "converged_deploy4d.npz")
WEIGHTS = os.path.join(ROOT, "timbir_amd", "weights", "unet_init.npz")
# the network trained on deployment-size scans of the same model (other seeds)
WEIGHTS_DEPLOY = os.path.join(ROOT, "timbir_amd", "weights", "unet_init_deploy4d.npz")


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


WORKLOADS = {
    "deploy4d": WORKLOAD,
    "interlaced4d": "4D interlaced-view MBIR, finest stage: 256x256x8 slices x 16 frames, 1024 "
                    "views K=8, 256 bins, I0=1e4, zingers 0.1%, rings N(0,0.01)",
    "precision3d": "3D MBIR, finest stage: 512x512x32 slices, 1440 views, 512 bins, I0=2e4, "
                   "rings N(0,0.005)",
}


def make_scan(small: bool, workload: str = "deploy4d"):
    from timbir_amd import datagen
    if small:
        return datagen.interlaced4d(N=64, nz=8, n_frames=4, n_views=256, K=8, ch_steps=16,
                                    frame_steps=2)
    return getattr(datagen, workload)()


def cpu_threads() -> int:
    """The CPUs this process may run on (the host cores the box gives it)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(scan, stage_kw, threads: int, budget_s: float = 15.0):
    """The oracle (the CPU restatement, colour schedule) on the same finest
    stage: HICD colour phases, one after the other, until budget_s of work;
    voxel updates per second."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import pyorc
    t0 = time.perf_counter()
    o = pyorc.Problem(scan, opt=pyorc.options(schedule=pyorc.SCHED_COLOUR, num_threads=threads),
                      **stage_kw)
    o.init_error_sino()
    setup = time.perf_counter() - t0
    o.L.orc_set_select_counts(o.h, o.info["U"])
    t0 = time.perf_counter()
    phases = 0
    for j in range(10):
        for col in (0, 1):
            o.run_phase(col, 1, j, 1)
            phases += 1
            if time.perf_counter() - t0 > budget_s:
                break
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    upd = o.update_count
    return dict(value=upd / el, unit="voxel updates/s", cores=threads, kind="port",
                sample=f"{phases} HICD colour phases (of 20 per ICD iteration) of the same "
                       f"finest stage, {upd} voxel updates in {el:.1f} s, {threads} OpenMP "
                       f"threads (set-up {setup:.1f} s not counted)")


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
    ap.add_argument("--workload", default="deploy4d", choices=sorted(WORKLOADS),
                    help="the scan (default: the deployment scan the metric is quoted on; the "
                         "others are for comparison, not bench lines)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline (default: the CPUs this process "
                         "may run on)")
    ap.add_argument("--no-converge", action="store_true",
                    help="skip the time-to-converged runs")
    ap.add_argument("--starts", default="zero,fbp_finest,learned",
                    help="starts timed to convergence (zero, fbp: back-projection at stage 0, "
                         "fbp_finest: back-projection at the finest stage, learned)")
    ap.add_argument("--cap", type=int, default=60,
                    help="finest-stage ICD iterations at most per start")
    ap.add_argument("--until", default="rel_change", choices=["rel_change", "map"],
                    help="rel_change: stop at the 1e-4 relative cost change; map: run on to "
                         "|C - C*| <= 1e-4 |C*| and 0.5%% RMSE to x*")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of known size (the scale for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--profile-steps", action="store_true",
                    help="only the timed steps (for counter and trace runs)")
    ap.add_argument("--stop-rule", action="store_true",
                    help="also time reconstruct()'s schedule on the workload from zero to the "
                         "reference's stop rule at every stage")
    ap.add_argument("--kernel", type=int, default=0,
                    help="voxel-update kernel (tb_options.kernel; 0: the fastest that applies)")
    ap.add_argument("--hicd-order", type=int, default=0, choices=[0, 1],
                    help="HICD chunk visit order (tb_options.hicd_order): 0 permutation order "
                         "(default), 1 raster order within each chunk's random subset")
    ap.add_argument("--row-lookahead", type=int, default=0,
                    help="voxels' (e, w) rows in flight in the deep-pipeline kernel (0: default)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1 "
                                   "(a z-slab split over GPUs is designed, not built)"}))
        sys.exit(2)

    import timbir_amd as tb
    from timbir_amd.batch import multires_schedule
    tb.set_options(verbose=0, write_files=0, kernel=args.kernel, row_lookahead=args.row_lookahead,
                   hicd_order=args.hicd_order)
    t0 = time.perf_counter()
    scan = make_scan(args.small, args.workload)
    log(f"scan generated in {time.perf_counter() - t0:.1f} s")
    sched = multires_schedule(scan["proj_rows"], scan["proj_cols"])
    stage_kw = dict(mult_xy=1, mult_z=1, stage_idx=len(sched) - 1, prev_mult_z=sched[-2][1])
    st = tb.Stage(scan, **stage_kw)
    info = st.info
    st.init_error_sino()
    it = 0
    for _ in range(args.warmup):
        it += 1
        st.icd_iteration(it)
    st.synchronize()
    st.reset_stats()
    log(f"{args.steps} timed ICD iterations ...")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        it += 1
        st.icd_iteration(it)
    st.synchronize()
    el = time.perf_counter() - t0
    s = st.stats()
    cost = st.cost()[0]
    st.close()
    value = s["updates"] / el
    vox_s = s["voxel_ms"] / 1e3
    achieved = s["alg_bytes"] / vox_s / 1e9 if vox_s > 0 else 0.0
    traffic = None if args.small else measured_traffic()
    fam = s["family_launches"]
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
            "workload": WORKLOADS[args.workload] if not args.small else "small check workload",
            "N_x": info["N_x"], "N_z": info["N_z"], "N_time": info["N_time"],
            "N_p": info["N_p"], "N_r": info["N_r"],
            "views_per_frame": info["N_p"] // info["N_time"],
        },
        "voxel_kernel": {"updates": s["updates"], "zero_skips": s["zero_skips"],
                         "launches": s["voxel_launches"], "time_ms": s["voxel_ms"],
                         "updates_per_s": s["updates"] / vox_s if vox_s > 0 else None,
                         "entries": s["entries"], "family_launches": fam},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ("k_icd_mwd (the deep-pipeline multi-wavefront voxel update, e in "
                                "global memory)" if fam[tb.K_MWD] == s["voxel_launches"] else
                                "k_icd_mwg (the multi-wavefront voxel update, e in global memory)"
                                if fam[tb.K_MWG] == s["voxel_launches"] else "mixed"),
                     "alg_bytes_per_launch": s["alg_bytes"] / max(s["voxel_launches"], 1),
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None},
        "cost_after": cost,
    }
    if args.profile_steps:
        print(json.dumps(line), flush=True)
        return
    if not args.no_converge:
        from timbir_amd import converge
        golden = (converge.load_golden(GOLDEN, scan)
                  if args.workload == "deploy4d" and not args.small else None)
        net = None
        starts = [x for x in args.starts.split(",") if x]
        if "learned" in starts:
            w = (WEIGHTS_DEPLOY if args.workload == "deploy4d" and os.path.exists(WEIGHTS_DEPLOY)
                 else WEIGHTS)
            if os.path.exists(w) and not args.small and args.workload != "precision3d":
                from timbir_amd.learned import UNet
                net = UNet(params=UNet.load_params(w))
            else:
                starts.remove("learned")
        until = ("rel_change",) if args.until == "rel_change" else ("rel_change", "map")
        ttc = {}
        for start in starts:
            log(f"time to converged from {start} ...")
            ttc[start] = converge.run(scan, start, tol=1e-4, cap=args.cap, golden=golden,
                                      net=net, until=until)
        ttc["golden"] = ({"cost": golden["cost"], "iterations": golden["iterations"],
                          "rel_stop": golden["rel_stop"], "file": os.path.relpath(GOLDEN, ROOT)}
                         if golden else None)
        if "learned" in ttc and "zero" in ttc:
            sp = {}
            for crit in ("rel_change", "map"):
                a, b = ttc["zero"].get(crit), ttc["learned"].get(crit)
                if a and b:
                    sp[crit] = a["seconds"] / b["seconds"]
            ttc["learned_speedup_vs_zero"] = sp
        line["time_to_converged"] = ttc
    if args.stop_rule:
        from timbir_amd import batch
        log("reference schedule to its stop rule ...")
        stages = []
        batch.reconstruct_stages(scan, timing=stages)
        line["time_to_stop_rule"] = {"seconds": sum(x["seconds"] for x in stages),
                                     "stages": stages}
    if not args.no_cpu and not args.small:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(scan, stage_kw, args.cpu_threads or cpu_threads())
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
