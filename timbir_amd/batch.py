"""Many independent reconstructions at once.

reconstruct() (src/include/mbir4d.h:34) takes one scan.  A small scan keeps
only a few CUs busy: its voxel kernels run one workgroup per cell (frame,
slice) of one colour, and a single-frame scan with 4 slices has 2 cells per
colour.  reconstruct_many() runs the same stage pipeline for many scans from
a pool of host threads; every stage owns a HIP stream, so the kernels of
different scans overlap on the GPU.  Each result equals, bit for bit, what
reconstruct() returns for that scan alone (no files are written).

How many streams the GPU runs side by side is capped by the hardware queues
HIP gives the process: GPU_MAX_HW_QUEUES (4 unless set before the first HIP
call) bounds the useful number of workers.

reconstruct_batched() instead advances all scans in lockstep through one
tb_batch (timbir_amd.Batch): the scans whose next stage has the coarsest
(mult_xy, mult_z) run that stage together, one ICD iteration of every
member per call, each voxel-update phase one launch per kernel
configuration over all their cells (a single-frame scan with 4 slices has 2
cells per colour; 64 of them fill 128 CUs in one launch), iterations 2 on
replayed from a HIP graph.  Every member keeps its own stop rule and drops
out when it fires; results equal reconstruct_stages(), bit for bit.
"""
from __future__ import annotations

from concurrent.futures import ThreadPoolExecutor, as_completed

import numpy as np

def multires_schedule(proj_rows: int, proj_cols: int) -> list:
    """(mult_xy, mult_z) per stage, coarse to fine, as reconstruct() picks
    them (XT_Engine.c:111-141 with one node): the library's own
    tb_multires_schedule, the code reconstruct() runs."""
    import ctypes as C
    from . import lib
    L = lib()
    mxy = (C.c_int32 * 8)()
    mz = (C.c_int32 * 8)()
    n = L.tb_multires_schedule(int(proj_rows), int(proj_cols), mxy, mz)
    if n < 2:
        raise ValueError(f"no multires schedule for {proj_rows} rows x {proj_cols} columns")
    return [(int(mxy[i]), int(mz[i])) for i in range(n)]


def reconstruct_stages(scan: dict, timing: list | None = None) -> np.ndarray:
    """reconstruct()'s schedule through the stage API, without files: each
    stage from the previous one in HBM (XT_Engine.c:156-182), run to the
    reference's stop rule.  Returns the finest object
    [recon_num][rows][cols][cols].  timing, when given, receives one dict
    per stage: mult, ICD iterations, seconds (set-up and hand-over included,
    the device synchronized at the end of the stage) and the kernel
    families' launch counts."""
    import time
    import timbir_amd as tb
    sched = multires_schedule(scan["proj_rows"], scan["proj_cols"])
    prev = None
    var_est = 1.0
    st = None
    try:
        for s, (mxy, mz) in enumerate(sched):
            t0 = time.perf_counter()
            st = tb.Stage(scan, mult_xy=mxy, mult_z=mz, stage_idx=s,
                          prev_mult_z=sched[s - 1][1] if s else 0, var_est=var_est)
            if prev is not None:
                st.init_from_stage(prev)
                st.set_offset(prev.offset())
                prev.close()
                prev = None
            costs = st.stage_run(0, with_cost=False)
            if timing is not None:
                st.synchronize()
                timing.append({"mult": [mxy, mz], "iterations": len(costs) - 1,
                               "seconds": time.perf_counter() - t0,
                               "family_launches": st.stats()["family_launches"]})
            if s < len(sched) - 1:
                prev, var_est, st = st, st.var_est, None
        return st.object()
    finally:
        for h in (prev, st):
            if h is not None:
                h.close()


def reconstruct_many(scans, workers: int = 4, progress=None) -> list:
    """reconstruct_stages() for every scan, `workers` scans at a time, each
    in its own host thread with its own streams.  Results in input order;
    progress(k), when given, is called as scan k finishes."""
    if workers <= 1:
        out = []
        for k, s in enumerate(scans):
            out.append(reconstruct_stages(s))
            if progress:
                progress(k)
        return out
    out = [None] * len(scans)
    with ThreadPoolExecutor(max_workers=workers) as pool:
        futs = {pool.submit(reconstruct_stages, s): k for k, s in enumerate(scans)}
        for f in as_completed(futs):
            k = futs[f]
            out[k] = f.result()
            if progress:
                progress(k)
    return out


def reconstruct_batched(scans, graph: bool = True, progress=None, stats: dict | None = None) -> list:
    """reconstruct_stages() for every scan, the scans advanced in lockstep
    (see the module docstring).  Results in input order; progress(k), when
    given, is called as scan k finishes; stats, when given, receives the
    batch's counters (timbir_amd.Batch.stats) and the stage groups run."""
    import timbir_amd as tb
    n = len(scans)
    scheds = [multires_schedule(s["proj_rows"], s["proj_cols"]) for s in scans]
    max_iter = tb.get_options().max_iterations
    nxt = [0] * n
    prev = [None] * n
    var = [1.0] * n
    out = [None] * n
    groups = []
    b = tb.Batch(graph=graph)
    live = {}
    try:
        while True:
            pend = [k for k in range(n) if nxt[k] < len(scheds[k])]
            if not pend:
                break
            # the coarsest stage any scan still has to run, for every scan
            # whose next stage it is
            key = max(scheds[k][nxt[k]] for k in pend)
            members = [k for k in pend if scheds[k][nxt[k]] == key]
            for k in members:
                s = nxt[k]
                mxy, mz = scheds[k][s]
                st = tb.Stage(scans[k], mult_xy=mxy, mult_z=mz, stage_idx=s,
                              prev_mult_z=scheds[k][s - 1][1] if s else 0, var_est=var[k])
                live[k] = st
                if prev[k] is not None:
                    st.init_from_stage(prev[k])
                    st.set_offset(prev[k].offset())
                    prev[k].close()
                    prev[k] = None
                st.init_error_sino()
            active = list(members)
            it = 1
            iters = {}
            while active and it <= max_iter:
                conv, _ = b.iteration([live[k] for k in active], it)
                still = []
                for k, c in zip(active, conv):
                    if c and it > 1:
                        iters[k] = it
                    else:
                        still.append(k)
                active = still
                it += 1
            for k in active:
                iters[k] = max_iter
            groups.append({"mult": list(key), "members": len(members),
                           "iterations": max(iters.values())})
            for k in members:
                st = live.pop(k)
                if nxt[k] < len(scheds[k]) - 1:
                    prev[k], var[k] = st, st.var_est
                else:
                    out[k] = st.object()
                    st.close()
                    if progress:
                        progress(k)
                nxt[k] += 1
        if stats is not None:
            stats.update(b.stats())
            stats["stages"] = groups
        return out
    finally:
        for st in list(live.values()) + [p for p in prev if p is not None]:
            st.close()
        b.close()
