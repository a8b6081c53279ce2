"""Time to the converged MAP point from three starts, on the GPU.

Starts:
  zero     reconstruct()'s multi-resolution schedule (XT_Engine.c:111-182)
           from a zero object (the reference's start, XT_ICD_update.c:549-553),
           the coarse stages stopped by the reference's rule
           (XT_ICD_update.c:1096-1104), then the finest stage;
  fbp      the same with stage 0 started from the filtered back-projection
           (tb_init_fbp);
  fbp_finest  the finest stage alone, started from its filtered
           back-projection;
  learned  the finest stage alone, started from the learned initialiser's
           output (timbir_amd.learned.UNet) on the stage's filtered
           back-projection, written into the stage's object in HBM.

Criteria, checked after every ICD iteration of the finest stage:
  rel_change  the iteration changed the MAP cost (computeCost,
              XT_ICD_update.c:186-319, FP64 reductions) by less than tol of
              its value (the workloads' own "ICD to 1e-4 relative cost
              change");
  map         |C_k - C*| <= tol |C*| and the RMSE between the object and x*
              at most rmse_frac of x*'s range, with C* and x* from a long
              run cached with the scan's SHA-256 (x* sampled on a fixed
              grid);
  target      the cost has come down to a given level C_k <= target (to
              compare starts at one cost, e.g. where another start met
              rel_change).

The clock runs from the creation of the first stage to the iteration that
meets a criterion, with the time spent evaluating the criteria (the cost,
the object download) taken out: a production run would not compute them.
"""
from __future__ import annotations

import time

import numpy as np

from . import Stage
from .batch import multires_schedule


class Clock:
    """Wall time with pauses."""

    def __init__(self):
        self.t0 = time.perf_counter()
        self.paused = 0.0
        self._p = None

    def pause(self):
        self._p = time.perf_counter()

    def resume(self):
        self.paused += time.perf_counter() - self._p
        self._p = None

    def now(self) -> float:
        ref = self._p if self._p is not None else time.perf_counter()
        return ref - self.t0 - self.paused


def _finest_loop(st, clock, cap, tol, golden, rmse_frac, trace, until, target):
    """ICD iterations of the finest stage until the criteria in `until`
    ('rel_change', 'map', 'target') are met or cap iterations; fills trace."""
    res = {"rel_change": None, "map": None, "target": None}
    st.synchronize()
    clock.pause()
    last = st.cost()[0]
    trace.append((clock.now(), last, None))
    clock.resume()
    for it in range(1, cap + 1):
        st.icd_iteration(it)
        st.synchronize()
        clock.pause()
        c = st.cost()[0]
        t = clock.now()
        rel = abs(last - c) / abs(c)
        gap = None
        if golden is not None:
            gap = (c - golden["cost"]) / abs(golden["cost"])
        trace.append((t, c, gap))
        if res["rel_change"] is None and rel < tol and it > 1:
            res["rel_change"] = dict(seconds=t, iterations=it, cost=c, gap=gap)
        if golden is not None and res["map"] is None and abs(gap) <= tol:
            x = st.object()[golden["sample"]].astype(np.float64)
            rmse = float(np.sqrt(np.mean((x - golden["x"]) ** 2)) / golden["range"])
            if rmse <= rmse_frac:
                res["map"] = dict(seconds=t, iterations=it, cost=c, gap=gap, rmse=rmse)
        if target is not None and res["target"] is None and c <= target:
            res["target"] = dict(seconds=t, iterations=it, cost=c, gap=gap)
        last = c
        clock.resume()
        if all(res[k] is not None for k in until):
            break
    return res, it


def run(scan: dict, start: str = "zero", tol: float = 1e-4, cap: int = 100,
        golden: dict | None = None, rmse_frac: float = 5e-3, net=None,
        until=("rel_change",), target: float | None = None) -> dict:
    """Time to the criteria from one start (see the module docstring).
    golden: {'cost', 'x' (sampled x*), 'sample' (index of the sample),
    'range'} or None.  net: a learned.UNet for start 'learned'.  target: the
    cost level of the 'target' criterion."""
    if "target" in until and target is None:
        raise ValueError("criterion 'target' needs a target cost")
    sched = multires_schedule(scan["proj_rows"], scan["proj_cols"])
    clock = Clock()
    stages = []
    trace = []
    prev, var = None, 1.0
    nres = len(sched)
    if start in ("learned", "fbp_finest"):
        if start == "learned" and net is None:
            raise ValueError("start 'learned' needs the network")
        s = nres - 1
        st = Stage(scan, mult_xy=1, mult_z=1, stage_idx=s,
                   prev_mult_z=sched[s - 1][1] if s else 0)
        st.init_fbp()
        if start == "learned":
            net.predict_into_stage(st)
        st.init_error_sino()
        stages.append(dict(mult=[1, 1], seconds=clock.now(),
                           start="network output on the back-projection" if start == "learned"
                           else "filtered back-projection"))
    else:
        for s, (mxy, mz) in enumerate(sched):
            ts = clock.now()
            st = Stage(scan, mult_xy=mxy, mult_z=mz, stage_idx=s,
                       prev_mult_z=sched[s - 1][1] if s else 0, var_est=var)
            if prev is not None:
                st.init_from_stage(prev)
                st.set_offset(prev.offset())
                prev.close()
                prev = None
            elif start == "fbp":
                st.init_fbp()
            elif start != "zero":
                raise ValueError(f"unknown start {start!r}")
            if s < nres - 1:
                n = len(st.stage_run(0, with_cost=False)) - 1
                st.synchronize()
                stages.append(dict(mult=[mxy, mz], iterations=n, seconds=clock.now() - ts))
                prev, var = st, st.var_est
            else:
                st.init_error_sino()
    res, its = _finest_loop(st, clock, cap, tol, golden, rmse_frac, trace, until, target)
    stats = st.stats()
    st.close()
    marks = [i for i in (0, 1, 2, 5, 10, 20, 50, 100, 200, 400) if i < len(trace)]
    return dict(start=start, tol=tol, cap=cap, finest_iterations_run=its,
                seconds_run=trace[-1][0], rel_change=res["rel_change"], map=res["map"],
                target=res["target"], final_cost=trace[-1][1],
                stages=stages, finest_cost_at={str(i): trace[i][1] for i in marks},
                finest_gap_at={str(i): trace[i][2] for i in marks},
                voxel_updates=stats["updates"], voxel_kernel_ms=stats["voxel_ms"])


def load_golden(path: str, scan: dict) -> dict | None:
    """C* and the sampled x* of a long run (tests/golden/make_converged_*.py)
    if they belong to this scan."""
    import os

    from . import datagen
    if not os.path.exists(path):
        return None
    z = np.load(path)
    if str(z["scan_sha256"]) != datagen.sha256(scan["projections"]):
        return None
    sample = (slice(None), slice(None, None, 4), slice(None, None, 8), slice(None, None, 8))
    x = z["x_sample"].astype(np.float64)
    return dict(cost=float(z["cost"]), x=x, sample=sample, range=float(z["x_range"]),
                iterations=int(z["iterations"]), rel_stop=float(z["rel_stop"]))
