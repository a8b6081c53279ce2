"""ctypes wrapper over the CPU oracle (oracle/orc.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

SCHED_REFERENCE = 0
SCHED_COLOUR = 1

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Options(C.Structure):
    _fields_ = [
        ("schedule", C.c_int32),
        ("num_threads", C.c_int32),
        ("compute_cost", C.c_int32),
        ("abort_on_cost_increase", C.c_int32),
        ("max_iterations", C.c_int32),
        ("write_files", C.c_int32),
        ("no_nhicd", C.c_int32),
        ("order_seed", C.c_uint32),
        ("hicd_order", C.c_int32),
    ]


def build(quiet: bool = True) -> None:
    """Compile the oracle libraries (and the reference harness when the
    reference tree is present)."""
    kw = dict(stdout=subprocess.DEVNULL) if quiet else {}
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True, **kw)
    if os.path.isdir("/root/reference/src/MBIR_4D"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True, **kw)


_LIBS: dict = {}


def lib(precision: str = "f32") -> C.CDLL:
    if precision in _LIBS:
        return _LIBS[precision]
    path = os.path.join(BUILD, f"liborc_{precision}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    vp = C.c_void_p
    L.orc_pm_next.restype = C.c_uint32
    L.orc_pm_next.argtypes = [C.c_uint32]
    L.orc_random2_stream.argtypes = [C.c_uint32, C.c_int64, _f64p]
    L.orc_perm_key.restype = C.c_uint32
    L.orc_perm_key.argtypes = [C.c_uint32, C.c_int32, C.c_int32, C.c_int32]
    L.orc_perm.restype = C.c_uint32
    L.orc_perm.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_perm_inv.restype = C.c_uint32
    L.orc_perm_inv.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_colour_chunk.restype = C.c_int32
    L.orc_colour_chunk.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   _i32p, _i32p]
    L.orc_detector_response.argtypes = [_f32p, C.c_int32, C.c_double, C.c_double,
                                        C.c_double, C.c_double, _f32p]
    L.orc_amatrix_column.restype = C.c_int32
    L.orc_amatrix_column.argtypes = [_f32p, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                     C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_int32, C.c_double,
                                     _f64p, _i32p]
    for n in ("orc_qggmrf_value", "orc_qggmrf_derivative"):
        getattr(L, n).restype = C.c_double
        getattr(L, n).argtypes = [C.c_double] * 4
    L.orc_qggmrf_second_derivative.restype = C.c_double
    L.orc_qggmrf_second_derivative.argtypes = [C.c_double] * 2
    L.orc_problem_new.restype = vp
    L.orc_problem_new.argtypes = ([_f32p, _f32p, _f32p, _f32p, _f32p] + [C.c_int32] * 4
                                  + [C.c_float] * 7 + [C.c_int32] * 2 + [C.c_float] * 2
                                  + [C.c_int32] * 4 + [C.c_double, C.POINTER(Options)])
    L.orc_problem_free.argtypes = [vp]
    L.orc_problem_info.argtypes = [vp, _i32p, _f64p]
    for n in ("orc_obj", "orc_err", "orc_wgt", "orc_proj", "orc_off"):
        getattr(L, n).restype = vp
        getattr(L, n).argtypes = [vp]
    for n in ("orc_sel", "orc_mask"):
        getattr(L, n).restype = C.POINTER(C.c_uint8)
        getattr(L, n).argtypes = [vp]
    for n in ("orc_hr", "orc_mag"):
        getattr(L, n).restype = C.POINTER(C.c_float)
        getattr(L, n).argtypes = [vp]
    for n in ("orc_view_off", "orc_view_idx"):
        getattr(L, n).restype = C.POINTER(C.c_int32)
        getattr(L, n).argtypes = [vp]
    L.orc_var_est.restype = C.c_double
    L.orc_var_est.argtypes = [vp]
    L.orc_set_var_est.argtypes = [vp, C.c_double]
    L.orc_zero_count.restype = C.c_int64
    L.orc_zero_count.argtypes = [vp]
    L.orc_update_count.restype = C.c_int64
    L.orc_update_count.argtypes = [vp]
    L.orc_rng_state.restype = C.c_uint32
    L.orc_rng_state.argtypes = [vp]
    L.orc_set_rng_state.argtypes = [vp, C.c_uint32]
    L.orc_set_select_counts.argtypes = [vp, C.c_int32]
    L.orc_forward_project.argtypes = [vp, vp]
    L.orc_init_error_sino.argtypes = [vp]
    L.orc_cost.restype = C.c_double
    L.orc_cost.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_offset_update.argtypes = [vp, C.c_int32]
    L.orc_variance_update.argtypes = [vp]
    L.orc_surrogate_update.restype = C.c_double
    L.orc_surrogate_update.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_double, C.c_double]
    L.orc_voxel_update.restype = C.c_double
    L.orc_voxel_update.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.orc_update_voxels.restype = C.c_double
    L.orc_update_voxels.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                    _i32p, _i32p, C.c_int32, C.c_int32]
    L.orc_icd_iteration.restype = C.c_int32
    L.orc_icd_iteration.argtypes = [vp, C.c_int32, C.POINTER(C.c_double)]
    L.orc_run_phase.restype = C.c_double
    L.orc_run_phase.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.orc_vsc_select.argtypes = [vp]
    L.orc_nhicd_count.restype = C.c_int32
    L.orc_nhicd_count.argtypes = [vp, C.c_int32, C.c_int32, vp, vp]
    L.orc_init_from_coarse.argtypes = [vp, vp, C.c_int32, vp, C.c_int32]
    L.orc_stage_run.restype = C.c_int32
    L.orc_stage_run.argtypes = [vp, C.c_int32, vp, vp, C.POINTER(C.c_int32)]
    L.orc_reconstruct.restype = C.c_int
    L.orc_reconstruct.argtypes = ([C.POINTER(vp), _f32p, _f32p, _f32p, _f32p, _f32p]
                                  + [C.c_int32] * 4 + [C.c_float] * 7 + [C.c_int32] * 2
                                  + [C.c_float] * 2 + [C.POINTER(Options),
                                                       C.POINTER(C.c_double),
                                                       C.POINTER(C.c_int32)])
    L.orc_free.argtypes = [vp]
    _LIBS[precision] = L
    return L


INFO_INTS = ["N_p", "N_r", "N_t", "N_x", "N_y", "N_z", "N_time", "mult_xy", "mult_z",
             "num_z_blocks", "updateProjOffset", "updateVar", "initICD", "U"]
INFO_DBLS = ["delta_r", "delta_t", "R0", "RMax", "T0", "x0", "y0", "delta_xy", "delta_z",
             "OffsetR", "radius_obj", "sigS_q", "sigS_qp", "sigT_q", "sigT_qp", "c_s", "c_t",
             "tfilt", "alpha", "T", "delta", "var_est", "delta_recon"]


def options(schedule=SCHED_COLOUR, num_threads=8, compute_cost=0, abort=0, max_iterations=0,
            write_files=0, no_nhicd=0, order_seed=0, hicd_order=0) -> Options:
    return Options(schedule, num_threads, compute_cost, abort, max_iterations, write_files,
                   no_nhicd, order_seed, hicd_order)


class Problem:
    """One multi-resolution stage held by the oracle."""

    def __init__(self, scan: dict, mult_xy=1, mult_z=1, stage_idx=0, prev_mult_z=0,
                 var_est=1.0, opt: Options | None = None, precision="f32"):
        self.L = lib(precision)
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.opt = opt if opt is not None else options()
        s = scan
        self._keep = [np.ascontiguousarray(s[k], dtype=np.float32) for k in
                      ("projections", "weights", "proj_angles", "proj_times", "recon_times")]
        self.h = self.L.orc_problem_new(
            *self._keep, s["proj_rows"], s["proj_cols"], s["proj_num"], s["recon_num"],
            s["vox_wid"], s["rot_center"], s["sig_s"], s["sig_t"], s["c_s"], s["c_t"],
            s["convg_thresh"], s["remove_rings"], s["quad_convex"], s["huber_delta"],
            s["huber_T"], mult_xy, mult_z, stage_idx, prev_mult_z, var_est, C.byref(self.opt))
        if not self.h:
            raise ValueError("oracle rejected the scan (times not sorted?)")
        ints = np.zeros(16, np.int32)
        dbls = np.zeros(64, np.float64)
        self.L.orc_problem_info(self.h, ints, dbls)
        self.info = {k: int(v) for k, v in zip(INFO_INTS, ints)}
        self.info.update({k: float(v) for k, v in zip(INFO_DBLS, dbls)})
        self.sfilt = dbls[len(INFO_DBLS):len(INFO_DBLS) + 27].reshape(3, 3, 3).copy()

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.L.orc_problem_free(self.h)
                self.h = None
        except Exception:
            pass

    # -- array views ------------------------------------------------------
    def _view(self, ptr, shape, dtype):
        n = int(np.prod(shape))
        ct = {np.float32: C.c_float, np.float64: C.c_double, np.uint8: C.c_uint8,
              np.int32: C.c_int32}[dtype]
        arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,))
        return arr.reshape(shape)

    @property
    def sino_shape(self):
        i = self.info
        return (i["N_p"], i["N_r"], i["N_t"])

    @property
    def obj_shape(self):
        i = self.info
        return (i["N_time"], i["N_z"], i["N_y"], i["N_x"])

    def obj(self):
        return self._view(self.L.orc_obj(self.h), self.obj_shape, self.dtype)

    def err(self):
        return self._view(self.L.orc_err(self.h), self.sino_shape, self.dtype)

    def wgt(self):
        return self._view(self.L.orc_wgt(self.h), self.sino_shape, self.dtype)

    def proj(self):
        return self._view(self.L.orc_proj(self.h), self.sino_shape, self.dtype)

    def sel(self):
        return self._view(self.L.orc_sel(self.h), self.sino_shape, np.uint8)

    def off(self):
        i = self.info
        return self._view(self.L.orc_off(self.h), (i["N_r"], i["N_t"]), self.dtype)

    def hr(self):
        return self._view(self.L.orc_hr(self.h), (self.info["N_p"], 65), np.float32)

    def mask(self):
        i = self.info
        return self._view(self.L.orc_mask(self.h), (i["N_y"], i["N_x"]), np.uint8)

    def mag(self):
        i = self.info
        return self._view(self.L.orc_mag(self.h),
                          (i["N_time"], i["num_z_blocks"], i["N_y"], i["N_x"]), np.float32)

    def views(self):
        nt = self.info["N_time"]
        off = np.ctypeslib.as_array(self.L.orc_view_off(self.h), shape=(nt + 1,)).copy()
        idx = np.ctypeslib.as_array(self.L.orc_view_idx(self.h), shape=(max(int(off[-1]), 1),)).copy()
        return off, idx[:off[-1]]

    # -- operations ---------------------------------------------------------
    @property
    def var_est(self):
        return self.L.orc_var_est(self.h)

    def init_error_sino(self):
        self.L.orc_init_error_sino(self.h)

    def forward_project(self):
        out = np.zeros(self.sino_shape, self.dtype)
        self.L.orc_forward_project(self.h, out.ctypes.data)
        return out

    def cost(self):
        f, p = C.c_double(), C.c_double()
        tot = self.L.orc_cost(self.h, C.byref(f), C.byref(p))
        return tot, f.value, p.value

    def offset_update(self, typ):
        self.L.orc_offset_update(self.h, typ)

    def variance_update(self):
        self.L.orc_variance_update(self.h)

    def voxel_update(self, f, z, y, x):
        return self.L.orc_voxel_update(self.h, f, z, y, x)

    def surrogate_update(self, f, z, y, x, th1, th2):
        return self.L.orc_surrogate_update(self.h, f, z, y, x, th1, th2)

    def icd_iteration(self, it):
        r = C.c_double()
        flag = self.L.orc_icd_iteration(self.h, it, C.byref(r))
        return flag, r.value

    def run_phase(self, phase, it, chunk, hicd=1):
        self.L.orc_set_select_counts(self.h, self.info["U"])
        return self.L.orc_run_phase(self.h, phase, it, chunk, hicd)

    def vsc_select(self):
        self.L.orc_vsc_select(self.h)

    def nhicd_list(self, f, b):
        nxy = self.info["N_x"] * self.info["N_y"]
        xs = np.zeros(nxy, np.int32)
        ys = np.zeros(nxy, np.int32)
        n = self.L.orc_nhicd_count(self.h, f, b, xs.ctypes.data, ys.ctypes.data)
        return xs[:n].copy(), ys[:n].copy()

    def init_from_coarse(self, prev: "Problem"):
        self.L.orc_init_from_coarse(self.h, prev.L.orc_obj(prev.h), prev.info["N_z"],
                                    C.cast(prev.L.orc_mag(prev.h), C.c_void_p),
                                    prev.info["num_z_blocks"])

    def stage_run(self, max_iter=0):
        n = max_iter if max_iter > 0 else 1000
        costs = np.zeros(n + 1, np.float64)
        ratios = np.zeros(n, np.float64)
        it = C.c_int32()
        st = self.L.orc_stage_run(self.h, max_iter, costs.ctypes.data, ratios.ctypes.data,
                                  C.byref(it))
        k = it.value
        return st, costs[:k + 1].copy(), ratios[:k].copy()

    @property
    def update_count(self):
        return self.L.orc_update_count(self.h)


def reconstruct(scan: dict, opt: Options | None = None, precision="f32"):
    """Whole multi-resolution reconstruct(); returns (object, final_cost, iters)."""
    L = lib(precision)
    opt = opt if opt is not None else options()
    s = scan
    keep = [np.ascontiguousarray(s[k], dtype=np.float32) for k in
            ("projections", "weights", "proj_angles", "proj_times", "recon_times")]
    out = C.c_void_p()
    cost = C.c_double()
    iters = C.c_int32()
    rc = L.orc_reconstruct(C.byref(out), *keep, s["proj_rows"], s["proj_cols"], s["proj_num"],
                           s["recon_num"], s["vox_wid"], s["rot_center"], s["sig_s"], s["sig_t"],
                           s["c_s"], s["c_t"], s["convg_thresh"], s["remove_rings"],
                           s["quad_convex"], s["huber_delta"], s["huber_T"], C.byref(opt),
                           C.byref(cost), C.byref(iters))
    if rc != 0:
        raise RuntimeError("oracle reconstruct failed")
    dt = C.c_float if precision == "f32" else C.c_double
    shape = (s["recon_num"], s["proj_rows"], s["proj_cols"], s["proj_cols"])
    arr = np.ctypeslib.as_array(C.cast(out, C.POINTER(dt)), shape=(int(np.prod(shape)),))
    obj = arr.reshape(shape).copy()
    L.orc_free(out)
    return obj, cost.value, iters.value


def ref_available() -> bool:
    return os.path.exists(REF_HARNESS)
