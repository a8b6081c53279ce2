"""timbir_amd: MI355X-native MBIR_4D (TIMBIR) ICD/NHICD hot path.

The product is the C-ABI library ``libtimbir_hip.so`` (declared in
``include/mbir4d.h`` and ``include/timbir_dev.h``); this module only binds it
with ctypes so Python callers and the tests can drive it.  There is no CPU
fallback: if the library is missing or cannot be loaded, every entry point
raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtimbir_hip.so")

EXPORTS = [
    "reconstruct", "tb_get_options", "tb_set_options", "tb_last_error", "tb_stage_create",
    "tb_stage_destroy", "tb_stage_info", "tb_set_object", "tb_get_object", "tb_set_offset",
    "tb_get_offset", "tb_set_magmap", "tb_get_magmap", "tb_init_from_coarse", "tb_init_from_stage",
    "tb_get_error_sino", "tb_get_weights", "tb_get_detector_response", "tb_get_var_est",
    "tb_init_error_sino", "tb_init_error_sino_static", "tb_forward_project", "tb_init_fbp",
    "tb_cost", "tb_offset_update",
    "tb_variance_update", "tb_voxel_update", "tb_update_colour", "tb_vsc_select",
    "tb_nhicd_list", "tb_icd_iteration", "tb_stage_run", "tb_get_stats", "tb_reset_stats",
    "tb_synchronize", "tb_stage_object_device", "tb_stage_abort", "tb_multires_schedule",
    # many problems in lockstep
    "tb_batch_create", "tb_batch_destroy", "tb_batch_iteration", "tb_batch_get_stats",
    # autotuner
    "tb_autotune", "tb_autotune_entries", "tb_autotune_put", "tb_autotune_clear",
    # learned initialiser kernels (tb_unet.hip)
    "tb_nn_last_error", "tb_dev_malloc", "tb_dev_free", "tb_dev_copy", "tb_dev_zero",
    "tb_dev_synchronize", "tb_nn_conv3d", "tb_nn_conv3d_flip", "tb_nn_conv3d_wgrad",
    "tb_nn_channel_sum", "tb_nn_relu_bwd", "tb_nn_avgpool2", "tb_nn_up2_concat",
    "tb_nn_up2_concat_bwd", "tb_nn_add", "tb_nn_make_input", "tb_nn_frame_mean", "tb_nn_head", "tb_nn_head_bwd",
    "tb_nn_mse", "tb_nn_sum_parts", "tb_nn_adam", "tb_nn_to_bf16",
]


class TbOptions(C.Structure):
    _fields_ = [("compute_cost", C.c_int32), ("abort_on_cost_increase", C.c_int32),
                ("max_iterations", C.c_int32), ("no_nhicd", C.c_int32),
                ("write_files", C.c_int32), ("order_seed", C.c_uint32), ("verbose", C.c_int32),
                ("kernel", C.c_int32), ("views_per_lane", C.c_int32),
                ("init_fbp", C.c_int32), ("no_column_table", C.c_int32),
                ("theta_fp", C.c_int32), ("row_lookahead", C.c_int32), ("autotune", C.c_int32),
                ("split_segments", C.c_int32), ("split_spin_limit", C.c_int32),
                ("hicd_order", C.c_int32)]


class TbStats(C.Structure):
    _fields_ = [("voxel_ms", C.c_double), ("other_ms", C.c_double),
                ("voxel_launches", C.c_int64), ("updates", C.c_int64),
                ("zero_skips", C.c_int64), ("entries", C.c_int64), ("alg_bytes", C.c_double),
                ("lds_launches", C.c_int64),
                ("global_launches", C.c_int64),
                ("family_launches", C.c_int64 * 12),
                ("column_builds", C.c_int64), ("column_bins", C.c_int32),
                ("reserved", C.c_int32)]


class TbBatchStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("launches", C.c_int64), ("captures", C.c_int64),
                ("graph_replays", C.c_int64), ("groups", C.c_int64), ("solo", C.c_int64),
                ("ms", C.c_double)]


BATCH_GRAPH = 1  # TB_BATCH_GRAPH


class TbTuneEntry(C.Structure):
    _fields_ = [("shape", C.c_int32 * 8), ("kernel", C.c_int32), ("views_per_lane", C.c_int32),
                ("ms", C.c_float), ("reserved", C.c_int32)]

# voxel-update kernel families (TB_K_* in timbir_dev.h)
K_ONE_VOXEL, K_LDS, K_MW, K_MWG, K_WIDE, K_WPASS, K_WLOOP, K_MWD, K_BAT, K_SPLIT = range(10)


class TimbirError(RuntimeError):
    pass


_LIB = None
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def lib() -> C.CDLL:
    """Load libtimbir_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise TimbirError(f"{LIB_PATH} is missing: build it with `python -c 'import "
                          f"__graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i32 = C.c_int32
    L.tb_last_error.restype = C.c_char_p
    L.tb_get_options.argtypes = [C.POINTER(TbOptions)]
    L.tb_set_options.argtypes = [C.POINTER(TbOptions)]
    L.tb_stage_create.restype = C.c_int
    L.tb_stage_create.argtypes = ([C.POINTER(vp)] + [_f32p] * 5 + [i32] * 4 + [C.c_float] * 7
                                  + [i32] * 2 + [C.c_float] * 2 + [i32] * 4 + [C.c_double])
    L.tb_stage_destroy.argtypes = [vp]
    L.tb_stage_info.argtypes = [vp, np.ctypeslib.ndpointer(np.int32), np.ctypeslib.ndpointer(np.float64)]
    for n in ("tb_set_object", "tb_get_object", "tb_set_offset", "tb_get_offset", "tb_set_magmap",
              "tb_get_magmap", "tb_get_error_sino", "tb_get_weights", "tb_get_detector_response",
              "tb_forward_project"):
        getattr(L, n).argtypes = [vp, _f32p]
        getattr(L, n).restype = C.c_int
    L.tb_init_from_coarse.argtypes = [vp, _f32p, i32, _f32p]
    L.tb_init_from_stage.argtypes = [vp, vp]
    L.tb_get_var_est.restype = C.c_double
    L.tb_get_var_est.argtypes = [vp]
    L.tb_init_error_sino.argtypes = [vp]
    L.tb_cost.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.tb_offset_update.argtypes = [vp, i32]
    L.tb_variance_update.argtypes = [vp]
    L.tb_voxel_update.argtypes = [vp, i32, i32, i32, i32, C.POINTER(C.c_float)]
    L.tb_update_colour.argtypes = [vp, i32, i32, i32, i32]
    L.tb_vsc_select.argtypes = [vp]
    L.tb_nhicd_list.argtypes = [vp, i32, i32, vp, vp]
    L.tb_icd_iteration.argtypes = [vp, i32, C.POINTER(i32), C.POINTER(C.c_double)]
    L.tb_stage_run.argtypes = [vp, i32, vp, C.POINTER(i32)]
    L.tb_get_stats.argtypes = [vp, C.POINTER(TbStats)]
    L.tb_reset_stats.argtypes = [vp]
    L.tb_synchronize.argtypes = [vp]
    L.tb_stage_abort.argtypes = [vp]
    L.tb_multires_schedule.restype = C.c_int32
    L.tb_multires_schedule.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32)]
    L.tb_stage_object_device.argtypes = [vp, C.POINTER(vp), C.POINTER(C.c_uint64)]
    L.tb_batch_create.argtypes = [C.POINTER(vp), i32]
    L.tb_batch_destroy.argtypes = [vp]
    L.tb_batch_destroy.restype = None
    L.tb_batch_iteration.argtypes = [vp, C.POINTER(vp), i32, i32, C.POINTER(i32),
                                     C.POINTER(C.c_double)]
    L.tb_batch_get_stats.argtypes = [vp, C.POINTER(TbBatchStats)]
    L.tb_autotune.argtypes = [vp, C.POINTER(TbTuneEntry)]
    L.tb_autotune_entries.argtypes = [C.POINTER(TbTuneEntry), i32]
    L.tb_autotune_put.argtypes = [C.POINTER(TbTuneEntry)]
    L.tb_autotune_clear.argtypes = []
    L.tb_autotune_clear.restype = None
    L.reconstruct.restype = C.c_int
    L.reconstruct.argtypes = ([C.POINTER(C.POINTER(C.c_float))] + [vp] * 5 + [i32] * 4
                              + [C.c_float] * 7 + [i32] * 2 + [C.c_float] * 2
                              + [C.c_uint8, vp])
    _LIB = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise TimbirError(lib().tb_last_error().decode())


def get_options() -> TbOptions:
    o = TbOptions()
    lib().tb_get_options(C.byref(o))
    return o


def set_options(**kw) -> TbOptions:
    o = get_options()
    for k, v in kw.items():
        setattr(o, k, v)
    lib().tb_set_options(C.byref(o))
    return o


INFO_INTS = ["N_p", "N_r", "N_t", "N_x", "N_y", "N_z", "N_time", "mult_xy", "mult_z",
             "num_z_blocks", "updateProjOffset", "updateVar", "initICD", "U"]
INFO_DBLS = ["delta_r", "delta_t", "R0", "RMax", "T0", "x0", "y0", "delta_xy", "delta_z",
             "OffsetR", "radius_obj", "sigS_q", "sigS_qp", "sigT_q", "sigT_qp", "c_s", "c_t",
             "tfilt", "alpha", "T", "delta", "var_est", "delta_recon"]


class Stage:
    """One multi-resolution stage held in GPU memory (tb_stage)."""

    def __init__(self, scan: dict, mult_xy=1, mult_z=1, stage_idx=0, prev_mult_z=0,
                 var_est=1.0):
        L = lib()
        s = scan
        keep = [np.ascontiguousarray(s[k], dtype=np.float32) for k in
                ("projections", "weights", "proj_angles", "proj_times", "recon_times")]
        h = C.c_void_p()
        _check(L.tb_stage_create(C.byref(h), *keep, s["proj_rows"], s["proj_cols"], s["proj_num"],
                                 s["recon_num"], s["vox_wid"], s["rot_center"], s["sig_s"],
                                 s["sig_t"], s["c_s"], s["c_t"], s["convg_thresh"],
                                 s["remove_rings"], s["quad_convex"], s["huber_delta"],
                                 s["huber_T"], mult_xy, mult_z, stage_idx, prev_mult_z, var_est))
        self.h = h
        self.L = L
        ints = np.zeros(16, np.int32)
        dbls = np.zeros(64, np.float64)
        L.tb_stage_info(h, ints, dbls)
        self.info = {k: int(v) for k, v in zip(INFO_INTS, ints)}
        self.info.update({k: float(v) for k, v in zip(INFO_DBLS, dbls)})
        self.sfilt = dbls[len(INFO_DBLS):len(INFO_DBLS) + 27].reshape(3, 3, 3).copy()

    def close(self):
        if getattr(self, "h", None):
            self.L.tb_stage_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def sino_shape(self):
        i = self.info
        return (i["N_p"], i["N_r"], i["N_t"])

    @property
    def obj_shape(self):
        i = self.info
        return (i["N_time"], i["N_z"], i["N_y"], i["N_x"])

    def _get(self, fn, shape):
        out = np.zeros(shape, np.float32)
        _check(fn(self.h, out))
        return out

    def object(self):
        return self._get(self.L.tb_get_object, self.obj_shape)

    def set_object(self, x):
        _check(self.L.tb_set_object(self.h, np.ascontiguousarray(x, np.float32)))

    def offset(self):
        return self._get(self.L.tb_get_offset, (self.info["N_r"], self.info["N_t"]))

    def set_offset(self, d):
        _check(self.L.tb_set_offset(self.h, np.ascontiguousarray(d, np.float32)))

    def magmap(self):
        return self._get(self.L.tb_get_magmap, self.obj_shape)

    def set_magmap(self, m):
        _check(self.L.tb_set_magmap(self.h, np.ascontiguousarray(m, np.float32)))

    def error_sino(self):
        return self._get(self.L.tb_get_error_sino, self.sino_shape)

    def weights(self):
        return self._get(self.L.tb_get_weights, self.sino_shape)

    def detector_response(self):
        return self._get(self.L.tb_get_detector_response, (self.info["N_p"], 65))

    def forward_project(self):
        return self._get(self.L.tb_forward_project, self.sino_shape)

    @property
    def var_est(self):
        return self.L.tb_get_var_est(self.h)

    def init_from_coarse(self, prev_obj, prev_nz, prev_mag):
        _check(self.L.tb_init_from_coarse(self.h, np.ascontiguousarray(prev_obj, np.float32),
                                          prev_nz, np.ascontiguousarray(prev_mag, np.float32)))

    def init_from_stage(self, prev: "Stage"):
        _check(self.L.tb_init_from_stage(self.h, prev.h))

    def init_fbp(self):
        _check(self.L.tb_init_fbp(self.h))

    def init_error_sino(self):
        _check(self.L.tb_init_error_sino(self.h))

    def cost(self):
        t, f, p = C.c_double(), C.c_double(), C.c_double()
        _check(self.L.tb_cost(self.h, C.byref(t), C.byref(f), C.byref(p)))
        return t.value, f.value, p.value

    def offset_update(self, typ):
        _check(self.L.tb_offset_update(self.h, typ))

    def variance_update(self):
        _check(self.L.tb_variance_update(self.h))

    def voxel_update(self, f, z, y, x):
        v = C.c_float()
        _check(self.L.tb_voxel_update(self.h, f, z, y, x, C.byref(v)))
        return v.value

    def update_colour(self, colour, it, chunk, nhicd=0):
        _check(self.L.tb_update_colour(self.h, colour, it, chunk, nhicd))

    def vsc_select(self):
        _check(self.L.tb_vsc_select(self.h))

    def nhicd_list(self, f, z):
        nxy = self.info["N_x"] * self.info["N_y"]
        xs = np.zeros(nxy, np.int32)
        ys = np.zeros(nxy, np.int32)
        n = self.L.tb_nhicd_list(self.h, f, z, xs.ctypes.data, ys.ctypes.data)
        if n < 0:
            raise TimbirError(self.L.tb_last_error().decode())
        return xs[:n].copy(), ys[:n].copy()

    def icd_iteration(self, it):
        conv = C.c_int32()
        r = C.c_double()
        _check(self.L.tb_icd_iteration(self.h, it, C.byref(conv), C.byref(r)))
        return conv.value, r.value

    def stage_run(self, max_iter=0, with_cost=True):
        n = max_iter if max_iter > 0 else 1000
        costs = np.zeros(n + 1, np.float64)
        it = C.c_int32()
        _check(self.L.tb_stage_run(self.h, max_iter, costs.ctypes.data if with_cost else None,
                                   C.byref(it)))
        return costs[:it.value + 1].copy()

    def stats(self) -> dict:
        s = TbStats()
        _check(self.L.tb_get_stats(self.h, C.byref(s)))
        out = {k: getattr(s, k) for k, _ in TbStats._fields_}
        out["family_launches"] = list(s.family_launches)
        return out

    def reset_stats(self):
        _check(self.L.tb_reset_stats(self.h))

    def synchronize(self):
        _check(self.L.tb_synchronize(self.h))

    def abort(self):
        """End the cross-workgroup waits of the split-cell launch running
        on this stage (tb_stage_abort); safe from another thread."""
        _check(self.L.tb_stage_abort(self.h))

    def object_device(self):
        """(device address, bytes) of the object in HBM."""
        p, n = C.c_void_p(), C.c_uint64()
        _check(self.L.tb_stage_object_device(self.h, C.byref(p), C.byref(n)))
        return p.value, n.value

    def autotune(self) -> dict:
        """The kernel configuration the autotuner finds fastest for this
        stage's shape (timed now unless cached), as a dict."""
        e = TbTuneEntry()
        _check(self.L.tb_autotune(self.h, C.byref(e)))
        return _tune_dict(e)


def _tune_dict(e) -> dict:
    return {"shape": list(e.shape), "kernel": e.kernel, "views_per_lane": e.views_per_lane,
            "ms": e.ms}


def autotune_entries() -> list:
    """The autotuner's cached choices, one dict per stage shape."""
    L = lib()
    n = L.tb_autotune_entries(None, 0)
    arr = (TbTuneEntry * max(n, 1))()
    n = L.tb_autotune_entries(arr, n)
    return [_tune_dict(arr[i]) for i in range(n)]


def autotune_put(entries) -> None:
    """Load choices (dicts as autotune_entries returns) into the cache."""
    L = lib()
    for d in entries:
        e = TbTuneEntry()
        for i, v in enumerate(d["shape"]):
            e.shape[i] = int(v)
        e.kernel = int(d["kernel"])
        e.views_per_lane = int(d["views_per_lane"])
        e.ms = float(d.get("ms", 0.0))
        _check(L.tb_autotune_put(C.byref(e)))


def autotune_clear() -> None:
    lib().tb_autotune_clear()


def autotune_save(path: str) -> None:
    """Write the cache as JSON (autotune_load reads it back)."""
    import json
    with open(path, "w") as f:
        json.dump(autotune_entries(), f, indent=1)


def autotune_load(path: str) -> None:
    import json
    with open(path) as f:
        autotune_put(json.load(f))


class Batch:
    """Stages of independent problems advanced in lockstep (tb_batch): one
    voxel-update launch per kernel configuration and phase over all
    members' cells, iterations 2 on replayed from a HIP graph when `graph`.
    A member's results equal, bit for bit, those of its own
    Stage.icd_iteration."""

    def __init__(self, graph: bool = True):
        self.L = lib()
        h = C.c_void_p()
        _check(self.L.tb_batch_create(C.byref(h), BATCH_GRAPH if graph else 0))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.tb_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def iteration(self, stages, it: int):
        """One ICD iteration of every stage; returns (converged, ratio)
        arrays, one entry per stage."""
        n = len(stages)
        arr = (C.c_void_p * n)(*[s.h.value for s in stages])
        conv = (C.c_int32 * n)()
        ratio = (C.c_double * n)()
        _check(self.L.tb_batch_iteration(self.h, arr, n, it, conv, ratio))
        return np.array(conv[:], np.int32), np.array(ratio[:], np.float64)

    def stats(self) -> dict:
        s = TbBatchStats()
        _check(self.L.tb_batch_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in TbBatchStats._fields_}


def reconstruct(scan: dict, restart: int = 0, debug=None):
    """Call the drop-in ``reconstruct()`` with a scan dict.  The C function
    frees its inputs, so copies are handed over (allocated with libc malloc).
    Returns the finest-stage object [recon_num][rows][cols][cols]."""
    L = lib()
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    libc.free.argtypes = [C.c_void_p]

    def mdup(a):
        a = np.ascontiguousarray(a, np.float32)
        p = libc.malloc(max(a.nbytes, 4))
        C.memmove(p, a.ctypes.data, a.nbytes)
        return p

    s = scan
    ptrs = [mdup(s[k]) for k in ("projections", "weights", "proj_angles", "proj_times",
                                   "recon_times")]
    obj = C.POINTER(C.c_float)()
    fp = None
    if debug is not None:
        libc.fdopen.restype = C.c_void_p
        libc.fdopen.argtypes = [C.c_int, C.c_char_p]
        fp = libc.fdopen(os.dup(debug.fileno()), b"w")
    rc = L.reconstruct(C.byref(obj), *ptrs, s["proj_rows"], s["proj_cols"], s["proj_num"],
                       s["recon_num"], s["vox_wid"], s["rot_center"], s["sig_s"], s["sig_t"],
                       s["c_s"], s["c_t"], s["convg_thresh"], s["remove_rings"],
                       s["quad_convex"], s["huber_delta"], s["huber_T"], restart, fp)
    if fp:
        libc.fflush.argtypes = [C.c_void_p]
        libc.fflush(fp)
    if rc != 0:
        raise TimbirError("reconstruct() returned -1: " + L.tb_last_error().decode())
    shape = (s["recon_num"], s["proj_rows"], s["proj_cols"], s["proj_cols"])
    n = int(np.prod(shape))
    out = np.ctypeslib.as_array(obj, shape=(n,)).reshape(shape).copy()
    libc.free(C.cast(obj, C.c_void_p))
    return out
