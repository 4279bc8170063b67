"""ctypes binding of ``libbug.so``, the gfx950 HIP library behind the BUG hot path.

Every numerical entry point of this package goes through here.  The library is loaded from the
package directory (built in-tree by ``bondupdategalerkin_amd.build``); if it is missing or fails
to load, :func:`lib` raises -- there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbug.so")

BUG_OK = 0
BUG_EINVAL = 1
BUG_ENONFINITE = 2
BUG_EINVARIANT = 3
BUG_EDEVICE = 4
BUG_ENOMEM = 5

PREC_FP64 = 0
PREC_FP32 = 1

OP_N, OP_T, OP_C, OP_R = 0, 1, 2, 3

_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


class BugLeg(ctypes.Structure):
    _fields_ = [("nsec", ctypes.c_int32), ("charge", _i32p), ("dim", _i32p)]


class BugBsm(ctypes.Structure):
    _fields_ = [
        ("row", BugLeg),
        ("col", BugLeg),
        ("nslice", ctypes.c_int32),
        ("shift", _i32p),
        ("offset", _i64p),
        ("data", ctypes.c_void_p),
        ("nelem", ctypes.c_int64),
    ]


class BugMpoSite(ctypes.Structure):
    _fields_ = [
        ("wl", ctypes.c_int32),
        ("wr", ctypes.c_int32),
        ("cl", _i32p),
        ("cr", _i32p),
        ("nent", ctypes.c_int32),
        ("a", _i32p),
        ("b", _i32p),
        ("sout", _i32p),
        ("sin", _i32p),
        ("coeff", _f64p),
    ]


class BugTerm(ctypes.Structure):
    _fields_ = [
        ("o", ctypes.c_int32),
        ("p", ctypes.c_int32),
        ("q", ctypes.c_int32),
        ("op_a", ctypes.c_int32),
        ("op_b", ctypes.c_int32),
        ("alpha_re", ctypes.c_double),
        ("alpha_im", ctypes.c_double),
    ]


class BugCopy(ctypes.Structure):
    _fields_ = [("src_off", ctypes.c_int64), ("dst_off", ctypes.c_int64),
                ("src_ld", ctypes.c_int32), ("dst_ld", ctypes.c_int32),
                ("rows", ctypes.c_int32), ("cols", ctypes.c_int32),
                ("mode", ctypes.c_int32), ("scale_mode", ctypes.c_int32),
                ("scale_off", ctypes.c_int64)]


class BugQrJob(ctypes.Structure):
    _fields_ = [("a_off", ctypes.c_int64), ("q_off", ctypes.c_int64), ("r_off", ctypes.c_int64),
                ("m", ctypes.c_int32), ("n", ctypes.c_int32), ("lda", ctypes.c_int32),
                ("ldq", ctypes.c_int32), ("ldr", ctypes.c_int32), ("pad", ctypes.c_int32)]


class BugSvdJob(ctypes.Structure):
    _fields_ = [("g_off", ctypes.c_int64), ("u_off", ctypes.c_int64), ("s_off", ctypes.c_int64),
                ("vh_off", ctypes.c_int64), ("m", ctypes.c_int32), ("n", ctypes.c_int32),
                ("ldg", ctypes.c_int32), ("ldu", ctypes.c_int32), ("ldvh", ctypes.c_int32),
                ("phase_on_v", ctypes.c_int32)]


class BugStepOpts(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int32), ("maxdim", ctypes.c_int32),
                ("hermitian", ctypes.c_int32), ("work_ceiling_rule", ctypes.c_int32),
                ("bond_ceiling_factor", ctypes.c_double), ("root_ceiling_factor", ctypes.c_double),
                ("tol_basis", ctypes.c_double), ("tol_trunc", ctypes.c_double),
                ("tol_expv", ctypes.c_double), ("tol_krylov", ctypes.c_double),
                ("root", ctypes.c_int32), ("reorth", ctypes.c_int32),
                ("truncate", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("parallel", ctypes.c_int32), ("seed", ctypes.c_uint64),
                ("fixed_width", ctypes.c_int32), ("tail_factor", ctypes.c_double),
                ("krylov_mode", ctypes.c_int32), ("mode", ctypes.c_int32),
                ("exact", ctypes.c_int32), ("fold_omega", ctypes.c_int32),
                ("krylov_rank_tol", ctypes.c_double), ("root_method", ctypes.c_int32),
                ("root_shift", ctypes.c_int32), ("root_maxiter", ctypes.c_int32)]


class BugBondLog(ctypes.Structure):
    _fields_ = [("bond", ctypes.c_int32), ("side", ctypes.c_int32),
                ("sample", ctypes.c_double), ("tail", ctypes.c_double),
                ("krylov_c", ctypes.c_double), ("krylov_d", ctypes.c_double),
                ("n_probes", ctypes.c_int32), ("n_holdout", ctypes.c_int32),
                ("n_selected", ctypes.c_int32), ("n_targeted", ctypes.c_int32),
                ("n_krylov", ctypes.c_int32), ("n_h1", ctypes.c_int32),
                ("workcap", ctypes.c_int32), ("met", ctypes.c_int32),
                ("unmet", ctypes.c_int32), ("width", ctypes.c_int32),
                ("r_old", ctypes.c_int32), ("r_final", ctypes.c_int32),
                ("discarded", ctypes.c_double)]


class BugStepInfo(ctypes.Structure):
    _fields_ = [("tolerance_met", ctypes.c_int32), ("unmet", ctypes.c_int32),
                ("max_sample", ctypes.c_double), ("max_tail", ctypes.c_double),
                ("max_krylov", ctypes.c_double), ("max_krylov_c", ctypes.c_double),
                ("max_krylov_d", ctypes.c_double), ("discarded", ctypes.c_double),
                ("worst_discarded", ctypes.c_double), ("expv_err", ctypes.c_double),
                ("expv_iters", ctypes.c_int32), ("proj_err", ctypes.c_double),
                ("defect", ctypes.c_double), ("leak", ctypes.c_double * 4),
                ("n_probes", ctypes.c_int64), ("n_selected", ctypes.c_int64),
                ("n_krylov", ctypes.c_int64), ("n_h1", ctypes.c_int64),
                ("n_h1_closure", ctypes.c_int64), ("n_closure_sweeps", ctypes.c_int64),
                ("workcap_reached", ctypes.c_int32), ("maxdim_reached", ctypes.c_int32),
                ("r_old", ctypes.c_int32), ("r_aug", ctypes.c_int32), ("r_final", ctypes.c_int32),
                ("root_link", ctypes.c_int32), ("root_rank", ctypes.c_int32),
                ("root_applications", ctypes.c_int32), ("seconds", ctypes.c_double * 6),
                ("expanded", _i32p), ("expanded_cap", ctypes.c_int32),
                ("log", ctypes.POINTER(BugBondLog)), ("log_cap", ctypes.c_int32),
                ("log_count", ctypes.c_int32)]


class BugExpvReport(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("substeps", ctypes.c_int32),
                ("err", ctypes.c_double)]


# int (*bug_apply_fn)(void* user, const void* x, void* y, void* stream)
APPLY_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                            ctypes.c_void_p)


# name -> (restype, argtypes); the list is the ABI of include/bug.h
_P = ctypes.c_void_p
_BSMP = ctypes.POINTER(BugBsm)
_MPOP = ctypes.POINTER(BugMpoSite)
SIGNATURES = {
    "bug_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_P)]),
    "bug_ctx_destroy": (ctypes.c_int, [_P]),
    "bug_last_error": (ctypes.c_char_p, [_P]),
    "bug_set_tile_config": (ctypes.c_int, [_P, ctypes.c_int]),
    "bug_tile_config_count": (ctypes.c_int, []),
    "bug_set_gemm_schedule": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "bug_set_autotune": (ctypes.c_int, [ctypes.c_int32]),
    "bug_autotune_count": (ctypes.c_int, [_i64p]),
    "bug_autotune_clear": (ctypes.c_int, []),
    "bug_autotune_save": (ctypes.c_int, [ctypes.c_char_p]),
    "bug_autotune_load": (ctypes.c_int, [ctypes.c_char_p]),
    "bug_last_schedules": (ctypes.c_int, [_P, ctypes.c_int32, _i32p, _i32p]),
    "bug_graph_capture_begin": (ctypes.c_int, [_P, _P]),
    "bug_graph_capture_end": (ctypes.c_int, [_P]),
    "bug_graph_release": (ctypes.c_int, [_P, _P]),
    "bug_ctx_stats": (ctypes.c_int, [_P, _i64p]),
    "bug_last_op_flops": (ctypes.c_int, [_P, _f64p]),
    "bug_bsm_product": (
        ctypes.c_int,
        [_P, _BSMP, _BSMP, _BSMP, ctypes.c_int32, ctypes.POINTER(BugTerm), ctypes.c_int32,
         ctypes.c_int32, _P],
    ),
    "bug_env_push_left": (ctypes.c_int, [_P, _BSMP, _BSMP, _BSMP, _MPOP, _BSMP, ctypes.c_int32, _P]),
    "bug_env_push_right": (ctypes.c_int, [_P, _BSMP, _BSMP, _BSMP, _MPOP, _BSMP, ctypes.c_int32, _P]),
    "bug_apply_one_site": (
        ctypes.c_int,
        [_P, _BSMP, _MPOP, _BSMP, _BSMP, _BSMP, ctypes.c_double, ctypes.c_int32, _P],
    ),
    "bug_apply_one_site_noalloc": (
        ctypes.c_int,
        [_P, _BSMP, _MPOP, _BSMP, _BSMP, _BSMP, ctypes.c_double, ctypes.c_int32, _P],
    ),
    "bug_apply_one_site_batch": (
        ctypes.c_int,
        [_P, ctypes.c_int32, _BSMP, _MPOP, _BSMP, _BSMP, _BSMP, _f64p, ctypes.c_int32, _P],
    ),
    "bug_apply_zero_site": (
        ctypes.c_int,
        [_P, _BSMP, _BSMP, _BSMP, _BSMP, ctypes.c_double, ctypes.c_int32, _P],
    ),
    "bug_apply_zero_site_batch": (
        ctypes.c_int,
        [_P, ctypes.c_int32, _BSMP, _BSMP, _BSMP, _BSMP, _f64p, ctypes.c_int32, _P],
    ),
    "bug_axpy": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _P, _P, _P]),
    "bug_scal": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _P, _P]),
    "bug_dotc_segments": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, _i64p, ctypes.c_int32,
                                         ctypes.POINTER(_P), _P, _f64p, _P]),
    "bug_axpy_segments": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, _i64p, ctypes.c_int32,
                                         _f64p, ctypes.POINTER(_P), _P, _P]),
    "bug_axpy_many": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, _f64p, ctypes.POINTER(_P),
                                     _P, _P]),
    "bug_dotc": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_P), _P, _f64p, _P]),
    "bug_copy_blocks": (
        ctypes.c_int,
        [_P, ctypes.c_int32, ctypes.POINTER(BugCopy), _P, ctypes.c_int64, _P, ctypes.c_int64,
         _f64p, ctypes.c_int32, _P],
    ),
    "bug_block_sqnorms": (
        ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(BugCopy), _P, ctypes.c_int64, _f64p, _P]),
    "bug_block_qr": (
        ctypes.c_int,
        [_P, ctypes.c_int32, ctypes.POINTER(BugQrJob), _P, ctypes.c_int64, _P, ctypes.c_int64, _P,
         ctypes.c_int64, _P],
    ),
    "bug_block_qrp": (
        ctypes.c_int,
        [_P, ctypes.c_int32, ctypes.POINTER(BugQrJob), _P, ctypes.c_int64, _P, ctypes.c_int64,
         _f64p, ctypes.c_int64, _P],
    ),
    "bug_set_svd_variant": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "bug_set_qr_variant": (ctypes.c_int, [_P, ctypes.c_int32]),
    "bug_set_jacobi_params": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "bug_set_gram_jacobi_params": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "bug_set_mixed_jacobi_params": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "bug_svd_sweeps": (ctypes.c_int, [_P, ctypes.c_int32, _i32p]),
    "bug_set_svd_floor": (ctypes.c_int, [_P, ctypes.c_double]),
    "bug_step_opts_default": (None, [ctypes.POINTER(BugStepOpts)]),
    "bug_two_site_sketch_left": (ctypes.c_int, [_P, _BSMP, _MPOP, _MPOP, _BSMP, _BSMP, _BSMP, _BSMP,
                                                _BSMP, ctypes.c_int32, _P]),
    "bug_two_site_sketch_right": (ctypes.c_int, [_P, _BSMP, _MPOP, _MPOP, _BSMP, _BSMP, _BSMP,
                                                 _BSMP, _BSMP, ctypes.c_int32, _P]),
    "bug_expv": (ctypes.c_int, [_P, APPLY_FN, _P, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                _P, _P, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                ctypes.c_int32, ctypes.c_double, ctypes.POINTER(BugExpvReport), _P]),
    "bug_svd_select": (ctypes.c_int, [ctypes.c_int32, _i32p, _f64p, ctypes.c_double, ctypes.c_int32,
                                      _i32p, _i32p, _i32p, _f64p]),
    "bug_mpo_create": (ctypes.c_int, [ctypes.c_int32, _MPOP, ctypes.POINTER(_P)]),
    "bug_mpo_destroy": (ctypes.c_int, [_P]),
    "bug_state_create": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, _i32p, _BSMP,
                                        ctypes.c_int32, _P, ctypes.POINTER(_P)]),
    "bug_state_create_device": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, _i32p, _BSMP,
                                               ctypes.c_int32, _P, ctypes.POINTER(_P)]),
    "bug_state_clone": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "bug_state_expect_diagonal": (ctypes.c_int, [_P, _P, _f64p, _f64p]),
    "bug_state_destroy": (ctypes.c_int, [_P]),
    "bug_state_info": (ctypes.c_int, [_P, _i32p, _i32p, _i32p]),
    "bug_state_site_shape": (ctypes.c_int, [_P, ctypes.c_int32, _i32p, _i32p, _i64p]),
    "bug_state_download_site": (ctypes.c_int, [_P, _P, ctypes.c_int32, _i32p, _i32p, _i32p, _i32p,
                                               _i64p, _P]),
    "bug_step": (ctypes.c_int, [_P, _P, _P, ctypes.c_double, ctypes.c_double,
                                ctypes.POINTER(BugStepOpts), ctypes.POINTER(BugStepInfo)]),
    "bug_step_batch": (ctypes.c_int, [_P, ctypes.c_int32, _P, _P, _f64p,
                                      ctypes.POINTER(BugStepOpts), ctypes.POINTER(BugStepInfo),
                                      ctypes.c_int32]),
    "bug_block_svd": (
        ctypes.c_int,
        [_P, ctypes.c_int32, ctypes.POINTER(BugSvdJob), _P, ctypes.c_int64, _P, ctypes.c_int64,
         _f64p, ctypes.c_int64, _P, ctypes.c_int64, ctypes.c_int32, _P],
    ),
}

_lib = None
_lock = threading.Lock()


def load_library(path: str = LIB_PATH):
    """Load ``libbug.so`` and attach the C signatures.  Raises if it is missing."""
    if not os.path.exists(path):
        raise RuntimeError(
            f"libbug.so not found at {path}: build it with bondupdategalerkin_amd.build.build() "
            "(the package has no CPU fallback)"
        )
    handle = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    return handle


def lib():
    global _lib
    with _lock:
        if _lib is None:
            _lib = load_library()
        return _lib


class BugError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"libbug error {code}: {message}")
        self.code = code


class Context:
    """One library context (plan cache + per-stream scratch) on one device."""

    def __init__(self, device: int = 0):
        self._lib = lib()
        handle = ctypes.c_void_p()
        rc = self._lib.bug_ctx_create(device, ctypes.byref(handle))
        if rc != BUG_OK:
            raise BugError(rc, f"cannot create a context on device {device}")
        self.handle = handle
        self.device = device

    def check(self, rc: int):
        if rc != BUG_OK:
            msg = self._lib.bug_last_error(self.handle)
            raise BugError(rc, msg.decode() if msg else "")

    def last_flops(self) -> float:
        v = ctypes.c_double()
        self.check(self._lib.bug_last_op_flops(self.handle, ctypes.byref(v)))
        return v.value

    def last_schedules(self):
        """[(tile configuration, split factor)] of the grouped-GEMM launches of the last call."""
        out = (ctypes.c_int32 * 64)()
        n = ctypes.c_int32()
        self.check(self._lib.bug_last_schedules(self.handle, 32, out, ctypes.byref(n)))
        return [(out[2 * i], out[2 * i + 1]) for i in range(min(n.value, 32))]

    def stats(self):
        out = (ctypes.c_int64 * 4)()
        self.check(self._lib.bug_ctx_stats(self.handle, out))
        return {"bytes": out[0], "plans": out[1], "hits": out[2], "misses": out[3]}

    def set_tile_config(self, config: int):
        self.check(self._lib.bug_set_tile_config(self.handle, config))

    def set_gemm_schedule(self, config: int = -1, splits: int = 0):
        """Force this context's grouped-GEMM tile configuration and split-k factor (-1 / 0:
        automatic)."""
        self.check(self._lib.bug_set_gemm_schedule(self.handle, config, splits))

    def close(self):
        if self.handle:
            self._lib.bug_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


AUTOTUNE_OFF, AUTOTUNE_USE, AUTOTUNE_TUNE = 0, 1, 2
# schedules the autotuner chose on MI355X for the XXZ L = 128, D = 1024 step (mixed and FP64)
SCHEDULE_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "schedules_gfx950.txt")


def set_autotune(mode: int):
    """Process-wide grouped-GEMM autotuning: AUTOTUNE_OFF (heuristic schedules), AUTOTUNE_USE
    (schedules from the tuning table where it has the plan's class) or AUTOTUNE_TUNE (as USE, and
    time every candidate on the first launch of a new class)."""
    if lib().bug_set_autotune(mode) != BUG_OK:
        raise ValueError(f"autotune mode must be 0, 1 or 2, not {mode}")


def autotune_count():
    """(classes in the tuning table, classes tuned by this process)."""
    out = (ctypes.c_int64 * 2)()
    lib().bug_autotune_count(out)
    return int(out[0]), int(out[1])


def autotune_clear():
    lib().bug_autotune_clear()


def autotune_save(path: str):
    if lib().bug_autotune_save(os.fsencode(path)) != BUG_OK:
        raise OSError(f"cannot write the tuning table to {path}")


def autotune_load(path: str):
    """Merge a saved tuning table (raises on a missing or malformed file)."""
    if lib().bug_autotune_load(os.fsencode(path)) != BUG_OK:
        raise ValueError(f"cannot read a tuning table from {path}")


_local = threading.local()


def context(device: int = 0) -> Context:
    """This thread's context on ``device``: contexts hold plan caches, arenas and pinned staging
    buffers that are not shared between host threads, so every thread gets its own (created on
    first use, destroyed with the thread)."""
    ctxs = getattr(_local, "contexts", None)
    if ctxs is None:
        ctxs = _local.contexts = {}
    ctx = ctxs.get(device)
    if ctx is None:
        ctx = ctxs[device] = Context(device)
    return ctx


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def ptr32(a: np.ndarray):
    return a.ctypes.data_as(_i32p)


def ptr64(a: np.ndarray):
    return a.ctypes.data_as(_i64p)


def ptrf(a: np.ndarray):
    return a.ctypes.data_as(_f64p)
