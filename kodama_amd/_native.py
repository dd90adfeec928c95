"""ctypes binding of libkodama_hip.so (the C ABI declared in include/kodama_hip.h).

The library is built in-tree (``make -C kodama_amd/csrc``); loading fails loudly
when it is missing, and every call raises when the library reports an error.
There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libkodama_hip.so")

OK, EINVAL, ERUNTIME, ENODEV = 0, 1, 2, 3
N_COUNTERS = 17
COUNTER_NAMES = (
    "cycles_completed", "proposals_evaluated", "best_state_updates",
    "current_state_accepts", "stochastic_state_attempts", "stochastic_state_accepts",
    "current_state_rejections", "coarsening_moves", "absorption_moves",
    "transition_attempted", "transition_accepted", "many_to_one_attempted",
    "many_to_one_accepted", "pls_coarsening_attempted", "pls_coarsening_accepted",
    "cv_evaluations", "success",
)
METRICS = {"cosine": 0, "inner_product": 1, "euclidean": 2}


class KodamaHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class InvalidArgument(KodamaHipError, ValueError):
    pass


class NoDevice(KodamaHipError):
    pass


class CoreOptions(C.Structure):
    """Mirror of ``kodama_hip_core_options``."""

    _fields_ = [
        ("cycles", C.c_int32), ("classifier", C.c_int32), ("seed", C.c_uint64),
        ("shake", C.c_int32), ("evolutionary_search", C.c_int32),
        ("guarded_diversity", C.c_int32), ("many_to_one_absorption", C.c_int32),
        ("auto_class_coarsening", C.c_int32), ("adaptive_proposal_size", C.c_int32),
        ("prediction_guidance", C.c_int32), ("policy_adaptive_proposal_size", C.c_int32),
        ("transition_proposal", C.c_int32), ("stochastic_acceptance", C.c_int32),
        ("diversity_multiplier", C.c_int32), ("pls_transition_coarsening", C.c_int32),
        ("pls_fragmentation_penalty", C.c_int32), ("reject_single_class", C.c_int32),
        ("folds", C.c_int32), ("stratified", C.c_int32), ("cv_seed", C.c_uint64),
        ("k", C.c_int32), ("metric", C.c_int32),
        ("max_components", C.c_int32), ("fixed_components", C.c_int32),
        ("center", C.c_int32), ("scale", C.c_int32), ("pls_fold_cache", C.c_int32),
        ("device", C.c_int32), ("n_threads", C.c_int32),
    ]


DEFAULTS = dict(cycles=100, classifier=1, seed=1, shake=0, evolutionary_search=0,
                guarded_diversity=0, many_to_one_absorption=0, auto_class_coarsening=0,
                adaptive_proposal_size=1, prediction_guidance=1,
                policy_adaptive_proposal_size=1, transition_proposal=1,
                stochastic_acceptance=1, diversity_multiplier=1,
                pls_transition_coarsening=1, pls_fragmentation_penalty=1,
                reject_single_class=1, folds=10, stratified=1, cv_seed=1, k=30, metric=2,
                max_components=10, fixed_components=0, center=1, scale=1,
                pls_fold_cache=1, device=0, n_threads=0)


def core_options(**kw) -> CoreOptions:
    """Options with kodama::CoreOptions defaults (kodama.hpp:294-314)."""
    d = dict(DEFAULTS)
    d.update(kw)
    o = CoreOptions()
    for key, v in d.items():
        setattr(o, key, int(v))
    return o


class MatrixOptions(C.Structure):
    """Mirror of ``kodama_hip_matrix_options``."""

    _fields_ = [
        ("runs", C.c_int32), ("cycles", C.c_int32), ("landmarks", C.c_int32), ("splitting", C.c_int32),
        ("folds", C.c_int32), ("k", C.c_int32), ("metric", C.c_int32), ("graph_neighbors", C.c_int32),
        ("seed", C.c_uint64), ("form", C.c_int32), ("device", C.c_int32), ("n_threads", C.c_int32),
        ("mlp_params", C.POINTER(C.c_float)), ("mlp_fallback", C.c_float),
        ("classifier", C.c_int32), ("components", C.c_int32),
    ]


class CoreResult(C.Structure):
    _fields_ = [
        ("clbest", C.POINTER(C.c_int32)), ("clbest_dirty", C.POINTER(C.c_int32)),
        ("cvpredbest", C.POINTER(C.c_int32)), ("fold_assignments", C.POINTER(C.c_int32)),
        ("vect_acc", C.POINTER(C.c_double)), ("vect_score", C.POINTER(C.c_double)),
        ("proposal_size", C.POINTER(C.c_int32)), ("active_classes", C.POINTER(C.c_int32)),
        ("accepted", C.POINTER(C.c_uint8)), ("improving_acceptance", C.POINTER(C.c_uint8)),
        ("temperature_acceptance", C.POINTER(C.c_uint8)), ("counters", C.POINTER(C.c_int32)),
        ("accbest", C.c_double), ("scorebest", C.c_double),
    ]


_CT = {np.int32: C.c_int32, np.float64: C.c_double, np.uint8: C.c_uint8, np.float32: C.c_float}


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(_CT[a.dtype.type]))


class CoreBuffers:
    """Caller-owned output arrays of one kodama_hip_core_result."""

    PER_CYCLE = ("vect_acc", "vect_score", "proposal_size", "active_classes", "accepted",
                 "improving_acceptance", "temperature_acceptance")

    def __init__(self, n: int, cycles: int):
        c = max(cycles, 1)
        self.cycles = cycles
        self.arrays = dict(
            clbest=np.zeros(n, np.int32), clbest_dirty=np.zeros(n, np.int32),
            cvpredbest=np.zeros(n, np.int32), fold_assignments=np.zeros(n, np.int32),
            vect_acc=np.zeros(c, np.float64), vect_score=np.zeros(c, np.float64),
            proposal_size=np.zeros(c, np.int32), active_classes=np.zeros(c, np.int32),
            accepted=np.zeros(c, np.uint8), improving_acceptance=np.zeros(c, np.uint8),
            temperature_acceptance=np.zeros(c, np.uint8), counters=np.zeros(N_COUNTERS, np.int32))
        self.struct = CoreResult()
        for k, a in self.arrays.items():
            setattr(self.struct, k, ptr(a))

    def result(self) -> dict:
        out = {}
        for k, a in self.arrays.items():
            if k in self.PER_CYCLE:
                out[k] = a[: self.cycles].copy()
            elif k == "counters":
                out.update({name: int(v) for name, v in zip(COUNTER_NAMES, a)})
            else:
                out[k] = a.copy()
        out["accbest"] = float(self.struct.accbest)
        out["scorebest"] = float(self.struct.scorebest)
        return out


_lib = None


def lib():
    """Load libkodama_hip.so once; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `make -C kodama_amd/csrc` "
                "(the HIP backend has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        L.kodama_hip_last_error.restype = C.c_char_p
        L.kodama_hip_version.restype = C.c_char_p
        L.kodama_hip_release_knn_vote_graph.restype = None
        _lib = L
    return _lib


def lib_available() -> bool:
    return os.path.exists(LIB_PATH)


def call(name: str, *args) -> None:
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != OK:
        msg = f"{name}: {L.kodama_hip_last_error().decode()}"
        if rc == EINVAL:
            raise InvalidArgument(rc, msg)
        if rc == ENODEV:
            raise NoDevice(rc, msg)
        raise KodamaHipError(rc, msg)


def device_count() -> int:
    n = C.c_int32(0)
    call("kodama_hip_device_count", C.byref(n))
    return n.value


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)
