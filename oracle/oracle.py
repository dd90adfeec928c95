"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (this project's C++ restatement of the reference
algorithms) and, where it was built, ``oracle/_ref/libkodama_ref.so`` (the
reference's own C++ core behind ``ref_driver.cpp``). Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
module, and only as the checker or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libkodama_ref.so")
REF_MATRIX_SO = os.path.join(HERE, "_ref", "libkodama_ref_matrix.so")

N_COUNTERS = 17
COUNTER_NAMES = (
    "cycles_completed", "proposals_evaluated", "best_state_updates",
    "current_state_accepts", "stochastic_state_attempts", "stochastic_state_accepts",
    "current_state_rejections", "coarsening_moves", "absorption_moves",
    "transition_attempted", "transition_accepted", "many_to_one_attempted",
    "many_to_one_accepted", "pls_coarsening_attempted", "pls_coarsening_accepted",
    "cv_evaluations", "success",
)


class CoreOptions(C.Structure):
    """Mirror of ``oracle_core_options`` (oracle/kodama_oracle.h)."""

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
        ("hnsw_ef_search", C.c_int32), ("n_threads", C.c_int32),
    ]


def core_options(**kw) -> CoreOptions:
    """Defaults follow kodama::CoreOptions (kodama.hpp:294-314) with KNN k=30."""
    d = dict(cycles=100, classifier=1, seed=1, shake=0, evolutionary_search=0,
             guarded_diversity=0, many_to_one_absorption=0, auto_class_coarsening=0,
             adaptive_proposal_size=1, prediction_guidance=1,
             policy_adaptive_proposal_size=1, transition_proposal=1,
             stochastic_acceptance=1, diversity_multiplier=1,
             pls_transition_coarsening=1, pls_fragmentation_penalty=1,
             reject_single_class=1, folds=10, stratified=1, cv_seed=1, k=30, metric=2,
             max_components=10, fixed_components=0, center=1, scale=1,
             pls_fold_cache=1, hnsw_ef_search=4096, n_threads=1)
    d.update(kw)
    o = CoreOptions()
    for key, v in d.items():
        setattr(o, key, int(v))
    return o


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


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr_or_null(a, ct):
    return None if a is None else _p(a, ct)


class _CoreBuffers:
    def __init__(self, n: int, cycles: int):
        self.arrays = dict(
            clbest=np.zeros(n, np.int32), clbest_dirty=np.zeros(n, np.int32),
            cvpredbest=np.zeros(n, np.int32), fold_assignments=np.zeros(n, np.int32),
            vect_acc=np.zeros(max(cycles, 1), np.float64),
            vect_score=np.zeros(max(cycles, 1), np.float64),
            proposal_size=np.zeros(max(cycles, 1), np.int32),
            active_classes=np.zeros(max(cycles, 1), np.int32),
            accepted=np.zeros(max(cycles, 1), np.uint8),
            improving_acceptance=np.zeros(max(cycles, 1), np.uint8),
            temperature_acceptance=np.zeros(max(cycles, 1), np.uint8),
            counters=np.zeros(N_COUNTERS, np.int32))
        self.cycles = cycles
        r = CoreResult()
        for k, a in self.arrays.items():
            ct = {np.int32: C.c_int32, np.float64: C.c_double, np.uint8: C.c_uint8}[a.dtype.type]
            setattr(r, k, _p(a, ct))
        self.struct = r

    def result(self) -> dict:
        out = {}
        for k, a in self.arrays.items():
            if k in ("vect_acc", "vect_score", "proposal_size", "active_classes", "accepted",
                     "improving_acceptance", "temperature_acceptance"):
                out[k] = a[: self.cycles].copy()
            elif k == "counters":
                out.update({name: int(v) for name, v in zip(COUNTER_NAMES, a)})
            else:
                out[k] = a.copy()
        out["accbest"] = float(self.struct.accbest)
        out["scorebest"] = float(self.struct.scorebest)
        return out


def build(with_reference: bool = False) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if with_reference and os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)


class _Lib:
    def __init__(self, path: str, prefix: str):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = C.CDLL(path)
        self.prefix = prefix
        err = getattr(self.lib, prefix + "last_error")
        err.restype = C.c_char_p
        self._err = err

    def call(self, name, *args):
        fn = getattr(self.lib, self.prefix + name)
        rc = fn(*args)
        if rc != 0:
            raise RuntimeError(f"{self.prefix}{name}: {self._err().decode()}")


class _Api:
    """Common calls that both the oracle and the reference driver export."""

    def __init__(self, lib: _Lib):
        self._l = lib

    def make_folds(self, labels, folds=10, stratified=True, seed=1, constrain=None):
        lab = _i32(labels)
        con = _i32(constrain)
        out = np.zeros(lab.size, np.int32)
        self._l.call("make_folds", _p(lab, C.c_int32), C.c_int32(lab.size),
                     _ptr_or_null(con, C.c_int32), C.c_int32(folds), C.c_int32(int(stratified)),
                     C.c_uint64(seed), _p(out, C.c_int32))
        return out

    def core_knn(self, x, labels, opt: CoreOptions, constrain=None, fixed=None):
        x = _f32(x)
        lab = _i32(labels)
        buf = _CoreBuffers(lab.size, opt.cycles)
        self._l.call("core_knn", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     _p(lab, C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32),
                     _ptr_or_null(_i32(fixed), C.c_int32), C.byref(opt), C.byref(buf.struct))
        return buf.result()

    def knncv(self, x, labels, opt: CoreOptions, constrain=None):
        x = _f32(x)
        lab = _i32(labels)
        pred = np.zeros(lab.size, np.int32)
        folds = np.zeros(lab.size, np.int32)
        self._l.call("knncv", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     _p(lab, C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32), C.byref(opt),
                     _p(pred, C.c_int32), _p(folds, C.c_int32))
        return pred, folds


class MatrixOptions(C.Structure):
    """Mirror of ``oracle_matrix_options``."""

    _fields_ = [
        ("runs", C.c_int32), ("cycles", C.c_int32), ("landmarks", C.c_int32), ("splitting", C.c_int32),
        ("folds", C.c_int32), ("k", C.c_int32), ("metric", C.c_int32), ("graph_neighbors", C.c_int32),
        ("seed", C.c_uint64), ("classifier", C.c_int32), ("components", C.c_int32), ("precision", C.c_int32),
    ]


def matrix_options(runs, cycles, landmarks, splitting=0, folds=5, k=30, metric=2, graph_neighbors=100,
                   seed=1234, classifier="knn", components=50, precision=64) -> MatrixOptions:
    """KODAMAMatrixOptions defaults (kodama.hpp:532-566); classifier "knn" or
    "pls_lda"."""
    return MatrixOptions(runs=runs, cycles=cycles, landmarks=landmarks, splitting=splitting, folds=folds, k=k,
                         metric=metric, graph_neighbors=graph_neighbors, seed=seed,
                         classifier=1 if classifier == "knn" else 0, components=components, precision=precision)


def _resolved_landmarks(n, landmarks):
    L = landmarks if 0 < landmarks < n else (int(np.ceil(0.75 * n)) if landmarks >= n else 10000)
    return max(2, min(L, n - 1))


class _MatrixBlocks:
    """k-means labels and quota landmarks, exported by the oracle and by the
    reference matrix driver under the same names."""

    def kmeans_labels(self, x, k, seed, discard=0, max_iter=10):
        x = _f32(x)
        out = np.zeros(x.shape[0], np.int32)
        self._l.call("kmeans_labels", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     C.c_int32(k), C.c_uint64(seed), C.c_int32(discard), C.c_int32(max_iter), _p(out, C.c_int32))
        return out

    def quota_sample(self, strata_labels, strata, target, seed, discard=0):
        lab = _i32(strata_labels)
        out = np.zeros(lab.size, np.int32)
        cnt = C.c_int32(0)
        self._l.call("quota_sample", _p(lab, C.c_int32), C.c_int32(lab.size), C.c_int32(strata),
                     C.c_int32(target), C.c_uint64(seed), C.c_int32(discard), _p(out, C.c_int32), C.byref(cnt))
        return out[: cnt.value].copy()

    def normalize_graph(self, indices, distances, index_base=-1, width=None):
        gi = _i32(indices)
        gd = _f32(distances)
        n, given = gi.shape
        width = given if width is None else width
        idx = np.zeros((n, width), np.int32)
        dist = np.zeros((n, width), np.float32)
        self._l.call("normalize_graph", _p(gi, C.c_int32), _p(gd, C.c_float), C.c_int32(n), C.c_int32(given),
                     C.c_int32(index_base), C.c_int32(width), _p(idx, C.c_int32), _p(dist, C.c_float))
        return idx, dist


class Oracle(_Api, _MatrixBlocks):
    def __init__(self):
        super().__init__(_Lib(ORACLE_SO, "oracle_"))
        L = self._l.lib
        L.oracle_distance.restype = C.c_float
        L.oracle_objective.restype = C.c_double

    def distance(self, a, b, metric=2):
        a = _f32(a)
        b = _f32(b)
        return float(self._l.lib.oracle_distance(_p(a, C.c_float), _p(b, C.c_float),
                                                  C.c_int32(a.size), C.c_int32(metric)))

    def exact_knn(self, train, query, k, metric=2, exclude=None):
        tr = _f32(train)
        q = _f32(query)
        idx = np.zeros((q.shape[0], k), np.int32)
        dist = np.zeros((q.shape[0], k), np.float32)
        self._l.call("exact_knn", _p(tr, C.c_float), C.c_int32(tr.shape[0]), _p(q, C.c_float),
                     C.c_int32(q.shape[0]), C.c_int32(tr.shape[1]), C.c_int32(k), C.c_int32(metric),
                     _ptr_or_null(_i32(exclude), C.c_int32), _p(idx, C.c_int32), _p(dist, C.c_float))
        return idx, dist

    def precompute_knn_cv(self, x, folds, k, metric=2):
        x = _f32(x)
        f = _i32(folds)
        nbr = np.zeros((x.shape[0], k), np.int32)
        score = np.zeros((x.shape[0], k), np.float32)
        self._l.call("precompute_knn_cv", _p(x, C.c_float), C.c_int32(x.shape[0]),
                     C.c_int32(x.shape[1]), _p(f, C.c_int32), C.c_int32(k), C.c_int32(metric),
                     _p(nbr, C.c_int32), _p(score, C.c_float))
        return nbr, score

    def knn_vote(self, nbr, score, folds, labels):
        nbr = _i32(nbr)
        score = _f32(score)
        f = _i32(folds)
        lab = _i32(labels)
        pred = np.zeros(lab.size, np.int32)
        self._l.call("knn_vote", _p(nbr, C.c_int32), _p(score, C.c_float), C.c_int32(nbr.shape[0]),
                     C.c_int32(nbr.shape[1]), _p(f, C.c_int32), _p(lab, C.c_int32), _p(pred, C.c_int32))
        return pred

    def objective(self, labels, accuracy, opt: CoreOptions):
        lab = _i32(labels)
        return float(self._l.lib.oracle_objective(_p(lab, C.c_int32), C.c_int32(lab.size),
                                                   C.c_double(accuracy), C.byref(opt)))

    def core_knn_tables(self, nbr, score, folds, labels, opt: CoreOptions, constrain=None, fixed=None):
        nbr = _i32(nbr)
        score = _f32(score)
        f = _i32(folds)
        lab = _i32(labels)
        buf = _CoreBuffers(lab.size, opt.cycles)
        self._l.call("core_knn_tables", _p(nbr, C.c_int32), _p(score, C.c_float),
                     C.c_int32(nbr.shape[0]), C.c_int32(nbr.shape[1]), _p(f, C.c_int32),
                     _p(lab, C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32),
                     _ptr_or_null(_i32(fixed), C.c_int32), C.byref(opt), C.byref(buf.struct))
        return buf.result()

    def knn_graph(self, x, k, metric=2):
        x = _f32(x)
        kk = max(1, min(k, x.shape[0] - 1))
        idx = np.zeros((x.shape[0], kk), np.int32)
        dist = np.zeros((x.shape[0], kk), np.float32)
        self._l.call("knn_graph", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     C.c_int32(k), C.c_int32(metric), _p(idx, C.c_int32), _p(dist, C.c_float))
        return idx, dist

    def dissimilarity(self, indices, distances, res, input_one_based=True, output_one_based=True):
        idx = np.array(indices, dtype=np.int32, copy=True, order="C")
        dist = np.array(distances, dtype=np.float32, copy=True, order="C")
        res = _i32(res)
        self._l.call("dissimilarity", _p(idx, C.c_int32), _p(dist, C.c_float), C.c_int32(idx.shape[0]),
                     C.c_int32(idx.shape[1]), _p(res, C.c_int32), C.c_int32(res.shape[0]),
                     C.c_int32(int(input_one_based)), C.c_int32(int(output_one_based)))
        return idx, dist

    def plsldacv(self, x, labels, opt: CoreOptions, precision=64, constrain=None):
        x = _f32(x)
        lab = _i32(labels)
        pred = np.zeros(lab.size, np.int32)
        folds = np.zeros(lab.size, np.int32)
        sel = C.c_int32(0)
        self._l.call("plsldacv", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     _p(lab, C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32), C.byref(opt),
                     C.c_int32(precision), _p(pred, C.c_int32), _p(folds, C.c_int32), C.byref(sel))
        return pred, folds, sel.value

    def pls_fold_detail(self, x, labels, folds, fold, opt: CoreOptions, precision=64):
        x = _f32(x)
        lab = _i32(labels)
        f = _i32(folds)
        n, p = x.shape
        A = opt.max_components
        w = np.zeros(p * A, np.float64)
        vs = np.zeros(n * A, np.float64)
        disc = np.zeros(n * n, np.float64)
        cls = np.zeros(n, np.int32)
        rank = C.c_int32(0)
        ncls = C.c_int32(0)
        nval = C.c_int32(0)
        self._l.call("pls_fold_detail", _p(x, C.c_float), C.c_int32(n), C.c_int32(p), _p(lab, C.c_int32),
                     _p(f, C.c_int32), C.c_int32(fold), C.byref(opt), C.c_int32(precision),
                     _p(w, C.c_double), C.byref(rank), _p(vs, C.c_double), _p(disc, C.c_double),
                     _p(cls, C.c_int32), C.byref(ncls), C.byref(nval))
        r, nc, nv = rank.value, ncls.value, nval.value
        a = min(opt.fixed_components or opt.max_components, r) if r > 0 else 0
        return dict(weights=w[: p * r].reshape(p, r) if r else np.zeros((p, 0)),
                    val_scores=vs[: nv * a].reshape(nv, a) if a else np.zeros((nv, 0)),
                    discriminant=disc[: nv * nc].reshape(nv, nc) if r else np.zeros((nv, 0)),
                    classes=cls[:nc].copy(), rank=r)

    def core_plslda(self, x, labels, opt: CoreOptions, precision=64, constrain=None, fixed=None):
        x = _f32(x)
        lab = _i32(labels)
        buf = _CoreBuffers(lab.size, opt.cycles)
        self._l.call("core_plslda", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     _p(lab, C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32),
                     _ptr_or_null(_i32(fixed), C.c_int32), C.byref(opt), C.c_int32(precision),
                     C.byref(buf.struct))
        return buf.result()


def _oracle_kodama_matrix(self, x, runs, cycles, landmarks, splitting=0, folds=5, k=30, metric=2,
                          graph_neighbors=100, seed=1234, graph=None, index_base=-1, classifier="knn",
                          components=50, precision=64, W=None, constrain=None, fixed=None):
    """KODAMA.matrix; with graph = (indices, distances), KODAMAMatrixFromGraphData.
    W / constrain / fixed: the reference's starting_labels, constrain and fixed
    (one per row, None for empty). diag [runs x 3]: FNV-1a hashes of each run's
    landmark rows, starting labels and fold ids."""
    x = _f32(x)
    n = x.shape[0]
    o = matrix_options(runs, cycles, landmarks, splitting, folds, k, metric, graph_neighbors, seed, classifier,
                       components, precision)
    L = _resolved_landmarks(n, landmarks)
    res = np.zeros((runs, n), np.int32)
    acc = np.zeros(runs, np.float64)
    v = np.zeros((runs, max(cycles, 1)), np.float64)
    land = np.zeros((runs, L), np.int32)
    diag = np.zeros((runs, 3), np.uint64)
    used = C.c_int32(0)
    extra = (_ptr_or_null(_i32(W), C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32),
             _ptr_or_null(_i32(fixed), C.c_int32))
    outs = (_p(res, C.c_int32), _p(acc, C.c_double), _p(v, C.c_double), _p(land, C.c_int32), C.byref(used),
            _p(diag, C.c_uint64))
    if graph is None:
        self._l.call("kodama_matrix", _p(x, C.c_float), C.c_int32(n), C.c_int32(x.shape[1]), C.byref(o),
                     *extra, *outs)
    else:
        gi, gd = _i32(graph[0]), _f32(graph[1])
        assert gi.shape == gd.shape and gi.shape[0] == n
        self._l.call("kodama_matrix_from_graph", _p(x, C.c_float), C.c_int32(n), C.c_int32(x.shape[1]),
                     C.byref(o), _p(gi, C.c_int32), _p(gd, C.c_float), C.c_int32(gi.shape[1]),
                     C.c_int32(index_base), *extra, *outs)
    assert used.value == L
    return dict(res=res, acc=acc, v=v[:, :cycles].copy(), landmarks=land, diag=diag)


Oracle.kodama_matrix = _oracle_kodama_matrix


def _pls_buffers(n, p, q, opt):
    A = max(1, opt.fixed_components if opt.fixed_components > 0 else opt.max_components)
    return dict(weights=np.zeros(p * A), response_loadings=np.zeros(q * A), scores=np.zeros(n * A),
                coefficients=np.zeros((p, q)), fitted=np.zeros((n, q)), x_center=np.zeros(p),
                x_scale=np.zeros(p), y_center=np.zeros(q))


def _pls_result(buf, n, p, q, comps):
    A = comps.value
    out = dict(buf)
    out["components"] = A
    out["weights"] = buf["weights"][: p * A].reshape(p, A)
    out["response_loadings"] = buf["response_loadings"][: q * A].reshape(q, A)
    out["scores"] = buf["scores"][: n * A].reshape(n, A)
    return out


def _pls_call(lib, name, x, y, opt, extra=()):
    x, y = _f32(x), _f32(y)
    if y.ndim == 1:
        y = y[:, None]
    n, p = x.shape
    q = y.shape[1]
    buf = _pls_buffers(n, p, q, opt)
    comps = C.c_int32(0)
    lib.call(name, _p(x, C.c_float), C.c_int32(n), C.c_int32(p), _p(y, C.c_float), C.c_int32(q), C.byref(opt),
             *extra, *(_p(buf[k], C.c_double) for k in ("weights", "response_loadings", "scores", "coefficients",
                                                         "fitted", "x_center", "x_scale", "y_center")),
             C.byref(comps))
    return _pls_result(buf, n, p, q, comps)


def _oracle_plsdacv(self, x, labels, opt, precision=64, constrain=None):
    """PLSDACV (plscv.cpp:2926-2928): the PLS-LDA CV folds and SIMPLS with the
    least-squares PLS-DA head."""
    x, lab = _f32(x), _i32(labels)
    pred = np.zeros(lab.size, np.int32)
    folds = np.zeros(lab.size, np.int32)
    sel = C.c_int32(0)
    self._l.call("plsdacv", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), _p(lab, C.c_int32),
                 _ptr_or_null(_i32(constrain), C.c_int32), C.byref(opt), C.c_int32(precision), _p(pred, C.c_int32),
                 _p(folds, C.c_int32), C.byref(sel))
    return pred, folds, sel.value


def _oracle_pls(self, x, y, opt, precision=64):
    """General PLS (PLS_CPU, plscv.cpp:2765-2901)."""
    return _pls_call(self._l, "pls", x, y, opt, (C.c_int32(precision),))


Oracle.plsdacv = _oracle_plsdacv
Oracle.pls = _oracle_pls


def _oracle_plslda_predict(self, train, labels, test, opt, precision=64):
    """PLSLDAPredict_CPU (plscv.cpp:2976-3005) restated as the validation fold
    of a two-fold split of [train; test]: its training rows are all of train,
    so the standardisation, SIMPLS fit (components capped by the columns and
    rows - 1) and LDA are the single fit of the reference; the prediction is
    the first class of largest discriminant. One training class predicts
    itself; a fit without a component raises (DegeneratePLSFit)."""
    tr, te, lab = _f32(train), _f32(test), _i32(labels)
    if tr.shape[0] == 0 or te.shape[0] == 0 or tr.shape[1] == 0:
        return np.zeros(0, np.int32)
    classes = np.unique(lab)
    if classes.size == 1:
        return np.full(te.shape[0], classes[0], np.int32)
    x = np.vstack([tr, te])
    lab_all = np.concatenate([lab, np.full(te.shape[0], lab[0], np.int32)])
    folds = np.concatenate([np.ones(tr.shape[0], np.int32), np.zeros(te.shape[0], np.int32)])
    d = self.pls_fold_detail(x, lab_all, folds, 0, opt, precision)
    if d["rank"] == 0:
        raise ValueError("PLSLDAPredict: degenerate PLS fit")
    return d["classes"][np.argmax(d["discriminant"], axis=1)].astype(np.int32)


Oracle.plslda_predict = _oracle_plslda_predict


class ReferenceMatrix(_MatrixBlocks):
    """The reference's own k-means and landmark sampling (oracle/_ref), where built."""

    def __init__(self):
        self._l = _Lib(REF_MATRIX_SO, "ref_matrix_")
        self._l.prefix = "ref_"
        self._l._err = self._l.lib.ref_matrix_last_error
        self._l._err.restype = C.c_char_p

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_MATRIX_SO)


class Reference(_Api):
    """The reference's own CPU core (oracle/_ref), where it was built."""

    def __init__(self):
        super().__init__(_Lib(REF_SO, "ref_"))

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)

    def plslda_predict(self, train, labels, test, opt: CoreOptions):
        """kodama::PLSLDAPredict_CPU (plscv.cpp:2976-3005)."""
        tr, te, lab = _f32(train), _f32(test), _i32(labels)
        pred = np.zeros(te.shape[0], np.int32)
        self._l.call("plslda_predict", _p(tr, C.c_float), C.c_int32(tr.shape[0]), _p(lab, C.c_int32),
                     _p(te, C.c_float), C.c_int32(te.shape[0]), C.c_int32(tr.shape[1]), C.byref(opt),
                     _p(pred, C.c_int32))
        return pred

    def hnsw_search(self, train, query, k, metric=2, ef_search=4096):
        tr = _f32(train)
        q = _f32(query)
        idx = np.zeros((q.shape[0], k), np.int32)
        dist = np.zeros((q.shape[0], k), np.float32)
        self._l.call("hnsw_search", _p(tr, C.c_float), C.c_int32(tr.shape[0]), _p(q, C.c_float),
                     C.c_int32(q.shape[0]), C.c_int32(tr.shape[1]), C.c_int32(k), C.c_int32(metric),
                     C.c_int32(ef_search), _p(idx, C.c_int32), _p(dist, C.c_float))
        return idx, dist

    def core_knn_graph(self, idx, dist, labels, opt: CoreOptions, constrain=None, fixed=None):
        idx = _i32(idx)
        dist = _f32(dist)
        lab = _i32(labels)
        buf = _CoreBuffers(lab.size, opt.cycles)
        self._l.call("core_knn_graph", _p(idx, C.c_int32), _p(dist, C.c_float), C.c_int32(idx.shape[0]),
                     C.c_int32(idx.shape[1]), _p(lab, C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32),
                     _ptr_or_null(_i32(fixed), C.c_int32), C.byref(opt), C.byref(buf.struct))
        return buf.result()

    def plsldacv(self, x, labels, opt: CoreOptions, constrain=None):
        x = _f32(x)
        lab = _i32(labels)
        pred = np.zeros(lab.size, np.int32)
        folds = np.zeros(lab.size, np.int32)
        sel = C.c_int32(0)
        self._l.call("plsldacv", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     _p(lab, C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32), C.byref(opt),
                     _p(pred, C.c_int32), _p(folds, C.c_int32), C.byref(sel))
        return pred, folds, sel.value

    def plsdacv(self, x, labels, opt: CoreOptions, constrain=None):
        """kodama::PLSDACV_CPU (plscv.cpp:2926-2928)."""
        x, lab = _f32(x), _i32(labels)
        pred = np.zeros(lab.size, np.int32)
        folds = np.zeros(lab.size, np.int32)
        sel = C.c_int32(0)
        self._l.call("plsdacv", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     _p(lab, C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32), C.byref(opt),
                     _p(pred, C.c_int32), _p(folds, C.c_int32), C.byref(sel))
        return pred, folds, sel.value

    def pls(self, x, y, opt: CoreOptions):
        """kodama::PLS_CPU (plscv.cpp:2899-2901), outputs widened to double."""
        return _pls_call(self._l, "pls", x, y, opt)

    def core_plslda(self, x, labels, opt: CoreOptions, constrain=None, fixed=None):
        x = _f32(x)
        lab = _i32(labels)
        buf = _CoreBuffers(lab.size, opt.cycles)
        self._l.call("core_plslda", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     _p(lab, C.c_int32), _ptr_or_null(_i32(constrain), C.c_int32),
                     _ptr_or_null(_i32(fixed), C.c_int32), C.byref(opt), C.byref(buf.struct))
        return buf.result()

    def dissimilarity(self, indices, distances, res):
        idx = np.array(indices, dtype=np.int32, copy=True, order="C")
        dist = np.array(distances, dtype=np.float32, copy=True, order="C")
        res = _i32(res)
        self._l.call("dissimilarity", _p(idx, C.c_int32), _p(dist, C.c_float), C.c_int32(idx.shape[0]),
                     C.c_int32(idx.shape[1]), _p(res, C.c_int32), C.c_int32(res.shape[0]))
        return idx, dist

    def kodama_matrix(self, x, runs, cycles, landmarks, splitting=0, folds=5, k=30, metric=2,
                      graph_neighbors=100, seed=1234, graph=None, index_base=-1, classifier="knn",
                      components=50, W=None, constrain=None, fixed=None):
        """kodama::KODAMAMatrix_CPU, or KODAMAMatrixFromGraphData_CPU with graph =
        (indices, distances); one thread. res, acc, v and the run diagnostics'
        hashes (diag [runs x 3])."""
        x = _f32(x)
        n = x.shape[0]
        o = matrix_options(runs, cycles, landmarks, splitting, folds, k, metric, graph_neighbors, seed,
                           classifier, components)
        res = np.zeros((runs, n), np.int32)
        acc = np.zeros(runs, np.float64)
        v = np.zeros((runs, max(cycles, 1)), np.float64)
        diag = np.zeros((runs, 3), np.uint64)
        if graph is None:
            gi, gd, width = None, None, 0
        else:
            gi, gd = _i32(graph[0]), _f32(graph[1])
            width = gi.shape[1]
        self._l.call("kodama_matrix", _p(x, C.c_float), C.c_int32(n), C.c_int32(x.shape[1]), C.byref(o),
                     _ptr_or_null(gi, C.c_int32), _ptr_or_null(gd, C.c_float), C.c_int32(width),
                     C.c_int32(index_base), _ptr_or_null(_i32(W), C.c_int32),
                     _ptr_or_null(_i32(constrain), C.c_int32), _ptr_or_null(_i32(fixed), C.c_int32),
                     _p(res, C.c_int32), _p(acc, C.c_double), _p(v, C.c_double), _p(diag, C.c_uint64))
        return dict(res=res, acc=acc, v=v[:, :cycles].copy(), diag=diag)

    def knn_graph(self, x, k, metric=2, n_threads=1):
        x = _f32(x)
        idx = np.zeros((x.shape[0], k), np.int32)
        dist = np.zeros((x.shape[0], k), np.float32)
        self._l.call("knn_graph", _p(x, C.c_float), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]),
                     C.c_int32(k), C.c_int32(metric), C.c_int32(n_threads), _p(idx, C.c_int32),
                     _p(dist, C.c_float))
        return idx, dist
