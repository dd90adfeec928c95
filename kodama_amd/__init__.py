"""kodama_amd -- the KODAMA.matrix hot path on AMD MI355X (gfx950).

Host-side mirror of the reference's user API for the cross-validated classifiers
and label-evolution loops (R/kodama.R:160-349: KNNCV, PLSLDACV, CoreKNN,
CorePLSLDA), the exact kNN graph (KODAMA.graph's builder) and the KODAMA
dissimilarity, with the same argument names, defaults and error behaviour. Every
function runs on the HIP backend through libkodama_hip.so; there is no CPU path.
Results are dicts with the fields of the reference's R lists (kodama_r.cpp).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import (CoreOptions, InvalidArgument, KodamaHipError, NoDevice,  # noqa: F401
                      core_options, device_count)

__all__ = [
    "KNNCV", "PLSLDACV", "CoreKNN", "CorePLSLDA", "CoreKNNGraph", "KODAMA_knn_graph",
    "KODAMA_dissimilarity", "core_knn", "core_plslda", "core_knn_batch", "exact_knn_search",
    "KnnVoteGraph", "core_options", "device_count", "version",
    "set_evolution_split", "last_evolution_workgroups", "set_cooperative_launch", "set_lds_flow_table",
    "last_evolution_lds_flows",
]


def version() -> str:
    return N.lib().kodama_hip_version().decode()


def _metric(metric) -> int:
    if isinstance(metric, (list, tuple)):
        metric = metric[0]
    if metric not in N.METRICS:
        raise ValueError(f"metric must be one of {tuple(N.METRICS)}")
    return N.METRICS[metric]


def _backend(backend) -> None:
    if backend not in (None, "hip", "auto"):
        raise ValueError("kodama_amd runs the 'hip' backend only")


def _matrix(data) -> np.ndarray:
    x = N.f32(data)
    if x.ndim != 2:
        raise ValueError("data must be a 2-D matrix")
    return x


# ---- exact kNN seam ---------------------------------------------------------

def exact_knn_search(train, query, k, metric="euclidean", exclude=None, gpu_device=0):
    """Exact kNN (native_cuda_exact_knn_search contract): internal distances."""
    tr, q = _matrix(train), _matrix(query)
    ex = N.i32(exclude)
    idx = np.zeros((q.shape[0], k), np.int32)
    dist = np.zeros((q.shape[0], k), np.float32)
    N.call("kodama_hip_exact_knn_search", N.ptr(tr), C.c_int32(tr.shape[0]), N.ptr(q),
           C.c_int32(q.shape[0]), C.c_int32(tr.shape[1]), C.c_int32(k), C.c_int32(_metric(metric)),
           C.c_int32(gpu_device), N.ptr(ex), N.ptr(idx), N.ptr(dist))
    return idx, dist


class KnnVoteGraph:
    """Resident vote graph (NativeCudaKNNVoteGraph contract)."""

    def __init__(self, neighbor_rows, scores, gpu_device=0):
        nb = N.i32(neighbor_rows)
        sc = N.f32(scores)
        self.samples, self.neighbors = nb.shape
        h = C.c_void_p()
        N.call("kodama_hip_build_knn_vote_graph", N.ptr(nb), N.ptr(sc), C.c_int32(self.samples),
               C.c_int32(self.neighbors), C.c_int32(gpu_device), C.byref(h))
        self._h = h

    def predict(self, label_codes, fallback_code=0):
        codes = N.i32(label_codes)
        out = np.zeros(self.samples, np.int32)
        N.call("kodama_hip_knn_vote_predict", self._h, N.ptr(codes), C.c_int32(fallback_code), N.ptr(out))
        return out

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            N.lib().kodama_hip_release_knn_vote_graph(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- cross-validated classifiers --------------------------------------------

def KNNCV(data, labels, constrain=None, folds=10, stratified=True, seed=1, k=10,
          metric=("cosine", "inner_product", "euclidean"), backend=None, n_cores=1, gpu_device=0):
    """k-nearest-neighbour cross-validation (R/kodama.R:160-186)."""
    _backend(backend)
    x = _matrix(data)
    lab = N.i32(labels)
    opt = core_options(k=k, metric=_metric(metric), folds=folds, stratified=int(bool(stratified)),
                       cv_seed=seed, device=gpu_device, n_threads=n_cores)
    pred = np.zeros(x.shape[0], np.int32)
    fold = np.zeros(x.shape[0], np.int32)
    acc = C.c_double(0.0)
    N.call("kodama_hip_knncv", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), N.ptr(lab),
           N.ptr(N.i32(constrain)), C.byref(opt), N.ptr(pred), N.ptr(fold), C.byref(acc))
    return dict(predicted_labels=pred, true_labels=lab.copy(), fold_assignments=fold,
                global_accuracy=acc.value, backend="hip")


def PLSLDACV(data, labels, constrain=None, folds=10, stratified=True, seed=1, ncomp=None,
             center=True, scale=True, backend=None, n_cores=1, gpu_device=0):
    """PLS-LDA cross-validation (R/kodama.R:212-239)."""
    _backend(backend)
    x = _matrix(data)
    lab = N.i32(labels)
    ncomp = min(50, x.shape[1]) if ncomp is None else int(ncomp)
    opt = core_options(classifier=0, folds=folds, stratified=int(bool(stratified)), cv_seed=seed,
                       max_components=ncomp, fixed_components=ncomp, center=int(bool(center)),
                       scale=int(bool(scale)), pls_fold_cache=0, device=gpu_device, n_threads=n_cores)
    pred = np.zeros(x.shape[0], np.int32)
    fold = np.zeros(x.shape[0], np.int32)
    sel = C.c_int32(0)
    N.call("kodama_hip_plsldacv", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), N.ptr(lab),
           N.ptr(N.i32(constrain)), C.byref(opt), N.ptr(pred), N.ptr(fold), C.byref(sel))
    return dict(predicted_labels=pred, true_labels=lab.copy(), fold_assignments=fold,
                selected_components=sel.value, global_accuracy=float(np.mean(pred == lab)),
                backend="hip")


def PLSDACV(data, labels, constrain=None, folds=10, stratified=True, seed=1, ncomp=None,
            center=True, scale=True, backend=None, n_cores=1, gpu_device=0, max_components=None):
    """PLS-DA cross-validation (PLSDACV, plscv.cpp:2926-2938): the folds and
    SIMPLS fits of PLSLDACV with the least-squares PLS-DA head. ncomp fixes
    the evaluated components (0: max_components)."""
    _backend(backend)
    x = _matrix(data)
    lab = N.i32(labels)
    ncomp = min(50, x.shape[1]) if ncomp is None else int(ncomp)
    mc = ncomp if max_components is None else int(max_components)
    opt = core_options(classifier=0, folds=folds, stratified=int(bool(stratified)), cv_seed=seed,
                       max_components=mc, fixed_components=ncomp, center=int(bool(center)),
                       scale=int(bool(scale)), pls_fold_cache=0, device=gpu_device, n_threads=n_cores)
    pred = np.zeros(x.shape[0], np.int32)
    fold = np.zeros(x.shape[0], np.int32)
    sel = C.c_int32(0)
    N.call("kodama_hip_plsdacv", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), N.ptr(lab),
           N.ptr(N.i32(constrain)), C.byref(opt), N.ptr(pred), N.ptr(fold), C.byref(sel))
    return dict(predicted_labels=pred, true_labels=lab.copy(), fold_assignments=fold,
                selected_components=sel.value, global_accuracy=float(np.mean(pred == lab)),
                backend="hip")


def PLS(x, y, ncomp=10, center=True, scale=True, max_components=None, gpu_device=0) -> dict:
    """General PLS (PLS_CUDA contract; fit_general_pls, plscv.cpp:2765-2878):
    weights [p x A], response_loadings [q x A], scores [n x A], coefficients
    [p x q], fitted [n x q], x_center, x_scale, y_center, components, and the
    device seconds of the cross-product kernel and of the whole fit."""
    xm = _matrix(x)
    ym = np.ascontiguousarray(y, np.float32)
    if ym.ndim == 1:
        ym = ym[:, None]
    n, p = xm.shape
    q = ym.shape[1]
    if ym.shape[0] != n:
        raise ValueError("PLS X and Y must have the same rows.")
    mc = int(ncomp) if max_components is None else int(max_components)
    opt = core_options(classifier=0, max_components=mc, fixed_components=int(ncomp), center=int(bool(center)),
                       scale=int(bool(scale)), device=gpu_device)
    A = max(1, int(ncomp) if int(ncomp) > 0 else mc)
    out = dict(weights=np.zeros(p * A, np.float32), response_loadings=np.zeros(q * A, np.float32),
               scores=np.zeros(n * A, np.float32), coefficients=np.zeros((p, q), np.float32),
               fitted=np.zeros((n, q), np.float32), x_center=np.zeros(p, np.float32),
               x_scale=np.zeros(p, np.float32), y_center=np.zeros(q, np.float32))
    comps = C.c_int32(0)
    timings = np.zeros(2, np.float64)
    N.call("kodama_hip_pls", N.ptr(xm), C.c_int32(n), C.c_int32(p), N.ptr(ym), C.c_int32(q), C.byref(opt),
           *(N.ptr(out[k]) for k in ("weights", "response_loadings", "scores", "coefficients", "fitted",
                                     "x_center", "x_scale", "y_center")), C.byref(comps), N.ptr(timings))
    a = comps.value
    out["weights"] = out["weights"][: p * a].reshape(p, a)
    out["response_loadings"] = out["response_loadings"][: q * a].reshape(q, a)
    out["scores"] = out["scores"][: n * a].reshape(n, a)
    out.update(components=a, cross_seconds=float(timings[0]), device_seconds=float(timings[1]))
    return out


def PLSLDAPredict(train, labels, test, ncomp=None, center=True, scale=True, backend=None, gpu_device=0):
    """PLSLDAPredict (plscv.cpp:2976-3005): fit PLS-LDA on the training rows
    and their labels, return the predicted labels of the test rows."""
    _backend(backend)
    tr = _matrix(train)
    te = _matrix(test)
    lab = N.i32(labels)
    if te.shape[1] != tr.shape[1] and te.size:
        raise ValueError("PLSLDAPredict train/test column mismatch.")
    ncomp = min(50, tr.shape[1]) if ncomp is None else int(ncomp)
    opt = core_options(classifier=0, max_components=ncomp, fixed_components=ncomp, center=int(bool(center)),
                       scale=int(bool(scale)), pls_fold_cache=0, device=gpu_device)
    pred = np.zeros(te.shape[0], np.int32)
    N.call("kodama_hip_plslda_predict", N.ptr(tr), C.c_int32(tr.shape[0]), N.ptr(lab), N.ptr(te),
           C.c_int32(te.shape[0]), C.c_int32(tr.shape[1]), C.byref(opt), N.ptr(pred))
    return pred


def plslda_fold_detail(data, labels, folds, fold, options: CoreOptions) -> dict:
    """One PLS-LDA fold evaluated on the device in FP64: SIMPLS weights,
    validation scores and LDA discriminants (validation rows ascending)."""
    x = _matrix(data)
    lab = N.i32(labels)
    fo = N.i32(folds)
    n, p = x.shape
    width = max(1, options.max_components)
    w = np.zeros(p * width, np.float64)
    vs = np.zeros(n * n, np.float64)
    disc = np.zeros(n * n, np.float64)
    cls = np.zeros(n, np.int32)
    rank, ncls, nval = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    N.call("kodama_hip_plslda_fold_detail", N.ptr(x), C.c_int32(n), C.c_int32(p), N.ptr(lab), N.ptr(fo),
           C.c_int32(fold), C.byref(options), N.ptr(w), C.byref(rank), N.ptr(vs), N.ptr(disc), N.ptr(cls),
           C.byref(ncls), C.byref(nval))
    r, nc, nv = rank.value, ncls.value, nval.value
    a = min(options.fixed_components or options.max_components, r) if r > 0 else 0
    return dict(weights=w[: p * r].reshape(p, r), val_scores=vs[: nv * a].reshape(nv, a),
                discriminant=disc[: nv * nc].reshape(nv, nc) if r else np.zeros((nv, 0)),
                classes=cls[:nc].copy(), rank=r)


# ---- label evolution ----------------------------------------------------------

def core_knn(data, labels, options: CoreOptions, constrain=None, fix=None) -> dict:
    """kodama::CoreKNN with full CoreOptions (core.cpp:1866-1876)."""
    x = _matrix(data)
    lab = N.i32(labels)
    buf = N.CoreBuffers(lab.size, options.cycles)
    N.call("kodama_hip_core_knn", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), N.ptr(lab),
           N.ptr(N.i32(constrain)), N.ptr(N.i32(fix)), C.byref(options), C.byref(buf.struct))
    out = buf.result()
    out["backend"] = "hip"
    return out


def core_plslda(data, labels, options: CoreOptions, constrain=None, fix=None) -> dict:
    """kodama::CorePLSLDA with full CoreOptions (core.cpp:1854-1864)."""
    x = _matrix(data)
    lab = N.i32(labels)
    buf = N.CoreBuffers(lab.size, options.cycles)
    N.call("kodama_hip_core_plslda", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), N.ptr(lab),
           N.ptr(N.i32(constrain)), N.ptr(N.i32(fix)), C.byref(options), C.byref(buf.struct))
    out = buf.result()
    out["backend"] = "hip"
    return out


def CoreKNN(data, labels, constrain=None, fix=None, cycles=100, folds=10, stratified=True, seed=1, k=30,
            metric=("euclidean", "cosine", "inner_product"), backend=None, n_cores=4, gpu_device=0):
    """kNN label evolution (R/kodama.R:266-296 -> core_impl_cpp, kodama_r.cpp:1118-1174)."""
    _backend(backend)
    opt = core_options(cycles=cycles, classifier=1, seed=seed, folds=folds, stratified=int(bool(stratified)),
                       cv_seed=seed, k=k, metric=_metric(metric), device=gpu_device, n_threads=n_cores)
    return core_knn(data, labels, opt, constrain=constrain, fix=fix)


def CorePLSLDA(data, labels, constrain=None, fix=None, cycles=100, folds=10, stratified=True, seed=1,
               ncomp=None, backend=None, n_cores=4, gpu_device=0):
    """PLS-LDA label evolution (R/kodama.R:322-349 -> core_impl_cpp)."""
    _backend(backend)
    x = _matrix(data)
    ncomp = min(50, x.shape[1]) if ncomp is None else int(ncomp)
    opt = core_options(cycles=cycles, classifier=0, seed=seed, folds=folds, stratified=int(bool(stratified)),
                       cv_seed=seed, max_components=ncomp, fixed_components=ncomp, device=gpu_device,
                       n_threads=n_cores)
    return core_plslda(x, labels, opt, constrain=constrain, fix=fix)


def CoreKNNGraph(graph_indices, graph_distances, labels, options: CoreOptions, constrain=None, fix=None,
                 index_base="auto") -> dict:
    """kodama::CoreKNNGraph on a caller graph (core.cpp:1684-1700)."""
    idx = N.i32(graph_indices)
    dist = N.f32(graph_distances)
    lab = N.i32(labels)
    base = {"auto": 0, "zero": 1, "one": 2}[index_base]
    buf = N.CoreBuffers(lab.size, options.cycles)
    N.call("kodama_hip_core_knn_graph", N.ptr(idx), N.ptr(dist), C.c_int32(idx.shape[0]),
           C.c_int32(idx.shape[1]), C.c_int32(base), N.ptr(lab), N.ptr(N.i32(constrain)),
           N.ptr(N.i32(fix)), C.byref(options), C.byref(buf.struct))
    out = buf.result()
    out["backend"] = "hip"
    return out


def core_knn_batch(data, run_rows, initial_labels, options: CoreOptions, mode="plain", fold_ids=None):
    """Label evolution of M runs in one batch (the M-run scheduler of KODAMAMatrix).

    run_rows / initial_labels / fold_ids: lists of per-run int arrays. mode
    "plain" replays the reference random streams per run (bit-identical to
    core_knn on that run); "device" keeps proposals, votes and acceptance on the
    device. Returns clbest per run, accbest, vect_acc [M x cycles] and timings.
    """
    x = _matrix(data)
    M = len(run_rows)
    offsets = np.zeros(M + 1, np.int32)
    offsets[1:] = np.cumsum([len(r) for r in run_rows])
    rows = N.i32(np.concatenate([np.asarray(r) for r in run_rows]))
    init = N.i32(np.concatenate([np.asarray(l) for l in initial_labels]))
    folds = None if fold_ids is None else N.i32(np.concatenate([np.asarray(f) for f in fold_ids]))
    clbest = np.zeros(int(offsets[-1]), np.int32)
    accbest = np.zeros(M, np.float64)
    vect = np.zeros(M * max(options.cycles, 1), np.float64)
    timings = np.zeros(5, np.float64)
    m = {"plain": 0, "device": 1}[mode]
    N.call("kodama_hip_core_knn_batch", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), C.c_int32(M),
           N.ptr(offsets), N.ptr(rows), N.ptr(init), N.ptr(folds), C.byref(options), C.c_int32(m),
           N.ptr(clbest), N.ptr(accbest), N.ptr(vect), N.ptr(timings))
    return dict(clbest=[clbest[offsets[r]:offsets[r + 1]].copy() for r in range(M)], accbest=accbest,
                vect_acc=vect[: M * options.cycles].reshape(M, options.cycles),
                precompute_seconds=float(timings[0]), evolution_seconds=float(timings[1]),
                steps=int(timings[2]), vote_seconds=float(timings[3]),
                evolution_seconds_with_copies=float(timings[4]))


def core_plslda_batch(data, run_rows, initial_labels, options: CoreOptions, fold_ids=None):
    """CorePLSLDA over M runs in one batch: host proposals with the reference
    random streams per run (seed + r, folds from cv_seed + r), one device
    evaluation of all runs x folds per cycle. Per run the result equals
    core_plslda on that run's rows."""
    x = _matrix(data)
    M = len(run_rows)
    offsets = np.zeros(M + 1, np.int32)
    offsets[1:] = np.cumsum([len(r) for r in run_rows])
    rows = N.i32(np.concatenate([np.asarray(r) for r in run_rows]))
    init = N.i32(np.concatenate([np.asarray(l) for l in initial_labels]))
    folds = None if fold_ids is None else N.i32(np.concatenate([np.asarray(f) for f in fold_ids]))
    clbest = np.zeros(int(offsets[-1]), np.int32)
    accbest = np.zeros(M, np.float64)
    vect = np.zeros(M * max(options.cycles, 1), np.float64)
    timings = np.zeros(5, np.float64)
    N.call("kodama_hip_core_plslda_batch", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), C.c_int32(M),
           N.ptr(offsets), N.ptr(rows), N.ptr(init), N.ptr(folds), C.byref(options), N.ptr(clbest),
           N.ptr(accbest), N.ptr(vect), N.ptr(timings))
    return dict(clbest=[clbest[offsets[r]:offsets[r + 1]].copy() for r in range(M)], accbest=accbest,
                vect_acc=vect[: M * options.cycles].reshape(M, options.cycles),
                precompute_seconds=float(timings[0]), evolution_seconds=float(timings[1]),
                steps=int(timings[2]), eval_kernel_seconds=float(timings[3]))


# ---- learned proposal ---------------------------------------------------------------

MLP_PARAMS = 10400


def mlp_init(seed=1) -> np.ndarray:
    """He-normal weights, zero biases (fp32 [10400])."""
    out = np.zeros(MLP_PARAMS, np.float32)
    N.call("kodama_hip_mlp_init", C.c_uint64(int(seed)), N.ptr(out))
    return out


def mlp_features(neighbors, labels, gpu_device=0):
    """Device features [rows x 64] and ranked voted codes [rows x 30] of a table."""
    nbr = N.i32(neighbors)
    lab = N.i32(labels)
    rows, k = nbr.shape
    feat = np.zeros((rows, 64), np.float32)
    ranked = np.zeros((rows, 30), np.int32)
    N.call("kodama_hip_mlp_features", N.ptr(nbr), N.ptr(lab), C.c_int32(rows), C.c_int32(k), C.c_int32(gpu_device),
           N.ptr(feat), N.ptr(ranked))
    return feat, ranked


def mlp_forward(params, features, gpu_device=0) -> np.ndarray:
    par = N.f32(params)
    feat = N.f32(features)
    probs = np.zeros((feat.shape[0], 32), np.float32)
    N.call("kodama_hip_mlp_forward", N.ptr(par), N.ptr(feat), C.c_int32(feat.shape[0]), C.c_int32(gpu_device),
           N.ptr(probs))
    return probs


def mlp_gradients(params, features, targets, gpu_device=0):
    """Mean cross-entropy gradient [10400] and loss."""
    par = N.f32(params)
    feat = N.f32(features)
    tgt = N.i32(targets)
    grads = np.zeros(MLP_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    N.call("kodama_hip_mlp_gradients", N.ptr(par), N.ptr(feat), N.ptr(tgt), C.c_int32(feat.shape[0]),
           C.c_int32(gpu_device), N.ptr(grads), N.ptr(loss))
    return grads, float(loss[0])


def _batch_args(data, run_rows, initial_labels, fold_ids):
    x = _matrix(data)
    M = len(run_rows)
    offsets = np.zeros(M + 1, np.int32)
    offsets[1:] = np.cumsum([len(r) for r in run_rows])
    rows = N.i32(np.concatenate([np.asarray(r) for r in run_rows]))
    init = N.i32(np.concatenate([np.asarray(l) for l in initial_labels]))
    folds = None if fold_ids is None else N.i32(np.concatenate([np.asarray(f) for f in fold_ids]))
    return x, M, offsets, rows, init, folds


def mlp_collect_traces(data, run_rows, initial_labels, options: CoreOptions, fold_ids=None, max_samples=200000,
                       targets="accept"):
    """Samples of the plain batched evolution: features [S x 64], targets [S].
    targets "accept": the rows a proposal relabels, target the rank of the new
    label when the proposal was accepted, 30 when rejected; "outcome": every
    row whose label differs from its prediction at every cycle, target the
    rank of its label in the run's final best labelling (30: unchanged)."""
    x, M, offsets, rows, init, folds = _batch_args(data, run_rows, initial_labels, fold_ids)
    fn = {"accept": "kodama_hip_mlp_collect_traces", "outcome": "kodama_hip_mlp_collect_outcomes"}[targets]
    feat = np.zeros((max_samples, 64), np.float32)
    tgt = np.zeros(max_samples, np.int32)
    count = C.c_int32(0)
    acc = np.zeros(M, np.float64)
    N.call(fn, N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), C.c_int32(M),
           N.ptr(offsets), N.ptr(rows), N.ptr(init), N.ptr(folds), C.byref(options), C.c_int32(max_samples),
           N.ptr(feat), N.ptr(tgt), C.byref(count), N.ptr(acc))
    c = count.value
    return feat[:c].copy(), tgt[:c].copy(), acc


def mlp_train(features, targets, params=None, epochs=10, batch=256, learning_rate=0.05, momentum=0.9, seed=1,
              gpu_device=0):
    """Device SGD; returns (params, mean loss per epoch)."""
    feat = N.f32(features)
    tgt = N.i32(targets)
    par = np.array(mlp_init(seed) if params is None else params, dtype=np.float32, copy=True)
    loss = np.zeros(max(int(epochs), 1), np.float64)
    N.call("kodama_hip_mlp_train", N.ptr(feat), N.ptr(tgt), C.c_int32(feat.shape[0]), C.c_int32(int(epochs)),
           C.c_int32(int(batch)), C.c_float(learning_rate), C.c_float(momentum), C.c_uint64(int(seed)),
           C.c_int32(gpu_device), N.ptr(par), N.ptr(loss))
    return par, loss[: int(epochs)]


def core_knn_batch_learned(data, run_rows, initial_labels, options: CoreOptions, params, fallback=0.25,
                           fold_ids=None):
    """core_knn_batch in the device-resident form with the learned proposal."""
    x, M, offsets, rows, init, folds = _batch_args(data, run_rows, initial_labels, fold_ids)
    par = N.f32(params)
    clbest = np.zeros(int(offsets[-1]), np.int32)
    accbest = np.zeros(M, np.float64)
    vect = np.zeros(M * max(options.cycles, 1), np.float64)
    timings = np.zeros(5, np.float64)
    N.call("kodama_hip_core_knn_batch_learned", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), C.c_int32(M),
           N.ptr(offsets), N.ptr(rows), N.ptr(init), N.ptr(folds), C.byref(options), N.ptr(par),
           C.c_float(fallback), N.ptr(clbest), N.ptr(accbest), N.ptr(vect), N.ptr(timings))
    return dict(clbest=[clbest[offsets[r]:offsets[r + 1]].copy() for r in range(M)], accbest=accbest,
                vect_acc=vect[: M * options.cycles].reshape(M, options.cycles),
                precompute_seconds=float(timings[0]), evolution_seconds=float(timings[1]),
                steps=int(timings[2]), device_seconds=float(timings[3]),
                evolution_seconds_with_copies=float(timings[4]))


# ---- KODAMA.matrix ----------------------------------------------------------------

def _matrix_call(x, graph, index_base, M, Tcycle, landmarks, splitting, folds, k, metric, graph_neighbors, seed,
                 form, gpu_device, n_threads, mlp_params, fallback, classifier, components, W, constrain, fixed):
    n = x.shape[0]
    if form == "learned" and mlp_params is None:
        raise ValueError('form "learned" needs mlp_params')
    if classifier not in ("knn", "pls_lda"):
        raise ValueError('classifier must be "knn" or "pls_lda"')
    par = None if mlp_params is None else np.ascontiguousarray(mlp_params, np.float32)
    extra = []
    for name, a in (("W", W), ("constrain", constrain), ("fixed", fixed)):
        a = N.i32(a)
        if a is not None and a.shape != (n,):
            raise ValueError(f"{name} must hold one value per row")
        extra.append(a)
    o = N.MatrixOptions(runs=int(M), cycles=int(Tcycle), landmarks=int(landmarks), splitting=int(splitting),
                        folds=int(folds), k=int(k), metric=_metric(metric), graph_neighbors=int(graph_neighbors),
                        seed=int(seed), form={"plain": 0, "device": 1, "learned": 2}[form], device=int(gpu_device),
                        n_threads=int(n_threads), mlp_params=N.ptr(par) if form == "learned" else None,
                        mlp_fallback=float(fallback), classifier=1 if classifier == "knn" else 0,
                        components=int(components))
    L = C.c_int32(0)
    Kr = C.c_int32(0)
    N.call("kodama_hip_matrix_sizes", C.c_int32(n), C.byref(o), C.byref(L), C.byref(Kr))
    K = Kr.value if graph is None else max(1, min(int(Kr.value), graph[0].shape[1]))
    res = np.zeros((M, n), np.int32)
    acc = np.zeros(M, np.float64)
    v = np.zeros((M, max(int(Tcycle), 1)), np.float64)
    land = np.zeros((M, L.value), np.int32)
    gi = np.zeros((n, K), np.int32)
    gd = np.zeros((n, K), np.float32)
    timings = np.zeros(6, np.float64)
    ptrs = [N.ptr(a) for a in extra]
    outs = (N.ptr(res), N.ptr(acc), N.ptr(v), N.ptr(land), N.ptr(gi), N.ptr(gd), N.ptr(timings))
    if graph is None:
        N.call("kodama_hip_kodama_matrix", N.ptr(x), C.c_int32(n), C.c_int32(x.shape[1]), C.byref(o), *ptrs, *outs)
    else:
        base = -1 if index_base == "auto" else int(index_base)
        N.call("kodama_hip_kodama_matrix_from_graph", N.ptr(x), C.c_int32(n), C.c_int32(x.shape[1]),
               N.ptr(graph[0]), N.ptr(graph[1]), C.c_int32(graph[0].shape[1]), C.c_int32(base), C.byref(o), *ptrs,
               *outs)
    return dict(res=res, acc=acc, v=v[:, : int(Tcycle)].copy(), landmarks=land, knn_indices=gi, knn_distances=gd,
                timings=dict(zip(("atlas", "graph", "landmarks", "core_precompute", "core_evolution",
                                  "projection"), timings.tolist())))


def KODAMA_matrix(data, M=100, Tcycle=20, landmarks=10000, splitting=0, folds=5, k=30,
                  metric="euclidean", graph_neighbors=100, seed=1234, form="plain", gpu_device=0,
                  n_threads=0, dissimilarity=False, mlp_params=None, fallback=0.25, classifier="knn",
                  components=50, W=None, constrain=None, fixed=None) -> dict:
    """KODAMA.matrix on the HIP backend (KODAMAMatrix_CUDA contract;
    kodama_matrix.cpp:3324-4051): shared k-means landmark atlas, per-run quota
    landmarks and k-means starting labels (or ``W``, the reference's
    starting_labels, which make the search greedy), ``constrain`` groups and
    ``fixed`` rows as the reference's, the Core on the landmarks for all M
    runs in one batch, projection of the other rows.

    classifier "knn": CoreKNN (form "plain" replays the reference streams,
    "device" keeps the evolution on the GPU, "learned" adds the learned
    proposal of ``mlp_params`` with ``fallback`` uniform cycles), projection
    through the exact global graph. classifier "pls_lda": CorePLSLDA with
    ``components`` components (plain form), projection by PLSLDAPredict.

    Returns res [M x n], acc [M], v [M x Tcycle], the landmark rows [M x L], the
    1-based global graph (indices, distances) and per-stage timings; with
    ``dissimilarity`` the graph is rewritten by the KODAMA dissimilarity.
    """
    x = _matrix(data)
    out = _matrix_call(x, None, "auto", M, Tcycle, landmarks, splitting, folds, k, metric, graph_neighbors, seed,
                       form, gpu_device, n_threads, mlp_params, fallback, classifier, components, W, constrain,
                       fixed)
    if dissimilarity:
        out["knn_indices"], out["knn_distances"] = KODAMA_dissimilarity(out["knn_indices"], out["knn_distances"],
                                                                        out["res"], gpu_device)
    return out


def KODAMA_matrix_from_graph(data, indices, distances, index_base="auto", M=100, Tcycle=20, landmarks=10000,
                             splitting=0, folds=5, k=30, metric="euclidean", graph_neighbors=100, seed=1234,
                             form="plain", gpu_device=0, n_threads=0, mlp_params=None, fallback=0.25,
                             classifier="knn", components=50, W=None, constrain=None, fixed=None) -> dict:
    """KODAMAMatrixFromGraphData: KODAMA_matrix with the caller's neighbour graph
    (indices / distances [n x neighbours], index_base "auto", 0 or 1) in place
    of the exact global graph. With the kNN classifier the graph drives the
    Core too: each run's landmark subgraph feeds CoreKNNGraph
    (kodama_matrix.cpp:3178-3192). Returns the same dictionary (graph outputs
    [n x K], K = min(neighbours, resolved width))."""
    x = _matrix(data)
    gi_in = N.i32(np.ascontiguousarray(indices))
    gd_in = np.ascontiguousarray(distances, np.float32)
    if gi_in.shape != gd_in.shape or gi_in.shape[0] != x.shape[0]:
        raise ValueError("indices and distances must both be [n x neighbours]")
    return _matrix_call(x, (gi_in, gd_in), index_base, M, Tcycle, landmarks, splitting, folds, k, metric,
                        graph_neighbors, seed, form, gpu_device, n_threads, mlp_params, fallback, classifier,
                        components, W, constrain, fixed)


# ---- graph and dissimilarity ----------------------------------------------------

def KODAMA_knn_graph(data, k=30, metric="euclidean", gpu_device=0):
    """Exact self-kNN graph (KODAMAKNNGraph_CUDA contract): 1-based ids, public distances."""
    x = _matrix(data)
    kk = max(1, min(int(k), x.shape[0] - 1))
    idx = np.zeros((x.shape[0], kk), np.int32)
    dist = np.zeros((x.shape[0], kk), np.float32)
    N.call("kodama_hip_knn_graph", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), C.c_int32(k),
           C.c_int32(_metric(metric)), C.c_int32(gpu_device), N.ptr(idx), N.ptr(dist))
    return dict(indices=idx, distances=dist, neighbors=kk, index_base="one")


def last_knn_fallbacks() -> int:
    """Queries of this thread's last exact kNN search / graph that the MFMA
    candidate path handed to the exact scan; -1 when the exact scan ran alone."""
    return int(N.lib().kodama_hip_last_knn_fallbacks())


_TABLE_PATHS = {"tuned": 0, "scan": 1, "mfma": 2}


def knn_fold_tables(data, run_rows, fold_ids, k, metric="euclidean", path="tuned", gpu_device=0):
    """Per-fold kNN tables of a batch of runs (precompute_knn_cv, core.cpp:694-782,
    for all runs at once). run_rows / fold_ids: one array per run. Returns
    (neighbours, scores, path used): lists of [L_r x k] run-local ids (-1
    missing) and scores (minus the internal distance)."""
    x = _matrix(data)
    M = len(run_rows)
    offsets = np.zeros(M + 1, np.int32)
    offsets[1:] = np.cumsum([len(r) for r in run_rows])
    rows = N.i32(np.concatenate([np.asarray(r) for r in run_rows]))
    folds = N.i32(np.concatenate([np.asarray(f) for f in fold_ids]))
    total = int(offsets[-1])
    nbr = np.zeros((total, k), np.int32)
    score = np.zeros((total, k), np.float32)
    used = C.c_int32(0)
    N.call("kodama_hip_knn_fold_tables", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), C.c_int32(M),
           N.ptr(offsets), N.ptr(rows), N.ptr(folds), C.c_int32(k), C.c_int32(_metric(metric)),
           C.c_int32(_TABLE_PATHS[path]), C.c_int32(gpu_device), N.ptr(nbr), N.ptr(score), C.byref(used))
    split = lambda a: [a[offsets[r]:offsets[r + 1]] for r in range(M)]
    return split(nbr), split(score), {1: "scan", 2: "mfma"}[used.value]


def set_fold_table_path(path="tuned"):
    """Process-wide path of the per-fold tables ("tuned", "scan" or "mfma")."""
    N.call("kodama_hip_set_fold_table_path", C.c_int32(_TABLE_PATHS[path]))


def set_fold_table_schedule(path="tuned", list=0, waves=0):
    """Process-wide schedule of the per-fold tables: path ("tuned", "scan",
    "mfma"), and for the MFMA path the candidate list length (8 or 16) and the
    waves per workgroup (4 or 8); 0 = tuned. Tables are identical for all."""
    N.call("kodama_hip_set_fold_table_schedule", C.c_int32(_TABLE_PATHS[path]), C.c_int32(int(list)),
           C.c_int32(int(waves)))


def set_mfma_packed_lists(on=True):
    """Process-wide form of the MFMA kNN candidate lists: packed 32-bit keys
    (value order bits and column index; the default, where the train rows
    number at most 32768) or separate values and indices. Same neighbours."""
    N.call("kodama_hip_set_mfma_packed_lists", C.c_int32(1 if on else 0))


def last_fold_table_schedule():
    """Schedule of this thread's last per-fold table build: dict(path, list, waves)."""
    p, l, w = C.c_int32(0), C.c_int32(0), C.c_int32(0)
    N.call("kodama_hip_last_fold_table_schedule", C.byref(p), C.byref(l), C.byref(w))
    return dict(path={0: None, 1: "scan", 2: "mfma"}[p.value], list=l.value, waves=w.value)


def last_fold_table_path():
    """Path of this thread's last per-fold table build ("scan", "mfma"; None before any)."""
    return {0: None, 1: "scan", 2: "mfma"}[int(N.lib().kodama_hip_last_fold_table_path())]


def set_evolution_split(max_group=0):
    """Process-wide workgroups per run of the device-resident evolution: 0 =
    automatic (large runs over up to 8 workgroups), 1 = one per run, 2-8 = at
    most that many. Results are identical for every setting."""
    N.call("kodama_hip_set_evolution_split", C.c_int32(int(max_group)))


def set_cooperative_launch(cooperative=True):
    """Split runs launch cooperatively (default) or as an ordinary grid, for
    profilers that cannot follow cooperative launches."""
    N.call("kodama_hip_set_cooperative_launch", C.c_int32(1 if cooperative else 0))


def last_evolution_workgroups():
    """Workgroups of this thread's last device-resident evolution (0 before any)."""
    return int(N.lib().kodama_hip_last_evolution_workgroups())


def set_lds_flow_table(enabled=True):
    """Process-wide: the device-resident evolution counts flows in an LDS
    table when it fits (default) or in a per-workgroup HBM table. Results are
    identical either way."""
    N.call("kodama_hip_set_lds_flow_table", C.c_int32(1 if enabled else 0))


def last_evolution_lds_flows():
    """Whether this thread's last device-resident evolution counted flows in LDS."""
    return bool(N.lib().kodama_hip_last_evolution_lds_flows())


class KODAMAGraphHandle:
    """KODAMAGraphHandle on the device: the exact self-kNN graph built once and
    kept on the GPU; materialize() returns it 1-based with public distances,
    dissimilarity(res) applies the KODAMA dissimilarity in place there."""

    def __init__(self, data, k=100, metric="euclidean", gpu_device=0):
        x = _matrix(data)
        self._h = C.c_void_p()
        N.call("kodama_hip_graph_build", N.ptr(x), C.c_int32(x.shape[0]), C.c_int32(x.shape[1]), C.c_int32(int(k)),
               C.c_int32(_metric(metric)), C.c_int32(gpu_device), C.byref(self._h))
        s, nb, dev = C.c_int32(), C.c_int32(), C.c_int32()
        N.call("kodama_hip_graph_info", self._h, C.byref(s), C.byref(nb), C.byref(dev))
        self.samples, self.neighbors, self.device = s.value, nb.value, dev.value

    def valid(self) -> bool:
        return bool(self._h)

    def materialize(self):
        self._check()
        idx = np.zeros((self.samples, self.neighbors), np.int32)
        dist = np.zeros((self.samples, self.neighbors), np.float32)
        N.call("kodama_hip_graph_materialize", self._h, N.ptr(idx), N.ptr(dist))
        return idx, dist

    def dissimilarity(self, res):
        self._check()
        r = N.i32(res)
        if r.ndim != 2 or r.shape[1] != self.samples:
            raise ValueError("res must be [runs x samples]")
        N.call("kodama_hip_graph_dissimilarity", self._h, N.ptr(r), C.c_int32(r.shape[0]))

    def release(self):
        if self._h:
            fn = N.lib().kodama_hip_graph_release
            fn.restype = None
            fn(self._h)
            self._h = C.c_void_p()

    def _check(self):
        if not self._h:
            raise ValueError("the graph handle was released")

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


def KODAMA_dissimilarity(indices, distances, res, gpu_device=0):
    """KODAMADissimilarityInPlace on a 1-based graph; returns corrected copies."""
    idx = np.array(indices, dtype=np.int32, order="C", copy=True)
    dist = np.array(distances, dtype=np.float32, order="C", copy=True)
    r = N.i32(res)
    N.call("kodama_hip_dissimilarity_inplace", N.ptr(idx), N.ptr(dist), C.c_int32(idx.shape[0]),
           C.c_int32(idx.shape[1]), N.ptr(r), C.c_int32(r.shape[0]), C.c_int32(gpu_device))
    return idx, dist
