"""KODAMA.matrix on the HIP backend against the reference's own CPU results
(fixtures) and the oracle's restatement of run_kodama_matrix / run_iteration:
the k-means atlas and starting labels (device k-means in the reference's float
order), the quota landmarks, the batched Core (kNN per-fold exact, kNN from a
caller graph, PLS-LDA), starting labels, constraint groups, fixed rows, and the
projection (graph majority, PLSLDAPredict). The plain form is bit-identical
(labels, accuracies, per-cycle accuracy, landmark rows); the device-resident
form shares the landmarks and starting labels and is held to the plain form's
accuracy."""
import numpy as np
import pytest

import kodama_amd as K
from _cases import MATRIX_CASES, matrix_case_inputs, matrix_case_options
from conftest import golden
from kodama_amd import datasets as D

pytestmark = pytest.mark.gpu

CASES = {
    "p10_euclidean": dict(n=1200, p=10, runs=3, cycles=8, landmarks=400, splitting=40, folds=5, k=10,
                          metric="euclidean", graph_neighbors=60, seed=7),
    "p50_cosine": dict(n=1500, p=50, runs=2, cycles=6, landmarks=500, splitting=30, folds=4, k=12,
                       metric="cosine", graph_neighbors=40, seed=1234),
    "default_graph_width": dict(n=2400, p=33, runs=2, cycles=5, landmarks=800, splitting=0, folds=5, k=30,
                                metric="euclidean", graph_neighbors=100, seed=3),
}


def _data(n, p, seed):
    x, _ = D.gaussian_mixture(n, p, 6, seed=seed)
    return x


@pytest.mark.parametrize("name", list(CASES))
def test_plain_matrix_matches_oracle(oracle, name):
    c = dict(CASES[name])
    x = _data(c.pop("n"), c.pop("p"), 5)
    metric = c.pop("metric")
    got = K.KODAMA_matrix(x, M=c["runs"], Tcycle=c["cycles"], landmarks=c["landmarks"], splitting=c["splitting"],
                          folds=c["folds"], k=c["k"], metric=metric, graph_neighbors=c["graph_neighbors"],
                          seed=c["seed"], form="plain")
    want = oracle.kodama_matrix(x, metric=K._native.METRICS[metric], **c)
    assert np.array_equal(got["landmarks"], want["landmarks"])
    assert np.array_equal(got["res"], want["res"])
    assert np.array_equal(got["acc"], want["acc"])
    assert np.array_equal(got["v"], want["v"])


def test_device_form_shares_landmarks_and_accuracy(oracle):
    c = dict(CASES["p10_euclidean"], runs=4, cycles=20)
    x = _data(c.pop("n"), c.pop("p"), 5)
    c.pop("metric")
    kw = dict(M=c["runs"], Tcycle=c["cycles"], landmarks=c["landmarks"], splitting=c["splitting"], folds=c["folds"],
              k=c["k"], graph_neighbors=c["graph_neighbors"], seed=c["seed"])
    plain = K.KODAMA_matrix(x, form="plain", **kw)
    dev = K.KODAMA_matrix(x, form="device", **kw)
    assert np.array_equal(plain["landmarks"], dev["landmarks"])
    assert np.array_equal(plain["knn_indices"], dev["knn_indices"])
    # same starting labels and CV evaluation, different proposal draws
    assert abs(dev["acc"].mean() - plain["acc"].mean()) < 0.05
    n = x.shape[0]
    for r in range(c["runs"]):
        assert dev["res"][r].min() >= 1 and np.unique(dev["res"][r]).size > 1
    assert dev["res"].shape == (c["runs"], n)


def test_learned_form_in_the_matrix():
    """form "learned": with every cycle falling back to the uniform proposal it
    is the device form exactly; with the network it keeps the landmarks, graph
    and the accuracy level of the device form."""
    c = dict(CASES["p10_euclidean"], runs=4, cycles=20)
    x = _data(c.pop("n"), c.pop("p"), 5)
    c.pop("metric")
    kw = dict(M=c["runs"], Tcycle=c["cycles"], landmarks=c["landmarks"], splitting=c["splitting"], folds=c["folds"],
              k=c["k"], graph_neighbors=c["graph_neighbors"], seed=c["seed"])
    dev = K.KODAMA_matrix(x, form="device", **kw)
    params = K.mlp_init(7)
    same = K.KODAMA_matrix(x, form="learned", mlp_params=params, fallback=1.0, **kw)
    assert np.array_equal(dev["res"], same["res"]) and np.array_equal(dev["v"], same["v"])
    lrn = K.KODAMA_matrix(x, form="learned", mlp_params=params, fallback=0.25, **kw)
    assert np.array_equal(dev["landmarks"], lrn["landmarks"])
    assert np.array_equal(dev["knn_indices"], lrn["knn_indices"])
    assert abs(lrn["acc"].mean() - dev["acc"].mean()) < 0.05
    assert np.all(np.diff(lrn["v"], axis=1) >= 0)


@pytest.mark.parametrize("name", list(MATRIX_CASES))
def test_matrix_cases_match_reference(oracle, name):
    """KODAMAMatrixFromGraphData with the kNN classifier (the caller's graph
    subset to each run's landmarks drives CoreKNNGraph) and KODAMA.matrix with
    the PLS-LDA classifier (PLSLDAPredict projection), with and without
    starting labels, constraint groups and fixed rows: res, acc and v equal
    the reference's own CPU results (tests/golden/matrix_runs.npz) and the
    oracle's (PLS-LDA: its FP64 form, as the device arithmetic)."""
    g = golden("matrix_runs.npz")
    x, graph, base, W, cons, fixed = matrix_case_inputs(name)
    o = matrix_case_options(name)
    metric = {0: "cosine", 1: "inner_product", 2: "euclidean"}[o["metric"]]
    kw = dict(M=o["runs"], Tcycle=o["cycles"], landmarks=o["landmarks"], splitting=o["splitting"], folds=o["folds"],
              k=o["k"], metric=metric, graph_neighbors=o["graph_neighbors"], seed=o["seed"],
              classifier=o["classifier"], components=o["components"], W=W, constrain=cons, fixed=fixed)
    if graph is None:
        got = K.KODAMA_matrix(x, **kw)
    else:
        got = K.KODAMA_matrix_from_graph(x, graph[0], graph[1], index_base="auto" if base < 0 else base, **kw)
    want = oracle.kodama_matrix(x, graph=graph, index_base=base, W=W, constrain=cons, fixed=fixed, precision=64, **o)
    for key in ("res", "acc", "v"):
        assert np.array_equal(got[key], g[f"{name}_{key}"]), (name, key, "reference")
        assert np.array_equal(got[key], want[key]), (name, key, "oracle")
    assert np.array_equal(got["landmarks"], want["landmarks"])


def test_matrix_from_graph_normalisation(oracle):
    """A narrower caller graph with gaps and non-finite distances: the
    normalised graph (first K valid neighbours, 1-based, caller distances)
    and the results equal the oracle's KODAMAMatrixFromGraphData."""
    c = dict(CASES["p10_euclidean"], runs=3, cycles=10)
    x = _data(c.pop("n"), c.pop("p"), 5)
    c.pop("metric")
    kw = dict(M=c["runs"], Tcycle=c["cycles"], landmarks=c["landmarks"], splitting=c["splitting"], folds=c["folds"],
              k=c["k"], graph_neighbors=c["graph_neighbors"], seed=c["seed"])
    gi, gd = oracle.knn_graph(x, c["graph_neighbors"], 2)
    narrow = gi[:, :7].copy()
    narrow[::5, 3] = -1  # missing entries: the row keeps its other neighbours in order
    ndist = gd[:, :7].copy()
    ndist[1::5, 2] = np.nan  # non-finite distances are dropped too
    nar = K.KODAMA_matrix_from_graph(x, narrow, ndist, index_base=1, **kw)
    gi_n, gd_n = nar["knn_indices"], nar["knn_distances"]
    assert gi_n.shape == (x.shape[0], 7)
    for rows, col in ((slice(0, None, 5), 3), (slice(1, None, 5), 2)):
        assert np.array_equal(gi_n[rows, col:6], narrow[rows, col + 1:7])
        assert np.all(gi_n[rows, 6] == -1) and np.all(np.isinf(gd_n[rows, 6]))
    want = oracle.kodama_matrix(x, metric=K._native.METRICS["euclidean"], graph=(narrow, ndist), index_base=1, **c)
    for key in ("res", "acc", "v", "landmarks"):
        assert np.array_equal(nar[key], want[key]), key
    oi, od = oracle.normalize_graph(narrow, ndist, 1, 7)
    assert np.array_equal(gi_n, np.where(oi >= 0, oi + 1, -1)) and np.array_equal(gd_n, od)
    zero = K.KODAMA_matrix_from_graph(x, np.where(narrow > 0, narrow - 1, narrow), ndist, index_base=0, **kw)
    for key in ("res", "acc", "v"):
        assert np.array_equal(zero[key], nar[key]), key


def test_matrix_device_forms_reject_unsupported_inputs():
    """The device-resident forms run the evolutionary kNN loop only: starting
    labels, constraint groups, fixed rows or the PLS-LDA classifier are EINVAL
    there (the plain form runs them)."""
    x = _data(600, 10, 5)
    kw = dict(M=2, Tcycle=3, landmarks=200, splitting=10, folds=5, k=10, graph_neighbors=30, seed=7, form="device")
    n = x.shape[0]
    for extra in (dict(W=np.ones(n, np.int32)), dict(constrain=np.arange(n) // 2), dict(fixed=np.ones(n, np.int32)),
                  dict(classifier="pls_lda")):
        with pytest.raises(K.InvalidArgument):
            K.KODAMA_matrix(x, **kw, **extra)


def test_matrix_graph_and_dissimilarity(oracle):
    c = dict(CASES["p10_euclidean"])
    x = _data(c["n"], c["p"], 5)
    out = K.KODAMA_matrix(x, M=2, Tcycle=4, landmarks=400, splitting=40, folds=5, k=10, graph_neighbors=60, seed=7,
                          dissimilarity=True)
    gi, gd = oracle.knn_graph(x, 60, 2)
    want_i, want_d = oracle.dissimilarity(gi, gd, out["res"])
    assert np.array_equal(out["knn_indices"], want_i) and np.array_equal(out["knn_distances"], want_d)
