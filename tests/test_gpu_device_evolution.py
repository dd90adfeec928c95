"""The device-resident evolution: same algorithm as the plain form, different
random draws. Checked by exact invariants and by accuracy against the plain form."""
import numpy as np
import pytest

import kodama_amd as K
from kodama_amd import datasets as D

pytestmark = pytest.mark.gpu


def _batch(n_runs=12, n=6000, seed=31):
    x, comp = D.config3(n=n, seed=seed)
    rows, init, folds = [], [], []
    for r in range(n_runs):
        L = int(D.randint(seed, 50 + r, 1, 300, 1200)[0])
        rw = np.sort(np.argsort(D.uniform(seed, 200 + r, n))[:L]).astype(np.int32)
        lab = D.randint(seed, 400 + r, L, 1, 40).astype(np.int32)
        f = (np.argsort(D.uniform(seed, 600 + r, L)) % 10).astype(np.int32)
        rows.append(rw)
        init.append(lab)
        folds.append(f)
    return x, rows, init, folds


OPTS = dict(cycles=25, k=15, folds=10, stratified=0, seed=3, cv_seed=3, evolutionary_search=1,
            guarded_diversity=1, many_to_one_absorption=1)


@pytest.mark.parametrize("k", [5, 15, 30, 40])
def test_device_accuracy_is_the_vote_of_the_returned_labels(oracle, k):
    x, rows, init, folds = _batch()
    opts = dict(OPTS, k=k)
    out = K.core_knn_batch(x, rows, init, K.core_options(**opts), mode="device", fold_ids=folds)
    for r in range(len(rows)):
        nbr, score = oracle.precompute_knn_cv(x[rows[r]], folds[r], k)
        cl = out["clbest"][r]
        pred = oracle.knn_vote(nbr, score, folds[r], cl)
        assert out["accbest"][r] == np.mean(pred == cl), r
        assert set(np.unique(cl)) <= set(np.unique(init[r]))
        assert out["vect_acc"][r][-1] == out["accbest"][r]


def _centre_batch(n_runs=24, n=6000, seed=31, centres=40):
    """Runs whose starting labels are the nearest of `centres` random rows (one
    k-means assignment step), as KODAMA starts from a clustering."""
    x, _ = D.config3(n=n, seed=seed)
    rows, init, folds = [], [], []
    for r in range(n_runs):
        L = int(D.randint(seed, 50 + r, 1, 300, 1200)[0])
        rw = np.sort(np.argsort(D.uniform(seed, 200 + r, n))[:L]).astype(np.int32)
        c = x[rw[np.argsort(D.uniform(seed, 800 + r, L))[:centres]]]
        idx, _ = K.exact_knn_search(c, x[rw], 1, "euclidean")
        rows.append(rw)
        init.append((idx[:, 0] + 1).astype(np.int32))
        folds.append((np.argsort(D.uniform(seed, 600 + r, L)) % 10).astype(np.int32))
    return x, rows, init, folds


def test_device_reaches_plain_accuracy():
    """Equal accuracy = the device form's mean best CV accuracy over 4 seeds x 24
    runs is no more than half a point below the plain form's. (From random
    starting labels single runs of either form scatter widely with the draws.)"""
    x, rows, init, folds = _centre_batch()
    means = {}
    for mode in ("plain", "device"):
        acc = []
        for seed in (1, 2, 3, 4):
            opts = dict(OPTS, seed=seed, cv_seed=seed)
            out = K.core_knn_batch(x, rows, init, K.core_options(**opts), mode=mode, fold_ids=folds)
            acc.append(np.mean(out["accbest"]))
        means[mode] = float(np.mean(acc))
    assert means["device"] >= means["plain"] - 0.005, means


def test_device_is_deterministic():
    x, rows, init, folds = _batch(n_runs=4)
    a = K.core_knn_batch(x, rows, init, K.core_options(**OPTS), mode="device", fold_ids=folds)
    b = K.core_knn_batch(x, rows, init, K.core_options(**OPTS), mode="device", fold_ids=folds)
    for r in range(4):
        assert np.array_equal(a["clbest"][r], b["clbest"][r])
    assert np.array_equal(a["accbest"], b["accbest"])


def test_device_runs_beyond_8191_rows(oracle):
    """Runs of 10,000 rows (the reference's default landmarks, kodama.hpp:536)
    and of 16,383 rows take the large-run layout (row queue and flow list in
    global scratch, two-pass absorption keys). The uniform and learned forms
    keep the invariants: the reported accuracy is the oracle's vote of the
    returned labels, labels stay within the starting classes, the history is
    monotone and ends at the reported value."""
    n = 40000
    x, _ = D.config3(n=n, seed=31)
    rows, init, folds = [], [], []
    for r, L in enumerate((10000, 16383)):
        rw = np.sort(np.argsort(D.uniform(31, 900 + r, n), kind="stable")[:L]).astype(np.int32)
        c = x[rw[np.argsort(D.uniform(31, 950 + r, L), kind="stable")[:60]]]
        idx, _ = K.exact_knn_search(c, x[rw], 1, "euclidean")
        rows.append(rw)
        init.append((idx[:, 0] + 1).astype(np.int32))
        folds.append((np.argsort(D.uniform(31, 980 + r, L), kind="stable") % 10).astype(np.int32))
    o = K.core_options(**dict(OPTS, cycles=12, k=30))
    dev = K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds)
    lrn = K.core_knn_batch_learned(x, rows[:1], init[:1], o, K.mlp_init(5), fallback=0.25, fold_ids=folds[:1])
    for out, count in ((dev, 2), (lrn, 1)):
        for r in range(count):
            nbr, score = oracle.precompute_knn_cv(x[rows[r]], folds[r], 30)
            cl = out["clbest"][r]
            assert out["accbest"][r] == np.mean(oracle.knn_vote(nbr, score, folds[r], cl) == cl), r
            assert set(np.unique(cl)) <= set(np.unique(init[r]))
            assert np.all(np.diff(out["vect_acc"][r]) >= 0)
            assert out["vect_acc"][r][-1] == out["accbest"][r]


def test_matrix_device_form_at_default_landmarks():
    """KODAMA_matrix(form="device") with the default landmarks = 10000 on
    20,000 rows: the same landmarks and graph as the plain form, labels in
    range, accuracy at the plain form's level."""
    x, _ = D.config3(n=20000, seed=41)
    kw = dict(M=2, Tcycle=6, folds=5, k=10, graph_neighbors=30, seed=5)
    plain = K.KODAMA_matrix(x, form="plain", **kw)
    dev = K.KODAMA_matrix(x, form="device", **kw)
    assert plain["landmarks"].shape == (2, 10000)
    assert np.array_equal(plain["landmarks"], dev["landmarks"])
    assert np.array_equal(plain["knn_indices"], dev["knn_indices"])
    assert dev["res"].shape == (2, 20000) and dev["res"].min() >= 1
    assert abs(dev["acc"].mean() - plain["acc"].mean()) < 0.05
