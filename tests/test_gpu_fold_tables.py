"""Per-fold kNN tables of a batch of runs (precompute_knn_cv, core.cpp:694-782):
the register scan and the split-BF16 MFMA candidates with the exact rerank
must both give the oracle's tables bit-exactly (run-local ids in (distance,
id) order, float32 scores), and the tuned choice must not change results."""
import numpy as np
import pytest

import kodama_amd as K
from kodama_amd import datasets as D

pytestmark = pytest.mark.gpu


def _runs(x, sizes, folds, seed, dup=False):
    """Landmark subsets of x with uneven folds; with dup, repeated rows (ties)."""
    rng = np.random.default_rng(seed)
    rows, fids = [], []
    for L in sizes:
        rw = np.sort(rng.choice(x.shape[0], size=L, replace=False)).astype(np.int32)
        if dup:
            rw[1::7] = rw[0::7][: rw[1::7].size]  # duplicated data rows: equal distances
        f = rng.integers(0, folds, size=L).astype(np.int32)
        f[:folds] = np.arange(folds)  # every fold present
        rows.append(rw)
        fids.append(f)
    return rows, fids


def _check(oracle, x, rows, fids, k, metric, paths=("scan", "mfma")):
    m = K._metric(metric)
    want = [oracle.precompute_knn_cv(x[rw], f, k, m) for rw, f in zip(rows, fids)]
    for path in paths:
        nbr, score, used = K.knn_fold_tables(x, rows, fids, k, metric, path=path)
        if path != "tuned":
            assert used == path
        for r, (wn, ws) in enumerate(want):
            assert np.array_equal(nbr[r], wn), (path, r)
            assert np.array_equal(score[r], ws), (path, r)


@pytest.mark.parametrize("p", [7, 50, 64])
@pytest.mark.parametrize("k", [5, 30, 40])
@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_fold_tables_match_oracle(oracle, p, k, metric):
    x, _ = D.config3(n=3000, p=p, seed=p)
    rows, fids = _runs(x, [700, 1300, 90, 400], folds=10, seed=k + p)
    _check(oracle, x, rows, fids, k, metric)


def test_fold_tables_ties_and_short_runs(oracle):
    """Duplicated rows (equal distances, smaller id first) and runs whose other
    folds hold fewer than k rows (-1 / -inf past the available rows)."""
    x, _ = D.config3(n=2000, p=50, seed=9)
    rows, fids = _runs(x, [600, 25, 12], folds=3, seed=5, dup=True)
    _check(oracle, x, rows, fids, 30, "euclidean")


def test_fold_tables_c3_shape(oracle):
    """Runs of 1000-5000 landmark rows on the 100k mixture with 10 uneven folds:
    the shape of the batched evolution, where the MFMA path is meant to win."""
    x, _ = D.config3(n=100_000, p=50)
    rng = np.random.default_rng(6)
    sizes = rng.integers(1000, 5001, size=3)
    rows, fids = _runs(x, sizes, folds=10, seed=7)
    _check(oracle, x, rows, fids, 30, "euclidean")


def test_tuned_path_is_stable_and_exact(oracle):
    x, _ = D.config3(n=5000, p=50, seed=4)
    rows, fids = _runs(x, [1500, 2500], folds=10, seed=3)
    first = K.knn_fold_tables(x, rows, fids, 30, "euclidean", path="tuned")
    second = K.knn_fold_tables(x, rows, fids, 30, "euclidean", path="tuned")
    assert first[2] == second[2]  # the bucket's choice is cached
    for a, b in zip(first[0], second[0]):
        assert np.array_equal(a, b)
    _check(oracle, x, rows, fids, 30, "euclidean", paths=("tuned",))


def test_batch_evolution_identical_under_both_paths():
    """The plain batched evolution reads the tables: its trajectories must not
    depend on how they were built."""
    x, _ = D.config3(n=6000, p=50, seed=12)
    rows, fids = _runs(x, [800, 1200, 500], folds=10, seed=2)
    init = [np.asarray(D.randint(12, r, L, 1, 8), np.int32) for r, L in enumerate([800, 1200, 500])]
    o = K.core_options(cycles=15, k=30, folds=10, seed=5, cv_seed=5, evolutionary_search=1,
                       guarded_diversity=1, many_to_one_absorption=1)
    out = {}
    try:
        for path in ("scan", "mfma"):
            K.set_fold_table_path(path)
            out[path] = K.core_knn_batch(x, rows, init, o, mode="plain", fold_ids=fids)
            assert K.last_fold_table_path() == path
    finally:
        K.set_fold_table_path("tuned")
    assert all(np.array_equal(a, b) for a, b in zip(out["scan"]["clbest"], out["mfma"]["clbest"]))
    assert np.array_equal(out["scan"]["vect_acc"], out["mfma"]["vect_acc"])


@pytest.mark.parametrize("list_,waves", [(8, 4), (8, 8), (16, 4), (16, 8)])
@pytest.mark.parametrize("k", [12, 30, 40])
def test_fold_table_schedules_match_oracle(oracle, list_, waves, k):
    """Every MFMA schedule the tuner chooses from (candidates per lane column,
    waves per workgroup) gives the oracle's tables; a list of 8 cannot serve
    k > 32, which then takes the scan."""
    x, _ = D.config3(n=5000, p=50, seed=21)
    rows, fids = _runs(x, [1500, 2400, 300], folds=10, seed=k + list_ + waves)
    want = [oracle.precompute_knn_cv(x[rw], f, k, K._metric("euclidean")) for rw, f in zip(rows, fids)]
    try:
        K.set_fold_table_schedule("tuned", list_, waves)
        nbr, score, used = K.knn_fold_tables(x, rows, fids, k, "euclidean", path="mfma")
        sched = K.last_fold_table_schedule()
    finally:
        K.set_fold_table_schedule("tuned", 0, 0)
    if list_ == 8 and k > 32:
        assert used == "scan" and sched["path"] == "scan"
    else:
        assert used == "mfma" and sched == dict(path="mfma", list=list_, waves=waves)
    for r, (wn, ws) in enumerate(want):
        assert np.array_equal(nbr[r], wn), r
        assert np.array_equal(score[r], ws), r


def test_tuner_times_schedules_and_keeps_its_choice(oracle):
    """With everything tuned, the first table of a shape bucket times the scan
    and the four MFMA schedules; the choice is kept for the bucket and the
    tables are the oracle's either way."""
    x, _ = D.config3(n=6000, p=40, seed=33)
    rows, fids = _runs(x, [1800, 2200], folds=8, seed=4)
    K.set_fold_table_schedule("tuned", 0, 0)
    first = K.knn_fold_tables(x, rows, fids, 20, "euclidean", path="tuned")
    s1 = K.last_fold_table_schedule()
    second = K.knn_fold_tables(x, rows, fids, 20, "euclidean", path="tuned")
    s2 = K.last_fold_table_schedule()
    assert s1 == s2 and s1["path"] in ("scan", "mfma")
    if s1["path"] == "mfma":
        assert s1["list"] in (8, 16) and s1["waves"] in (4, 8)
    want = [oracle.precompute_knn_cv(x[rw], f, 20, K._metric("euclidean")) for rw, f in zip(rows, fids)]
    for out in (first, second):
        for r, (wn, ws) in enumerate(want):
            assert np.array_equal(out[0][r], wn) and np.array_equal(out[1][r], ws)


def test_tuned_choice_is_kept_per_k_bucket(oracle):
    """A schedule tuned for k <= 32 (possibly 8 candidates per lane column) is
    not reused for a k above 32 on the same shape: each k bucket is tuned on
    its own and both tables are the oracle's."""
    x, _ = D.config3(n=6000, p=50, seed=47)
    rows, fids = _runs(x, [2600, 3100], folds=10, seed=9)
    K.set_fold_table_schedule("tuned", 0, 0)
    for k in (30, 40):
        nbr, score, used = K.knn_fold_tables(x, rows, fids, k, "euclidean", path="tuned")
        sched = K.last_fold_table_schedule()
        if sched["path"] == "mfma" and k > 32:
            assert sched["list"] == 16
        want = [oracle.precompute_knn_cv(x[rw], f, k, K._metric("euclidean")) for rw, f in zip(rows, fids)]
        for r, (wn, ws) in enumerate(want):
            assert np.array_equal(nbr[r], wn) and np.array_equal(score[r], ws), (k, r)
