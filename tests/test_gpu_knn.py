"""GPU parity of the kNN half of the hot path against the oracle and the
reference's golden vectors (bit-exact: integer, index and float32 distance work)."""
import numpy as np
import pytest

import kodama_amd as K
from kodama_amd import datasets as D

from _cases import CORE_API, MATRIX_RUN, assert_same_core, check_invariants, golden_core
from conftest import golden

pytestmark = pytest.mark.gpu


def _oopt(O, **kw):
    import oracle
    return oracle.core_options(**kw)


@pytest.mark.parametrize("p", [7, 32, 50, 100])
@pytest.mark.parametrize("k", [1, 10, 30])
@pytest.mark.parametrize("metric", ["euclidean", "inner_product"])
def test_exact_knn_matches_oracle(oracle, p, k, metric):
    rng = np.random.default_rng(p * 100 + k)
    train = rng.normal(size=(700, p)).astype(np.float32)
    query = rng.normal(size=(300, p)).astype(np.float32)
    m = K._native.METRICS[metric]
    idx, dist = K.exact_knn_search(train, query, k, metric)
    oi, od = oracle.exact_knn(train, query, k, m)
    assert np.array_equal(idx, oi)
    assert np.array_equal(dist, od)


def test_exact_knn_matches_reference_golden():
    g = golden("knn.npz")
    x, _, _ = D.config1(n=400)
    train, query = x[:320], x[320:]
    for m, name in ((2, "euclidean"), (1, "inner_product")):
        idx, dist = K.exact_knn_search(train, query, 16, name)
        assert np.array_equal(idx, g[f"idx_m{m}"])
        assert np.array_equal(dist, g[f"dist_m{m}"])
    dup = np.repeat(x[:40], 3, axis=0)
    idx, dist = K.exact_knn_search(dup, x[:10], 8, "euclidean")
    assert np.array_equal(idx, g["dup_idx"]) and np.array_equal(dist, g["dup_dist"])


def test_exact_knn_exclusion_and_padding(oracle):
    rng = np.random.default_rng(5)
    x = rng.normal(size=(20, 50)).astype(np.float32)
    ex = np.arange(20, dtype=np.int32)
    idx, dist = K.exact_knn_search(x, x, 25, "euclidean", exclude=ex)
    oi, od = oracle.exact_knn(x, x, 25, 2, exclude=ex)
    assert np.array_equal(idx, oi) and np.array_equal(dist, od)
    assert np.all(idx[:, 19:] == -1) and np.all(np.isinf(dist[:, 19:]))
    assert not np.any(idx[:, :19] == np.arange(20)[:, None])


@pytest.mark.parametrize("metric", ["euclidean", "inner_product"])
def test_sliced_scan_merge_with_ties_across_slices(oracle, metric):
    """Few queries against many train rows: the scan cuts the train rows into
    slices and merges their lists. 40 distinct rows repeated 75 times, copies
    spread over every slice, so equal distances meet in the merge and only the
    index breaks the tie; k = 64 fills each slice list."""
    rng = np.random.default_rng(11)
    base = rng.normal(size=(40, 90)).astype(np.float32)
    train = np.tile(base, (75, 1))
    query = np.concatenate([base[:20], rng.normal(size=(13, 90)).astype(np.float32)])
    ex = (np.arange(33, dtype=np.int32) * 37) % train.shape[0]
    idx, dist = K.exact_knn_search(train, query, 64, metric, exclude=ex)
    oi, od = oracle.exact_knn(train, query, 64, K._native.METRICS[metric], exclude=ex)
    assert np.array_equal(idx, oi) and np.array_equal(dist, od)


@pytest.mark.parametrize("queries", [40, 1500])
def test_exact_knn_beyond_64_neighbours(oracle, queries):
    """k = 100 (the default global-graph width of KODAMA.matrix): the
    query-vs-train scan keeps up to 128 neighbours for up to 128 dimensions;
    40 queries take the sliced scan with its merge, 1500 the plain one."""
    rng = np.random.default_rng(queries)
    train = rng.normal(size=(2600, 50)).astype(np.float32)
    train[1300:1400] = train[:100]  # exact duplicates: ties broken by index
    query = rng.normal(size=(queries, 50)).astype(np.float32)
    for metric in ("euclidean", "inner_product"):
        idx, dist = K.exact_knn_search(train, query, 100, metric)
        oi, od = oracle.exact_knn(train, query, 100, K._native.METRICS[metric])
        assert np.array_equal(idx, oi) and np.array_equal(dist, od)


def test_graph_with_100_neighbours(oracle):
    x, _, _ = D.config1(n=1800, seed=23)
    g = K.KODAMA_knn_graph(x, k=100)
    idx, dist = oracle.knn_graph(x, 100, 2)
    assert np.array_equal(g["indices"], idx) and np.array_equal(g["distances"], dist)


def test_vote_graph_matches_oracle_with_ties(oracle):
    rng = np.random.default_rng(3)
    n, k = 900, 12
    x = rng.normal(size=(n, 50)).astype(np.float32)
    folds = rng.integers(0, 10, size=n).astype(np.int32)
    nbr, score = oracle.precompute_knn_cv(x, folds, k)
    # Round scores to a coarse grid so that count and score-sum ties happen.
    score = np.round(score).astype(np.float32)
    g = K.KnnVoteGraph(nbr, score)
    for trial in range(4):
        labels = rng.integers(0, 3 + trial, size=n).astype(np.int32)
        ref = oracle.knn_vote(nbr, score, folds, labels)
        got = g.predict(labels)
        assert np.array_equal(got, ref), f"trial {trial}"
    g.close()


@pytest.mark.parametrize("metric", [0, 2])
def test_knncv_matches_reference_golden(metric):
    g = golden("knncv_c1.npz")
    x, _, comp = D.config1()
    name = {0: "cosine", 2: "euclidean"}[metric]
    r = K.KNNCV(x, comp, k=10, folds=10, stratified=True, seed=3, metric=name)
    assert np.array_equal(r["fold_assignments"], g[f"folds_m{metric}"])
    assert np.array_equal(r["predicted_labels"], g[f"pred_m{metric}"])


@pytest.mark.parametrize("variant", ["api", "matrix"])
@pytest.mark.parametrize("run", [0, 1, 2])
def test_core_knn_trajectory_matches_reference_golden(variant, run):
    g = golden("core_knn_c1.npz")
    x, labels, _ = D.config1()
    flags = CORE_API if variant == "api" else MATRIX_RUN
    opt = K.core_options(cycles=20, k=10, folds=10, seed=1 + run, cv_seed=1 + run, **flags)
    r = K.core_knn(x, labels, opt)
    assert_same_core(r, golden_core(g, f"{variant}_r{run}_"), f"{variant} run {run}")
    check_invariants(r, 20, flags["evolutionary_search"] == 1)


CASES = {
    "greedy": dict(),
    "evolutionary": dict(evolutionary_search=1, guarded_diversity=1, many_to_one_absorption=1, stratified=0),
    "no_guidance": dict(prediction_guidance=0, evolutionary_search=1, guarded_diversity=1),
    "fixed_budget": dict(policy_adaptive_proposal_size=0, many_to_one_absorption=1),
    "greedy_accept": dict(evolutionary_search=1, stochastic_acceptance=0),
    "raw_score": dict(evolutionary_search=1, guarded_diversity=1, diversity_multiplier=0),
    "shake": dict(shake=1, evolutionary_search=1),
    "cosine": dict(metric=0, evolutionary_search=1),
    "inner": dict(metric=1, guarded_diversity=1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_core_knn_matches_oracle(oracle, case):
    import oracle as O
    x, _, comp = D.config1(n=500, seed=9)
    labels = D.randint(4, 1, 500, 1, 9).astype(np.int32)
    kw = dict(cycles=15, k=10, folds=5, seed=7, cv_seed=8)
    kw.update(CASES[case])
    r = K.core_knn(x, labels, K.core_options(**kw))
    o = oracle.core_knn(x, labels, O.core_options(**kw))
    assert_same_core(r, o, case)


def test_core_knn_groups_and_fixed_rows(oracle):
    import oracle as O
    x, _, _ = D.config1(n=360, seed=4)
    labels = D.randint(6, 1, 360, 1, 6).astype(np.int32)
    groups = (np.arange(360) // 4).astype(np.int32)
    fixed = (D.uniform(6, 2, 360) < 0.2).astype(np.int32)
    kw = dict(cycles=12, k=7, folds=4, seed=3, cv_seed=3, evolutionary_search=1, guarded_diversity=1,
              many_to_one_absorption=1, stratified=1)
    r = K.core_knn(x, labels, K.core_options(**kw), constrain=groups, fix=fixed)
    o = oracle.core_knn(x, labels, O.core_options(**kw), constrain=groups, fixed=fixed)
    assert_same_core(r, o, "groups+fixed")
    assert np.array_equal(r["clbest"][fixed == 1], labels[fixed == 1])


def test_core_knn_k_larger_than_training_fold(oracle):
    import oracle as O
    x, _, _ = D.config1(n=40, seed=2)
    labels = D.randint(2, 3, 40, 1, 3).astype(np.int32)
    kw = dict(cycles=5, k=60, folds=2, seed=1, cv_seed=1)
    r = K.core_knn(x, labels, K.core_options(**kw))
    o = oracle.core_knn(x, labels, O.core_options(**kw))
    assert_same_core(r, o, "k>train")


def test_core_knn_graph_matches_oracle_tables(oracle):
    """CoreKNNGraph: same-fold neighbours dropped, score -d^2 (core.cpp:624-692)."""
    import oracle as O
    x, labels, _ = D.config1(n=400, seed=6)
    idx1, dist = oracle.knn_graph(x, 25)           # 1-based public graph
    kw = dict(cycles=10, k=10, folds=5, seed=2, cv_seed=2, evolutionary_search=1, guarded_diversity=1)
    r = K.CoreKNNGraph(idx1, dist, labels, K.core_options(**kw))
    folds = r["fold_assignments"]
    n = x.shape[0]
    nbr = -np.ones((n, 10), np.int32)
    score = np.full((n, 10), -np.inf, np.float32)
    for i in range(n):
        c = 0
        for j in range(idx1.shape[1]):
            nb = idx1[i, j] - 1
            if nb < 0 or nb == i or folds[nb] == folds[i] or not np.isfinite(dist[i, j]):
                continue
            d = np.float32(max(np.float32(0), dist[i, j]))
            nbr[i, c], score[i, c] = nb, -(d * d)
            c += 1
            if c == 10:
                break
    o = oracle.core_knn_tables(nbr, score, folds, labels, O.core_options(**kw))
    assert_same_core(r, o, "graph")


def test_batch_plain_equals_independent_runs():
    x, labels, _ = D.config1(n=1500, seed=12)
    rows = [np.sort(np.random.default_rng(r).choice(1500, size=s, replace=False)).astype(np.int32)
            for r, s in enumerate((300, 517, 256))]
    init = [labels[rw] for rw in rows]
    base = dict(cycles=10, k=10, folds=10, seed=40, cv_seed=90, evolutionary_search=1, guarded_diversity=1,
                many_to_one_absorption=1, stratified=0)
    b = K.core_knn_batch(x, rows, init, K.core_options(**base), mode="plain")
    for r in range(3):
        kw = dict(base, seed=40 + r, cv_seed=90 + r)
        single = K.core_knn(x[rows[r]], init[r], K.core_options(**kw))
        assert np.array_equal(b["clbest"][r], single["clbest"]), r
        assert b["accbest"][r] == single["accbest"]
        assert np.array_equal(b["vect_acc"][r], single["vect_acc"])
