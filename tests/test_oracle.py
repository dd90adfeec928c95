"""The oracle (the CPU restatement in oracle/) pinned before it is trusted as the
checker of the HIP path:

* against the golden fixtures that the compiled reference core wrote
  (tests/golden/make_golden.py): folds, exhaustive-search neighbours, CoreKNN and
  CorePLSLDA trajectories, KNNCV / PLSLDACV predictions, dissimilarity rewrites;
* against hand-built cases and closed forms chosen so that a dropped term, a
  wrong tie order or a transposed operand fails (vote ties, the diversity
  objective, brute-force neighbours with exact integer distances, co-assignment
  counts, the two-class PLS weight and LDA discriminant);
* where oracle/_ref is built (the build container), against the reference core
  on fresh seeded inputs.

CPU only: nothing here touches a GPU.
"""
import os
import re
import sys

import numpy as np
import pytest

import oracle as O
from _cases import CORE_API, MATRIX_RUN, assert_same_core, check_invariants, golden_core
from conftest import golden
from kodama_amd import datasets as D


# ---- folds (common.cpp:108-214) --------------------------------------------------

def test_folds_match_golden(oracle):
    g = golden("folds.npz")
    labels, groups = g["labels"], g["groups"]
    keys = [k for k in g.files if k.startswith("f")]
    assert len(keys) == 24
    for key in keys:
        m = re.fullmatch(r"f(\d+)_s(\d)_seed(\d+)(_grouped)?", key)
        folds, strat, seed, grouped = int(m[1]), int(m[2]), int(m[3]), m[4] is not None
        got = oracle.make_folds(labels, folds, bool(strat), seed, groups if grouped else None)
        assert np.array_equal(got, g[key]), key


@pytest.mark.parametrize("folds", [2, 3, 10])
def test_fold_balance_invariants(oracle, folds):
    lab = D.randint(5, 1, 503, 1, 6).astype(np.int32)
    plain = oracle.make_folds(lab, folds, False, 9)
    sizes = np.bincount(plain, minlength=folds)
    assert sizes.max() - sizes.min() <= 1
    strat = oracle.make_folds(lab, folds, True, 9)
    for c in np.unique(lab):
        per = np.bincount(strat[lab == c], minlength=folds)
        assert per.max() - per.min() <= 1, c
    # grouped rows share their fold
    groups = (np.arange(lab.size) // 4).astype(np.int32)
    gf = oracle.make_folds(lab, folds, False, 9, groups)
    for gid in np.unique(groups):
        assert np.unique(gf[groups == gid]).size == 1


# ---- exact kNN (native_knn.cpp:121-123, 1005-1035) --------------------------------

def test_exact_knn_matches_reference_exhaustive_search(oracle):
    g = golden("knn.npz")
    x, _, _ = D.config1(n=400)
    train, query = x[:320], x[320:]
    for metric in (2, 1):
        idx, dist = oracle.exact_knn(train, query, 16, metric)
        assert np.array_equal(idx, g[f"idx_m{metric}"]), metric
        assert np.array_equal(dist, g[f"dist_m{metric}"]), metric
    dup = np.repeat(x[:40], 3, axis=0)
    idx, dist = oracle.exact_knn(dup, x[:10], 8, 2)
    assert np.array_equal(idx, g["dup_idx"])
    assert np.array_equal(dist, g["dup_dist"])


def test_exact_knn_brute_force_with_ties(oracle):
    """Integer coordinates make every squared distance exact in float32, so the
    (distance, index) order is fixed and ties are frequent."""
    rng = np.random.default_rng(3)
    train = rng.integers(0, 3, size=(60, 7)).astype(np.float32)
    query = rng.integers(0, 3, size=(15, 7)).astype(np.float32)
    k = 9
    idx, dist = oracle.exact_knn(train, query, k, 2)
    d2 = ((query[:, None, :].astype(np.int64) - train[None, :, :].astype(np.int64)) ** 2).sum(-1)
    for q in range(query.shape[0]):
        order = np.lexsort((np.arange(train.shape[0]), d2[q]))[:k]
        assert np.array_equal(idx[q], order), q
        assert np.array_equal(dist[q], d2[q, order].astype(np.float32)), q
    # exclusion: a row never lists itself
    ex = np.arange(15, dtype=np.int32)
    idx, _ = oracle.exact_knn(train, train[:15], k, 2, exclude=ex)
    assert not np.any(idx == ex[:, None])


# ---- vote (core.cpp:784-914, fallbacks 258-309) ----------------------------------

def test_vote_hand_built_ties(oracle):
    labels = np.array([1, 1, 2, 2, 3, 3], np.int32)
    folds = np.array([0, 1, 0, 0, 0, 1], np.int32)
    nbr = np.full((6, 4), -1, np.int32)
    score = np.zeros((6, 4), np.float32)
    # row 0: count decides (label 2 twice, label 1 once)
    nbr[0, :3], score[0, :3] = [2, 3, 1], [-5, -6, -0.5]
    # row 1: count tie, the larger score sum (label 1) wins
    nbr[1, :2], score[1, :2] = [0, 2], [-1, -2]
    # row 2: count tie, the larger sum is on the farther neighbour's label
    nbr[2, :2], score[2, :2] = [4, 0], [-3, -1]
    # row 3: count and sum tie, the smaller label wins
    nbr[3, :2], score[3, :2] = [4, 0], [-2, -2]
    # row 4: two labels with two votes each and equal sums, the smaller label wins
    nbr[4], score[4] = [0, 2, 1, 3], [-1, -1, -1, -1]
    # row 5: no neighbour: the training-fold majority (fold 1 holds rows 1 and 5,
    # the other rows' labels are 1, 2, 2, 3)
    pred = oracle.knn_vote(nbr, score, folds, labels)
    assert pred.tolist() == [2, 1, 1, 1, 1, 2]


def test_vote_sums_in_double(oracle):
    """Two labels with two votes each. In float32 both sums round to -1 and the
    smaller label (1) would win; in double, label 2's sum (-1 - 2^-30) exceeds
    label 1's (-1 - 2^-29), so the answer is 2 only if the sums are doubles."""
    labels = np.array([1, 1, 2, 2, 2], np.int32)
    nbr = np.full((5, 4), -1, np.int32)
    score = np.zeros((5, 4), np.float32)
    nbr[0] = [1, 3, 1, 4]
    score[0] = [-1.0, -1.0, -(2.0 ** -29), -(2.0 ** -30)]
    assert np.float32(-1.0) + np.float32(-(2.0 ** -29)) == np.float32(-1.0)
    got = oracle.knn_vote(nbr, score, np.zeros(5, np.int32), labels)
    assert got[0] == 2


# ---- objective (core.cpp:916-972) ------------------------------------------------

def test_objective_closed_form(oracle):
    guarded = O.core_options(guarded_diversity=1, diversity_multiplier=1) if "diversity_multiplier" in \
        dict(O.CoreOptions._fields_) else O.core_options(guarded_diversity=1)
    plain = O.core_options(guarded_diversity=0)
    lab = np.array([1] * 6 + [2] * 3 + [7] * 1, np.int32)
    acc = 0.8
    p = np.array([6, 3, 1]) / 10.0
    assert oracle.objective(lab, acc, guarded) == pytest.approx(acc * np.sqrt(1.0 - (p ** 2).sum()), rel=1e-15)
    assert oracle.objective(lab, acc, plain) == acc
    assert oracle.objective(np.ones(5, np.int32), 1.0, guarded) == -np.inf


# ---- trajectories and CV against the reference's fixtures ------------------------

@pytest.mark.parametrize("name,flags", [("api", CORE_API), ("matrix", MATRIX_RUN)])
def test_core_knn_trajectories_match_golden(oracle, name, flags):
    g = golden("core_knn_c1.npz")
    x, labels, _ = D.config1()
    assert np.array_equal(labels, g["initial"])
    for run in range(3):
        opt = O.core_options(cycles=20, k=10, folds=10, seed=1 + run, cv_seed=1 + run, **flags)
        got = oracle.core_knn(x, labels, opt)
        assert_same_core(got, golden_core(g, f"{name}_r{run}_"), f"{name} run {run}")
        check_invariants(got, 20, flags["evolutionary_search"] == 1)


def test_knncv_matches_golden(oracle):
    g = golden("knncv_c1.npz")
    x, _, comp = D.config1()
    for metric in (0, 2):
        opt = O.core_options(k=10, folds=10, stratified=1, cv_seed=3, metric=metric)
        pred, folds = oracle.knncv(x, comp, opt)
        assert np.array_equal(pred, g[f"pred_m{metric}"]), metric
        assert np.array_equal(folds, g[f"folds_m{metric}"]), metric


def test_plsldacv_fp32_matches_golden(oracle):
    g = golden("plslda_c2.npz")
    x2, truth2 = D.config2()
    x2, truth2 = x2[:2000], truth2[:2000]
    for strat in (1, 0):
        opt = O.core_options(folds=10, stratified=strat, cv_seed=3, max_components=5, fixed_components=5)
        pred, folds, sel = oracle.plsldacv(x2, truth2, opt, precision=32)
        assert np.array_equal(pred, g[f"cv_pred_s{strat}"]), strat
        assert np.array_equal(folds, g[f"cv_folds_s{strat}"]), strat
        assert sel == int(g[f"cv_sel_s{strat}"])


def test_core_plslda_fp32_trajectories_match_golden(oracle):
    g = golden("plslda_c2.npz")
    x2, _ = D.config2()
    x2 = x2[:2000]
    init = g["init"]
    for run in range(2):
        opt = O.core_options(cycles=20, classifier=0, folds=10, stratified=0, seed=1 + run, cv_seed=1 + run,
                             max_components=5, fixed_components=5, evolutionary_search=1,
                             guarded_diversity=1, many_to_one_absorption=1, auto_class_coarsening=1)
        got = oracle.core_plslda(x2, init, opt, precision=32)
        assert_same_core(got, golden_core(g, f"core_r{run}_"), f"run {run}")


# ---- PLS-LDA closed forms (plscv.cpp:816-955, 1419-1510) -------------------------

def _standardize(x, rows):
    """Training mean and scale as plscv.cpp:669-698 computes them: the scale
    accumulator starts at 1, so sd = sqrt((1 + sum (x - mean)^2) / (n - 1))."""
    xt = x[rows].astype(np.float64)
    mean = xt.mean(0)
    sd = np.sqrt((1.0 + ((xt - mean) ** 2).sum(0)) / max(1, rows.size - 1))
    sd[~(sd > 0)] = 1.0
    return mean, sd


def test_pls_two_class_weight_and_discriminant_closed_form(oracle):
    """Two classes make S = X'(Y - ybar) rank one, so one power step gives
    w = s / sqrt(s'Gs) up to sign (G = X'X of the standardized training rows);
    with one component the LDA discriminant is
    t*mu_c/v - mu_c^2/(2v) + log(n_c/n), v = pooled variance * (1 + 1e-8)."""
    x, _ = D.config2(n=600)
    x = x.astype(np.float32)
    lab = np.where(x[:, 0] + 0.5 * x[:, 1] > 0, 3, 8).astype(np.int32)
    folds = (np.arange(600) % 5).astype(np.int32)
    opt = O.core_options(max_components=1, fixed_components=1, stratified=0, pls_fold_cache=1)
    d = oracle.pls_fold_detail(x, lab, folds, 2, opt, precision=64)
    tr = np.flatnonzero(folds != 2)
    va = np.flatnonzero(folds == 2)
    mean, sd = _standardize(x, tr)
    xt = (x[tr].astype(np.float64) - mean) / sd
    xv = (x[va].astype(np.float64) - mean) / sd
    y = (lab[tr] == 3).astype(np.float64)
    s = xt.T @ (y - y.mean())
    gmat = xt.T @ xt
    w = s / np.sqrt(s @ gmat @ s)
    got = d["weights"][:, 0]
    w = w * np.sign(w @ got)
    np.testing.assert_allclose(got, w, rtol=1e-9, atol=1e-12)
    assert float(got @ gmat @ got) == pytest.approx(1.0, rel=1e-12)
    # LDA with one score
    t = xt @ got
    classes = np.array([3, 8])
    mu = np.array([t[lab[tr] == c].mean() for c in classes])
    cnt = np.array([(lab[tr] == c).sum() for c in classes])
    pooled = ((t ** 2).sum() - (cnt * mu ** 2).sum()) / (tr.size - 2)
    v = pooled * (1.0 + 1e-8)
    tv = xv @ got
    np.testing.assert_allclose(d["val_scores"][:, 0], tv, rtol=1e-10, atol=1e-12)
    disc = tv[:, None] * mu[None, :] / v - 0.5 * mu[None, :] ** 2 / v + np.log(cnt / tr.size)[None, :]
    np.testing.assert_allclose(d["discriminant"], disc, rtol=1e-9, atol=1e-9)
    assert d["classes"].tolist() == [3, 8]


def test_pls_weights_are_g_orthonormal(oracle):
    x, truth = D.config2(n=800)
    folds = (np.arange(800) % 4).astype(np.int32)
    opt = O.core_options(max_components=5, fixed_components=5, stratified=0, pls_fold_cache=1)
    d = oracle.pls_fold_detail(x, truth, folds, 1, opt, precision=64)
    tr = np.flatnonzero(folds != 1)
    mean, sd = _standardize(x, tr)
    xt = (x[tr].astype(np.float64) - mean) / sd
    w = d["weights"]
    assert w.shape[1] == 5
    np.testing.assert_allclose(np.diag(w.T @ (xt.T @ xt) @ w), 1.0, rtol=1e-10)


# ---- dissimilarity (kodama_matrix.cpp:2685-2776) ----------------------------------

def dissimilarity_brute_force(idx1, dist, res):
    """1-based graph in, 1-based graph out: co-assignment counts over runs, then
    d' = (1 + d) / (same / valid)^2 in double, rounded to float (inf when no run
    agrees), rows sorted stably by (d', id)."""
    n, K = idx1.shape
    out_i = np.empty_like(idx1)
    out_d = np.empty_like(dist)
    for i in range(n):
        vals = []
        for j in range(K):
            nb = idx1[i, j] - 1 if idx1[i, j] >= 0 else idx1[i, j]
            if nb < 0 or nb >= n or not np.isfinite(dist[i, j]):
                vals.append((np.float32(np.inf), idx1[i, j]))  # entry kept, pushed last
                continue
            a, b = res[:, i], res[:, nb]
            valid = int(((a != 0) & (b != 0)).sum())
            same = int(((a == b) & (a != 0)).sum())
            if same == 0:
                v = np.float32(np.inf)
            else:
                r = float(same) / float(valid)
                v = np.float32((1.0 + float(dist[i, j])) / (r * r))
            vals.append((v, idx1[i, j]))
        vals.sort(key=lambda t: (t[0], t[1]))
        out_d[i] = [v for v, _ in vals]
        out_i[i] = [c for _, c in vals]
    return out_i, out_d


def test_dissimilarity_matches_golden_and_counts(oracle):
    g = golden("dissimilarity.npz")
    idx, dist = oracle.dissimilarity(g["idx"], g["dist"], g["res"])
    assert np.array_equal(idx, g["out_idx"])
    assert np.array_equal(dist, g["out_dist"])
    bi, bd = dissimilarity_brute_force(g["idx"], g["dist"], g["res"])
    assert np.array_equal(idx, bi)
    assert np.array_equal(dist, bd)


# ---- the reference core on fresh inputs (build container only) --------------------

def test_fresh_core_knn_matches_reference(oracle, reference):
    x, labels, _ = D.config1(n=700, seed=19)
    for flags in (CORE_API, MATRIX_RUN):
        opt = O.core_options(cycles=12, k=7, folds=5, seed=4, cv_seed=4, **flags)
        assert_same_core(oracle.core_knn(x, labels, opt), reference.core_knn(x, labels, opt), str(flags))


def test_fresh_plsldacv_matches_reference(oracle, reference):
    x, truth = D.config2(n=900, seed=23)
    for strat in (0, 1):
        opt = O.core_options(folds=6, stratified=strat, cv_seed=5, max_components=4, fixed_components=4)
        a = oracle.plsldacv(x, truth, opt, precision=32)
        b = reference.plsldacv(x, truth, opt)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_fresh_dissimilarity_matches_reference(oracle, reference):
    x, _, _ = D.config1(n=250, seed=29)
    idx, dist = oracle.knn_graph(x, 9)
    res = np.stack([D.randint(60 + r, 1, 250, 0, 5) for r in range(6)]).astype(np.int32)
    a = oracle.dissimilarity(idx, dist, res)
    b = reference.dissimilarity(idx, dist, res)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---- KODAMA.matrix building blocks (kodama_matrix.cpp:91-124, 535-944) -----------

def _matrix_golden():
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_golden as MG
    return MG, golden("matrix_blocks.npz")


def test_kmeans_matches_reference_golden(oracle):
    """Seeding permutation, SSE-order distances, sequential centroid sums,
    empty-cluster splits (the duplicates case) and the early stop all show in
    the labels; the fixtures come from the reference's own kmeans_labels."""
    MG, g = _matrix_golden()
    for name, (x, k, seed, discard, iters) in MG.matrix_inputs().items():
        assert np.array_equal(oracle.kmeans_labels(x, k, seed, discard, iters), g[f"kmeans_{name}"]), name


def test_quota_landmarks_match_reference_golden(oracle):
    MG, g = _matrix_golden()
    labels = {name: g[f"kmeans_{name}"] for name in MG.matrix_inputs()}
    for i, (name, strata, target, seed, discard) in enumerate(MG.QUOTA_CASES):
        rows = oracle.quota_sample(labels[name], strata, target, seed, discard)
        assert np.array_equal(rows, g[f"quota_{i}"]), i
        # invariants: sorted distinct rows, every stratum within one of its share
        assert rows.size == target and np.all(np.diff(rows) > 0)
        lab = labels[name]
        share = target * np.bincount(lab, minlength=strata + 1)[1:] / lab.size
        got = np.bincount(lab[rows], minlength=strata + 1)[1:]
        assert np.all(got >= np.floor(share)) and np.all(got <= np.floor(share) + 1)


def test_kmeans_closed_cases(oracle):
    """Two groups 100 standard deviations apart split exactly into two
    clusters from any seeding (two seeds in one group pull the second centroid
    over within a step); k = n puts every distinct row in a class of its own;
    k = 1 gives one class."""
    rng = np.random.default_rng(3)
    truth = np.repeat([0, 1], [70, 130])
    centres = rng.normal(size=(2, 12)) * 100.0
    x = (centres[truth] + rng.normal(size=(200, 12))).astype(np.float32)
    for seed in range(1, 9):
        lab = oracle.kmeans_labels(x, 2, seed, 0, 20)
        assert len(set(zip(truth.tolist(), lab.tolist()))) == 2, seed
    y = rng.normal(size=(40, 5)).astype(np.float32)
    assert sorted(oracle.kmeans_labels(y, 40, 3).tolist()) == list(range(1, 41))
    assert np.all(oracle.kmeans_labels(y, 1, 3) == 1)


def test_identity_constraint_equals_no_constraint(oracle):
    """KODAMA.matrix passes constrain = 1..L to each Core run; one group per
    row in row order draws the same streams as singleton groups, so the batched
    engine (which takes no constraint) replays it."""
    x, y, _ = D.config1(n=300, seed=4)
    init = (y % 4 + 1).astype(np.int32)
    opt = O.core_options(cycles=6, k=8, folds=5, seed=7, cv_seed=7, **MATRIX_RUN)
    a = oracle.core_knn(x, init, opt)
    b = oracle.core_knn(x, init, opt, constrain=np.arange(1, 301, dtype=np.int32))
    for key in ("clbest", "vect_acc", "accepted", "fold_assignments"):
        assert np.array_equal(a[key], b[key]), key


def test_fresh_kmeans_and_quota_match_reference(oracle):
    if not O.ReferenceMatrix.available():
        pytest.skip("oracle/_ref (the compiled reference core) is not built here")
    ref = O.ReferenceMatrix()
    x, _, _ = D.config1(n=900, seed=13)
    for k, seed, discard, iters in ((50, 11, 0, 10), (9, 12, 1, 10), (300, 13, 0, 4)):
        a = oracle.kmeans_labels(x, k, seed, discard, iters)
        assert np.array_equal(a, ref.kmeans_labels(x, k, seed, discard, iters)), k
    lab = oracle.kmeans_labels(x, 50, 5)
    for target in (10, 200, 777):
        assert np.array_equal(oracle.quota_sample(lab, 50, target, 3, 1), ref.quota_sample(lab, 50, target, 3, 1))


PREDICT_CASES = [(5, 1, 1), (2, 1, 1), (8, 1, 0)]


def _predict_inputs():
    x2, _ = D.config2()
    return x2[:1500], x2[1500:2000]


@pytest.mark.parametrize("ncomp,center,scale", PREDICT_CASES)
def test_plslda_predict_matches_reference_golden(oracle, ncomp, center, scale):
    """The oracle's PLSLDAPredict (one validation fold of [train; test], FP32
    as the reference) against the reference's own PLSLDAPredict_CPU outputs,
    with the true labels and with random 12-class labels."""
    g = golden("plslda_predict.npz")
    train, test = _predict_inputs()
    opt = O.core_options(classifier=0, max_components=ncomp, fixed_components=ncomp, center=center, scale=scale)
    for name in ("truth", "noisy"):
        got = oracle.plslda_predict(train, g[f"labels_{name}"], test, opt, precision=32)
        assert np.array_equal(got, g[f"pred_{name}_{ncomp}_{center}{scale}"]), name


def test_plslda_predict_closed_cases(oracle):
    """One training class predicts itself; constant data has no PLS component
    (the reference throws DegeneratePLSFit); two well-separated classes on one
    axis are told apart."""
    rng = np.random.default_rng(3)
    train = rng.normal(size=(40, 4)).astype(np.float32)
    test = rng.normal(size=(7, 4)).astype(np.float32)
    opt = O.core_options(classifier=0, max_components=2, fixed_components=2)
    assert np.array_equal(oracle.plslda_predict(train, np.full(40, 9), test, opt), np.full(7, 9))
    with pytest.raises(ValueError):
        oracle.plslda_predict(np.ones((40, 4), np.float32), np.arange(40) % 2 + 1, test, opt)
    train[:20, 0] += 20.0
    lab = np.r_[np.full(20, 3), np.full(20, 5)]
    probe = np.zeros((2, 4), np.float32)
    probe[0, 0] = 20.0
    assert np.array_equal(oracle.plslda_predict(train, lab, probe, opt), [3, 5])
    if O.Reference.available():
        ref = O.Reference()
        assert np.array_equal(ref.plslda_predict(train, lab, probe, opt), [3, 5])
        assert np.array_equal(oracle.plslda_predict(train, lab, test, opt, precision=32),
                              ref.plslda_predict(train, lab, test, opt))


def test_normalize_graph_closed_case(oracle):
    """normalize_external_graph on a hand-worked 4-row graph: self entries, an
    out-of-range id, a NaN and a negative distance (clamped at 0), rows left
    shorter than the width, and how auto detection picks the base."""
    gi = np.array([[2, 1, 4, 3], [1, 5, 3, 2], [4, 3, -1, 1], [3, 1, 2, 4]], np.int32)
    gd = np.array([[0.5, 0.0, np.nan, 2.0], [0.25, 9.0, -1.0, 7.0], [1.0, 0.0, 3.0, 4.0],
                   [0.1, 0.2, 0.3, 0.4]], np.float32)
    inf = np.inf
    idx, dist = oracle.normalize_graph(gi, gd, 1, 3)  # one-based: 5 is out of range
    assert np.array_equal(idx, [[1, 2, -1], [0, 2, -1], [3, 0, -1], [2, 0, 1]])
    assert np.array_equal(dist, np.array([[0.5, 2.0, inf], [0.25, 0.0, inf], [1.0, 4.0, inf],
                                          [0.1, 0.2, 0.3]], np.float32))
    # Auto: the 5 lies outside 1..4, so the graph reads as zero-based.
    a_idx, a_dist = oracle.normalize_graph(gi, gd, -1, 3)
    assert np.array_equal(a_idx[0], [2, 1, 3]) and np.array_equal(a_dist[0], np.float32([0.5, 0.0, 2.0]))
    # Auto with every id in 1..4: one-based.
    gi4 = gi.copy()
    gi4[1, 1] = 4
    b_idx, b_dist = oracle.normalize_graph(gi4, gd, -1, 3)
    assert np.array_equal(b_idx[1], [0, 3, 2]) and np.array_equal(b_dist[1], np.float32([0.25, 9.0, 0.0]))
    # Forced zero-based on the shifted graph: 4 is out of range, -1 is a gap.
    z_idx, _ = oracle.normalize_graph(gi - 1, gd, 0, 4)
    assert np.array_equal(z_idx[:2], [[1, 2, -1, -1], [0, 2, -1, -1]])


def test_normalize_graph_matches_reference(oracle):
    if not O.ReferenceMatrix.available():
        pytest.skip("oracle/_ref (the compiled reference core) is not built here")
    ref = O.ReferenceMatrix()
    rng = np.random.default_rng(21)
    n, w = 300, 12
    for case in range(8):
        base = (-1, 0, 1)[case % 3]
        lo = 1 if case in (0, 3, 6) else 0  # cases 0, 3, 6: a one-based graph
        gi = rng.integers(lo - 2, n + lo + 2, size=(n, w)).astype(np.int32)
        gi[np.arange(n) % 7 == 0, 1] = np.arange(n)[np.arange(n) % 7 == 0] + lo  # self entries
        gd = rng.normal(1.0, 1.0, size=(n, w)).astype(np.float32)
        gd[rng.random((n, w)) < 0.05] = np.nan
        gd[rng.random((n, w)) < 0.03] = np.inf
        if case >= 6:
            gi = np.clip(gi, lo, n - 1 + lo)  # every id in range: auto detection decides
        for width in (1, 7, w):
            a = oracle.normalize_graph(gi, gd, base, width)
            b = ref.normalize_graph(gi, gd, base, width)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (case, width)


# ---- learned proposal (no reference counterpart): the float64 restatement ----------

def test_mlp_features_hand_built():
    import mlp_oracle as MO
    lab = np.array([5, 2, 2, 5, 7, 5, 2, 9], np.int32)
    nbr = np.array([[1, 2, 3, 4, 5, 6],     # votes 2,2,5,7,5,2 own 5 -> (2:3) (5:2) (7:1)
                    [7, -1, -1, 0, 0, 0],   # votes 9,5,5,5 own 2 (not voted)
                    [1, 1, 1, 1, 1, 1],     # all 2 (own): one rank, margin 0
                    [-1, -1, -1, -1, -1, -1],
                    [0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0]], np.int32)
    f, r = MO.features(nbr, lab)
    assert np.array_equal(r[0, :4], [2, 5, 7, -1])
    assert np.allclose(f[0, :3], [3 / 6, 2 / 6, 1 / 6]) and f[0, 3] == 0 and f[0, 31] == 1 and f[0, 61] == np.float32(1 / 6)
    assert np.array_equal(r[1, :3], [5, 9, -1]) and f[1, 60] == 1 and f[1, 61] == np.float32(3 / 6)
    assert f[2, 30] == 1 and f[2, 61] == 0 and np.allclose(f[2, :2], [1, 0])
    assert np.all(f[3, :30] == 0) and f[3, 60] == 1 and f[3, 61] == 0
    assert np.all(f[:, 62] == 1) and np.all(f[:, 30:61].sum(axis=1) == 1)
    v = MO.valid_outputs(f)
    assert v[0, 0] and not v[0, 1] and v[0, 2] and not v[0, 3] and v[0, 30] and not v[0, 31]


def test_mlp_oracle_gradient_matches_finite_differences():
    """Central differences of the float64 loss with respect to the weights the
    network computes with (bf16-rounded values taken as the variables) and the
    biases: a dropped ReLU mask, a transposed product or a missing 1/n fails."""
    import mlp_oracle as MO
    rng = np.random.default_rng(0)
    nbr = rng.integers(-1, 60, size=(60, 12)).astype(np.int32)
    lab = rng.integers(0, 7, size=60).astype(np.int32)
    feat, _ = MO.features(nbr, lab)
    valid = MO.valid_outputs(feat)
    tgt = np.array([rng.choice(np.flatnonzero(v)) for v in valid], np.int32)
    params = (rng.normal(size=MO.PARAMS) * 0.3).astype(np.float32)
    w = MO.unpack(params)
    loss, g = MO.loss_and_grads_of(w, feat, tgt)
    names = (("w1", MO.W1), ("b1", MO.B1), ("w2", MO.W2), ("b2", MO.B2), ("w3", MO.W3), ("b3", MO.B3))
    for name, off in names:
        flat = w[name].ravel()
        for idx in rng.choice(flat.size, size=6, replace=False):
            h = 1e-6
            wp = {k: v.copy() for k, v in w.items()}
            wm = {k: v.copy() for k, v in w.items()}
            wp[name].ravel()[idx] += h
            wm[name].ravel()[idx] -= h
            fd = (MO.loss_and_grads_of(wp, feat, tgt)[0] - MO.loss_and_grads_of(wm, feat, tgt)[0]) / (2 * h)
            assert abs(fd - g[off + idx]) <= 1e-6 + 1e-5 * abs(fd), (name, idx, fd, g[off + idx])
    assert np.isfinite(loss)
