"""HIP PLS-LDA (FP64 on the device) against the FP64 oracle -- scores and LDA
discriminants within 1e-4 relative, predictions and CorePLSLDA trajectories
identical -- and against the reference's FP32 golden fixtures."""
import numpy as np
import pytest

import kodama_amd as K
import oracle as O
from _cases import assert_same_core, golden_core
from conftest import golden
from kodama_amd import datasets as D

pytestmark = pytest.mark.gpu

RTOL = 1e-4  # PLS scores and LDA outputs vs the FP64 oracle


def _opts(**kw):
    d = dict(classifier=0, folds=10, stratified=0, cv_seed=3, max_components=5, fixed_components=5,
             center=1, scale=1, pls_fold_cache=1)
    d.update(kw)
    return d


def _rel_close(a, b, rtol=RTOL):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    scale = np.maximum(np.abs(b), np.max(np.abs(b)) * 1e-6 + 1e-300)
    err = np.max(np.abs(a - b) / scale) if a.size else 0.0
    assert err <= rtol, err


def _c2(n=2000):
    x, truth = D.config2()
    return x[:n], truth[:n]


@pytest.mark.parametrize("strat", [0, 1])
def test_plsldacv_predictions_match_oracle_fp64(oracle, strat):
    x, truth = _c2()
    o = _opts(stratified=strat, pls_fold_cache=0)  # the options PLSLDACV passes
    got = K.PLSLDACV(x, truth, folds=10, stratified=bool(strat), seed=3, ncomp=5)
    pred, folds, sel = oracle.plsldacv(x, truth, O.core_options(**o), precision=64)
    assert np.array_equal(got["fold_assignments"], folds)
    assert np.array_equal(got["predicted_labels"], pred)
    assert got["selected_components"] == sel


@pytest.mark.parametrize("strat", [1, 0])
def test_plsldacv_matches_reference_golden(strat):
    """The reference computes in FP32; the device in FP64. Their predictions agree
    on this workload (no validation row sits within float rounding of a tie)."""
    g = golden("plslda_c2.npz")
    x, truth = _c2()
    got = K.PLSLDACV(x, truth, folds=10, stratified=bool(strat), seed=3, ncomp=5)
    assert np.array_equal(got["fold_assignments"], g[f"cv_folds_s{strat}"])
    assert np.array_equal(got["predicted_labels"], g[f"cv_pred_s{strat}"])
    assert got["selected_components"] == int(g[f"cv_sel_s{strat}"])


@pytest.mark.parametrize("ncomp,classes", [(5, 8), (10, 3), (50, 8), (3, 2)])
def test_fold_scores_and_discriminants_within_tolerance(oracle, ncomp, classes):
    x, truth = D.config2(n=1500, seed=61)
    lab = (truth % classes + 1).astype(np.int32)
    folds = (np.argsort(D.uniform(9, 1, lab.size)) % 5).astype(np.int32)
    o = _opts(max_components=ncomp, fixed_components=ncomp, folds=5)
    for fold in (0, 3):
        dev = K.plslda_fold_detail(x, lab, folds, fold, K.core_options(**o))
        ref = oracle.pls_fold_detail(x, lab, folds, fold, O.core_options(**o), precision=64)
        assert dev["rank"] == ref["rank"]
        assert np.array_equal(dev["classes"], ref["classes"])
        _rel_close(dev["weights"], ref["weights"])
        _rel_close(dev["val_scores"], ref["val_scores"])
        _rel_close(dev["discriminant"], ref["discriminant"])
        # LDA posteriors (softmax of the discriminants)
        def post(d):
            e = np.exp(d - d.max(1, keepdims=True))
            return e / e.sum(1, keepdims=True)
        np.testing.assert_allclose(post(dev["discriminant"]), post(ref["discriminant"]), rtol=RTOL, atol=1e-12)


def test_core_plslda_trajectories_match_oracle_fp64_and_reference(oracle):
    g = golden("plslda_c2.npz")
    x, _ = _c2()
    init = g["init"]
    for run in range(2):
        kw = dict(cycles=20, classifier=0, folds=10, stratified=0, seed=1 + run, cv_seed=1 + run,
                  max_components=5, fixed_components=5, evolutionary_search=1, guarded_diversity=1,
                  many_to_one_absorption=1, auto_class_coarsening=1)
        dev = K.core_plslda(x, init, K.core_options(**kw))
        ref64 = oracle.core_plslda(x, init, O.core_options(**kw), precision=64)
        assert_same_core(dev, ref64, f"fp64 run {run}")
        assert_same_core(dev, golden_core(g, f"core_r{run}_"), f"reference run {run}")


def test_core_plslda_stratified_api_defaults(oracle):
    x, truth = D.config2(n=600, seed=71)
    init = (truth % 4 + 1).astype(np.int32)
    kw = dict(cycles=8, classifier=0, folds=5, stratified=1, seed=2, cv_seed=2, max_components=4,
              fixed_components=4)
    dev = K.core_plslda(x, init, K.core_options(**kw))
    ref = oracle.core_plslda(x, init, O.core_options(**kw), precision=64)
    assert_same_core(dev, ref, "stratified")


def test_single_class_and_degenerate_folds(oracle):
    # Constant data: no PLS direction, the selected component count is 1
    # (tests/testthat/test-pls-degenerate.R:4-19).
    x = np.ones((60, 4), np.float32)
    lab = np.repeat([1, 2, 3], 20).astype(np.int32)
    got = K.PLSLDACV(x, lab, folds=3, stratified=True, seed=1, ncomp=3)
    pred, folds, sel = oracle.plsldacv(x, lab, O.core_options(**_opts(folds=3, stratified=1, cv_seed=1,
                                                                       max_components=3, fixed_components=3,
                                                                       pls_fold_cache=0)),
                                       precision=64)
    assert got["selected_components"] == 1 == sel
    assert np.array_equal(got["predicted_labels"], pred)
    # one class only
    one = np.full(60, 7, np.int32)
    xs, _ = D.config2(n=60, seed=3)
    got = K.PLSLDACV(xs, one, folds=3, stratified=False, seed=1, ncomp=2)
    assert np.all(got["predicted_labels"] == 7)


def test_more_components_than_columns(oracle):
    x, truth = D.config2(n=400, seed=81)
    x = x[:, :6]
    o = _opts(folds=4, max_components=10, fixed_components=10, pls_fold_cache=0)
    got = K.PLSLDACV(x, truth, folds=4, stratified=False, seed=3, ncomp=10)
    pred, folds, sel = oracle.plsldacv(x, truth, O.core_options(**o), precision=64)
    assert np.array_equal(got["predicted_labels"], pred)
    assert got["selected_components"] == sel


def test_core_plslda_batch_equals_independent_runs():
    """The M-run batch is the per-run CorePLSLDA with seed + r and cv_seed + r."""
    x, truth = D.config2(n=3000, seed=91)
    rows, init = [], []
    for r in range(4):
        L = 400 + 150 * r
        rw = np.sort(np.argsort(D.uniform(91, 10 + r, 3000))[:L]).astype(np.int32)
        rows.append(rw)
        init.append(((truth[rw] + r) % 6 + 1).astype(np.int32))
    kw = dict(cycles=6, classifier=0, folds=5, stratified=0, seed=7, cv_seed=11, max_components=5,
              fixed_components=5, evolutionary_search=1, guarded_diversity=1, many_to_one_absorption=1,
              auto_class_coarsening=1)
    batch = K.core_plslda_batch(x, rows, init, K.core_options(**kw))
    for r in range(4):
        one = K.core_plslda(x[rows[r]], init[r], K.core_options(**dict(kw, seed=7 + r, cv_seed=11 + r)))
        assert np.array_equal(batch["clbest"][r], one["clbest"]), r
        assert batch["accbest"][r] == one["accbest"], r
        assert np.array_equal(batch["vect_acc"][r], one["vect_acc"]), r


@pytest.mark.parametrize("ncomp,center,scale", [(5, 1, 1), (2, 1, 1), (8, 1, 0)])
def test_plslda_predict_matches_reference_and_oracle(oracle, ncomp, center, scale):
    """PLSLDAPredict on the device: the reference's PLSLDAPredict_CPU outputs
    (golden, FP32) and the oracle's FP64 restatement, with the true labels and
    with random 12-class labels."""
    g = golden("plslda_predict.npz")
    x2, _ = D.config2()
    train, test = x2[:1500], x2[1500:2000]
    opt = O.core_options(classifier=0, max_components=ncomp, fixed_components=ncomp, center=center, scale=scale)
    for name in ("truth", "noisy"):
        lab = g[f"labels_{name}"]
        got = K.PLSLDAPredict(train, lab, test, ncomp=ncomp, center=bool(center), scale=bool(scale))
        assert np.array_equal(got, g[f"pred_{name}_{ncomp}_{center}{scale}"]), name
        assert np.array_equal(got, oracle.plslda_predict(train, lab, test, opt, precision=64)), name


def test_plslda_predict_edge_cases():
    """One training class predicts itself, an empty test set predicts
    nothing, constant data is a degenerate fit (EINVAL), mismatched columns
    are rejected."""
    rng = np.random.default_rng(5)
    train = rng.normal(size=(30, 6)).astype(np.float32)
    test = rng.normal(size=(9, 6)).astype(np.float32)
    assert np.array_equal(K.PLSLDAPredict(train, np.full(30, 4), test, ncomp=3), np.full(9, 4))
    assert K.PLSLDAPredict(train, np.arange(30) % 3, test[:0], ncomp=3).shape == (0,)
    with pytest.raises(K.InvalidArgument):
        K.PLSLDAPredict(np.ones((30, 6), np.float32), np.arange(30) % 2, test, ncomp=3)
    with pytest.raises(ValueError):
        K.PLSLDAPredict(train, np.arange(30) % 2, test[:, :5], ncomp=3)



# ---- PLSDACV and general PLS ----------------------------------------------------

def test_plsdacv_matches_reference_and_oracle(oracle):
    """PLSDACV on the device (FP64): predictions and selected components equal
    the reference's PLSDACV_CPU fixtures and the FP64 oracle."""
    from _cases import pls_general_inputs
    g = golden("pls_general.npz")
    x, truth, noisy, _ = pls_general_inputs()
    for name, lab in (("truth", truth), ("noisy", noisy)):
        for strat in (1, 0):
            for nc in (2, 5):
                key = f"dacv_{name}_s{strat}_c{nc}"
                got = K.PLSDACV(x, lab, folds=10, stratified=bool(strat), seed=4, ncomp=nc)
                opt = O.core_options(folds=10, stratified=strat, cv_seed=4, max_components=nc, fixed_components=nc)
                want, folds, sel = oracle.plsdacv(x, lab, opt, precision=64)
                assert np.array_equal(got["fold_assignments"], folds), key
                assert np.array_equal(got["predicted_labels"], want), key
                assert np.array_equal(got["predicted_labels"], g[key]), key
                assert got["selected_components"] == sel == int(g[key + "_sel"]), key


def test_plsdacv_edge_cases(oracle):
    """One training class per fold predicts it; a constant matrix gives the
    degenerate fit's training majority (components 1); both as the oracle."""
    x = np.random.default_rng(3).standard_normal((60, 5)).astype(np.float32)
    one = np.full(60, 7, np.int32)
    got = K.PLSDACV(x, one, folds=5, seed=2, ncomp=3)
    assert np.all(got["predicted_labels"] == 7)
    const = np.ones((60, 5), np.float32)
    lab = (np.arange(60) % 3 + 1).astype(np.int32)
    got = K.PLSDACV(const, lab, folds=5, seed=2, ncomp=3)
    opt = O.core_options(folds=5, stratified=1, cv_seed=2, max_components=3, fixed_components=3)
    want, _, sel = oracle.plsdacv(const, lab, opt, precision=64)
    assert np.array_equal(got["predicted_labels"], want)
    assert got["selected_components"] == sel == 1


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_pls_general_matches_oracle_fp64_and_reference(oracle, case):
    """General PLS on the device (cross-products on the matrix cores, FP64):
    every output within 1e-4 relative of the FP64 oracle (observed ~1e-12) and
    of the reference's FP32 PLS_CPU fixture; the same component count."""
    from _cases import PLS_FIELDS, PLS_GENERAL_CASES, pls_general_inputs, rel_err
    g = golden("pls_general.npz")
    x, _, _, y = pls_general_inputs()
    nc, center, scale = PLS_GENERAL_CASES[case]
    got = K.PLS(x, y, ncomp=nc, center=bool(center), scale=bool(scale))
    opt = O.core_options(max_components=nc, fixed_components=nc, center=center, scale=scale)
    want = oracle.pls(x, y, opt, precision=64)
    tag = f"pls_{nc}_{center}{scale}"
    assert got["components"] == want["components"] == int(g[tag + "_components"])
    for key in PLS_FIELDS:
        assert rel_err(got[key], want[key]) < 1e-6, (tag, key, rel_err(got[key], want[key]))
        assert rel_err(got[key], g[f"{tag}_{key}"].reshape(np.asarray(got[key]).shape)) < RTOL, (tag, key)


def test_pls_general_large_and_wide(oracle):
    """200,000 rows (many row chunks of the cross-product kernel) and 70
    predictors with 20 responses (several 16 x 16 output tiles per wave and
    padded edges): within 1e-6 relative of the FP64 oracle."""
    from _cases import PLS_FIELDS, rel_err
    rng = np.random.default_rng(8)
    for n, p, q, nc in ((200000, 50, 3, 10), (3000, 70, 20, 6)):
        x = rng.standard_normal((n, p)).astype(np.float32)
        y = (x[:, :5] @ rng.standard_normal((5, q)) + rng.standard_normal((n, q))).astype(np.float32)
        got = K.PLS(x, y, ncomp=nc)
        want = oracle.pls(x, y, O.core_options(max_components=nc, fixed_components=nc, center=1, scale=1),
                          precision=64)
        assert got["components"] == want["components"]
        for key in PLS_FIELDS:
            assert rel_err(got[key], want[key]) < 1e-6, (n, p, q, key, rel_err(got[key], want[key]))
        assert got["cross_seconds"] > 0.0
