"""The learned proposal: device features, forward pass and gradients against
the float64 restatement (oracle/mlp_oracle.py, itself pinned by finite
differences), deterministic device training, accept/reject traces that leave
the plain trajectory untouched, the uniform fallback, what the trained
network gains over the uniform proposal, and an ablation that separates the
candidate list from the network's scores."""
import numpy as np
import pytest

import kodama_amd as K
import mlp_oracle as MO

from test_gpu_device_evolution import OPTS, _centre_batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [6, 12, 30, 64])
def test_features_match_oracle(k):
    """Bit-identical features and ranked labels: missing neighbours, ties of
    counts (smaller code first), own label voted or not, more than 30 voted labels."""
    rng = np.random.default_rng(k)
    rows = 700
    nbr = rng.integers(-1, rows, size=(rows, k)).astype(np.int32)
    lab = rng.integers(0, 45, size=rows).astype(np.int32)
    lab[:50] = 7  # a frequent label: wide counts
    f, r = K.mlp_features(nbr, lab)
    fo, ro = MO.features(nbr, lab)
    assert np.array_equal(f, fo) and np.array_equal(r, ro)


def _samples(rows=400, seed=2):
    rng = np.random.default_rng(seed)
    nbr = rng.integers(-1, rows, size=(rows, 30)).astype(np.int32)
    lab = rng.integers(0, 9, size=rows).astype(np.int32)
    feat, _ = MO.features(nbr, lab)
    valid = MO.valid_outputs(feat)
    tgt = np.array([rng.choice(np.flatnonzero(v)) for v in valid], np.int32)
    return feat, tgt


def test_forward_matches_oracle():
    feat, _ = _samples()
    for seed in (1, 2):
        params = K.mlp_init(seed)
        params[MO.B1:MO.W2] = 0.1  # non-zero biases
        got = K.mlp_forward(params, feat)
        want = MO.forward(params, feat)
        assert np.max(np.abs(got - want)) < 1e-5
        assert np.all(got[~MO.valid_outputs(feat)] == 0.0)


def test_gradients_match_oracle():
    """FP32 device sums against the float64 restatement: within 1e-4 of the
    largest gradient entry (observed about 1e-7 relative)."""
    feat, tgt = _samples()
    params = K.mlp_init(4)
    g, loss = K.mlp_gradients(params, feat, tgt)
    lo, go = MO.loss_and_grads(params, feat, tgt)
    assert abs(loss - lo) < 1e-5 * max(1.0, lo)
    assert np.max(np.abs(g - go)) <= 1e-4 * np.max(np.abs(go))


def test_training_step_and_determinism():
    """One epoch with the whole set as the minibatch and no momentum is one SGD
    step p - lr * grad; two trainings with one seed are bit-identical."""
    feat, tgt = _samples()
    p0 = K.mlp_init(9)
    p1, loss = K.mlp_train(feat, tgt, params=p0, epochs=1, batch=feat.shape[0], learning_rate=0.1, momentum=0.0)
    _, go = MO.loss_and_grads(p0, feat, tgt)
    assert np.max(np.abs(p1 - (p0 - 0.1 * go))) <= 1e-5 * max(1.0, np.max(np.abs(p0)))
    a, la = K.mlp_train(feat, tgt, params=p0, epochs=3, batch=64, seed=3)
    b, lb = K.mlp_train(feat, tgt, params=p0, epochs=3, batch=64, seed=3)
    assert np.array_equal(a, b) and np.array_equal(la, lb)
    assert la[-1] < la[0]


def test_traces_leave_the_plain_trajectory_unchanged():
    x, rows, init, folds = _centre_batch(n_runs=6)
    o = K.core_options(**OPTS)
    feat, tgt, acc = K.mlp_collect_traces(x, rows, init, o, fold_ids=folds)
    plain = K.core_knn_batch(x, rows, init, o, mode="plain", fold_ids=folds)
    assert np.array_equal(acc, plain["accbest"])
    assert tgt.size > 0
    assert np.all(feat[:, 30:61].sum(axis=1) == 1) and np.all(feat[:, 62] == 1)
    valid = MO.valid_outputs(feat)
    assert np.all(valid[np.arange(tgt.size), tgt])
    assert 0 < np.mean(tgt != 30) < 1  # accepted and rejected samples both present


def test_fallback_one_is_the_uniform_device_form():
    """With every cycle falling back to the uniform proposal, the learned
    variant of the persistent kernel reproduces the uniform one exactly (its
    extra per-cycle draw is a separate counter stream)."""
    x, rows, init, folds = _centre_batch(n_runs=6)
    o = K.core_options(**OPTS)
    uni = K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds)
    lrn = K.core_knn_batch_learned(x, rows, init, o, K.mlp_init(1), fallback=1.0, fold_ids=folds)
    assert all(np.array_equal(a, b) for a, b in zip(uni["clbest"], lrn["clbest"]))
    assert np.array_equal(uni["vect_acc"], lrn["vect_acc"])


@pytest.mark.parametrize("k", [5, 15, 30, 40])
def test_learned_accuracy_is_the_vote_of_the_returned_labels(oracle, k):
    """Every width of the learned variant (8, 16, 32 and 64 neighbour slots)."""
    x, rows, init, folds = _centre_batch(n_runs=6)
    o = K.core_options(**dict(OPTS, k=k))
    out = K.core_knn_batch_learned(x, rows, init, o, K.mlp_init(2), fallback=0.25, fold_ids=folds)
    for r in range(len(rows)):
        nbr, score = oracle.precompute_knn_cv(x[rows[r]], folds[r], k)
        cl = out["clbest"][r]
        pred = oracle.knn_vote(nbr, score, folds[r], cl)
        assert out["accbest"][r] == np.mean(pred == cl), r
        assert set(np.unique(cl)) <= set(np.unique(init[r]))
        assert np.all(np.diff(out["vect_acc"][r]) >= 0)


def test_learned_is_deterministic():
    """The candidate list is filled by atomics in any order; each row's score
    depends only on its own features, so two launches agree exactly."""
    x, rows, init, folds = _centre_batch(n_runs=6)
    o = K.core_options(**OPTS)
    a = K.core_knn_batch_learned(x, rows, init, o, K.mlp_init(3), fallback=0.25, fold_ids=folds)
    b = K.core_knn_batch_learned(x, rows, init, o, K.mlp_init(3), fallback=0.25, fold_ids=folds)
    assert all(np.array_equal(u, v) for u, v in zip(a["clbest"], b["clbest"]))
    assert np.array_equal(a["vect_acc"], b["vect_acc"])


def test_learned_proposal_gains_over_uniform():
    """What the trained network adds: trained on the plain form's traces of 24
    other runs, the learned form's mean best CV accuracy after 10 cycles (4
    proposal seeds x 24 runs, KODAMA-style starting labels) beats the uniform
    device form by more than half a point, and is not below the plain form
    (observed 0.9959 against 0.9826 and 0.9825)."""
    tx, trows, tinit, tfolds = _centre_batch(seed=77)
    feat, tgt, _ = K.mlp_collect_traces(tx, trows, tinit, K.core_options(**OPTS), fold_ids=tfolds)
    params, losses = K.mlp_train(feat, tgt, epochs=20, batch=256, learning_rate=0.05, momentum=0.9, seed=5)
    assert losses[-1] < losses[0]
    x, rows, init, folds = _centre_batch(seed=31)
    means = {"plain": [], "device": [], "learned": []}
    for seed in (1, 2, 3, 4):
        o = K.core_options(**dict(OPTS, cycles=10, seed=seed, cv_seed=seed))
        means["plain"].append(K.core_knn_batch(x, rows, init, o, mode="plain", fold_ids=folds)["accbest"].mean())
        means["device"].append(K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds)["accbest"].mean())
        out = K.core_knn_batch_learned(x, rows, init, o, params, fallback=0.25, fold_ids=folds)
        means["learned"].append(out["accbest"].mean())
    m = {k: float(np.mean(v)) for k, v in means.items()}
    assert m["learned"] > m["device"] + 0.005, m
    assert m["learned"] >= m["plain"], m


def _ablation_params():
    """Parameters whose scores are uniform over the candidate list: all weights
    zero, so every candidate gets the same keep probability (bias -80 on the
    keep slot) and its top-voted other label as the target."""
    p = np.zeros(MO.PARAMS, np.float32)
    p[MO.B3 + MO.KEEP] = -80.0
    return p


def test_learned_gain_splits_into_candidates_and_network():
    """Ablation of the learned proposal (3 cycles, no uniform fallback, 8
    proposal seeds x 24 runs). Restricting the proposal to the rows whose label
    differs from their prediction, with uniform scores over that list, is most
    of the gain over the uniform device form (observed about +7 points); a
    network trained on the outcome traces (the rank of each candidate's label
    in its run's final best labelling) adds a smaller gain on top of the same
    list, checked here per run, paired on seed and run (observed device
    0.8390, list with uniform scores 0.9170, list with the network 0.9194:
    +0.24 points, better in 107 of 192 runs, worse in 72)."""
    tx, trows, tinit, tfolds = _centre_batch(seed=77)
    feat, tgt, _ = K.mlp_collect_traces(tx, trows, tinit, K.core_options(**OPTS), fold_ids=tfolds, targets="outcome")
    params, losses = K.mlp_train(feat, tgt, epochs=20, batch=256, learning_rate=0.05, momentum=0.9, seed=5)
    assert losses[-1] < losses[0]
    x, rows, init, folds = _centre_batch(seed=31)
    acc = {"device": [], "ablation": [], "network": []}
    for seed in range(1, 9):
        o = K.core_options(**dict(OPTS, cycles=3, seed=seed, cv_seed=seed))
        acc["device"].append(K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds)["accbest"])
        acc["ablation"].append(K.core_knn_batch_learned(x, rows, init, o, _ablation_params(), fallback=0.0,
                                                        fold_ids=folds)["accbest"])
        acc["network"].append(K.core_knn_batch_learned(x, rows, init, o, params, fallback=0.0,
                                                       fold_ids=folds)["accbest"])
    a = {k: np.concatenate(v) for k, v in acc.items()}
    m = {k: float(v.mean()) for k, v in a.items()}
    diff = a["network"] - a["ablation"]
    print(f"ablation: device {m['device']:.5f} candidates+uniform {m['ablation']:.5f} "
          f"candidates+network {m['network']:.5f}; network - uniform per run: mean {diff.mean():+.5f}, "
          f"better in {int((diff > 0).sum())} of {diff.size}, worse in {int((diff < 0).sum())}")
    assert m["ablation"] > m["device"] + 0.02, m
    assert diff.mean() > 0.0, m
