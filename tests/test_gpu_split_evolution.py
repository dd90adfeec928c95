"""Runs of the device-resident evolution split over several workgroups: the
members replay the same proposal, absorption and acceptance and each votes a
slice of the rows, exchanging predictions and match counts through device
memory. The results must be bit-identical to one workgroup per run, for the
uniform and learned proposals and for runs beyond 8191 rows."""
import numpy as np
import pytest

import kodama_amd as K
from kodama_amd import datasets as D

from test_gpu_device_evolution import OPTS

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore_split():
    yield
    K.set_evolution_split(0)


def _batch(sizes, n=40000, seed=13):
    x, _ = D.config3(n=n, seed=seed)
    rows, init, folds = [], [], []
    for r, L in enumerate(sizes):
        rw = np.sort(np.argsort(D.uniform(seed, 100 + r, n), kind="stable")[:L]).astype(np.int32)
        c = x[rw[np.argsort(D.uniform(seed, 300 + r, L), kind="stable")[:60]]]
        idx, _ = K.exact_knn_search(c, x[rw], 1, "euclidean")
        rows.append(rw)
        init.append((idx[:, 0] + 1).astype(np.int32))
        folds.append((np.argsort(D.uniform(seed, 500 + r, L), kind="stable") % 10).astype(np.int32))
    return x, rows, init, folds


def _both(run, runs):
    """The same evolution with one workgroup per run and split automatically."""
    K.set_evolution_split(1)
    one = run()
    assert K.last_evolution_workgroups() == runs
    K.set_evolution_split(0)
    split = run()
    workgroups = K.last_evolution_workgroups()
    return one, split, workgroups


def _assert_equal(one, split):
    for a, b in zip(one["clbest"], split["clbest"]):
        assert np.array_equal(a, b)
    assert np.array_equal(one["accbest"], split["accbest"])
    assert np.array_equal(one["vect_acc"], split["vect_acc"])


def test_split_runs_equal_single_workgroup_runs():
    sizes = [900, 1500, 2600, 3800, 5000, 4400, 2048, 3100]
    x, rows, init, folds = _batch(sizes)
    o = K.core_options(**dict(OPTS, cycles=10, k=30))
    one, split, workgroups = _both(
        lambda: K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds), len(rows))
    # Every run of 1024 rows or more has at least two members.
    assert workgroups >= len(rows) + sum(1 for L in sizes if L >= 1024)
    _assert_equal(one, split)


@pytest.mark.parametrize("k", [8, 40])
def test_split_other_neighbour_counts(k):
    sizes = [2500, 4100, 1200]
    x, rows, init, folds = _batch(sizes, seed=17)
    o = K.core_options(**dict(OPTS, cycles=6, k=k))
    one, split, workgroups = _both(
        lambda: K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds), len(rows))
    assert workgroups > len(rows)
    _assert_equal(one, split)


def test_split_learned_proposal_equal():
    sizes = [3000, 4700, 1800]
    x, rows, init, folds = _batch(sizes, seed=19)
    o = K.core_options(**dict(OPTS, cycles=8, k=30))
    params = K.mlp_init(5)
    one, split, workgroups = _both(
        lambda: K.core_knn_batch_learned(x, rows, init, o, params, fallback=0.25, fold_ids=folds), len(rows))
    assert workgroups > len(rows)
    _assert_equal(one, split)


def test_split_large_runs_equal_and_capped():
    """Runs beyond 8191 rows (global-scratch layout); a cap of two members per
    run is honoured and gives the same results again."""
    sizes = [10000, 12500]
    x, rows, init, folds = _batch(sizes, seed=23)
    o = K.core_options(**dict(OPTS, cycles=5, k=30))
    run = lambda: K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds)
    one, split, workgroups = _both(run, len(rows))
    assert workgroups == 2 * 8
    _assert_equal(one, split)
    K.set_evolution_split(2)
    capped = run()
    assert K.last_evolution_workgroups() == 2 * 2
    _assert_equal(one, capped)


def test_split_control_rejects_out_of_range():
    with pytest.raises(ValueError):
        K.set_evolution_split(9)
