"""The workload the headline rate is measured on, at its full size: X =
100,000 x 50 (20-cluster Zipf mixture), kNN k = 30, 100 runs of U[1000, 5000]
landmarks with 10 uneven folds, starting labels the nearest of up to 300
landmark centres (bench.make_workload). The device-resident form, with its
large runs split over several workgroups, keeps the invariants the algorithm
fixes on sampled runs, matches one workgroup per run bit for bit, and reaches
the plain form's mean best CV accuracy within half a point."""
import numpy as np
import pytest

import bench
import kodama_amd as K

pytestmark = pytest.mark.gpu

OPTS = dict(cycles=10, k=30, folds=10, stratified=0, seed=1, cv_seed=1, evolutionary_search=1,
            guarded_diversity=1, many_to_one_absorption=1)


def test_c3_batch_device_form(oracle):
    x, rows, folds, init = bench.make_workload(100, 100000, 50, 30)
    o = K.core_options(**OPTS)
    try:
        K.set_evolution_split(0)
        dev = K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds)
        assert K.last_evolution_workgroups() > len(rows)
        K.set_evolution_split(1)
        one = K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds)
    finally:
        K.set_evolution_split(0)
    for a, b in zip(dev["clbest"], one["clbest"]):
        assert np.array_equal(a, b)
    assert np.array_equal(dev["vect_acc"], one["vect_acc"])
    plain = K.core_knn_batch(x, rows, init, o, mode="plain", fold_ids=folds)
    for r in range(len(rows)):
        assert set(np.unique(dev["clbest"][r])) <= set(np.unique(init[r]))
        assert np.all(np.diff(dev["vect_acc"][r]) >= 0)
        assert dev["vect_acc"][r][-1] == dev["accbest"][r]
    for r in (0, 37, 64, 99):
        nbr, score = oracle.precompute_knn_cv(x[rows[r]], folds[r], 30)
        cl = dev["clbest"][r]
        assert dev["accbest"][r] == np.mean(oracle.knn_vote(nbr, score, folds[r], cl) == cl), r
    assert abs(float(np.mean(dev["accbest"])) - float(np.mean(plain["accbest"]))) < 0.005


def test_c3_member_assignment_from_measured_times():
    """The second evolution of a batch shape assigns the split runs' members
    from the first one's measured run times (256 workgroups across 100 runs,
    so the costs decide which runs get them); the results are identical
    whatever the assignment, and a third evolution refines it again."""
    x, rows, folds, init = bench.make_workload(100, 100000, 50, 30, salt=7)
    o = K.core_options(**dict(OPTS, cycles=4))
    K.set_evolution_split(0)
    outs = [K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds) for _ in range(3)]
    assert K.last_evolution_workgroups() > len(rows)
    for out in outs[1:]:
        for a, b in zip(outs[0]["clbest"], out["clbest"]):
            assert np.array_equal(a, b)
        assert np.array_equal(outs[0]["accbest"], out["accbest"])
        assert np.array_equal(outs[0]["vect_acc"], out["vect_acc"])
