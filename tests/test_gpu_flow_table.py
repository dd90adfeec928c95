"""Flow counts of the device-resident evolution's absorption in the LDS table
(open addressing, more slots than the run has rows) against the per-workgroup
HBM table: the results must be bit-identical, for one and several workgroups
per run, the uniform and learned proposals, k up to 64 and many classes (long
probe chains). The large-run layout and states that leave no LDS room take
the HBM table."""
import numpy as np
import pytest

import kodama_amd as K
from kodama_amd import datasets as D

from test_gpu_device_evolution import OPTS
from test_gpu_split_evolution import _assert_equal, _batch

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore():
    yield
    K.set_lds_flow_table(True)
    K.set_evolution_split(0)


def _both(run):
    K.set_lds_flow_table(True)
    lds = run()
    used = K.last_evolution_lds_flows()
    K.set_lds_flow_table(False)
    hbm = run()
    assert not K.last_evolution_lds_flows()
    return lds, hbm, used


@pytest.mark.parametrize("k", [15, 30, 40])
@pytest.mark.parametrize("split", [1, 0])
def test_lds_flow_table_equals_hbm_table(k, split):
    sizes = [900, 1500, 2600, 3800, 5000, 4400]
    x, rows, init, folds = _batch(sizes, seed=29)
    K.set_evolution_split(split)
    o = K.core_options(**dict(OPTS, cycles=12, k=k))
    lds, hbm, used = _both(lambda: K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds))
    assert used
    _assert_equal(lds, hbm)


def test_lds_flow_table_many_classes():
    """Random starting labels over 700 classes: thousands of distinct flow
    entries per cycle, many sharing a probe chain."""
    sizes = [3000, 4000, 2500]
    x, rows, _, folds = _batch(sizes, seed=31)
    init = [D.randint(31, 900 + r, len(rw), 1, 700).astype(np.int32) for r, rw in enumerate(rows)]
    o = K.core_options(**dict(OPTS, cycles=10, k=30))
    lds, hbm, used = _both(lambda: K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds))
    assert used
    _assert_equal(lds, hbm)


def test_lds_flow_table_learned_proposal():
    sizes = [3000, 4700, 1800]
    x, rows, init, folds = _batch(sizes, seed=37)
    o = K.core_options(**dict(OPTS, cycles=8, k=30))
    params = K.mlp_init(5)
    lds, hbm, _ = _both(lambda: K.core_knn_batch_learned(x, rows, init, o, params, fallback=0.25, fold_ids=folds))
    _assert_equal(lds, hbm)


def test_large_runs_take_the_hbm_table():
    sizes = [10000, 3000]
    x, rows, init, folds = _batch(sizes, seed=41)
    o = K.core_options(**dict(OPTS, cycles=4, k=30))
    K.set_lds_flow_table(True)
    out = K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=folds)
    assert not K.last_evolution_lds_flows()
    assert np.all(out["accbest"] > 0.5)


def test_flow_table_control_rejects_out_of_range():
    with pytest.raises(ValueError):
        K.N.call("kodama_hip_set_lds_flow_table", K.C.c_int32(2))
