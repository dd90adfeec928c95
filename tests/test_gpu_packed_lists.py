"""MFMA kNN candidate lists as packed 32-bit keys (the candidate value's order
bits with the row's index within its lane column in the low bits) and as
separate values and indices: both must give the oracle's neighbours bit for
bit, for the global search, the per-fold tables of a batch of runs (the
rerank finds each query's run), every list length, negative candidate values
(inner product), and train sets at the widest packed index (32768 rows) and
past it (8-entry lists packed per 8192-row segment and merged into (value,
row) pairs at each segment's end; 16-entry lists separate)."""
import numpy as np
import pytest

import kodama_amd as K
from kodama_amd import datasets as D

from test_gpu_fold_tables import _runs

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore():
    yield
    K.set_mfma_packed_lists(True)
    K.set_fold_table_schedule("tuned", 0, 0)


@pytest.mark.parametrize("packed", [True, False])
@pytest.mark.parametrize("k,metric,name", [(20, 2, "euclidean"), (20, 1, "inner_product"), (100, 0, "cosine")])
def test_graph_lists_match_oracle(oracle, packed, k, metric, name):
    x, _ = D.config3(n=6000, p=50, seed=k + metric)
    K.set_mfma_packed_lists(packed)
    g = K.KODAMA_knn_graph(x, k=k, metric=name)
    assert K.last_knn_fallbacks() < 600
    idx, dist = oracle.knn_graph(x, k, metric)
    assert np.array_equal(g["indices"], idx)
    assert np.array_equal(g["distances"], dist)


@pytest.mark.parametrize("packed", [True, False])
@pytest.mark.parametrize("list_", [8, 16])
def test_fold_table_lists_match_oracle(oracle, packed, list_):
    x, _ = D.config3(n=6000, p=50, seed=3)
    rows, fids = _runs(x, [2500, 700, 1800, 64], folds=10, seed=list_)
    want = [oracle.precompute_knn_cv(x[rw], f, 30, K._metric("euclidean")) for rw, f in zip(rows, fids)]
    K.set_mfma_packed_lists(packed)
    K.set_fold_table_schedule("tuned", list_, 8)
    nbr, score, used = K.knn_fold_tables(x, rows, fids, 30, "euclidean", path="mfma")
    assert used == "mfma"
    for r, (wn, ws) in enumerate(want):
        assert np.array_equal(nbr[r], wn), r
        assert np.array_equal(score[r], ws), r


@pytest.mark.parametrize("train_rows", [32768, 32769])
def test_widest_packed_index_and_past_it(oracle, train_rows):
    """32768 train rows fill 11 index bits (column rows 0..2047); one more takes
    the segments (five, the last of one row). Queries near the last train rows
    make the top column indices decode into neighbours."""
    x, _ = D.config3(n=train_rows + 400, p=32, seed=train_rows)
    train = x[:train_rows]
    rng = np.random.default_rng(7)
    query = np.concatenate([train[-300:] + rng.normal(0.0, 1e-3, size=(300, 32)).astype(np.float32),
                            x[train_rows:train_rows + 100]]).astype(np.float32)
    want_idx, want_dist = oracle.exact_knn(train, query, 16, 2)
    for packed in (True, False):
        K.set_mfma_packed_lists(packed)
        idx, dist = K.exact_knn_search(train, query, 16, "euclidean")
        assert K.last_knn_fallbacks() < 100, packed
        assert np.array_equal(idx, want_idx), packed
        assert np.array_equal(dist, want_dist), packed
    assert np.any(want_idx[:300] >= train_rows - 16)


@pytest.mark.parametrize("k,metric,name", [(20, 1, "inner_product"), (32, 2, "euclidean"), (40, 2, "euclidean")])
def test_segmented_lists_with_exclusion(oracle, k, metric, name):
    """40000 train rows: k <= 32 merges segment keys into pairs (negative
    values for the inner product), k = 40 keeps separate 16-entry lists; every
    query excludes one train row."""
    x, _ = D.config3(n=40600, p=50, seed=k)
    train, query = x[:40000], x[40000:]
    ex = (np.arange(600, dtype=np.int32) * 61) % 40000
    want_idx, want_dist = oracle.exact_knn(train, query, k, metric, exclude=ex)
    for packed in (True, False):
        K.set_mfma_packed_lists(packed)
        idx, dist = K.exact_knn_search(train, query, k, name, exclude=ex)
        assert K.last_knn_fallbacks() < 300, packed
        assert np.array_equal(idx, want_idx), packed
        assert np.array_equal(dist, want_dist), packed
        assert not np.any(idx == ex[:, None])


def test_duplicates_fall_back_with_packed_lists(oracle):
    """Equal rows make equal candidate values; their keys stay distinct by
    their column index, every list fills inside the bound and each query
    takes the exact scan."""
    base, _, _ = D.config1(n=5, seed=11)
    x = np.repeat(base, 1000, axis=0)
    K.set_mfma_packed_lists(True)
    g = K.KODAMA_knn_graph(x, k=24)
    assert K.last_knn_fallbacks() == x.shape[0]
    idx, dist = oracle.knn_graph(x, 24, 2)
    assert np.array_equal(g["indices"], idx)
    assert np.array_equal(g["distances"], dist)


@pytest.mark.parametrize("packed", [True, False])
def test_second_pass_certifies_what_eight_entries_cannot(oracle, packed):
    """150 copies of each of 40 rows: every lane column holds about 9 copies
    at distance 0, so every 8-entry list fills inside the bound and all 6000
    queries are flagged; the second pass with 16-entry lists certifies them
    (ties at distance 0 ranked by index) and nothing reaches the scan."""
    base, _ = D.config3(n=40, p=24, seed=5)
    x = np.repeat(base, 150, axis=0)
    K.set_mfma_packed_lists(packed)
    g = K.KODAMA_knn_graph(x, k=20)
    assert K.last_knn_fallbacks() == 0
    idx, dist = oracle.knn_graph(x, 20, 2)
    assert np.array_equal(g["indices"], idx)
    assert np.array_equal(g["distances"], dist)


def test_packed_list_control_rejects_out_of_range():
    with pytest.raises(ValueError):
        K.N.call("kodama_hip_set_mfma_packed_lists", K.C.c_int32(2))
