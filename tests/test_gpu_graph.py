"""HIP exact self-kNN graph and KODAMA dissimilarity against the oracle, the
reference's golden fixture and a brute force, bit-exact."""
import numpy as np
import pytest

import kodama_amd as K
from conftest import golden
from kodama_amd import datasets as D
from test_oracle import dissimilarity_brute_force

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("metric,name", [(2, "euclidean"), (0, "cosine"), (1, "inner_product")])
def test_knn_graph_matches_oracle(oracle, metric, name):
    x, _, _ = D.config1(n=900, seed=41)
    g = K.KODAMA_knn_graph(x, k=20, metric=name)
    idx, dist = oracle.knn_graph(x, 20, metric)
    assert np.array_equal(g["indices"], idx)
    assert np.array_equal(g["distances"], dist)
    assert g["indices"].min() >= 1 and not np.any(g["indices"] == np.arange(1, 901)[:, None])


def test_knn_graph_small_and_duplicates(oracle):
    x, _, _ = D.config1(n=60, seed=5)
    x = np.repeat(x[:20], 3, axis=0)  # exact duplicates: ties resolved by index
    g = K.KODAMA_knn_graph(x, k=100)  # clamps to n - 1
    idx, dist = oracle.knn_graph(x, 100, 2)
    assert g["neighbors"] == 59
    assert np.array_equal(g["indices"], idx) and np.array_equal(g["distances"], dist)


def test_dissimilarity_matches_golden():
    g = golden("dissimilarity.npz")
    idx, dist = K.KODAMA_dissimilarity(g["idx"], g["dist"], g["res"])
    assert np.array_equal(idx, g["out_idx"])
    assert np.array_equal(dist, g["out_dist"])


@pytest.mark.parametrize("runs", [1, 5, 64, 301])
def test_dissimilarity_edge_cases(oracle, runs):
    """Runs not a multiple of 4 and above 256 (two label chunks), unassigned
    labels (0), invalid graph entries (-1 ids, out-of-range ids, +inf
    distances) and exact (distance, id) ties."""
    n, K_ = 400, 17
    rng = np.random.default_rng(runs)
    x, _, _ = D.config1(n=n, seed=7)
    idx, dist = oracle.knn_graph(x, K_)
    idx = idx.copy()
    dist = dist.copy()
    idx[3, 5] = -1
    idx[8, 0] = n + 5
    dist[11, 2] = np.inf
    dist[20, :4] = dist[20, 0]  # equal distances: ids decide
    res = rng.integers(0, 4, size=(runs, n)).astype(np.int32)
    got_i, got_d = K.KODAMA_dissimilarity(idx, dist, res)
    ref_i, ref_d = oracle.dissimilarity(idx, dist, res)
    assert np.array_equal(got_i, ref_i)
    assert np.array_equal(got_d, ref_d)
    if runs <= 64:
        bi, bd = dissimilarity_brute_force(idx, dist, res)
        assert np.array_equal(got_i, bi)
        assert np.array_equal(got_d, bd)


@pytest.mark.parametrize("metric", ["euclidean", "cosine"])
def test_graph_handle_materialize_and_dissimilarity(oracle, metric):
    """The device-resident graph handle against the oracle: materialised it is
    the oracle's exact public graph (1-based, public distances); after its
    in-place dissimilarity it is the oracle's dissimilarity rewrite of that
    graph; released twice is harmless."""
    x, _, _ = D.config1(n=900, seed=17)
    m = K._native.METRICS[metric]
    want_idx, want_dist = oracle.knn_graph(x, 12, m)
    h = K.KODAMAGraphHandle(x, k=12, metric=metric)
    assert (h.samples, h.neighbors) == (900, 12)
    idx, dist = h.materialize()
    assert np.array_equal(idx, want_idx) and np.array_equal(dist, want_dist)
    res = np.stack([D.randint(60 + r, 1, 900, 0, 5) for r in range(6)]).astype(np.int32)
    ref_idx, ref_dist = oracle.dissimilarity(want_idx, want_dist, res)
    h.dissimilarity(res)
    idx2, dist2 = h.materialize()
    assert np.array_equal(idx2, ref_idx) and np.array_equal(dist2, ref_dist)
    h.release()
    h.release()
    assert not h.valid()
    with pytest.raises(ValueError):
        h.materialize()
