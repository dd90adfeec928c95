"""Split-BF16 MFMA candidates + exact FP32 rerank: neighbour ids and distances
bit-identical to the exact scan (the oracle restates the reference's order),
including the queries the candidate lists cannot certify (exact fallback)."""
import numpy as np
import pytest

import kodama_amd as K
from kodama_amd import datasets as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("metric,name", [(2, "euclidean"), (0, "cosine"), (1, "inner_product")])
def test_graph_via_mfma_matches_oracle(oracle, metric, name):
    x, _, _ = D.config1(n=6000, seed=17)
    g = K.KODAMA_knn_graph(x, k=20, metric=name)
    fb = K.last_knn_fallbacks()
    assert 0 <= fb < 60, fb  # the MFMA path ran and certified almost every query
    idx, dist = oracle.knn_graph(x, 20, metric)
    assert np.array_equal(g["indices"], idx)
    assert np.array_equal(g["distances"], dist)


@pytest.mark.parametrize("k,metric,name", [(33, 2, "euclidean"), (100, 2, "euclidean"), (100, 0, "cosine")])
def test_wide_graph_via_mfma_matches_oracle(oracle, k, metric, name):
    """More than 32 neighbours: 16-entry candidate lists per lane column (256
    candidates per query), the same certification and exact rerank."""
    x, _ = D.config3(n=6000, p=50, seed=k)
    g = K.KODAMA_knn_graph(x, k=k, metric=name)
    fb = K.last_knn_fallbacks()
    assert 0 <= fb < 600, fb  # the MFMA path ran and certified most queries
    idx, dist = oracle.knn_graph(x, k, metric)
    assert np.array_equal(g["indices"], idx)
    assert np.array_equal(g["distances"], dist)


def test_search_with_exclusion_and_dimension_tail(oracle):
    x, _ = D.config3(n=5700, p=40, seed=5)
    train, query = x[:5000], x[5000:]
    ex = np.arange(700, dtype=np.int32) * 7
    idx, dist = K.exact_knn_search(train, query, 32, "euclidean", exclude=ex)
    assert K.last_knn_fallbacks() >= 0
    ridx, rdist = oracle.exact_knn(train, query, 32, 2, exclude=ex)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(dist, rdist)
    assert not np.any(idx == ex[:, None])


def test_duplicate_rows_take_the_exact_fallback(oracle):
    """1000 copies of each of 5 rows: every column list fills with (near-)zero
    candidate values inside the bound, so no query can be certified and the
    exact scan answers them all (ties resolved by index)."""
    base, _, _ = D.config1(n=5, seed=9)
    x = np.repeat(base, 1000, axis=0)
    g = K.KODAMA_knn_graph(x, k=24)
    assert K.last_knn_fallbacks() == x.shape[0]
    idx, dist = oracle.knn_graph(x, 24, 2)
    assert np.array_equal(g["indices"], idx)
    assert np.array_equal(g["distances"], dist)


@pytest.mark.parametrize("p,k,mfma", [(64, 32, True), (65, 16, False), (30, 33, True), (64, 128, True)])
def test_path_boundaries(oracle, p, k, mfma):
    x, _ = D.config3(n=4200, p=p, seed=p + k)
    idx, dist = K.exact_knn_search(x[:4100], x[4100:], k, "euclidean")
    assert (K.last_knn_fallbacks() >= 0) == mfma
    ridx, rdist = oracle.exact_knn(x[:4100], x[4100:], k, 2)
    assert np.array_equal(idx, ridx) and np.array_equal(dist, rdist)
