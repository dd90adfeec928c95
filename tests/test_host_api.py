"""Host-side parts of the C ABI that need no GPU: option resolution of
KODAMA.matrix (run_kodama_matrix's landmark and graph-width rules), argument
errors reported before any device work, and the network initialisation."""
import math

import numpy as np
import pytest

import kodama_amd as K
from kodama_amd import _native as N


def _sizes(n, **kw):
    import ctypes as C
    o = N.MatrixOptions(**dict(dict(runs=1, cycles=1, landmarks=10000, splitting=0, folds=5, k=30, metric=2,
                                    graph_neighbors=100, seed=1234, form=0, device=0, n_threads=0), **kw))
    L = C.c_int32(0)
    G = C.c_int32(0)
    N.call("kodama_hip_matrix_sizes", C.c_int32(n), C.byref(o), C.byref(L), C.byref(G))
    return L.value, G.value


def _reference_rule(n, landmarks=10000, graph_neighbors=100):
    """kodama_matrix.cpp:3340-3346 and 3463-3467, restated."""
    L = landmarks if landmarks > 0 else 10000
    if L >= n:
        L = math.ceil(n * 0.75)
    L = max(2, min(L, n - 1))
    req = graph_neighbors if graph_neighbors > 0 else 100
    return L, max(1, math.floor(min(L, n * 0.75 - 1.0, req)))


@pytest.mark.parametrize("n,kw", [(100000, {}), (1000, {}), (50, {}), (12, {}), (3, {}),
                                  (5000, dict(landmarks=300, graph_neighbors=0)),
                                  (5000, dict(landmarks=20, graph_neighbors=64)),
                                  (40000, dict(landmarks=-1, graph_neighbors=10))])
def test_matrix_sizes_follow_the_reference_rules(n, kw):
    if not N.lib_available():
        pytest.skip("libkodama_hip.so is not built")
    assert _sizes(n, **kw) == _reference_rule(n, **kw)


@pytest.mark.parametrize("kw,n,p", [(dict(k=0), 100, 5), (dict(k=65), 100, 5), (dict(folds=1), 100, 5),
                                    (dict(runs=0), 100, 5), (dict(cycles=-1), 100, 5),
                                    ({}, 2, 5), ({}, 100, 300)])
def test_matrix_argument_errors_are_einval(kw, n, p):
    if not N.lib_available():
        pytest.skip("libkodama_hip.so is not built")
    x = np.zeros((n, p), np.float32)
    with pytest.raises(K.InvalidArgument):
        K.KODAMA_matrix(x, **{dict(runs="M", cycles="Tcycle").get(a, a): v for a, v in kw.items()})


@pytest.mark.parametrize("case", ["base", "empty", "rows", "shape"])
def test_matrix_from_graph_argument_errors(case):
    """KODAMAMatrixFromGraphData rejects a bad graph before any device work:
    an index base other than 0, 1 or auto, a graph without columns, too few
    rows (EINVAL from the C ABI), and mismatched indices / distances."""
    if not N.lib_available():
        pytest.skip("libkodama_hip.so is not built")
    n = 2 if case == "rows" else 50
    x = np.zeros((n, 4), np.float32)
    width = 0 if case == "empty" else 5
    gi = np.ones((n, width), np.int32)
    gd = np.ones((n, width + (1 if case == "shape" else 0)), np.float32)
    if case == "shape":
        with pytest.raises(ValueError):
            K.KODAMA_matrix_from_graph(x, gi, gd)
        return
    with pytest.raises(K.InvalidArgument):
        K.KODAMA_matrix_from_graph(x, gi, gd, index_base=2 if case == "base" else "auto")


def test_graph_handle_null_is_einval():
    """A null graph handle is an argument error, and releasing it does nothing."""
    if not N.lib_available():
        pytest.skip("libkodama_hip.so is not built")
    import ctypes as C
    with pytest.raises(K.InvalidArgument):
        N.call("kodama_hip_graph_info", C.c_void_p(), None, None, None)
    with pytest.raises(K.InvalidArgument):
        N.call("kodama_hip_graph_dissimilarity", C.c_void_p(), None, C.c_int32(1))
    fn = N.lib().kodama_hip_graph_release
    fn.restype = None
    fn(C.c_void_p())


def test_mlp_init_is_he_normal_and_deterministic():
    if not N.lib_available():
        pytest.skip("libkodama_hip.so is not built")
    import mlp_oracle as MO
    a, b = K.mlp_init(7), K.mlp_init(7)
    assert np.array_equal(a, b) and not np.array_equal(a, K.mlp_init(8))
    assert np.all(a[MO.B1:MO.W2] == 0) and np.all(a[MO.B2:MO.W3] == 0) and np.all(a[MO.B3:] == 0)
    w1 = a[MO.W1:MO.B1]
    assert abs(w1.std() - math.sqrt(2 / 64)) < 0.01 and abs(w1.mean()) < 0.01
