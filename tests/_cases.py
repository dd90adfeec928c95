"""Shared option sets and comparison helpers of the test-suite."""
import numpy as np

CORE_FIELDS = ("clbest", "clbest_dirty", "cvpredbest", "fold_assignments", "vect_acc", "vect_score",
               "proposal_size", "active_classes", "accepted", "improving_acceptance",
               "temperature_acceptance")
COUNTERS = (
    "cycles_completed", "proposals_evaluated", "best_state_updates",
    "current_state_accepts", "stochastic_state_attempts", "stochastic_state_accepts",
    "current_state_rejections", "coarsening_moves", "absorption_moves",
    "transition_attempted", "transition_accepted", "many_to_one_attempted",
    "many_to_one_accepted", "pls_coarsening_attempted", "pls_coarsening_accepted",
    "cv_evaluations", "success",
)
# R Core API defaults (kodama_r.cpp:1118-1174) and the per-run flags of
# KODAMA.matrix (kodama_matrix.cpp:3108-3128).
CORE_API = dict(stratified=1, evolutionary_search=0, guarded_diversity=0, many_to_one_absorption=0)
MATRIX_RUN = dict(stratified=0, evolutionary_search=1, guarded_diversity=1, many_to_one_absorption=1)


def assert_same_core(a: dict, b: dict, where: str = "") -> None:
    """Bit-exact equality of two CoreResult dicts."""
    for f in CORE_FIELDS:
        assert np.array_equal(np.asarray(a[f]), np.asarray(b[f])), f"{where}: field {f} differs"
    for c in COUNTERS:
        assert a[c] == b[c], f"{where}: counter {c}: {a[c]} != {b[c]}"
    assert a["accbest"] == b["accbest"], where
    assert a["scorebest"] == b["scorebest"] or (np.isinf(a["scorebest"]) and np.isinf(b["scorebest"])), where


def golden_core(g, prefix: str) -> dict:
    out = {f: g[prefix + f] for f in CORE_FIELDS}
    out.update({c: int(v) for c, v in zip(COUNTERS, g[prefix + "counters"])})
    out["accbest"] = float(g[prefix + "accbest"])
    out["scorebest"] = float(g[prefix + "scorebest"])
    return out


def check_invariants(r: dict, cycles: int, evolutionary: bool) -> None:
    """Counter invariants of tests/testthat/test-kodama-matrix.R:262-267, 627."""
    assert r["proposals_evaluated"] == r["cycles_completed"] == cycles
    assert r["cv_evaluations"] == cycles + 1
    assert r["stochastic_state_accepts"] <= r["stochastic_state_attempts"]
    if not evolutionary:
        assert r["current_state_accepts"] + r["current_state_rejections"] == 0
    vscore = np.asarray(r["vect_score"])
    assert np.all(np.diff(vscore) >= 0), "best score trace must be monotone"


# ---- KODAMA.matrix end-to-end cases pinned against the reference's own
# KODAMAMatrix_CPU / KODAMAMatrixFromGraphData_CPU (deterministic there: the
# caller's graph drives the kNN Core, and the PLS-LDA classifier uses no graph).

def matrix_case_inputs(name):
    """Seeded inputs of a matrix case: (x, graph or None, index_base, W,
    constrain, fixed). Graphs are the oracle's exact public graph (1-based)."""
    from kodama_amd import datasets as D
    import oracle as O
    spec = MATRIX_CASES[name]
    if spec.get("data") == "pls":
        x, _ = D.config2(n=2000)
        x = x[:900]
    else:
        x, _, _ = D.config1(n=spec.get("n", 600), seed=21)
    graph, base = None, -1
    if spec.get("graph"):
        gi, gd = O.Oracle().knn_graph(x, spec["graph"], spec.get("metric", 2))
        if spec.get("gaps"):
            gi = gi.copy()
            gd = gd.copy()
            gi[::5, 3] = -1
            gd[1::5, 2] = np.nan
            gi[2::7, 0] = gi[2::7, 1]  # a repeated neighbour
        if spec.get("zero_based"):
            gi = np.where(gi > 0, gi - 1, gi)
            base = 0
        graph = (gi, gd)
    n = x.shape[0]
    rng = np.random.default_rng(spec.get("seed", 7) + 100)
    W = rng.integers(1, 6, n).astype(np.int32) if spec.get("W") else None
    cons = (np.arange(n) // 3 + 1000).astype(np.int32) if spec.get("constrain") else None
    fixed = (rng.random(n) < 0.1).astype(np.int32) if spec.get("fixed") else None
    return x, graph, base, W, cons, fixed


def matrix_case_options(name):
    spec = MATRIX_CASES[name]
    return dict(runs=spec.get("runs", 3), cycles=spec.get("cycles", 8), landmarks=spec.get("landmarks", 300),
                splitting=spec.get("splitting", 20), folds=spec.get("folds", 5), k=spec.get("k", 10),
                metric=spec.get("metric", 2), graph_neighbors=spec.get("graph_neighbors", 40),
                seed=spec.get("seed", 7), classifier=spec.get("classifier", "knn"),
                components=spec.get("components", 5))


MATRIX_CASES = {
    "graph_euclidean": dict(graph=40),
    "graph_cosine_zero_based": dict(graph=30, metric=0, zero_based=True, k=12, seed=11, folds=4),
    "graph_narrow_k_above_width": dict(graph=8, k=12, graph_neighbors=40, seed=5),
    "graph_gaps": dict(graph=20, gaps=True, seed=13, cycles=10),
    "graph_starting_labels": dict(graph=40, W=True, seed=17),
    "graph_constrain_fixed": dict(graph=40, constrain=True, fixed=True, seed=19),
    "graph_all_inputs": dict(graph=25, W=True, constrain=True, fixed=True, seed=23, metric=1),
    "pls": dict(classifier="pls_lda", data="pls", landmarks=400, seed=3),
    "pls_starting_labels_constrain": dict(classifier="pls_lda", data="pls", landmarks=300, W=True,
                                          constrain=True, fixed=True, seed=29, components=3),
    "pls_more_components": dict(classifier="pls_lda", data="pls", landmarks=500, components=8, folds=4,
                                cycles=6, seed=31),
}


# ---- PLSDACV and general PLS fixtures (reference PLSDACV_CPU / PLS_CPU) ------
PLS_GENERAL_CASES = ((5, 1, 1), (3, 0, 1), (8, 1, 0), (2, 1, 1))  # (components, center, scale)
PLS_FIELDS = ("weights", "response_loadings", "scores", "coefficients", "fitted", "x_center", "x_scale",
              "y_center")


def pls_general_inputs():
    """Data of the PLSDACV / PLS fixtures: 1500 rows of the PLS workload, its
    labels and random 12-class labels, and a 4-response Y linear in three
    columns plus seeded noise."""
    from kodama_amd import datasets as D
    x2, truth2 = D.config2(n=2000)
    x = x2[:1500]
    noisy = D.randint(41, 1, 1500, 1, 12).astype(np.int32)
    mix = D.normal(42, 1, 3 * 4).reshape(3, 4)
    y = (x[:, :3].astype(np.float64) @ mix + 0.5 * D.normal(42, 2, 1500 * 4).reshape(1500, 4)).astype(np.float32)
    return x, truth2[:1500].astype(np.int32), noisy, y


def rel_err(a, b):
    """max |a - b| / max |b| (0 for two zero arrays)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b), initial=0.0))
