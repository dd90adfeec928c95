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
