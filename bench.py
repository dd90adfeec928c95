"""Benchmark: CV label-evolution steps/s of the KODAMA.matrix kNN loop on one MI355X.

Workload (the batched-M-runs configuration): X = 100,000 x 50 float32 from a
20-cluster Gaussian mixture (Zipf(1.1) cluster sizes, sigma 1, means
U[-4,4]^50); classifier kNN with k = 30; M = 100 runs, each on its own landmark
subset of U[1000, 5000] rows with 10 uneven folds (Dirichlet(1) fold shares);
starting labels = nearest of min(300, L) random landmark centres (one k-means
assignment step); Tcycle = --steps cycles. A step is one (run, cycle) CV
evaluation with its proposal and accept/reject; steps/s = M * cycles / wall time
of the label-evolution phase with its inputs in device memory (the one-time
per-fold exact kNN precompute and the device forms' label-independent set-up
are excluded and reported apart as precompute_seconds). The device forms'
upload of the starting labels and download of the results are outside that
clock and reported beside it (steps_per_s_with_copies); the plain form moves
labels between host and device every cycle by construction (host proposals
with the reference's streams, device votes), so its loop includes them.

Three forms run on the same inputs:
  plain   host proposals with the reference's random streams (per run
          bit-identical to CoreKNN), one batched device vote per cycle;
  device  proposal, absorption, vote, objective and acceptance on the device:
          one persistent launch (one workgroup per run, run state in LDS) for
          all cycles, counter-based random streams;
  learned the device form with the learned proposal, also one persistent
          launch: in each learned cycle the network scores the rows whose
          label differs from their prediction inside the kernel (a quarter of
          the cycles stay uniform). The network is first trained, untimed, on
          the plain form's accept/reject traces over 16 other runs.
A fourth form, pls, runs the batched CorePLSLDA (host proposals with the
reference streams, one FP64 PLS-LDA evaluation of all runs x folds per cycle)
on the PLS-LDA workload (X = 10,000 x 50, 100 runs of 2000 landmarks,
10 folds, 5 components); it is reported apart in `pls_lda_form`.
`value` is the device form's steps/s; every form's mean final CV accuracy is
reported ("equal accuracy" = within 0.5 points), with the cycles each form
needs to reach the plain form's final accuracy minus half a point.

The roofline object covers the dominant kernel of the reported form, timed with
HIP events on its launch stream. Algorithmic bytes follow the per-step cost of
the vote and accept/commit: per landmark row k neighbour ids and k scores of 4
bytes plus 7 words of label traffic, L * (8k + 28) bytes per evaluation:
  device  evo_persistent_kernel: one launch evaluates the starting labels and
          then runs every cycle, so bytes per launch = (cycles + 1) * sum(L) *
          (8k + 28); the kernel's own layout reads 2k bytes per row (16-bit
          ids, labels in LDS), reported beside it as *_16bit_ids;
  plain   vote_kernel: bytes per launch = sum(L) * (8k + 28) over the average
          launch duration.

`traffic` (HBM bytes per launch of that kernel) comes from the committed
counter summary profiles/evo_persistent_kernel_traffic.txt when it was
measured on this same workload (two rocprofv3 --pmc runs, FETCH_SIZE and
WRITE_SIZE, scaled by what the counters read for a plain copy of known size:
`python bench.py --copy-probe`); otherwise null.

cpu_baseline: the reference's own CPU label-evolution loop (oracle/_ref,
CoreKNNGraph_CPU on exact 64-NN landmark graphs of the same runs), run on
`cores` host threads, one run per thread at a time, in passes over all runs
until at least --cpu-seconds of wall time.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--runs 100]
    python bench.py --copy-probe     # the 256 MiB copy that calibrates the counters
    python bench.py --graph-rows 1000000   # exact 30-NN graph of 1M x 50 (own JSON line)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CV label-evolution steps/s (n=100k,k=30), plain vs tuned at equal CV accuracy"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(runs: int, n: int, p: int, k: int, folds: int = 10, salt: int = 0, data=None,
                  labels=None):
    """The batched-runs workload; `salt` moves the landmark, fold and
    starting-label streams (the data matrix stays the same). `data` replaces
    the 20-cluster mixture, `labels` (one per data row) the nearest-centre
    starting labels."""
    import kodama_amd as K
    from kodama_amd import datasets as D
    x = data if data is not None else D.config3(n=n, p=p)[0]
    n = x.shape[0]
    sizes = D.randint(6 + salt, 1, runs, 1000, 5000)
    rows, fold_ids, init = [], [], []
    for r in range(runs):
        L = int(sizes[r])
        keys = D.uniform(6 + salt, 100 + r, n)
        rw = np.sort(np.argsort(keys, kind="stable")[:L]).astype(np.int32)
        # Uneven folds: Dirichlet(1) shares, at least one row per fold.
        e = -np.log1p(-D.uniform(7 + salt, r, folds))
        share = e / e.sum()
        cnt = np.maximum(1, np.floor(share * (L - folds)).astype(np.int64) + 1)
        cnt[np.argmax(cnt)] += L - cnt.sum()
        perm = np.argsort(D.uniform(7 + salt, 1000 + r, L), kind="stable")
        f = np.empty(L, np.int32)
        f[perm] = np.repeat(np.arange(folds, dtype=np.int32), cnt)
        rows.append(rw)
        fold_ids.append(f)
        if labels is not None:
            init.append(np.asarray(labels[rw], np.int32))
            continue
        # Starting labels: nearest of min(300, L) random landmark centres.
        nc = min(300, L)
        centres = x[rw[np.argsort(D.uniform(8 + salt, r, L), kind="stable")[:nc]]]
        idx, _ = K.exact_knn_search(centres, x[rw], 1, "euclidean")
        init.append((idx[:, 0] + 1).astype(np.int32))
    return x, rows, fold_ids, init


def cpu_baseline(x, rows, init, k, cycles, cores, min_seconds):
    """Reference CoreKNNGraph_CPU label evolution, `cores` runs at a time, in
    passes over every run until min_seconds have elapsed."""
    sample_runs = len(rows)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import kodama_amd as K
    import oracle as O
    kind = "reference" if O.Reference.available() else "port"
    eng = O.Reference() if kind == "reference" else O.Oracle()
    graphs = []
    for r in range(sample_runs):
        xl = x[rows[r]]
        ex = np.arange(xl.shape[0], dtype=np.int32)
        idx, d2 = K.exact_knn_search(xl, xl, 64, "euclidean", exclude=ex)
        graphs.append((idx, np.sqrt(np.maximum(d2, 0)).astype(np.float32)))
    opt = O.core_options(cycles=cycles, k=k, folds=10, stratified=0, evolutionary_search=1,
                         guarded_diversity=1, many_to_one_absorption=1)

    def one(r):
        o = O.core_options(**{f: getattr(opt, f) for f, _ in O.CoreOptions._fields_})
        o.seed = 1 + r
        o.cv_seed = 1 + r
        idx, d = graphs[r]
        if kind == "reference":
            return eng.core_knn_graph(idx, d, init[r], o)["accbest"]
        nbr = np.where(idx >= 0, idx, -1)[:, :k]
        score = -(d[:, :k] ** 2)
        folds = O.Oracle().make_folds(init[r], 10, 0, 1 + r)
        return eng.core_knn_tables(nbr, score, folds, init[r], o)["accbest"]

    passes = 0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as pool:
        while True:
            accs = list(pool.map(one, range(sample_runs)))
            passes += 1
            if time.perf_counter() - t0 >= min_seconds:
                break
    wall = time.perf_counter() - t0
    steps = passes * sample_runs * cycles
    return dict(value=steps / wall, unit="steps/s", cores=cores, kind=kind,
                sample=f"{passes} passes over the {sample_runs} runs x {cycles} cycles of this workload "
                       f"(CoreKNNGraph_CPU on exact 64-NN landmark graphs), {cores} threads",
                wall_seconds=wall, mean_accbest=float(np.mean(accs)))


def pls_cpu_baseline(x, rows, init, kw, cycles, cores, min_seconds):
    """Reference CorePLSLDA_CPU (oracle/_ref) on the runs of the PLS-LDA
    workload, `cores` runs at a time, until min_seconds have elapsed (at least
    one group of runs)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if not O.Reference.available():
        return None
    eng = O.Reference()

    def one(r):
        o = O.core_options(cycles=cycles, **kw)
        o.seed = 1 + r
        o.cv_seed = 1 + r
        return eng.core_plslda(x[rows[r]], init[r], o)["accbest"]

    done, t0 = 0, time.perf_counter()
    accs = []
    with ThreadPoolExecutor(max_workers=cores) as pool:
        while True:
            group = [(done + i) % len(rows) for i in range(cores)]
            accs += list(pool.map(one, group))
            done += cores
            if time.perf_counter() - t0 >= min_seconds:
                break
    wall = time.perf_counter() - t0
    return dict(value=done * cycles / wall, unit="steps/s", cores=cores, kind="reference",
                sample=f"{done} runs x {cycles} cycles of this workload (CorePLSLDA_CPU), {cores} threads",
                wall_seconds=wall, mean_accbest=float(np.mean(accs)))


def copy_probe() -> int:
    """A plain device-to-device copy of 256 MiB, three times: the counter
    calibration (FETCH_SIZE / WRITE_SIZE per KiB actually moved)."""
    import torch
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3}), flush=True)
    return 0


def graph_bench(rows: int, k: int, lists_ab: bool = False) -> int:
    """Exact self-kNN graph of the 20-cluster mixture (config-4 generator,
    p = 50) through the split-BF16 MFMA path; one JSON line. With lists_ab,
    the graph is built with packed candidate lists, separate ones and packed
    again (one line each, and whether the graphs are identical)."""
    import kodama_amd as K
    from kodama_amd import datasets as D
    x, _ = D.config4(n=rows)
    K.KODAMA_knn_graph(x[:8192], k=k)  # warm-up (code objects, allocator)
    forms = (True, False, True) if lists_ab else (True,)
    first = None
    try:
        for packed in forms:
            K.set_mfma_packed_lists(packed)
            t0 = time.perf_counter()
            g = K.KODAMA_knn_graph(x, k=k)
            wall = time.perf_counter() - t0
            fb = K.last_knn_fallbacks()
            flops = 2.0 * rows * rows * 50 * 3  # three bf16 products per MAC (hi.hi, hi.lo, lo.hi)
            line = {"graph_rows": rows, "k": k, "p": 50, "seconds": wall, "fallback_queries": fb,
                    "candidate_bf16_tflops_per_s": flops / wall / 1e12, "neighbors": int(g["neighbors"])}
            if lists_ab:
                line["packed_lists"] = packed
                if first is None:
                    first = (g["indices"], g["distances"])
                else:
                    line["identical"] = bool(np.array_equal(first[0], g["indices"]) and
                                             np.array_equal(first[1], g["distances"]))
            print(json.dumps(line), flush=True)
    finally:
        K.set_mfma_packed_lists(True)
    return 0


def train_network(k: int, cycles: int = 50):
    """The learned proposal's network, trained (untimed) on the plain form's
    accept/reject traces of 16 runs of the 20-cluster workload."""
    import kodama_amd as K
    base = dict(k=k, folds=10, stratified=0, seed=1, cv_seed=1, evolutionary_search=1, guarded_diversity=1,
                many_to_one_absorption=1, n_threads=0)
    tx, trows, tfolds, tinit = make_workload(16, 100000, 50, k, salt=100)
    feat, tgt, _ = K.mlp_collect_traces(tx, trows, tinit, K.core_options(cycles=cycles, **base), fold_ids=tfolds)
    params, _ = K.mlp_train(feat, tgt, epochs=20, batch=256, learning_rate=0.05, momentum=0.9, seed=5)
    return params


def matrix_bench(rows: int, runs: int, cycles: int, k: int, forms) -> int:
    """End-to-end KODAMA.matrix (kNN) of the 20-cluster mixture (config-3
    generator, p = 50): L = 3000 landmarks per run, splitting 300, global graph
    of 100 neighbours; one JSON line per form with the stage timings."""
    import kodama_amd as K
    from kodama_amd import datasets as D
    x, _ = D.config3(n=rows)
    params = train_network(k) if "learned" in forms else None
    K.KODAMA_matrix(x[:4000], M=2, Tcycle=2, landmarks=1000, k=k)  # warm-up (code objects, allocator)
    for form in forms:
        t0 = time.perf_counter()
        out = K.KODAMA_matrix(x, M=runs, Tcycle=cycles, landmarks=3000, splitting=300, folds=5, k=k,
                              graph_neighbors=100, seed=1234, form=form, mlp_params=params)
        wall = time.perf_counter() - t0
        print(json.dumps({"matrix_rows": rows, "runs": runs, "cycles": cycles, "k": k, "form": form,
                          "seconds": wall, "mean_acc": float(np.mean(out["acc"])),
                          "stage_seconds": out["timings"]}), flush=True)
    return 0


def tables_bench(runs: int, n: int, k: int) -> int:
    """Per-fold kNN tables of the batched-runs workload built by every schedule
    the tuner chooses from (register scan; split-BF16 MFMA + exact rerank with
    8 or 16 candidates per lane column and 4 or 8 waves per workgroup, the
    lists as packed keys or as separate values and indices), twice each, then
    the tuned build; one JSON line per build with its wall time, the rows the
    MFMA path rescanned, and whether the tables equal the scan's."""
    import kodama_amd as K
    x, rows, fids, _ = make_workload(runs, n, 50, k)
    ref = None
    builds = [("scan", 0, 0, True)] + [("mfma", l, w, pk) for l in (8, 16) if l == 16 or k <= 32 for w in (8, 4)
                                       for pk in (True, False)]
    try:
        for path, lst, waves, packed in builds * 2 + [("tuned", 0, 0, True)]:
            K.set_fold_table_schedule("tuned", lst, waves)
            K.set_mfma_packed_lists(packed)
            t0 = time.perf_counter()
            nbr, score, used = K.knn_fold_tables(x, rows, fids, k, "euclidean", path=path)
            wall = time.perf_counter() - t0
            same = None
            if ref is None:
                ref = (nbr, score)
            else:
                same = all(np.array_equal(a, b) for a, b in zip(nbr, ref[0])) and \
                    all(np.array_equal(a, b) for a, b in zip(score, ref[1]))
            print(json.dumps({"tables": path, "schedule": K.last_fold_table_schedule(), "packed_lists": packed,
                              "runs": runs, "n": n, "k": k, "seconds": wall,
                              "mfma_fallback_rows": K.last_knn_fallbacks(), "equal_to_scan": same}), flush=True)
    finally:
        K.set_fold_table_schedule("tuned", 0, 0)
        K.set_mfma_packed_lists(True)
    return 0


def learned_curve_bench(runs: int, cycles: int, k: int) -> int:
    """The learned-proposal configuration: X = 200,000 x 50 (30-cluster Zipf
    mixture), starting labels the true components with 5 % label noise, M runs
    on landmark subsets of U[1000, 5000] rows with 10 uneven folds, `cycles`
    cycles, kNN k. The network is trained (untimed) on the plain form's traces
    of the 20-cluster workload (16 runs). One JSON line per form (uniform
    device form, learned form): mean best CV accuracy against wall time, at
    cycles 1, 2, 5, 10, 20, 50, 100 (wall time pro rata of the single launch)."""
    import kodama_amd as K
    from kodama_amd import datasets as D
    base = dict(k=k, folds=10, stratified=0, seed=1, cv_seed=1, evolutionary_search=1, guarded_diversity=1,
                many_to_one_absorption=1, n_threads=0)
    params = train_network(k)
    x, noisy, comp = D.config5()
    x, rows, fids, init = make_workload(runs, x.shape[0], 50, k, salt=200, data=x, labels=noisy)
    start = float(np.mean([np.mean(i == comp[r]) for i, r in zip(init, rows)]))  # agreement with the components
    for form in ("device", "learned"):
        run = (lambda o: K.core_knn_batch_learned(x, rows, init, o, params, fallback=0.25, fold_ids=fids)) \
            if form == "learned" else (lambda o: K.core_knn_batch(x, rows, init, o, mode="device", fold_ids=fids))
        run(K.core_options(cycles=5, **base))  # warm-up
        out = run(K.core_options(cycles=cycles, **base))
        curve = out["vect_acc"].mean(axis=0)
        pts = [c for c in (1, 2, 5, 10, 20, 50, 100) if c <= cycles]
        print(json.dumps({"learned_curve": form, "runs": runs, "cycles": cycles, "k": k,
                          "landmarks_total": int(sum(len(r) for r in rows)), "label_noise": 0.05,
                          "initial_agreement_with_components": start, "evolution_seconds": out["evolution_seconds"],
                          "steps_per_s": runs * cycles / out["evolution_seconds"],
                          "mean_accbest": float(np.mean(out["accbest"])),
                          "curve": [{"cycle": c, "seconds": out["evolution_seconds"] * c / cycles,
                                     "mean_accbest": float(curve[c - 1])} for c in pts]}), flush=True)
    return 0


def pls_workload(runs: int = 100, landmarks: int = 2000, classes: int = 8, seed: int = 13):
    """The PLS-LDA batched workload: X = 10,000 x 50 latent Gaussian (variance
    1/i per column) with 8 class-shifted clusters, M runs of 2000 landmarks,
    10 folds, starting labels the k-means-like nearest of 8 random landmark
    centres."""
    import kodama_amd as K
    from kodama_amd import datasets as D
    x, _ = D.config2()
    n = x.shape[0]
    rows, init = [], []
    for r in range(runs):
        rw = np.sort(np.argsort(D.uniform(seed, 100 + r, n), kind="stable")[:landmarks]).astype(np.int32)
        centres = x[rw[np.argsort(D.uniform(seed, 300 + r, landmarks), kind="stable")[:classes]]]
        idx, _ = K.exact_knn_search(centres, x[rw], 1, "euclidean")
        rows.append(rw)
        init.append((idx[:, 0] + 1).astype(np.int32))
    return x, rows, init


def pls_form(args) -> dict:
    """The batched CorePLSLDA form on the PLS-LDA workload (10 folds, 5
    components, the KODAMA.matrix per-run flags): steps/s and the fused
    evaluation kernel against its per-step bytes (one read of each landmark
    row of p floats plus three words per row)."""
    import kodama_amd as K
    x, rows, init = pls_workload(args.runs)
    kw = dict(classifier=0, folds=10, stratified=0, seed=1, cv_seed=1, max_components=5, fixed_components=5,
              center=1, scale=1, pls_fold_cache=1, evolutionary_search=1, guarded_diversity=1,
              many_to_one_absorption=1, auto_class_coarsening=1, n_threads=0)
    if args.warmup > 0:
        K.core_plslda_batch(x, rows, init, K.core_options(cycles=min(args.warmup, 2), **kw))
    out = K.core_plslda_batch(x, rows, init, K.core_options(cycles=args.steps, **kw))
    L = sum(len(r) for r in rows)
    kern = out["eval_kernel_seconds"]
    cpu = None
    if not args.no_cpu_baseline:
        ckw = {k: v for k, v in kw.items() if k != "n_threads"}
        cpu = pls_cpu_baseline(x, rows, init, ckw, args.steps, min(16, os.cpu_count() or 1), 5.0)
    algo = float(L) * (50 * 4 + 12)  # bytes per evaluation of all runs
    return dict(workload=f"PLS-LDA: X=10000x50 latent + 8 shifted clusters, M={args.runs}, 2000 landmarks, "
                         f"10 folds, ncomp=5, Tcycle={args.steps}",
                steps_per_s=args.runs * args.steps / out["evolution_seconds"],
                evolution_seconds=out["evolution_seconds"], mean_accbest=float(np.mean(out["accbest"])),
                eval_kernel_ms=kern * 1e3,
                roofline={"kernel": "pls_fold_sums_kernel + pls_eval_kernel",
                          "bound": "latency (one wave per fold, FP64)",
                          "achieved": algo / kern / 1e9 if kern > 0 else 0.0, "unit": "GB/s",
                          "peak": HBM_PEAK_GBS, "frac": algo / kern / 1e9 / HBM_PEAK_GBS if kern > 0 else 0.0,
                          "algorithmic_bytes_per_launch": algo, "mfma": "none (class sums)"},
                cpu_baseline=cpu)


def pls_fit_bench(rows: int, ncomp: int) -> int:
    """General PLS of the one-hot mixture components (20 responses) on rows of
    the 20-cluster mixture (the 1M x 50 generator): the FP64 matrix-core
    cross-product kernel against the FP64 matrix peak, and the whole fit."""
    import kodama_amd as K
    from kodama_amd import datasets as D
    x, comp = D.config4(n=rows)
    y = np.zeros((rows, 20), np.float32)
    y[np.arange(rows), comp - 1] = 1.0
    K.PLS(x[:4096], y[:4096], ncomp=ncomp)  # warm-up
    t0 = time.perf_counter()
    out = K.PLS(x, y, ncomp=ncomp)
    wall = time.perf_counter() - t0
    flops = 2.0 * rows * 50 * (50 + 20)
    print(json.dumps({"pls_fit_rows": rows, "p": 50, "q": 20, "ncomp": ncomp, "components": out["components"],
                      "seconds": wall, "device_seconds": out["device_seconds"],
                      "cross_kernel_ms": out["cross_seconds"] * 1e3,
                      "cross_tflops": flops / out["cross_seconds"] / 1e12,
                      "cross_frac_of_fp64_matrix_peak": flops / out["cross_seconds"] / 78.6e12}), flush=True)
    return 0


def measured_traffic(kernel: str, runs: int, cycles: int, k: int, landmarks: int):
    """HBM bytes per launch from the committed counter summary when it covers
    this kernel on this workload, and the basis of the figure. (None, None)
    when the summary covers another workload."""
    path = os.path.join(ROOT, "profiles", "evo_persistent_kernel_traffic.txt")
    if not os.path.exists(path):
        return None, None
    kv = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            key, val = line.split("=", 1)
            kv[key.strip()] = val.strip()
    try:
        same = (kv["kernel"] == kernel and int(kv["runs"]) == runs and int(kv["k"]) == k
                and int(kv["landmarks_total"]) == landmarks)
        if not same:
            return None, None
        measured, at = float(kv["hbm_bytes_per_launch"]), int(kv["cycles"])
    except (KeyError, ValueError):
        return None, None
    if at == cycles:
        return measured, f"rocprofv3 FETCH_SIZE + WRITE_SIZE at {at} cycles"
    # Measured at another cycle count: the split runs' exchange grows with the
    # cycles, so the figure is scaled per evaluation (cycles + 1 per launch).
    return measured * (cycles + 1) / (at + 1), f"rocprofv3 at {at} cycles, scaled per evaluation"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="label-evolution cycles per run (Tcycle)")
    ap.add_argument("--warmup", type=int, default=5, help="cycles of an untimed warm-up evolution")
    ap.add_argument("--runs", type=int, default=100)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--k", type=int, default=30)
    ap.add_argument("--forms", default="plain,device,learned,pls")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="minimum wall time of the CPU baseline")
    ap.add_argument("--copy-probe", action="store_true", help="run the counter-calibration copy and exit")
    ap.add_argument("--ordinary-launch", action="store_true",
                    help="launch split evolution runs as an ordinary grid instead of cooperatively "
                         "(rocprofv3 ends with a fault after tracing cooperative launches and writes nothing)")
    ap.add_argument("--graph-rows", type=int, default=0, help="time the exact k-NN graph of this many rows and exit")
    ap.add_argument("--lists-ab", action="store_true",
                    help="with --graph-rows: build the graph with packed and with separate candidate lists")
    ap.add_argument("--learned-curve", action="store_true",
                    help="accuracy against wall time of the uniform and learned proposals on the 200k"
                         " label-noise configuration (--steps cycles, default 100 here) and exit")
    ap.add_argument("--tables", action="store_true",
                    help="time the per-fold kNN tables of the workload by each path and exit")
    ap.add_argument("--pls-fit-rows", type=int, default=0,
                    help="time the general PLS fit (FP64 matrix-core cross-products) on this many rows and exit")
    ap.add_argument("--matrix-rows", type=int, default=0,
                    help="time end-to-end KODAMA.matrix on this many rows (runs, steps, forms apply) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe()
    if args.graph_rows > 0:
        return graph_bench(args.graph_rows, args.k, args.lists_ab)
    if args.tables:
        return tables_bench(args.runs, args.n, args.k)
    if args.learned_curve:
        return learned_curve_bench(args.runs, 100 if args.steps == 20 else args.steps, args.k)
    if args.pls_fit_rows > 0:
        return pls_fit_bench(args.pls_fit_rows, 10)
    if args.matrix_rows > 0:
        return matrix_bench(args.matrix_rows, args.runs, args.steps, args.k, [f for f in args.forms.split(",") if f])
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus above 1 is not supported"}))
        return 2

    import kodama_amd as K
    if args.ordinary_launch:
        K.set_cooperative_launch(False)
    log(f"{K.version()}; devices={K.device_count()}")
    t0 = time.perf_counter()
    x, rows, fold_ids, init = make_workload(args.runs, args.n, 50, args.k)
    L = np.array([len(r) for r in rows])
    log(f"workload ready in {time.perf_counter() - t0:.1f}s: runs={args.runs} landmarks total={L.sum()}")
    base = dict(k=args.k, folds=10, stratified=0, seed=1, cv_seed=1, evolutionary_search=1,
                guarded_diversity=1, many_to_one_absorption=1, n_threads=0)
    forms = [f for f in args.forms.split(",") if f]
    pls = None
    if "pls" in forms:
        forms.remove("pls")
        pls = pls_form(args)
        log(f"pls: {pls['steps_per_s']:,.0f} steps/s, eval kernel {pls['eval_kernel_ms']:.3f} ms")
        if not forms:
            # PLS-LDA alone (profiling runs): its own record, no kNN forms.
            print(json.dumps({"pls_lda_form": pls}), flush=True)
            return 0
    res = {}
    curves = {}
    params = None
    for form in forms:
        extra = {}
        if form == "learned":
            # The network is trained (untimed) on the accept/reject traces of
            # the plain form over 16 other runs (other landmarks, folds and
            # starting labels of the same data).
            t0 = time.perf_counter()
            tx, trows, tfolds, tinit = make_workload(16, args.n, 50, args.k, salt=100)
            feat, tgt, _ = K.mlp_collect_traces(tx, trows, tinit, K.core_options(cycles=args.steps, **base),
                                                fold_ids=tfolds)
            params, losses = K.mlp_train(feat, tgt, epochs=20, batch=256, learning_rate=0.05, momentum=0.9, seed=5)
            extra = dict(trace_samples=int(tgt.size), accepted_share=float(np.mean(tgt != 30)),
                         final_train_loss=float(losses[-1]), training_seconds=time.perf_counter() - t0)
        run = (lambda o, rr, ii, ff: K.core_knn_batch_learned(x, rr, ii, o, params, fallback=0.25, fold_ids=ff)) \
            if form == "learned" else \
            (lambda o, rr, ii, ff, form=form: K.core_knn_batch(x, rr, ii, o, mode=form, fold_ids=ff))
        if args.warmup > 0:
            run(K.core_options(cycles=args.warmup, **base), rows, init, fold_ids)
        out = run(K.core_options(cycles=args.steps, **base), rows, init, fold_ids)
        sps = args.runs * args.steps / out["evolution_seconds"]
        curves[form] = out["vect_acc"].mean(axis=0)
        wall = out.get("evolution_seconds_with_copies", out["evolution_seconds"])
        # steps_per_s: the evolution with its inputs in device memory (the
        # device forms' label upload and result download excluded; the plain
        # form moves labels every cycle by construction, so both are its loop).
        res[form] = dict(steps_per_s=sps, evolution_seconds=out["evolution_seconds"],
                         steps_per_s_with_copies=args.runs * args.steps / wall,
                         evolution_seconds_with_copies=wall,
                         precompute_seconds=out["precompute_seconds"],
                         mean_accbest=float(np.mean(out["accbest"])),
                         vote_kernel_ms=out.get("vote_seconds", out.get("device_seconds", 0.0)) * 1e3, **extra)
        log(f"{form}: {sps:,.0f} steps/s ({args.runs * args.steps / wall:,.0f} with host copies), "
            f"evolution {out['evolution_seconds'] * 1e3:.3f} ms, "
            f"precompute {out['precompute_seconds']:.3f}s, mean accbest {res[form]['mean_accbest']:.4f}, "
            f"device kernel {res[form]['vote_kernel_ms']:.3f} ms")
    # Cycles each form needs before its mean best accuracy reaches the plain
    # form's final mean best accuracy minus half a point.
    if "plain" in res:
        goal = res["plain"]["mean_accbest"] - 0.005
        for form, c in curves.items():
            hit = np.flatnonzero(c >= goal)
            res[form]["cycles_to_plain_accuracy"] = int(hit[0]) + 1 if hit.size else None
            # Wall time to that accuracy, pro rata of the evolution call.
            res[form]["seconds_to_plain_accuracy"] = \
                res[form]["evolution_seconds"] * (int(hit[0]) + 1) / args.steps if hit.size else None
    head = "device" if "device" in res else forms[0]
    kern_s = res[head]["vote_kernel_ms"] * 1e-3
    evaluations = float(args.steps + 1) if head == "device" else 1.0
    if head == "device":
        kmax = 8 if args.k <= 8 else 16 if args.k <= 16 else 32 if args.k <= 32 else 64
        kernel = f"evo_persistent_kernel<{kmax},{1024 if kmax <= 32 else 512}>"
    else:
        kernel = "vote_kernel"
    algo_bytes = evaluations * float(L.sum()) * (8 * args.k + 28)
    algo_bytes_16 = evaluations * float(L.sum()) * (2 * args.k)
    achieved = algo_bytes / kern_s / 1e9 if kern_s > 0 else 0.0
    achieved_16 = algo_bytes_16 / kern_s / 1e9 if kern_s > 0 else 0.0
    traffic, traffic_basis = (measured_traffic(kernel, args.runs, args.steps, args.k, int(L.sum()))
                              if head == "device" else (None, None))
    line = {
        "metric": METRIC,
        "value": res[head]["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * res[head]["evolution_seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched M runs: X={args.n}x50 20-cluster Zipf(1.1) mixture, kNN k={args.k}, "
                        f"M={args.runs}, landmarks U[1000,5000], 10 uneven folds, Tcycle={args.steps}",
            "n": args.n, "p": 50, "k": args.k, "runs": args.runs, "cycles": args.steps,
            "landmarks_total": int(L.sum()), "form": head,
        },
        "forms": res,
        "roofline": {"bound": "hbm", "kernel": kernel, "kernel_ms": kern_s * 1e3, "achieved": achieved,
                     "launch": "ordinary" if args.ordinary_launch else "cooperative",
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "traffic_basis": traffic_basis,
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "bytes_per_row_evaluation": 8 * args.k + 28, "evaluations_per_launch": evaluations,
                     "achieved_16bit_ids": achieved_16, "frac_16bit_ids": achieved_16 / HBM_PEAK_GBS},
    }
    if "plain" in res and "device" in res:
        line["accuracy_gap_pt"] = 100.0 * (res["device"]["mean_accbest"] - res["plain"]["mean_accbest"])
        line["speedup_device_vs_plain"] = res["device"]["steps_per_s"] / res["plain"]["steps_per_s"]
    if "plain" in res and "learned" in res:
        line["accuracy_gap_learned_pt"] = 100.0 * (res["learned"]["mean_accbest"] - res["plain"]["mean_accbest"])
        line["speedup_learned_vs_plain"] = res["learned"]["steps_per_s"] / res["plain"]["steps_per_s"]
        # Time to the plain form's final accuracy minus half a point: the plain
        # form's own time to it against the learned form's.
        tp, tl = res["plain"].get("seconds_to_plain_accuracy"), res["learned"].get("seconds_to_plain_accuracy")
        line["time_to_accuracy_learned_vs_plain"] = tp / tl if tp and tl else None
        # The same with the network's trace collection and training counted
        # on the learned side (a one-time cost per trained network).
        tt = res["learned"].get("training_seconds")
        line["time_to_accuracy_learned_incl_training_vs_plain"] = tp / (tl + tt) if tp and tl and tt else None
    if pls is not None:
        line["pls_lda_form"] = pls
    if not args.no_cpu_baseline:
        cores = min(16, os.cpu_count() or 1)
        cb = cpu_baseline(x, rows, init, args.k, args.steps, cores, args.cpu_seconds)
        log(f"cpu baseline ({cb['kind']}): {cb['value']:,.0f} steps/s on {cores} threads")
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
