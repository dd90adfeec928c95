// C ABI of libkodama_hip.so, PLS-LDA half: PLSLDACV, CorePLSLDA and the fold
// detail used to check scores and discriminants against an FP64 restatement.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/kodama_hip.h"
#include "capi_common.h"
#include "engine.h"
#include "pls_engine.h"
#include "batch_api.h"

using namespace kodama_hip;

namespace {

void require(bool ok, const char* msg) {
  if (!ok) throw InvalidArgument(msg);
}

void validate(const float* x, int n, int p, const int* labels, const kodama_hip_core_options* o) {
  require(x != nullptr, "MatrixView data pointer is null.");
  require(n > 0 && p > 0, "Input matrix must have at least one row and one column.");
  require(labels != nullptr, "labels size must match number of rows.");
  require(o != nullptr, "null options.");
  require(o->max_components >= 1, "PLSOptions::max_components must be positive.");
  require(o->folds >= 2, "PLS cross-validation needs at least two folds.");
}

IVec vec_or_empty(const int* p, int n) { return p ? IVec(p, p + n) : IVec(); }

IVec values_of(const IVec& codes, const IVec& values) {
  IVec out(codes.size());
  for (size_t i = 0; i < codes.size(); ++i) out[i] = values[codes[i]];
  return out;
}

// Class statistics of a labelling for the evolution engine.
Tally tally_of(const IVec& codes, const int* pred, int classes, IVec& counts, IVec& firsts) {
  counts.assign(static_cast<size_t>(classes), 0);
  firsts.assign(static_cast<size_t>(classes), -1);
  Tally t;
  for (size_t i = 0; i < codes.size(); ++i) {
    const int c = codes[i];
    if (firsts[c] < 0) firsts[c] = static_cast<int>(i);
    ++counts[c];
    t.matches += pred[i] == c ? 1 : 0;
  }
  t.counts = counts.data();
  t.firsts = firsts.data();
  return t;
}

}  // namespace

namespace {

// One F-fold PLS CV with the LDA or the PLS-DA head (run_plscv_host,
// plscv.cpp:2293-2507).
void plscv_impl(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                const kodama_hip_core_options* options, bool da, int32_t* predicted_labels,
                int32_t* fold_assignments, int32_t* selected_components) {
  validate(x, n, p, labels, options);
  require(predicted_labels != nullptr, "null output.");
  require(options->fixed_components >= 0, "PLSOptions::fixed_components must be non-negative.");
  require(options->fixed_components <= options->max_components,
          "PLSOptions::fixed_components cannot exceed max_components.");
  select_device(options->device);
  IVec values, codes;
  encode_labels(labels, n, values, codes);
  const IVec lab(labels, labels + n);
  const IVec folds = make_folds(lab, vec_or_empty(constrain, n), options->folds, options->stratified != 0,
                                options->cv_seed);
  PlsCv pls(options->device, x, n, p, *options, static_cast<int>(values.size()));
  pls.set_da(da);
  IVec pred(static_cast<size_t>(n));
  int sel = 0;
  pls.evaluate(codes, folds, pred.data(), &sel);
  for (int i = 0; i < n; ++i) predicted_labels[i] = values[pred[i]];
  if (fold_assignments) std::copy(folds.begin(), folds.end(), fold_assignments);
  if (selected_components) *selected_components = sel;
}

}  // namespace

extern "C" {

int kodama_hip_plsldacv(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                        const kodama_hip_core_options* options, int32_t* predicted_labels,
                        int32_t* fold_assignments, int32_t* selected_components) {
  return guarded([&] {
    plscv_impl(x, n, p, labels, constrain, options, false, predicted_labels, fold_assignments, selected_components);
  });
}

// PLSDACV_CUDA (plscv.cpp:2934-2938): the same folds and SIMPLS fits with the
// PLS-DA head: the least-squares regression of the centred one-hot labels on
// the training scores, the class of largest fitted value.
int kodama_hip_plsdacv(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                       const kodama_hip_core_options* options, int32_t* predicted_labels,
                       int32_t* fold_assignments, int32_t* selected_components) {
  return guarded([&] {
    plscv_impl(x, n, p, labels, constrain, options, true, predicted_labels, fold_assignments, selected_components);
  });
}

// CorePLSLDA: maximize_core with the PLS predictor (core.cpp:1720-1749). With
// unstratified folds the fold cache keeps the folds of the first evaluation
// (plscv.cpp:2312); stratified folds follow the labels of every evaluation.
int kodama_hip_core_plslda(const float* x, int32_t n, int32_t p, const int32_t* labels, const int32_t* constrain,
                           const int32_t* fixed, const kodama_hip_core_options* options,
                           kodama_hip_core_result* result) {
  return guarded([&] {
    validate(x, n, p, labels, options);
    require(result != nullptr, "null result.");
    require(options->cycles >= 0, "Core cycles must be non-negative.");
    select_device(options->device);
    CoreFlags f = CoreFlags::from(*options);
    f.knn = false;
    IVec values, codes;
    encode_labels(labels, n, values, codes);
    const IVec cons = vec_or_empty(constrain, n);
    const IVec init(labels, labels + n);
    const bool cache = options->stratified == 0;
    const IVec first_folds = make_folds(init, cons, options->folds, options->stratified != 0, options->cv_seed);
    const int C = static_cast<int>(values.size());
    PlsCv pls(options->device, x, n, p, *options, C);
    RunEvolution run(values, codes, cons, vec_or_empty(fixed, n), f);
    IVec pred(static_cast<size_t>(n)), counts, firsts;
    auto evaluate = [&](const IVec& cl) {
      const IVec folds = cache ? first_folds
                               : make_folds(values_of(cl, values), cons, options->folds, true, options->cv_seed);
      pls.evaluate(cl, folds, pred.data(), nullptr);
      return tally_of(cl, pred.data(), C, counts, firsts);
    };
    {
      const Tally t = evaluate(run.initial());
      run.begin(pred.data(), t);
    }
    for (int cycle = 0; cycle < options->cycles; ++cycle) {
      const IVec& cl = run.propose(cycle);
      const Tally t = evaluate(cl);
      run.finish(cycle, pred.data(), t);
    }
    run.export_result(result);
    if (result->fold_assignments) std::copy(first_folds.begin(), first_folds.end(), result->fold_assignments);
  });
}

}  // extern "C"

namespace kodama_hip {

// The M-run form of CorePLSLDA: run r evolves its landmark rows with seed + r
// and cv_seed + r (kodama_matrix.cpp:3113, 3123-3135); every run's proposals
// come from its own host engine with the reference's random streams, and one
// group + evaluation launch per cycle serves all runs x folds: the landmark
// rows of all runs are concatenated and fold f of run r is fold r * F + f of
// that matrix. Per run the result equals kodama_hip_core_plslda on its rows
// (with the run's constraint groups and fixed rows, when given).
void run_plslda_batch(const BatchRequest& q, int32_t* clbest, double* accbest, double* vect_acc,
                      double* timings) {
  const float* x = q.x;
  const int n = q.n, p = q.p, runs = q.runs;
  const kodama_hip_core_options* options = q.options;
  const int32_t* fold_ids = q.fold_ids;
  require(x != nullptr && n > 0 && p > 0, "invalid data matrix.");
  require(runs >= 1, "runs must be positive.");
  require(q.run_offsets != nullptr && q.landmark_rows != nullptr && q.initial_labels != nullptr, "null input.");
  require(options != nullptr && clbest != nullptr && accbest != nullptr, "null options or output.");
  require(options->max_components >= 1, "PLSOptions::max_components must be positive.");
  require(options->folds >= 2, "PLS cross-validation needs at least two folds.");
  require(options->cycles >= 0, "Core cycles must be non-negative.");
  require(options->stratified == 0 || fold_ids == nullptr, "explicit folds are unstratified.");
  select_device(options->device);
  const auto t0 = std::chrono::steady_clock::now();
  const IVec offsets(q.run_offsets, q.run_offsets + runs + 1);
  require(offsets[0] == 0, "run_offsets[0] must be 0.");
  for (int r = 0; r < runs; ++r) require(offsets[r + 1] > offsets[r] + 1, "every run needs at least two rows.");
  const int total = offsets[runs];
  std::vector<float> xcat(static_cast<size_t>(total) * p);
  for (int i = 0; i < total; ++i) {
    const int row = q.landmark_rows[i];
    require(row >= 0 && row < n, "landmark row outside the data matrix.");
    std::copy(x + static_cast<size_t>(row) * p, x + static_cast<size_t>(row + 1) * p,
              xcat.begin() + static_cast<size_t>(i) * p);
  }
  std::vector<IVec> values(static_cast<size_t>(runs)), codes(static_cast<size_t>(runs)),
    cons(static_cast<size_t>(runs)), fixed(static_cast<size_t>(runs));
  int classes = 1;
  for (int r = 0; r < runs; ++r) {
    encode_labels(q.initial_labels + offsets[r], offsets[r + 1] - offsets[r], values[r], codes[r]);
    classes = std::max(classes, static_cast<int>(values[r].size()));
    if (q.constrain) cons[r].assign(q.constrain + offsets[r], q.constrain + offsets[r + 1]);
    if (q.fixed) fixed[r].assign(q.fixed + offsets[r], q.fixed + offsets[r + 1]);
  }
  const int F = options->folds;
  const bool cache = options->stratified == 0;
  auto run_folds = [&](int r, const IVec& labs) {
    IVec fo = fold_ids ? IVec(fold_ids + offsets[r], fold_ids + offsets[r + 1])
                       : make_folds(labs, cons[r], F, options->stratified != 0,
                                    options->cv_seed + static_cast<std::uint64_t>(r));
    for (int v : fo) require(v >= 0 && v < std::max(F, 1 << 20), "fold id out of range.");
    return fo;
  };
  int fold_span = F;
  if (fold_ids) {
    for (int i = 0; i < total; ++i) fold_span = std::max(fold_span, fold_ids[i] + 1);
  }
  IVec gfolds(static_cast<size_t>(total));
  auto set_folds = [&](int r, const IVec& fo) {
    for (size_t i = 0; i < fo.size(); ++i) gfolds[offsets[r] + i] = r * fold_span + fo[i];
  };
  for (int r = 0; r < runs; ++r) {
    const IVec fo = run_folds(r, IVec(q.initial_labels + offsets[r], q.initial_labels + offsets[r + 1]));
    const auto mm = std::minmax_element(fo.begin(), fo.end());
    require(*mm.first != *mm.second, "every run needs at least two distinct folds.");
    set_folds(r, fo);
  }
  PlsCv pls(options->device, xcat.data(), total, p, *options, classes, offsets);
  std::vector<RunEvolution> ev;
  ev.reserve(static_cast<size_t>(runs));
  for (int r = 0; r < runs; ++r) {
    CoreFlags f = CoreFlags::from(*options);
    f.knn = false;
    f.seed = options->seed + static_cast<std::uint64_t>(r);
    ev.emplace_back(values[r], codes[r], cons[r], fixed[r], f);
  }
  // Folds that do not depend on the labels: the fold cache is part of the
  // precompute.
  if (cache) pls.prepare(gfolds);
  const double t_pre = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const auto t1 = std::chrono::steady_clock::now();
  const int threads = resolve_threads(options->n_threads, runs);
  // The proposals are written straight into the engine's pinned staging and
  // the predictions read from it.
  int* ccat = pls.staged_codes();
  const int* pred = pls.staged_pred();
  auto evaluate_all = [&](int cycle) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int r = 0; r < runs; ++r) {
      const IVec& cl = cycle < 0 ? ev[r].initial() : ev[r].propose(cycle);
      std::copy(cl.begin(), cl.end(), ccat + offsets[r]);
      if (!cache) set_folds(r, run_folds(r, values_of(cl, values[r])));
    }
    pls.evaluate_staged(gfolds, nullptr);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int r = 0; r < runs; ++r) {
      const IVec cl(ccat + offsets[r], ccat + offsets[r + 1]);
      IVec counts, firsts;
      const Tally t = tally_of(cl, pred + offsets[r], static_cast<int>(values[r].size()), counts, firsts);
      if (cycle < 0) ev[r].begin(pred + offsets[r], t);
      else ev[r].finish(cycle, pred + offsets[r], t);
    }
  };
  evaluate_all(-1);
  for (int cycle = 0; cycle < options->cycles; ++cycle) evaluate_all(cycle);
  const double t_evo = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  for (int r = 0; r < runs; ++r) {
    const IVec& b = ev[r].best_codes();
    for (size_t i = 0; i < b.size(); ++i) clbest[offsets[r] + i] = values[r][b[i]];
    accbest[r] = ev[r].accbest();
    if (vect_acc) {
      std::copy(ev[r].vect_acc().begin(), ev[r].vect_acc().end(), vect_acc + static_cast<size_t>(r) * options->cycles);
    }
  }
  if (timings) {
    timings[0] = t_pre;
    timings[1] = t_evo;
    timings[2] = static_cast<double>(runs) * options->cycles;
    timings[3] = pls.eval_kernel_ms_average() * 1e-3;
    timings[4] = t_evo;  // the labels cross PCIe every cycle by construction
  }
}

// PLSLDAPredict_CPU (plscv.cpp:2976-3005) for a batch of fits: fit r trains
// on the rows train_rows[train_off[r] .. train_off[r+1]) of x with their
// labels and predicts the rows test_rows[test_off[r] .. test_off[r+1]) into
// pred (same indexing). Each fit is the validation fold of a two-fold split of
// [train; test]; fits go through the fused evaluation in groups whose rows fit
// a bounded staging matrix. A fit with one training class predicts it; a fit
// without a usable component is EINVAL (DegeneratePLSFit).
void plslda_predict_batch(const float* x, int n, int p, int fits, const IVec& train_off, const int* train_rows,
                          const int* labels, const IVec& test_off, const int* test_rows,
                          const kodama_hip_core_options& options, int* pred) {
  const size_t budget = size_t(1) << 26;  // floats staged per group (256 MiB)
  int r0 = 0;
  while (r0 < fits) {
    size_t rows = 0;
    int r1 = r0;
    while (r1 < fits) {
      const size_t add = static_cast<size_t>(train_off[r1 + 1] - train_off[r1] + test_off[r1 + 1] - test_off[r1]);
      if (r1 > r0 && (rows + add) * p > budget) break;
      rows += add;
      ++r1;
    }
    // Stage the group's fits: [train; test] per fit, folds 2 g + 1 (train)
    // and 2 g (test), label codes of each fit's own classes.
    IVec offsets{0}, all_codes, folds, ncls;
    std::vector<IVec> values(static_cast<size_t>(r1 - r0));
    std::vector<float> xs(rows * p);
    size_t at = 0;
    int classes = 1;
    std::vector<int> live;
    for (int r = r0; r < r1; ++r) {
      const int g = r - r0;
      const int nt = train_off[r + 1] - train_off[r], nq = test_off[r + 1] - test_off[r];
      IVec codes;
      encode_labels(labels + train_off[r], nt, values[g], codes);
      if (nq == 0) continue;
      require(nt > 0, "PLSLDAPredict: no training rows.");
      if (values[g].size() == 1) {
        std::fill(pred + test_off[r], pred + test_off[r + 1], values[g][0]);
        continue;
      }
      classes = std::max(classes, static_cast<int>(values[g].size()));
      for (int i = 0; i < nt + nq; ++i) {
        const int row = i < nt ? train_rows[train_off[r] + i] : test_rows[test_off[r] + i - nt];
        std::copy_n(x + static_cast<size_t>(row) * p, p, xs.begin() + (at + i) * p);
        all_codes.push_back(i < nt ? codes[i] : 0);
        folds.push_back(i < nt ? 2 * static_cast<int>(live.size()) + 1 : 2 * static_cast<int>(live.size()));
      }
      at += static_cast<size_t>(nt + nq);
      offsets.push_back(static_cast<int>(at));
      live.push_back(r);
    }
    if (!live.empty()) {
      PlsCv pls(options.device, xs.data(), static_cast<int>(at), p, options, classes, offsets);
      IVec pc(at);
      pls.evaluate(all_codes, folds, pc.data(), nullptr);
      for (size_t j = 0; j < live.size(); ++j) {
        const int r = live[j];
        require(pls.detail(static_cast<int>(2 * j)).rank > 0, "PLSLDAPredict: the PLS fit has no usable component.");
        const int nt = train_off[r + 1] - train_off[r];
        const IVec& vals = values[r - r0];
        for (int i = test_off[r]; i < test_off[r + 1]; ++i) {
          pred[i] = vals[pc[static_cast<size_t>(offsets[j]) + nt + (i - test_off[r])]];
        }
      }
    }
    r0 = r1;
  }
}

}  // namespace kodama_hip

extern "C" {

int kodama_hip_core_plslda_batch(const float* x, int32_t n, int32_t p, int32_t runs, const int32_t* run_offsets,
                                 const int32_t* landmark_rows, const int32_t* initial_labels,
                                 const int32_t* fold_ids, const kodama_hip_core_options* options,
                                 int32_t* clbest, double* accbest, double* vect_acc, double* timings) {
  return guarded([&] {
    BatchRequest q{x, n, p, runs, run_offsets, landmark_rows, initial_labels, fold_ids, options};
    run_plslda_batch(q, clbest, accbest, vect_acc, timings);
  });
}

// PLSLDAPredict_CPU (plscv.cpp:2976-3005): one PLS-LDA fit on all training
// rows, applied to the test rows. It is the validation fold of a two-fold
// evaluation of [train; test] in which the training rows form the other fold:
// the training rows' mean and sd standardise both, the SIMPLS fit takes the
// requested components capped by the columns and rows - 1, and a single
// training class predicts itself; a fit without a usable component is an
// error, as the reference's DegeneratePLSFit.
int kodama_hip_plslda_predict(const float* train, int32_t n_train, const int32_t* labels, const float* test,
                              int32_t n_test, int32_t p, const kodama_hip_core_options* options, int32_t* pred) {
  return guarded([&] {
    require(train != nullptr && test != nullptr, "PLSLDAPredict input matrix pointer is null.");
    require(options != nullptr && labels != nullptr && (pred != nullptr || n_test == 0), "null pointer.");
    require(n_train >= 0 && n_test >= 0 && p >= 0, "negative size.");
    if (n_train == 0 || n_test == 0 || p == 0) return;
    IVec values, codes;
    encode_labels(labels, n_train, values, codes);
    if (values.size() == 1) {
      std::fill(pred, pred + n_test, values[0]);
      return;
    }
    const int n = n_train + n_test;
    std::vector<float> x(static_cast<size_t>(n) * p);
    std::copy_n(train, static_cast<size_t>(n_train) * p, x.begin());
    std::copy_n(test, static_cast<size_t>(n_test) * p, x.begin() + static_cast<size_t>(n_train) * p);
    IVec all_codes(static_cast<size_t>(n), 0), folds(static_cast<size_t>(n), 0);
    std::copy(codes.begin(), codes.end(), all_codes.begin());
    std::fill(folds.begin(), folds.begin() + n_train, 1);
    std::vector<int32_t> lab_all(static_cast<size_t>(n), labels[0]);
    std::copy_n(labels, n_train, lab_all.begin());
    validate(x.data(), n, p, lab_all.data(), options);
    select_device(options->device);
    PlsCv pls(options->device, x.data(), n, p, *options, static_cast<int>(values.size()));
    IVec pred_codes(static_cast<size_t>(n));
    pls.evaluate(all_codes, folds, pred_codes.data(), nullptr);
    require(pls.detail(0).rank > 0, "PLSLDAPredict: the PLS fit has no usable component.");
    for (int i = 0; i < n_test; ++i) pred[i] = values[pred_codes[static_cast<size_t>(n_train) + i]];
  });
}

int kodama_hip_plslda_fold_detail(const float* x, int32_t n, int32_t p, const int32_t* labels,
                                  const int32_t* folds, int32_t fold, const kodama_hip_core_options* options,
                                  double* weights, int32_t* rank, double* val_scores, double* discriminants,
                                  int32_t* classes, int32_t* n_classes, int32_t* n_val) {
  return guarded([&] {
    validate(x, n, p, labels, options);
    require(folds != nullptr && weights != nullptr && rank != nullptr && val_scores != nullptr &&
              discriminants != nullptr && classes != nullptr && n_classes != nullptr && n_val != nullptr,
            "null pointer.");
    select_device(options->device);
    IVec values, codes;
    encode_labels(labels, n, values, codes);
    const IVec fv(folds, folds + n);
    IVec ids = fv;
    std::sort(ids.begin(), ids.end());
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    const auto it = std::find(ids.begin(), ids.end(), fold);
    require(it != ids.end(), "no such fold.");
    PlsCv pls(options->device, x, n, p, *options, static_cast<int>(values.size()));
    IVec pred(static_cast<size_t>(n));
    pls.evaluate(codes, fv, pred.data(), nullptr);
    const PlsFoldDetail d = pls.detail(static_cast<int>(it - ids.begin()));
    *rank = d.rank;
    *n_classes = static_cast<int>(d.classes.size());
    *n_val = static_cast<int>(d.val_rows.size());
    std::copy(d.weights.begin(), d.weights.end(), weights);
    std::copy(d.val_scores.begin(), d.val_scores.end(), val_scores);
    std::copy(d.discriminants.begin(), d.discriminants.end(), discriminants);
    for (size_t c = 0; c < d.classes.size(); ++c) classes[c] = values[d.classes[c]];
  });
}

// PLS_CUDA (plscv.cpp:2899-2905 -> fit_general_pls 2765-2878): general PLS
// of y [n x q] on x [n x p], FP64 on the device (cross-products on the
// matrix cores), float outputs as PLSFitResult.
int kodama_hip_pls(const float* x, int32_t n, int32_t p, const float* y, int32_t q,
                   const kodama_hip_core_options* options, float* weights, float* response_loadings,
                   float* scores, float* coefficients, float* fitted, float* x_center, float* x_scale,
                   float* y_center, int32_t* components, double* timings) {
  return guarded([&] {
    require(x != nullptr && y != nullptr && n >= 2 && p >= 1 && q >= 1,
            "PLS requires non-empty X and Y matrices with at least two rows.");
    require(options != nullptr, "null options.");
    require(weights && response_loadings && scores && coefficients && fitted && x_center && x_scale && y_center &&
              components,
            "null output.");
    for (size_t e = 0; e < static_cast<size_t>(n) * q; ++e) {
      require(std::isfinite(y[e]), "PLS Y must contain only finite values.");
    }
    select_device(options->device);
    const PlsGeneralResult r = pls_general_fit(options->device, x, y, n, p, q, *options);
    *components = r.components;
    std::copy(r.weights.begin(), r.weights.end(), weights);
    std::copy(r.response_loadings.begin(), r.response_loadings.end(), response_loadings);
    std::copy(r.scores.begin(), r.scores.end(), scores);
    std::copy(r.coefficients.begin(), r.coefficients.end(), coefficients);
    std::copy(r.fitted.begin(), r.fitted.end(), fitted);
    std::copy(r.x_center.begin(), r.x_center.end(), x_center);
    std::copy(r.x_scale.begin(), r.x_scale.end(), x_scale);
    std::copy(r.y_center.begin(), r.y_center.end(), y_center);
    if (timings) {
      timings[0] = r.cross_ms * 1e-3;
      timings[1] = r.total_ms * 1e-3;
    }
  });
}

}  // extern "C"
