/*
 * kodama_hip -- MI355X (gfx950) HIP backend of the KODAMA.matrix hot path.
 *
 * C ABI of libkodama_hip.so: plain pointers, sizes and integer status codes,
 * no exceptions and no C++ or framework types across the boundary. Every entry
 * point returns KODAMA_HIP_OK (0) or an error code; kodama_hip_last_error()
 * returns the message of the last failure on the calling thread (the pattern of
 * the reference's embedding launchers, visualization.cpp:2594-2596).
 *
 * Host pointers unless a name says otherwise. Labels are arbitrary int32 class
 * values as in the reference; neighbour ids are 0-based unless the function name
 * says one_based. Entry points are re-entrant: concurrent callers on different
 * threads each get their own device workspaces (kodama_matrix.cpp:2651-2682 runs
 * *_CUDA from many lane threads at once).
 *
 * Each declaration cites the reference interface it replaces (file:line in
 * tkcaccia/KODAMA).
 */
#ifndef KODAMA_HIP_H
#define KODAMA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KODAMA_HIP_OK 0
#define KODAMA_HIP_EINVAL 1   /* std::invalid_argument in the reference */
#define KODAMA_HIP_ERUNTIME 2 /* std::runtime_error / device failure */
#define KODAMA_HIP_ENODEV 3   /* no usable gfx950 device */

/* DistanceMetric order (kodama.hpp:27-31). */
#define KODAMA_HIP_COSINE 0
#define KODAMA_HIP_INNER_PRODUCT 1
#define KODAMA_HIP_EUCLIDEAN 2

/* Flat mirror of kodama::CoreOptions with its KNNOptions / CorePLSLDAOptions
 * (kodama.hpp:149-179, 294-314) and EvolutionPolicy (kodama.hpp:49-61),
 * restricted to the fields the hot path reads. */
typedef struct kodama_hip_core_options {
  int32_t cycles;                 /* CoreOptions::cycles */
  int32_t classifier;             /* 0 = PLS_LDA, 1 = KNN (CoreClassifier) */
  uint64_t seed;                  /* CoreOptions::seed */
  int32_t shake;
  int32_t evolutionary_search;
  int32_t guarded_diversity;
  int32_t many_to_one_absorption;
  int32_t auto_class_coarsening;
  int32_t adaptive_proposal_size;
  int32_t prediction_guidance;    /* EvolutionPolicy fields ... */
  int32_t policy_adaptive_proposal_size;
  int32_t transition_proposal;
  int32_t stochastic_acceptance;
  int32_t diversity_multiplier;
  int32_t pls_transition_coarsening;
  int32_t pls_fragmentation_penalty;
  int32_t reject_single_class;
  int32_t folds;                  /* FoldOptions::folds */
  int32_t stratified;             /* FoldOptions::stratified */
  uint64_t cv_seed;               /* FoldOptions::seed */
  int32_t k;                      /* KNNOptions::k */
  int32_t metric;                 /* KNNOptions::metric */
  int32_t max_components;         /* CorePLSLDAOptions::max_components */
  int32_t fixed_components;       /* CorePLSLDAOptions::fixed_components */
  int32_t center;
  int32_t scale;
  int32_t pls_fold_cache;         /* 1 = data_epoch != 0 (folds cached per run) */
  int32_t device;                 /* KNNOptions::gpu_device / PLSOptions::gpu_device */
  int32_t n_threads;              /* host threads for the batched engine (0 = auto) */
} kodama_hip_core_options;

/* Outputs of one label-evolution run (kodama::CoreResult, kodama.hpp:316-349).
 * Caller-allocated; per-sample arrays hold n entries, per-cycle arrays `cycles`.
 * counters[KODAMA_HIP_N_COUNTERS]: cycles_completed, proposals_evaluated,
 * best_state_updates, current_state_accepts, stochastic_state_attempts,
 * stochastic_state_accepts, current_state_rejections, coarsening_moves,
 * absorption_moves, transition_attempted, transition_accepted,
 * many_to_one_attempted, many_to_one_accepted, pls_coarsening_attempted,
 * pls_coarsening_accepted, cv_evaluations, success. */
typedef struct kodama_hip_core_result {
  int32_t* clbest;
  int32_t* clbest_dirty;
  int32_t* cvpredbest;
  int32_t* fold_assignments;      /* may be NULL */
  double* vect_acc;
  double* vect_score;
  int32_t* proposal_size;
  int32_t* active_classes;
  uint8_t* accepted;
  uint8_t* improving_acceptance;
  uint8_t* temperature_acceptance;
  int32_t* counters;
  double accbest;
  double scorebest;
} kodama_hip_core_result;

#define KODAMA_HIP_N_COUNTERS 17

/* ---- library -------------------------------------------------------------- */
const char* kodama_hip_last_error(void);
/* Number of visible gfx950 devices (hipGetDeviceCount + arch check). */
int kodama_hip_device_count(int32_t* count);
/* Library version string, e.g. "kodama_hip 0.1.0 gfx950". */
const char* kodama_hip_version(void);

/* No counterpart in the reference: for callers that check which path the last
 * exact kNN search of the calling thread took (kodama_hip_exact_knn_search,
 * kodama_hip_knn_graph, and the per-fold tables of the kNN classifier calls).
 * With at most 64 dimensions, k <= 128 and (for a search) at least 4096 train
 * rows the search runs split-BF16 MFMA candidates with an exact FP32 rerank and
 * returns the number of queries (table rows) that fell back to the exact scan;
 * -1 when the exact scan served every query. Results are identical either way. */
int kodama_hip_last_knn_fallbacks(void);

/* Per-fold kNN tables of a batch of runs: replaces precompute_knn_cv_cuda
 * (core.cpp:694-782), which builds one run's tables, for `runs` runs at once.
 * Run r owns rows[run_offsets[r] .. run_offsets[r + 1]) (data row ids) with
 * fold ids fold_ids (same indexing); each row's k nearest rows of its run in
 * another fold, rows prepared as CoreKNN prepares them (cosine rows
 * normalised, core.cpp:159-193). Output [total x k] row-major: run-local row
 * ids in (distance, id) order and scores = minus the internal distance, -1 /
 * -inf past the available rows -- the reference's per-fold exact tables,
 * bit-identical. path: 1 the register scan in the reference's summation
 * order; 2 split-BF16 MFMA candidates with the exact FP32 rerank (p <= 64,
 * k <= 128; otherwise the scan); 0 the process-wide setting below. Both paths
 * give identical tables. *path_used (may be NULL) = the path that ran. */
int kodama_hip_knn_fold_tables(const float* x, int32_t n, int32_t p, int32_t runs, const int32_t* run_offsets,
                               const int32_t* rows, const int32_t* fold_ids, int32_t k, int32_t metric,
                               int32_t path, int32_t device, int32_t* neighbors, float* scores,
                               int32_t* path_used);

/* No counterpart in the reference: the process-wide path of the per-fold
 * tables that every kNN classifier call builds (KNNCV, CoreKNN, the batches,
 * KODAMA.matrix): 1 scan, 2 MFMA, 0 (default) tuned -- the first table of a
 * shape bucket (columns, k class, metric, log2 of the largest run and of all
 * rows, folds) runs both paths, timed, and the faster serves the bucket for
 * the rest of the process. kodama_hip_last_fold_table_path returns the path
 * the calling thread's last table build used (0 before any). */
int kodama_hip_set_fold_table_path(int32_t path);
int kodama_hip_last_fold_table_path(void);
/* The tuner's full schedule (no reference counterpart): path as above; for
 * the MFMA path the candidates kept per (query, lane column), list = 8 (k <=
 * 32) or 16, and the waves of 16 queries per workgroup, waves = 4 or 8; 0 =
 * tuned. With path 0 the first table of a shape bucket times the scan (below
 * the bucket sizes where MFMA is known to win) and every MFMA schedule the
 * forced values leave, and the fastest serves the bucket from then on; all
 * schedules give identical tables. A forced list that excludes the job's k
 * leaves the scan. kodama_hip_last_fold_table_schedule reports the calling
 * thread's last build (list and waves 0 for the scan). */
int kodama_hip_set_fold_table_schedule(int32_t path, int32_t list, int32_t waves);
int kodama_hip_last_fold_table_schedule(int32_t* path, int32_t* list, int32_t* waves);
/* Candidate lists of the MFMA kNN path (no reference counterpart), process
 * wide: on = 1 (the default) keeps each candidate as one 32-bit key, the
 * value's order bits with the row's index within its lane column in the low
 * bits, wherever a search's train rows (a batch's largest run) number at most
 * 32768; on = 0 keeps separate values and indices. Both give the same
 * neighbours. */
int kodama_hip_set_mfma_packed_lists(int32_t on);

/* Tuning control of the device-resident evolution (no reference counterpart:
 * the reference runs one CUDA stream per run). Runs of at least 1024 rows are
 * split over up to max_group workgroups (0 = automatic, up to 8; 1 = one
 * workgroup per run) while all the workgroups fit the device at once; the
 * members replay the same proposal, absorption and acceptance and each votes
 * a slice of the rows, so the results are identical for every setting.
 * kodama_hip_last_evolution_workgroups returns the workgroups of the calling
 * thread's last device-resident evolution (0 before any). */
int kodama_hip_set_evolution_split(int32_t max_group);
int kodama_hip_last_evolution_workgroups(void);
/* Split runs wait on each other, so their workgroups must be resident
 * together: by default they take a cooperative launch, which guarantees it.
 * cooperative = 0 launches them as an ordinary grid of the same size (for
 * profilers that cannot follow cooperative launches); the members' waits are
 * bounded, so a member that is not resident makes the call fail with
 * ERUNTIME instead of hanging. No reference counterpart. */
int kodama_hip_set_cooperative_launch(int32_t cooperative);
/* Flow counts of the device-resident evolution's absorption: in an LDS table
 * when it fits beside the run state (enabled = 1, default) or in a
 * per-workgroup HBM table (0); the results are identical.
 * kodama_hip_last_evolution_lds_flows returns 1 when the calling thread's
 * last device-resident evolution used the LDS table. No reference
 * counterpart. */
int kodama_hip_set_lds_flow_table(int32_t enabled);
int kodama_hip_last_evolution_lds_flows(void);

/* ---- device seam: exact kNN ------------------------------------------------
 * Replaces detail::native_cuda_exact_knn_search (native_cuda_backend.hpp:165-175).
 * train [train_rows x dimensions], query [query_rows x dimensions] row-major
 * float32 (rows already normalised for cosine). Output [query_rows x k]:
 * ascending internal distance (squared L2, or minus the inner product), ties by
 * smaller train index, -1 / +inf beyond min(k, available). query_train_indices
 * (may be NULL) excludes one train index per query. Distances are summed in the
 * reference CPU kernel's order (native_knn.cpp:41-95), so they are bit-identical
 * to the reference's exhaustive search. k <= 128 up to 128 dimensions, k <= 64
 * beyond. */
int kodama_hip_exact_knn_search(const float* train, int32_t train_rows, const float* query,
                                int32_t query_rows, int32_t dimensions, int32_t k,
                                int32_t metric, int32_t device,
                                const int32_t* query_train_indices, int32_t* out_indices,
                                float* out_distances);

/* ---- device seam: resident kNN vote graph ----------------------------------
 * Replace NativeCudaKNNVoteGraph, native_cuda_build_knn_vote_graph and
 * native_cuda_knn_vote_predict_into (native_cuda_backend.hpp:92-131, 277-296).
 * neighbor_rows/scores: [samples x neighbors], -1 / -inf padded. The graph stays
 * resident on the device; each predict uploads dense label codes (0..C-1) and
 * returns the winning code per row (count, then score sum, then smaller code),
 * or fallback_code for rows without neighbours. */
typedef struct kodama_hip_vote_graph kodama_hip_vote_graph;
int kodama_hip_build_knn_vote_graph(const int32_t* neighbor_rows, const float* scores,
                                    int32_t samples, int32_t neighbors, int32_t device,
                                    kodama_hip_vote_graph** out_graph);
int kodama_hip_knn_vote_predict(kodama_hip_vote_graph* graph, const int32_t* label_codes,
                                int32_t fallback_code, int32_t* prediction_codes);
void kodama_hip_release_knn_vote_graph(kodama_hip_vote_graph* graph);

/* ---- kodama-cpp entry points (Backend::HIP twins) -------------------------- */

/* kodama::KNNCV_CUDA (kodama.hpp:871-876; knncv.cpp:422-488) with the exact
 * search. x [n x p] row-major; constrain may be NULL. Outputs predicted labels
 * and fold assignments [n]; *global_accuracy may be NULL. */
int kodama_hip_knncv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                     const int32_t* constrain, const kodama_hip_core_options* options,
                     int32_t* predicted_labels, int32_t* fold_assignments,
                     double* global_accuracy);

/* kodama::CoreKNN_CUDA (kodama.hpp:1010-1016; core.cpp:1782-1812): exact per-fold
 * kNN tables on the device, then maximize_core with the device vote. constrain
 * and fixed may be NULL. */
int kodama_hip_core_knn(const float* x, int32_t n, int32_t p, const int32_t* labels,
                        const int32_t* constrain, const int32_t* fixed,
                        const kodama_hip_core_options* options,
                        kodama_hip_core_result* result);

/* kodama::CoreKNNGraph_CUDA (kodama.hpp:1043-1050; core.cpp:1684-1700): the
 * caller's graph [samples x neighbors] (0- or 1-based per index_base: 0 auto,
 * 1 zero-based, 2 one-based; GraphIndexBase order) with distances. */
int kodama_hip_core_knn_graph(const int32_t* graph_indices, const float* graph_distances,
                              int32_t samples, int32_t neighbors, int32_t index_base,
                              const int32_t* labels, const int32_t* constrain,
                              const int32_t* fixed, const kodama_hip_core_options* options,
                              kodama_hip_core_result* result);

/* kodama::PLSLDACV_CUDA (kodama.hpp:920-925; plscv.cpp:2510-2717): F-fold
 * PLS-LDA CV. Outputs predictions, folds and the selected component count. */
int kodama_hip_plsldacv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                        const int32_t* constrain, const kodama_hip_core_options* options,
                        int32_t* predicted_labels, int32_t* fold_assignments,
                        int32_t* selected_components);

/* kodama::PLSDACV_CUDA (kodama.hpp:913-918; plscv.cpp:2934-2938): the F-fold
 * CV of kodama_hip_plsldacv with the PLS-DA head of run_plscv_host
 * (plscv.cpp:1210-1296, 2430-2442): per fold, the least-squares regression of
 * the centred one-hot training labels on the training scores, solved by
 * Gauss-Jordan elimination with partial pivoting, and the class of largest
 * fitted value. FP64 on the device. */
int kodama_hip_plsdacv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                       const int32_t* constrain, const kodama_hip_core_options* options,
                       int32_t* predicted_labels, int32_t* fold_assignments,
                       int32_t* selected_components);

/* kodama::PLS_CUDA (kodama.hpp:1377-1381; fit_general_pls, plscv.cpp:2765-2878):
 * general PLS of y [n x q] on x [n x p]. X centred / scaled as
 * train_center_scale_float (options center, scale), Y centred; S = X'Y and
 * the Gram matrix X'X on the matrix cores (FP64), power-refresh SIMPLS with
 * A = max(1, min(fixed or max components, p, n - 1)) at most, response
 * weights Q = S'W, coefficients W Q', scores X W and fitted values
 * ybar + X W Q'. Outputs (float, row-major, PLSFitResult): weights [p x A],
 * response_loadings [q x A], scores [n x A] (buffers for the requested A),
 * coefficients [p x q], fitted [n x q], x_center / x_scale [p], y_center [q];
 * *components = A. timings[2] (may be NULL): the cross-product kernel and the
 * whole device fit (seconds, HIP events). A fit without a component is
 * ERUNTIME (the reference's DegeneratePLSFit). Limits (EINVAL beyond): p <= 256,
 * q x p <= 8192, p + q <= 448, at most 64 components. */
int kodama_hip_pls(const float* x, int32_t n, int32_t p, const float* y, int32_t q,
                   const kodama_hip_core_options* options, float* weights,
                   float* response_loadings, float* scores, float* coefficients, float* fitted,
                   float* x_center, float* x_scale, float* y_center, int32_t* components,
                   double* timings);

/* kodama::CorePLSLDA_CUDA (kodama.hpp:1002-1008; core.cpp:1720-1749). */
int kodama_hip_core_plslda(const float* x, int32_t n, int32_t p, const int32_t* labels,
                           const int32_t* constrain, const int32_t* fixed,
                           const kodama_hip_core_options* options,
                           kodama_hip_core_result* result);

/* The M-run form of CorePLSLDA (the scheduler of KODAMAMatrix_CUDA,
 * kodama_matrix.cpp:2651-2682, around CorePLSLDA_CUDA): same indexing as
 * kodama_hip_core_knn_batch; proposals on the host with the reference's random
 * streams (seed + r, folds from cv_seed + r unless fold_ids is given), one
 * device evaluation of all runs x folds per cycle. Per run the result equals
 * kodama_hip_core_plslda on the run's rows. timings[5] as for the kNN batch
 * (plain form), timings[3] = average duration of the PLS-LDA evaluation
 * kernels. */
int kodama_hip_core_plslda_batch(const float* x, int32_t n, int32_t p, int32_t runs,
                                 const int32_t* run_offsets, const int32_t* landmark_rows,
                                 const int32_t* initial_labels, const int32_t* fold_ids,
                                 const kodama_hip_core_options* options, int32_t* clbest,
                                 double* accbest, double* vect_acc, double* timings);

/* No counterpart in the reference (its fold internals are not exported): the
 * detail of one PLS-LDA fold, for callers that check the FP64 device results
 * (tests against an FP64 restatement). Evaluates `labels` under the caller's
 * fold ids and reports fold `fold` (a fold id): SIMPLS weights [p x rank]
 * row-major (rank 0 for single-class or degenerate folds), validation scores
 * [n_val x components] and LDA discriminants [n_val x n_classes] for the
 * fold's validation rows in ascending order, and the fold's training classes
 * (label values, ascending). Buffers: weights p x max(max_components, 1),
 * val_scores and discriminants n x n, classes n. */
/* kodama::PLSLDAPredict_CUDA (kodama.hpp:941-946; CPU semantics plscv.cpp:2976-3005):
 * fit PLS-LDA on train [n_train x p] with labels, predict test [n_test x p]
 * into pred [n_test]. options: max_components / fixed_components, center,
 * scale, device. Empty inputs predict nothing; one training class predicts
 * itself; a degenerate fit is EINVAL (the reference throws DegeneratePLSFit). */
int kodama_hip_plslda_predict(const float* train, int32_t n_train, const int32_t* labels, const float* test,
                              int32_t n_test, int32_t p, const kodama_hip_core_options* options, int32_t* pred);

int kodama_hip_plslda_fold_detail(const float* x, int32_t n, int32_t p, const int32_t* labels,
                                  const int32_t* folds, int32_t fold,
                                  const kodama_hip_core_options* options, double* weights,
                                  int32_t* rank, double* val_scores, double* discriminants,
                                  int32_t* classes, int32_t* n_classes, int32_t* n_val);

/* kodama::KODAMAKNNGraph_CUDA (kodama.hpp:1524-1527; graph_cluster.cpp:790-837):
 * exact self-kNN graph, k = min(k, n-1), self excluded; 1-based ids and public
 * distances (sqrt of squared L2 for Euclidean). Output [n x min(k, n-1)];
 * min(k, n-1) <= 128 up to 128 dimensions, <= 64 beyond. */
int kodama_hip_knn_graph(const float* x, int32_t n, int32_t p, int32_t k, int32_t metric,
                         int32_t device, int32_t* indices_one_based, float* distances);

/* kodama::KODAMADissimilarityInPlace with Backend::CUDA (kodama.hpp:1181-1189;
 * kodama_matrix.cpp:2685-2776, 4383-4463): 1-based graph [samples x neighbors]
 * rewritten in place from the run labels res [runs x samples]. */
int kodama_hip_dissimilarity_inplace(int32_t* indices_one_based, float* distances,
                                     int32_t samples, int32_t neighbors,
                                     const int32_t* run_labels, int32_t runs,
                                     int32_t device);

/* No counterpart in the reference: the allocation-free twin of
 * kodama_hip_dissimilarity_inplace, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same rewrite, on caller-owned DEVICE buffers:
 * d_indices_one_based / d_distances [samples x neighbors] in place,
 * d_run_labels [runs x samples]; enqueued on `stream` (a hipStream_t; NULL =
 * the null stream) of `device`, which the call makes current, and not waited
 * for. The transposed run labels (samples x runs rounded up to 4, int32) live
 * in a library-owned static device array of 64 MiB per GPU, held until the
 * process ends: larger inputs are EINVAL. The call allocates nothing (no
 * hipMalloc, hipFree or pool allocation). Calls that overlap on one GPU share
 * that array and race; serialise them (one stream, or wait between calls). */
int kodama_hip_dissimilarity_inplace_noalloc(int32_t* d_indices_one_based, float* d_distances,
                                             int32_t samples, int32_t neighbors,
                                             const int32_t* d_run_labels, int32_t runs,
                                             int32_t device, void* stream);

/* ---- device-resident graph handle --------------------------------------------
 * kodama::KODAMAGraphHandle + KODAMAGraphMaterialize (kodama.hpp:481-530;
 * kodama_matrix.cpp:2422-2490, 4364-4381) for the exact self-kNN graph: the
 * graph stays on the device in its public form (1-based ids, public
 * distances) until materialised; the KODAMA dissimilarity (kodama_matrix.cpp:
 * 2685-2776) rewrites it there. A handle serialises its own calls; release it
 * once. */
typedef struct kodama_hip_graph kodama_hip_graph;
/* k neighbours per row, capped at n - 1 (as kodama_hip_knn_graph). */
int kodama_hip_graph_build(const float* x, int32_t n, int32_t p, int32_t k, int32_t metric, int32_t device,
                           kodama_hip_graph** out);
int kodama_hip_graph_info(const kodama_hip_graph* graph, int32_t* samples, int32_t* neighbors, int32_t* device);
/* indices_one_based / distances [samples x neighbors]. */
int kodama_hip_graph_materialize(kodama_hip_graph* graph, int32_t* indices_one_based, float* distances);
/* run_labels [runs x samples]; runs <= 0 leaves the graph as is. */
int kodama_hip_graph_dissimilarity(kodama_hip_graph* graph, const int32_t* run_labels, int32_t runs);
void kodama_hip_graph_release(kodama_hip_graph* graph);

/* ---- batched label evolution over M runs ----------------------------------
 * The M-run scheduler of KODAMAMatrix_CUDA (kodama_matrix.cpp:2651-2682,
 * 3823-3850) around CoreKNN_CUDA (3141-3144), extended to a different landmark
 * count per run. x [n x p]; run r evolves the rows
 * landmark_rows[run_offsets[r] .. run_offsets[r+1]) starting from
 * initial_labels (same indexing); folds are make_folds(options->cv_seed + r)
 * unless fold_ids is non-NULL (same indexing as landmark_rows). Seeds:
 * options->seed + r for the proposals. mode 0 = plain (proposals on the host
 * with the reference's random streams: per run bit-identical to
 * kodama_hip_core_knn); mode 1 = device-resident (proposal, transition, vote
 * and acceptance on the device, one persistent launch for all cycles with one
 * workgroup per run, counter-based random streams; same algorithm, different
 * random draws; runs of up to 16383 rows and 1024 classes, no constraint
 * groups or fixed rows).
 * Outputs: clbest (same indexing as landmark_rows), accbest[runs],
 * vect_acc[runs x cycles]; timings[5] (may be NULL): precompute seconds,
 * evolution seconds, steps completed, the device seconds of the dominant
 * kernel (HIP events on its launch stream: mode 0 the average batched vote
 * launch, mode 1 the persistent evolution launch), and the evolution's wall
 * seconds with its host copies. Mode 1 times the evolution with its inputs in
 * device memory (from the launch, the starting labels already uploaded, to
 * the end of the launch); timings[4] adds the label upload and the result
 * download and conversion. Mode 0 moves labels between host and device every
 * cycle by construction, and both figures are the whole loop. */
int kodama_hip_core_knn_batch(const float* x, int32_t n, int32_t p, int32_t runs,
                              const int32_t* run_offsets, const int32_t* landmark_rows,
                              const int32_t* initial_labels, const int32_t* fold_ids,
                              const kodama_hip_core_options* options, int32_t mode,
                              int32_t* clbest, double* accbest, double* vect_acc,
                              double* timings);

/* ---- KODAMA.matrix orchestration -------------------------------------------
 * Flat subset of kodama::KODAMAMatrixOptions (kodama.hpp:532-566) for data
 * without spatial coordinates. */
typedef struct kodama_hip_matrix_options {
  int32_t runs;             /* M */
  int32_t cycles;           /* Tcycle */
  int32_t landmarks;        /* <= 0: 10000; >= n: ceil(0.75 n); then clamped to [2, n-1] */
  int32_t splitting;        /* k-means classes; <= 0: 100 below 40000 rows, else 300 */
  int32_t folds;            /* CV folds of every Core run (reference default 5) */
  int32_t k;                /* knn.k: Core neighbours and projection votes, 1..64 */
  int32_t metric;           /* KODAMA_HIP_COSINE / _INNER_PRODUCT / _EUCLIDEAN */
  int32_t graph_neighbors;  /* global graph width; <= 0: 100 */
  uint64_t seed;            /* KODAMAMatrixOptions::seed (reference default 1234) */
  int32_t form;             /* 0 = plain (reference streams per run), 1 = device-resident,
                               2 = device-resident with the learned proposal (kNN only) */
  int32_t device;
  int32_t n_threads;        /* host threads of the plain form (0 = automatic) */
  const float* mlp_params;  /* form 2: the network, fp32 [10400] (kodama_hip_mlp_train) */
  float mlp_fallback;       /* form 2: share of cycles on the uniform proposal */
  int32_t classifier;       /* KODAMAMatrixOptions::classifier: 1 = KNN, 0 = PLS_LDA
                               (CoreClassifier order) */
  int32_t components;       /* KODAMAMatrixOptions::components (reference default 50),
                               capped at min(n, p): the PLS-LDA components */
} kodama_hip_matrix_options;

/* Landmarks per run and global graph width that kodama_hip_kodama_matrix
 * resolves for n rows (the sizes of its optional outputs). */
int kodama_hip_matrix_sizes(int32_t n, const kodama_hip_matrix_options* options, int32_t* landmarks,
                            int32_t* graph_neighbors);

/* kodama::KODAMAMatrix_CUDA (kodama.hpp:1077-1083; run_kodama_matrix,
 * kodama_matrix.cpp:3324-4051; run_iteration, 2870-3322), without spatial
 * data. A shared k-means atlas of all rows (mt19937_64(seed + 1)); for run
 * r = 1..M an mt19937_64(seed + r) that skips one draw and quota-samples L
 * landmarks over the atlas strata; the landmarks' starting labels are
 * starting_labels (the reference's starting_labels: the search is then greedy,
 * plain form only) or the k-means labels of the landmarks. Constraint groups
 * (constrain, [n], any values; NULL: every row its own group) give their
 * members their majority starting label, move together in the Core and take
 * their majority label in the result; fixed rows ([n], nonzero = fixed; NULL:
 * none) keep their label. Core on the landmarks (non-stratified folds, seed +
 * r): kNN with per-fold exact neighbours, or PLS-LDA with `components` fixed
 * components and class coarsening. Every other row takes, for kNN, the
 * majority label of the first k landmarks among its global graph neighbours
 * (exact self-kNN graph), for PLS-LDA, the PLSLDAPredict label of a fit on the
 * run's landmarks. The k-means runs on the device in the reference's float
 * order, so atlas, landmarks and starting labels equal the reference CPU
 * build's. The device-resident forms (1, 2) run the kNN classifier without
 * starting labels, constraint groups or fixed rows.
 * Outputs: res [runs x n] labels, acc [runs] (accbest), v [runs x cycles]
 * (may be NULL), landmark_rows [runs x L] (may be NULL), the global graph
 * [n x graph_neighbors] 1-based with public distances (both NULL to skip; the
 * PLS-LDA classifier builds it only when asked), timings[6] (may be NULL):
 * atlas, graph, landmarks + starting labels, Core precompute, Core evolution,
 * projection (seconds). */
int kodama_hip_kodama_matrix(const float* x, int32_t n, int32_t p,
                             const kodama_hip_matrix_options* options, const int32_t* starting_labels,
                             const int32_t* constrain, const int32_t* fixed, int32_t* res, double* acc,
                             double* v, int32_t* landmark_rows, int32_t* graph_indices_one_based,
                             float* graph_distances, double* timings);

/* kodama::KODAMAMatrixFromGraphData_CUDA (kodama.hpp:1215-1222;
 * kodama_matrix.cpp:4672-4690): kodama_hip_kodama_matrix with the caller's
 * neighbour graph in place of the exact global graph. graph_indices /
 * graph_distances_in [n x graph_neighbors] row-major; index_base 0, 1, or -1
 * (auto: one-based when every id lies in 1..n), normalised as
 * normalize_external_graph (kodama_matrix.cpp:1292-1330): per row the first
 * K = min(graph_neighbors, resolved graph width) valid neighbours (in range,
 * not the row itself, finite distance), distances clamped at 0. With the kNN
 * classifier the graph also drives the Core: each run's landmark subgraph
 * (subset_graph_to_rows, kodama_matrix.cpp:1332-1366) feeds CoreKNNGraph
 * (core.cpp:624-692): per landmark row the first min(k, K) landmark
 * neighbours outside its fold, score -d^2 (Euclidean) or 1 - d. The graph
 * outputs (may be NULL) are [n x K], 1-based, the caller's distances. */
int kodama_hip_kodama_matrix_from_graph(const float* x, int32_t n, int32_t p, const int32_t* graph_indices,
                                        const float* graph_distances_in, int32_t graph_neighbors,
                                        int32_t index_base, const kodama_hip_matrix_options* options,
                                        const int32_t* starting_labels, const int32_t* constrain,
                                        const int32_t* fixed, int32_t* res, double* acc, double* v,
                                        int32_t* landmark_rows, int32_t* graph_indices_one_based,
                                        float* graph_distances, double* timings);

/* ---- learned proposal -------------------------------------------------------
 * No counterpart in the reference: the learned proposal that the device-resident
 * evolution can use in place of the uniform one, for callers that train it on
 * the plain method's accept/reject traces and run it. A 64-64-64-32 MLP (ReLU)
 * over per-row features of the neighbour vote: [0,30) the vote counts of the
 * row's 30 most voted labels (count desc, code asc) over k, [30,61) one-hot of
 * its own label's rank (60: not voted), 61 the margin (top - own count) / k,
 * 62 a constant 1. Outputs: j < 30 relabel to the label of rank j, 30 keep;
 * ranks that do not exist or hold the row's own label are masked. Weights
 * and inputs are rounded to bf16, sums are FP32 (the forward pass runs 16 rows
 * per matrix-core tile, hidden activations entering as three bf16 parts, which
 * gives the FP32 sums up to their order). params: fp32 [10400] =
 * W1[64x64] b1[64] W2[64x64] b2[64] W3[64x32] b3[32], W row-major [in][out]. */
int kodama_hip_mlp_init(uint64_t seed, float* params);

/* Features of a table's rows: neighbors [rows x k] row ids (-1 missing),
 * labels [rows] codes. features [rows x 64], ranked [rows x 30] (-1 past the
 * voted labels). */
int kodama_hip_mlp_features(const int32_t* neighbors, const int32_t* labels, int32_t rows, int32_t k,
                            int32_t device, float* features, int32_t* ranked);

/* Forward pass: probs [rows x 32] (masked outputs 0). */
int kodama_hip_mlp_forward(const float* params, const float* features, int32_t rows, int32_t device,
                           float* probs);

/* Gradient of the mean cross-entropy over the rows (targets 0..30) with
 * respect to params (the bf16-rounded weights' gradient, applied to the fp32
 * master); *loss = the mean loss (loss may be NULL). */
int kodama_hip_mlp_gradients(const float* params, const float* features, const int32_t* targets,
                             int32_t rows, int32_t device, float* grads, float* loss);

/* The plain batched evolution (kodama_hip_core_knn_batch mode 0) recording its
 * accept/reject samples: every row a proposal relabels (to its prediction),
 * its features under the labels the proposal started from, and its target
 * (the rank of the new label when the proposal was accepted, 30 otherwise).
 * features [max_samples x 64], targets [max_samples]; *samples = recorded;
 * accbest [runs] may be NULL. */
int kodama_hip_mlp_collect_traces(const float* x, int32_t n, int32_t p, int32_t runs,
                                  const int32_t* run_offsets, const int32_t* landmark_rows,
                                  const int32_t* initial_labels, const int32_t* fold_ids,
                                  const kodama_hip_core_options* options, int32_t max_samples,
                                  float* features, int32_t* targets, int32_t* samples, double* accbest);

/* The same run recording outcome samples: at every cycle every row whose
 * label differs from its prediction, its features under the current labels,
 * and its target: the rank of the label it holds in the run's final best
 * labelling (30 when that is its current label; rows whose final label is
 * not among the voted ranks are skipped). Same buffers and outputs. */
int kodama_hip_mlp_collect_outcomes(const float* x, int32_t n, int32_t p, int32_t runs,
                                    const int32_t* run_offsets, const int32_t* landmark_rows,
                                    const int32_t* initial_labels, const int32_t* fold_ids,
                                    const kodama_hip_core_options* options, int32_t max_samples,
                                    float* features, int32_t* targets, int32_t* samples,
                                    double* accbest);

/* Minibatch SGD with momentum on the device (deterministic for a seed):
 * params in/out; loss_per_epoch [epochs] may be NULL. */
int kodama_hip_mlp_train(const float* features, const int32_t* targets, int32_t samples, int32_t epochs,
                         int32_t batch, float learning_rate, float momentum, uint64_t seed,
                         int32_t device, float* params, double* loss_per_epoch);

/* kodama_hip_core_knn_batch in the device-resident form with the learned
 * proposal, one persistent launch for all cycles: in each cycle the kernel
 * scores the rows whose label differs from their prediction with the network,
 * chooses rows with probability take x score / (sum of scores) and relabels
 * them to the network's target; with probability `fallback` a cycle uses the
 * uniform proposal instead. The CV accuracy stays the only acceptance
 * criterion. EINVAL when the network (21 KB) and the per-wave tiles (24 KB)
 * do not fit one CU's LDS beside the largest run's state. Outputs as the
 * batch. */
int kodama_hip_core_knn_batch_learned(const float* x, int32_t n, int32_t p, int32_t runs,
                                      const int32_t* run_offsets, const int32_t* landmark_rows,
                                      const int32_t* initial_labels, const int32_t* fold_ids,
                                      const kodama_hip_core_options* options, const float* params,
                                      float fallback, int32_t* clbest, double* accbest, double* vect_acc,
                                      double* timings);

#ifdef __cplusplus
}
#endif

#endif /* KODAMA_HIP_H */
