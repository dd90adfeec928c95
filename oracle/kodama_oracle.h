/*
 * CPU oracle for the KODAMA label-evolution hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This header declares the C ABI of liboracle.so (a C++ restatement of the
 * reference algorithms, kodama_oracle.cpp) and of oracle/_ref/libkodama_ref.so
 * (the reference's own vendored C++ core driven through ref_driver.cpp).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * either library, and only as a checker or as the timed CPU baseline. The product
 * (kodama_amd/) never links, loads or calls anything here.
 */
#ifndef KODAMA_ORACLE_H
#define KODAMA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flat mirror of kodama::CoreOptions + KNNOptions + CorePLSLDAOptions
 * (kodama.hpp:149-179, 294-314) restricted to the fields the hot path reads. */
typedef struct oracle_core_options {
  int32_t cycles;                 /* CoreOptions::cycles */
  int32_t classifier;             /* 0 = PLS_LDA, 1 = KNN (CoreClassifier order) */
  uint64_t seed;                  /* CoreOptions::seed (mt19937_64 of the proposals) */
  int32_t shake;
  int32_t evolutionary_search;
  int32_t guarded_diversity;
  int32_t many_to_one_absorption;
  int32_t auto_class_coarsening;
  int32_t adaptive_proposal_size;
  /* EvolutionPolicy (kodama.hpp:49-61) */
  int32_t prediction_guidance;
  int32_t policy_adaptive_proposal_size;
  int32_t transition_proposal;
  int32_t stochastic_acceptance;
  int32_t diversity_multiplier;
  int32_t pls_transition_coarsening;
  int32_t pls_fragmentation_penalty;
  int32_t reject_single_class;
  /* cross-validation (FoldOptions) shared by the kNN and PLS classifiers */
  int32_t folds;
  int32_t stratified;
  uint64_t cv_seed;
  /* kNN */
  int32_t k;
  int32_t metric;                 /* 0 = Cosine, 1 = InnerProduct, 2 = Euclidean */
  /* PLS-LDA */
  int32_t max_components;
  int32_t fixed_components;
  int32_t center;
  int32_t scale;
  int32_t pls_fold_cache;         /* 1 = data_epoch != 0 (CorePLSLDA), 0 = fresh folds */
  /* reference-only knobs (ignored by the oracle) */
  int32_t hnsw_ef_search;
  int32_t n_threads;
} oracle_core_options;

/* Outputs of one maximize_core run (kodama::CoreResult, kodama.hpp:316-349).
 * Every pointer is caller-allocated: per-sample arrays hold n entries, per-cycle
 * arrays hold `cycles` entries. counters[] holds, in this order:
 * cycles_completed, proposals_evaluated, best_state_updates, current_state_accepts,
 * stochastic_state_attempts, stochastic_state_accepts, current_state_rejections,
 * coarsening_moves, absorption_moves, transition_attempted, transition_accepted,
 * many_to_one_attempted, many_to_one_accepted, pls_coarsening_attempted,
 * pls_coarsening_accepted, cv_evaluations, success. */
typedef struct oracle_core_result {
  int32_t* clbest;
  int32_t* clbest_dirty;
  int32_t* cvpredbest;
  int32_t* fold_assignments;
  double* vect_acc;
  double* vect_score;
  int32_t* proposal_size;
  int32_t* active_classes;
  uint8_t* accepted;
  uint8_t* improving_acceptance;
  uint8_t* temperature_acceptance;
  int32_t* counters;              /* ORACLE_N_COUNTERS entries */
  double accbest;
  double scorebest;
} oracle_core_result;

#define ORACLE_N_COUNTERS 17

/* ---- oracle (liboracle.so) ------------------------------------------------ */
int oracle_make_folds(const int32_t* labels, int32_t n, const int32_t* constrain,
                      int32_t folds, int32_t stratified, uint64_t seed, int32_t* out);
float oracle_distance(const float* a, const float* b, int32_t p, int32_t metric);
void oracle_prepare_rows(const float* x, int32_t n, int32_t p, int32_t metric,
                         int32_t variant, float* out);
int oracle_exact_knn(const float* train, int32_t nt, const float* query, int32_t nq,
                     int32_t p, int32_t k, int32_t metric, const int32_t* exclude,
                     int32_t* out_idx, float* out_dist);
int oracle_precompute_knn_cv(const float* x, int32_t n, int32_t p, const int32_t* folds,
                             int32_t k, int32_t metric, int32_t* nbr, float* score);
int oracle_knn_vote(const int32_t* nbr, const float* score, int32_t n, int32_t width,
                    const int32_t* folds, const int32_t* labels, int32_t* pred);
double oracle_objective(const int32_t* labels, int32_t n, double accuracy,
                        const oracle_core_options* opt);
int oracle_core_knn(const float* x, int32_t n, int32_t p, const int32_t* labels,
                    const int32_t* constrain, const int32_t* fixed,
                    const oracle_core_options* opt, oracle_core_result* out);
int oracle_core_knn_tables(const int32_t* nbr, const float* score, int32_t n,
                           int32_t width, const int32_t* folds, const int32_t* labels,
                           const int32_t* constrain, const int32_t* fixed,
                           const oracle_core_options* opt, oracle_core_result* out);
int oracle_knncv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                 const int32_t* constrain, const oracle_core_options* opt,
                 int32_t* pred, int32_t* folds_out);
int oracle_knn_graph(const float* x, int32_t n, int32_t p, int32_t k, int32_t metric,
                     int32_t* idx_one_based, float* dist);
int oracle_dissimilarity(int32_t* indices, float* distances, int32_t samples,
                         int32_t neighbors, const int32_t* res, int32_t runs,
                         int32_t input_one_based, int32_t output_one_based);
int oracle_plsldacv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                    const int32_t* constrain, const oracle_core_options* opt,
                    int32_t precision, int32_t* pred, int32_t* folds_out,
                    int32_t* selected_components);
/* PLSDACV: the F-fold PLS-DA CV (same folds and SIMPLS, least-squares head). */
int oracle_plsdacv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                   const int32_t* constrain, const oracle_core_options* opt,
                   int32_t precision, int32_t* pred, int32_t* folds_out,
                   int32_t* selected_components);
/* PLS_CPU: general PLS of y [n x q] on x [n x p]; outputs in double, A =
 * *components columns: weights [p x A], response_loadings [q x A], scores
 * [n x A], coefficients [p x q], fitted [n x q], x_center / x_scale [p],
 * y_center [q]. */
int oracle_pls(const float* x, int32_t n, int32_t p, const float* y, int32_t q,
               const oracle_core_options* opt, int32_t precision, double* weights,
               double* response_loadings, double* scores, double* coefficients, double* fitted,
               double* x_center, double* x_scale, double* y_center, int32_t* components);
int ref_plsdacv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                const int32_t* constrain, const oracle_core_options* opt,
                int32_t* pred, int32_t* folds_out, int32_t* selected_components);
int ref_pls(const float* x, int32_t n, int32_t p, const float* y, int32_t q,
            const oracle_core_options* opt, double* weights, double* response_loadings,
            double* scores, double* coefficients, double* fitted, double* x_center,
            double* x_scale, double* y_center, int32_t* components);
int oracle_pls_fold_detail(const float* x, int32_t n, int32_t p, const int32_t* labels,
                           const int32_t* folds, int32_t fold, const oracle_core_options* opt,
                           int32_t precision, double* weights, int32_t* rank,
                           double* val_scores, double* discriminants, int32_t* classes,
                           int32_t* n_classes, int32_t* n_val);
int oracle_core_plslda(const float* x, int32_t n, int32_t p, const int32_t* labels,
                       const int32_t* constrain, const int32_t* fixed,
                       const oracle_core_options* opt, int32_t precision,
                       oracle_core_result* out);
/* KODAMA.matrix building blocks (kodama_matrix.cpp): k-means labels (1-based)
 * from an mt19937_64(rng_seed) that first skips `discard` draws; quota-sampled
 * landmark rows of the strata given as 1-based labels. */
int oracle_kmeans_labels(const float* x, int32_t n, int32_t p, int32_t k, uint64_t rng_seed,
                         int32_t discard, int32_t max_iter, int32_t* labels);
int oracle_quota_sample(const int32_t* strata_labels, int32_t n, int32_t strata, int32_t target,
                        uint64_t rng_seed, int32_t discard, int32_t* rows_out, int32_t* count_out);

/* KODAMAMatrixOptions subset without spatial data (kodama.hpp:532-566). */
typedef struct oracle_matrix_options {
  int32_t runs;
  int32_t cycles;
  int32_t landmarks;
  int32_t splitting;
  int32_t folds;
  int32_t k;
  int32_t metric;
  int32_t graph_neighbors;
  uint64_t seed;
  int32_t classifier;   /* 1 = KNN, 0 = PLS_LDA (CoreClassifier order) */
  int32_t components;   /* PLS-LDA components (KODAMAMatrixOptions::components) */
  int32_t precision;    /* PLS-LDA arithmetic: 32 (the reference's float) or 64 */
} oracle_matrix_options;

/* res [runs x n], acc [runs], v [runs x cycles], landmark_rows [runs x L] (may
 * be NULL), *landmarks_used = L; starting_labels / constrain / fixed [n] (each
 * may be NULL: the reference's empty vectors); diag [runs x 3] (may be NULL):
 * FNV-1a hashes of each run's landmark rows, starting labels and fold ids. */
int oracle_kodama_matrix(const float* x, int32_t n, int32_t p, const oracle_matrix_options* options,
                         const int32_t* starting_labels, const int32_t* constrain, const int32_t* fixed,
                         int32_t* res, double* acc, double* v, int32_t* landmark_rows,
                         int32_t* landmarks_used, uint64_t* diag);

/* The same with the caller's graph [n x graph_neighbors] (index_base 0, 1 or
 * -1 = auto) in place of the exact global graph (KODAMAMatrixFromGraphData). */
int oracle_kodama_matrix_from_graph(const float* x, int32_t n, int32_t p, const oracle_matrix_options* options,
                                    const int32_t* graph_indices, const float* graph_distances,
                                    int32_t graph_neighbors, int32_t index_base, const int32_t* starting_labels,
                                    const int32_t* constrain, const int32_t* fixed, int32_t* res, double* acc,
                                    double* v, int32_t* landmark_rows, int32_t* landmarks_used, uint64_t* diag);

/* The reference's own KODAMAMatrix_CPU / KODAMAMatrixFromGraphData_CPU
 * (oracle/_ref, one thread, no visual initialisation, no dissimilarity):
 * res, acc, v and diag as above; graph_indices NULL for KODAMAMatrix_CPU. */
int ref_kodama_matrix(const float* x, int32_t n, int32_t p, const oracle_matrix_options* options,
                      const int32_t* graph_indices, const float* graph_distances, int32_t graph_neighbors,
                      int32_t index_base, const int32_t* starting_labels, const int32_t* constrain,
                      const int32_t* fixed, int32_t* res, double* acc, double* v, uint64_t* diag);

/* normalize_external_graph: outputs [n x width], 0-based, -1 / +inf padded. */
int oracle_normalize_graph(const int32_t* graph_indices, const float* graph_distances, int32_t n,
                           int32_t graph_neighbors, int32_t index_base, int32_t width, int32_t* indices_out,
                           float* distances_out);
int ref_normalize_graph(const int32_t* graph_indices, const float* graph_distances, int32_t n,
                        int32_t graph_neighbors, int32_t index_base, int32_t width, int32_t* indices_out,
                        float* distances_out);

const char* oracle_last_error(void);

/* ---- reference driver (oracle/_ref/libkodama_ref.so) ---------------------- */
int ref_make_folds(const int32_t* labels, int32_t n, const int32_t* constrain,
                   int32_t folds, int32_t stratified, uint64_t seed, int32_t* out);
int ref_hnsw_search(const float* train, int32_t nt, const float* query, int32_t nq,
                    int32_t p, int32_t k, int32_t metric, int32_t ef_search,
                    int32_t* out_idx, float* out_dist);
int ref_core_knn(const float* x, int32_t n, int32_t p, const int32_t* labels,
                 const int32_t* constrain, const int32_t* fixed,
                 const oracle_core_options* opt, oracle_core_result* out);
int ref_core_knn_graph(const int32_t* idx, const float* dist, int32_t samples,
                       int32_t neighbors, const int32_t* labels, const int32_t* constrain,
                       const int32_t* fixed, const oracle_core_options* opt,
                       oracle_core_result* out);
int ref_knncv(const float* x, int32_t n, int32_t p, const int32_t* labels,
              const int32_t* constrain, const oracle_core_options* opt,
              int32_t* pred, int32_t* folds_out);
int ref_plslda_predict(const float* train, int32_t n_train, const int32_t* labels, const float* test,
                       int32_t n_test, int32_t p, const oracle_core_options* opt, int32_t* pred);
int ref_plsldacv(const float* x, int32_t n, int32_t p, const int32_t* labels,
                 const int32_t* constrain, const oracle_core_options* opt,
                 int32_t* pred, int32_t* folds_out, int32_t* selected_components);
int ref_core_plslda(const float* x, int32_t n, int32_t p, const int32_t* labels,
                    const int32_t* constrain, const int32_t* fixed,
                    const oracle_core_options* opt, oracle_core_result* out);
int ref_dissimilarity(int32_t* indices, float* distances, int32_t samples,
                      int32_t neighbors, const int32_t* res, int32_t runs);
int ref_knn_graph(const float* x, int32_t n, int32_t p, int32_t k, int32_t metric,
                  int32_t n_threads, int32_t* idx_one_based, float* dist);
const char* ref_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* KODAMA_ORACLE_H */
