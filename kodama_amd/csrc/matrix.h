// Host side of the KODAMA.matrix orchestration (matrix.cpp): the landmark atlas
// and per-run landmarks, k-means on the device, the batched Core and the
// projection of the other rows.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <random>
#include <vector>

#include "engine.h"

namespace kodama_hip {

// Lloyd k-means with the reference's seeding, float order, empty-cluster
// splitting and stopping rule (kodama_matrix.cpp:664-799); one draw of `rng`.
// h_x / d_x: the same rows [n x p] on the host and on the device. Returns
// 1-based labels.
IVec device_kmeans(const float* h_x, const float* d_x, int n, int p, int k, std::mt19937_64& rng, int max_iter,
                   hipStream_t s);

// device_kmeans of `runs` problems of n rows each, in lockstep: problem r's
// rows are data rows rows[r * n .. (r + 1) * n) of h_x (host, for the seeds)
// gathered into d_rows [runs x n x p] (device), its draw comes from rngs[r].
// Every launch serves all problems still iterating; the results equal runs
// separate device_kmeans calls.
std::vector<IVec> device_kmeans_batch(const float* h_x, const int* rows, const float* d_rows, int runs, int n, int p,
                                      int k, std::vector<std::mt19937_64>& rngs, int max_iter, hipStream_t s);

// Rows of each 1-based stratum in ascending order (kodama_matrix.cpp:800-824).
void index_strata(const IVec& labels, int strata, IVec& offsets, IVec& rows);

// Quota sample of `target` rows over shuffled strata, sorted
// (kodama_matrix.cpp:856-944).
IVec quota_sample_landmarks(const IVec& offsets, const IVec& rows, int target, std::mt19937_64& rng);

// Constraint groups of KODAMA.matrix (kodama_matrix.cpp:169-216, 988-1011):
// normalize_constrain renumbers the caller's values 1.. by first occurrence
// (none: every row its own group); is_identity_constrain; factor_subset
// renumbers the groups of `rows` 1.. by first occurrence; constrained_majority
// gives every member of a group the group's majority label (ties: smaller).
IVec normalize_constrain(const int32_t* constrain, int n);
bool is_identity_constrain(const IVec& groups);
IVec factor_subset(const IVec& values, const IVec& rows);
int majority_label(const IVec& labels);
IVec constrained_majority(const IVec& labels, const IVec& groups);

}  // namespace kodama_hip
