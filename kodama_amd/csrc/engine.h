// Host engine of the HIP backend: cross-validation folds, the label-evolution
// state machine of one run (split into propose / finish so that the votes of
// many runs share one device launch per cycle) and the device-resident kNN
// tables with their batched vote.
#pragma once

#include <cstdint>
#include <functional>
#include <random>
#include <vector>

#include "../../include/kodama_hip.h"
#include "device.h"

namespace kodama_hip {

using IVec = std::vector<int>;

struct InvalidArgument : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};

// detail::make_folds (common.cpp:108-214).
IVec make_folds(const IVec& labels, const IVec& constrain, int folds, bool stratified, std::uint64_t seed);

// Row preparation for cosine search: variant 0 as core.cpp:159-193 (CoreKNN),
// variant 1 as native_knn.cpp:1005-1035 (KNNCV, graphs). Other metrics copy.
std::vector<float> prepare_rows(const float* x, int n, int p, int metric, int variant);

// Options of maximize_core (core.cpp:1268-1549) in boolean form.
struct CoreFlags {
  int cycles = 0;
  bool knn = true;
  std::uint64_t seed = 1;
  bool shake = false, evolutionary = false, guarded = false, absorption = false, coarsening = false;
  bool adaptive = true;
  bool guidance = true, policy_adaptive = true, transition = true, stochastic = true;
  bool diversity = true, pls_coarsening = true, fragmentation = true, reject_single = true;
  static CoreFlags from(const kodama_hip_core_options& o);
};

// Label statistics of one evaluated labelling, reduced on the device.
struct Tally {
  int matches = 0;          // rows whose prediction equals the label
  const int* counts = nullptr;  // [classes] rows per label code
  const int* firsts = nullptr;  // [classes] first row of each label code
};

// core_objective_score (core.cpp:916-972) from class counts. The class
// frequencies are summed in the iteration order of the reference's
// unordered_map (filled in sample order after reserve(n)): the map is rebuilt
// here by inserting the present labels in order of first occurrence, which
// yields the same bucket chains and so the same double result.
double objective_from_counts(const int* counts, const int* firsts, const IVec& label_values, int n,
                             double accuracy, const CoreFlags& f);

// One run of maximize_core over label codes 0..C-1 (code order = label order).
class RunEvolution {
 public:
  RunEvolution(IVec label_values, IVec initial_codes, const IVec& constrain, IVec fixed, const CoreFlags& f);

  int rows() const { return n_; }
  int classes() const { return static_cast<int>(label_values_.size()); }
  const IVec& initial() const { return best_; }
  void begin(const int* pred, const Tally& t);
  const IVec& propose(int cycle);
  void finish(int cycle, const int* pred, const Tally& t);
  void export_result(kodama_hip_core_result* out) const;
  double accbest() const { return accbest_; }
  const IVec& best_codes() const { return best_; }
  const std::vector<double>& vect_acc() const { return vacc_; }
  const IVec& label_values() const { return label_values_; }
  // Trace access for the learned proposal: the current state a proposal starts
  // from, the groups it sampled (the first proposal_size(cycle) entries of
  // sampled_groups()) and whether the cycle's proposal became the current state.
  const IVec& current_labels() const { return cur_; }
  const IVec& current_predictions() const { return curpred_; }
  const IVec& sampled_groups() const { return order_; }
  int proposal_size(int cycle) const { return psize_[cycle]; }
  bool accepted(int cycle) const { return accepted_[cycle] != 0; }
  bool fixed_row(int i) const { return fixed_[i] == 1; }
  bool singleton_groups() const { return singletons_; }

 private:
  struct Flows;  // per-class transition counts
  void build_flows(const IVec& cl, const IVec& guide);
  bool coarsen(IVec& cl);
  bool absorb(IVec& cl);

  CoreFlags f_;
  int n_ = 0;
  IVec label_values_;
  IVec fixed_;
  bool singletons_ = true;
  std::vector<IVec> members_;
  int groups_ = 0;
  std::mt19937_64 rng_;

  IVec best_, best_dirty_, bestpred_, cur_, curpred_;
  IVec best_counts_, cur_counts_;
  double accbest_ = 0.0, scorebest_ = 0.0, cur_acc_ = 0.0, cur_score_ = 0.0;
  bool success_ = false;

  IVec cl_, cl_dirty_, order_;
  bool coarsened_ = false, absorbed_ = false;

  // transition statistics scratch
  IVec sizes_, movable_, present_, flow_start_, flow_dest_, flow_count_, tmp_, touched_, bucket_, bucket_start_;

  std::vector<double> vacc_, vscore_;
  IVec psize_, active_;
  std::vector<uint8_t> accepted_, improving_, temperature_;
  int counters_[KODAMA_HIP_N_COUNTERS] = {0};
};

// Device-resident exact kNN tables of a batch of runs (per-fold neighbours,
// core.cpp:463-584) and the batched vote (core.cpp:784-914).
class KnnBatch {
 public:
  // x: prepared rows [n x p] (host). Run r owns rows[run_offsets[r] ..
  // run_offsets[r+1]) (data row ids) with fold ids `folds` (same indexing).
  // table_path: how the per-fold tables are built (device_fold_tables).
  KnnBatch(int device, const float* x, int n, int p, const IVec& run_offsets, const IVec& rows,
           const IVec& folds, int k, int metric, int classes, int table_path = 0);
  // Runs from caller tables [run_offsets.back() x k] row-major: run r owns
  // rows [run_offsets[r], run_offsets[r+1]) with run-local neighbour ids
  // (-1 / -inf padded).
  KnnBatch(int device, const int* neighbors, const float* scores, const IVec& run_offsets, int k, int classes);

  int runs() const { return runs_; }
  int k() const { return k_; }
  int pad_offset(int r) const { return pad_offsets_[r]; }
  int pad_total() const { return pad_total_; }
  int max_rows() const { return max_rows_; }
  int run_rows(int r) const { return run_offsets_[r + 1] - run_offsets_[r]; }
  int classes() const { return classes_; }
  hipStream_t stream() const { return stream_.get(); }

  // Pinned staging: label codes in (padded layout), predictions / reductions out.
  int* labels_host() { return h_labels_.get(); }
  // Pinned staging of the device-resident evolution (allocated with the batch,
  // outside its timed region): the upload image and the download of its
  // results. Null when `bytes` exceeds the staging (the caller then stages in
  // pageable memory). One evolution of a batch at a time: its upload is
  // complete once its results are exported (the export synchronises).
  unsigned char* stage_up(size_t bytes) { return bytes <= h_stage_up_.size() ? h_stage_up_.get() : nullptr; }
  unsigned char* stage_down(size_t bytes) { return bytes <= h_stage_down_.size() ? h_stage_down_.get() : nullptr; }
  const int* predictions_host() const { return h_pred_.get(); }
  const int* counts_host() const { return h_counts_.get(); }
  const int* firsts_host() const { return h_firsts_.get(); }
  const int* matches_host() const { return h_matches_.get(); }

  // Upload labels_host(), vote all runs, download the results; blocks.
  void vote_all(bool reduce = true);
  // From the next vote_all(true) on, the upload, the vote and the downloads
  // run as one captured graph (hipGraph), replayed once per call; the first
  // such call runs eagerly (timed) and captures. The staging buffers keep
  // their addresses, so every replay reads the labels of its call.
  void use_graph(bool on) { graph_on_ = on; }
  // Rows (local to run r) whose neighbour list is empty.
  const IVec& empty_rows(int r) const { return empty_rows_[r]; }
  // Download a run's tables as [L x k] row-major with local ids (tests, APIs).
  void download_tables(int r, int* nbr, float* score);
  float vote_ms_total() const { return vote_ms_; }
  int vote_calls() const { return vote_calls_; }

  DeviceBuffer<int>& d_neighbors() { return d_nbr_; }
  DeviceBuffer<float>& d_scores() { return d_score_; }
  DeviceBuffer<int>& d_labels() { return d_labels_; }
  DeviceBuffer<int>& d_pred() { return d_pred_; }
  DeviceBuffer<int>& d_block_run() { return d_block_run_; }
  DeviceBuffer<int>& d_run_offsets() { return d_run_off_; }
  DeviceBuffer<int>& d_pad_offsets() { return d_pad_off_; }
  const IVec& run_offsets() const { return run_offsets_; }
  const IVec& pad_offsets() const { return pad_offsets_; }

 private:
  void allocate(const IVec& run_offsets);
  void find_empty_rows();

  Stream stream_;
  int runs_ = 0, k_ = 0, pad_total_ = 0, classes_ = 0, blocks_ = 0, max_rows_ = 0;
  IVec run_offsets_, pad_offsets_;
  DeviceBuffer<int> d_nbr_, d_labels_, d_pred_, d_block_run_, d_run_off_, d_pad_off_;
  DeviceBuffer<float> d_score_;
  DeviceBuffer<int> d_counts_, d_firsts_, d_matches_;
  PinnedBuffer<int> h_labels_, h_pred_, h_counts_, h_firsts_, h_matches_;
  PinnedBuffer<unsigned char> h_stage_up_, h_stage_down_;
  std::vector<IVec> empty_rows_;
  EventTimer timer_;
  float vote_ms_ = 0.0f;
  int vote_calls_ = 0;
  bool graph_on_ = false;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  void enqueue_vote(bool reduce, bool timed);

 public:
  ~KnnBatch();
  KnnBatch(const KnnBatch&) = delete;
  KnnBatch& operator=(const KnnBatch&) = delete;
};

// Sorted distinct labels and the code of each row.
void encode_labels(const int* labels, int n, IVec& values, IVec& codes);

// Drive plain (host-proposal) evolution of every run in `runs` against the
// batch's device vote: one vote launch per cycle for all runs. run i of `runs`
// is batch run i. Rows without neighbours get the training-fold majority label
// (core.cpp:274-309). n_threads host workers run the proposals. `hook`
// (optional) is called on the host thread once per cycle after the vote of the
// proposals (after_finish false: the runs still hold their pre-proposal
// state) and once after every run has accepted or rejected (true).
using CycleHook = std::function<void(int cycle, bool after_finish)>;
void evolve_plain(std::vector<RunEvolution>& runs, KnnBatch& batch, const std::vector<IVec>& folds,
                  int cycles, int n_threads, const CycleHook& hook = CycleHook());

int resolve_threads(int requested, int runs);

}  // namespace kodama_hip
