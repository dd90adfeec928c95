// Host orchestration of the device-resident evolution: state allocation, the
// evaluation of the starting labels, one captured hipGraph per cycle replayed
// `cycles` times, and result export.
#include "evolve_device.h"

#include <algorithm>
#include <cmath>

#include "evolve_launch.h"
#include "launch.h"

namespace kodama_hip {

namespace {
uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
}  // namespace

struct DeviceEvolution::Impl {
  KnnBatch& batch;
  std::vector<IVec> values;
  int runs = 0, classes = 0, cycles_cap = 0, blocks = 0;
  DeviceBuffer<int> cur, curpred, best, bestpred, sizes, flows, remap, counts, firsts, matches, decision, cycle,
    perm_half;
  DeviceBuffer<uint64_t> run_key;
  DeviceBuffer<double> state, history;
  EvoParams prm;
  VoteParams vote;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int done_cycles = 0;

  explicit Impl(KnnBatch& b) : batch(b) {}
  ~Impl() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }

  void enqueue_cycle(hipStream_t s) {
    check(launch_evo_clear(prm, s), "evo_clear");
    check(launch_evo_propose(prm, blocks, s), "evo_propose");
    check(launch_evo_absorb(prm, s), "evo_absorb");
    check(launch_evo_remap(prm, blocks, s), "evo_remap");
    check(launch_vote(vote, blocks, s), "vote");
    check(launch_evo_accept(prm, 0, s), "evo_accept");
    check(launch_evo_commit(prm, blocks, s), "evo_commit");
  }
};

DeviceEvolution::DeviceEvolution(KnnBatch& batch, const std::vector<IVec>& values, const std::vector<IVec>& codes,
                                 const std::vector<IVec>& folds, const kodama_hip_core_options& options) {
  (void)folds;
  if (options.evolutionary_search == 0) {
    throw InvalidArgument("the device-resident evolution implements the evolutionary KODAMA.matrix loop.");
  }
  auto impl = std::make_unique<Impl>(batch);
  Impl& m = *impl;
  m.values = values;
  m.runs = batch.runs();
  m.classes = batch.classes();
  m.cycles_cap = std::max(1, options.cycles);
  const size_t flow_items = static_cast<size_t>(m.runs) * m.classes * m.classes;
  if (flow_items > (static_cast<size_t>(1) << 28)) {
    throw InvalidArgument("device-resident evolution: runs x classes^2 exceeds the flow-table limit (2^28).");
  }
  const int pad = batch.pad_total();
  const hipStream_t s = batch.stream();
  m.blocks = pad / 256;
  for (auto* b : {&m.cur, &m.curpred, &m.best, &m.bestpred}) b->resize(static_cast<size_t>(pad));
  const size_t rc = static_cast<size_t>(m.runs) * m.classes;
  m.sizes.resize(rc);
  m.remap.resize(rc);
  m.counts.resize(rc);
  m.firsts.resize(rc);
  m.flows.resize(flow_items);
  m.matches.resize(static_cast<size_t>(m.runs));
  m.decision.resize(static_cast<size_t>(m.runs));
  m.cycle.resize(1);
  m.perm_half.resize(static_cast<size_t>(m.runs));
  m.run_key.resize(static_cast<size_t>(m.runs));
  m.state.resize(static_cast<size_t>(m.runs) * kEvoStateStride);
  m.history.resize(static_cast<size_t>(m.runs) * m.cycles_cap);

  std::vector<int> h_init(static_cast<size_t>(pad), 0), h_half(static_cast<size_t>(m.runs));
  std::vector<uint64_t> h_key(static_cast<size_t>(m.runs));
  for (int r = 0; r < m.runs; ++r) {
    std::copy(codes[r].begin(), codes[r].end(), h_init.begin() + batch.pad_offset(r));
    const int n = static_cast<int>(codes[r].size());
    int bits = 1;
    while ((1 << bits) < n) ++bits;
    h_half[r] = (bits + 1) / 2;
    h_key[r] = splitmix(options.seed + static_cast<uint64_t>(r) * 0xD1B54A32D192ED03ull);
  }
  m.cur.upload(h_init.data(), h_init.size(), s);
  m.best.upload(h_init.data(), h_init.size(), s);
  batch.d_labels().upload(h_init.data(), h_init.size(), s);
  m.perm_half.upload(h_half.data(), h_half.size(), s);
  m.run_key.upload(h_key.data(), h_key.size(), s);
  const int zero = 0;
  m.cycle.upload(&zero, 1, s);

  EvoParams& p = m.prm;
  p.runs = m.runs;
  p.classes = m.classes;
  p.cycles = m.cycles_cap;
  p.guarded = options.guarded_diversity != 0 && options.diversity_multiplier != 0;
  p.block_run = batch.d_block_run().get();
  p.run_offsets = batch.d_run_offsets().get();
  p.pad_offsets = batch.d_pad_offsets().get();
  p.cycle = m.cycle.get();
  p.run_key = m.run_key.get();
  p.perm_half = m.perm_half.get();
  p.cur = m.cur.get();
  p.curpred = m.curpred.get();
  p.best = m.best.get();
  p.bestpred = m.bestpred.get();
  p.prop = batch.d_labels().get();
  p.proppred = batch.d_pred().get();
  p.sizes = m.sizes.get();
  p.flows = m.flows.get();
  p.remap = m.remap.get();
  p.counts = m.counts.get();
  p.firsts = m.firsts.get();
  p.matches = m.matches.get();
  p.state = m.state.get();
  p.decision = m.decision.get();
  p.history = m.history.get();

  VoteParams& v = m.vote;
  v.neighbors = batch.d_neighbors().get();
  v.scores = batch.d_scores().get();
  v.labels = p.prop;
  v.block_run = p.block_run;
  v.run_offsets = p.run_offsets;
  v.pad_offsets = p.pad_offsets;
  v.pad_total = pad;
  v.k = batch.k();
  v.classes = m.classes;
  v.predictions = p.proppred;
  v.counts = p.counts;
  v.firsts = p.firsts;
  v.matches = p.matches;

  // Evaluate the starting labels: predictions become the current and best
  // predictions, scores the current and best scores.
  check(launch_evo_clear(p, s), "evo_clear");
  check(launch_vote(v, m.blocks, s), "vote");
  check(launch_evo_accept(p, 1, s), "evo_accept");
  m.curpred.resize(static_cast<size_t>(pad));
  check(hipMemcpyAsync(m.curpred.get(), p.proppred, sizeof(int) * pad, hipMemcpyDeviceToDevice, s), "copy");
  check(hipMemcpyAsync(m.bestpred.get(), p.proppred, sizeof(int) * pad, hipMemcpyDeviceToDevice, s), "copy");
  check(hipStreamSynchronize(s), "init sync");

  // Capture one cycle.
  check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
  m.enqueue_cycle(s);
  check(hipStreamEndCapture(s, &m.graph), "end capture");
  check(hipGraphInstantiate(&m.exec, m.graph, nullptr, nullptr, 0), "graph instantiate");
  impl_ = impl.release();
}

DeviceEvolution::~DeviceEvolution() { delete impl_; }

void DeviceEvolution::run(int cycles) {
  Impl& m = *impl_;
  if (m.done_cycles + cycles > m.cycles_cap) throw InvalidArgument("more cycles than configured.");
  const hipStream_t s = m.batch.stream();
  for (int c = 0; c < cycles; ++c) check(hipGraphLaunch(m.exec, s), "hipGraphLaunch");
  check(hipStreamSynchronize(s), "evolution sync");
  m.done_cycles += cycles;
  // Average duration of the vote kernel at this batch's shape, timed on the
  // batch stream outside the graph (the evolution state is not modified: the
  // vote reads the proposal buffer and rewrites predictions/reductions only).
  EventTimer t;
  const int reps = 10;
  t.start(s);
  for (int i = 0; i < reps; ++i) check(launch_vote(m.vote, m.blocks, s), "vote timing");
  t.stop(s);
  vote_seconds_ = t.ms() * 1e-3 / reps;
}

void DeviceEvolution::export_results(int* clbest, double* accbest, double* vect_acc) {
  Impl& m = *impl_;
  const hipStream_t s = m.batch.stream();
  std::vector<int> best(static_cast<size_t>(m.batch.pad_total()));
  std::vector<double> state(static_cast<size_t>(m.runs) * kEvoStateStride);
  std::vector<double> hist(static_cast<size_t>(m.runs) * m.cycles_cap);
  m.best.download(best.data(), best.size(), s);
  m.state.download(state.data(), state.size(), s);
  m.history.download(hist.data(), hist.size(), s);
  check(hipStreamSynchronize(s), "export sync");
  const IVec& off = m.batch.run_offsets();
  for (int r = 0; r < m.runs; ++r) {
    const int L = off[r + 1] - off[r];
    for (int i = 0; i < L; ++i) clbest[off[r] + i] = m.values[r][best[m.batch.pad_offset(r) + i]];
    accbest[r] = state[static_cast<size_t>(r) * kEvoStateStride + 2];
    if (vect_acc) {
      for (int c = 0; c < m.done_cycles; ++c) vect_acc[static_cast<size_t>(r) * m.done_cycles + c] =
        hist[static_cast<size_t>(r) * m.cycles_cap + c];
    }
  }
}

}  // namespace kodama_hip
