// Host side of the device-resident evolution: state allocation, the 16-bit
// neighbour table, one persistent launch per run() (all runs, all requested
// cycles; the first launch also evaluates the starting labels) and result
// export.
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
  int runs = 0, classes = 0, cycles_cap = 0;
  DeviceBuffer<unsigned short> nbr16, cur, curpred, best, bestpred;
  DeviceBuffer<int> flows, perm_bits;
  DeviceBuffer<uint64_t> run_key;
  DeviceBuffer<double> state, history;
  EvoParams prm;
  int done_cycles = 0;
  bool initialized = false;

  explicit Impl(KnnBatch& b) : batch(b) {}
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
  if (m.classes > 1024) throw InvalidArgument("device-resident evolution supports up to 1024 classes per run.");
  if (batch.max_rows() > kEvoMaxRows) {
    throw InvalidArgument("device-resident evolution supports up to 8191 rows per run.");
  }
  if (evo_lds_bytes(m.classes, batch.max_rows()) > kEvoMaxLds) {
    throw InvalidArgument("device-resident evolution: the run state does not fit the LDS of one compute unit.");
  }
  const size_t flow_items = static_cast<size_t>(m.runs) * m.classes * m.classes;
  const int pad = batch.pad_total();
  const hipStream_t s = batch.stream();
  m.nbr16.resize(static_cast<size_t>(batch.k()) * pad);
  for (auto* b : {&m.cur, &m.curpred, &m.best, &m.bestpred}) b->resize(static_cast<size_t>(pad));
  m.flows.resize(flow_items);
  m.perm_bits.resize(static_cast<size_t>(m.runs));
  m.run_key.resize(static_cast<size_t>(m.runs));
  m.state.resize(static_cast<size_t>(m.runs) * kEvoStateStride);
  m.history.resize(static_cast<size_t>(m.runs) * m.cycles_cap);

  std::vector<unsigned short> h_init(static_cast<size_t>(pad), 0);
  std::vector<int> h_bits(static_cast<size_t>(m.runs));
  std::vector<uint64_t> h_key(static_cast<size_t>(m.runs));
  for (int r = 0; r < m.runs; ++r) {
    for (size_t i = 0; i < codes[r].size(); ++i) {
      h_init[batch.pad_offset(r) + i] = static_cast<unsigned short>(codes[r][i]);
    }
    const int n = static_cast<int>(codes[r].size());
    int bits = 1;
    while ((1 << bits) < n) ++bits;
    h_bits[r] = bits;
    h_key[r] = splitmix(options.seed + static_cast<uint64_t>(r) * 0xD1B54A32D192ED03ull);
  }
  m.cur.upload(h_init.data(), h_init.size(), s);
  m.curpred.upload(h_init.data(), h_init.size(), s);
  m.perm_bits.upload(h_bits.data(), h_bits.size(), s);
  m.run_key.upload(h_key.data(), h_key.size(), s);
  check(launch_evo_zero(m.flows.get(), m.flows.size(), s), "evo_zero");
  check(launch_evo_narrow(batch.d_neighbors().get(), batch.d_run_offsets().get(), batch.d_pad_offsets().get(),
                          batch.d_block_run().get(), m.nbr16.get(), m.nbr16.size(), pad, s),
        "evo_narrow");

  EvoParams& p = m.prm;
  p.runs = m.runs;
  p.classes = m.classes;
  p.cycles = m.cycles_cap;
  p.guarded = options.guarded_diversity != 0 && options.diversity_multiplier != 0;
  p.k = batch.k();
  p.max_rows = batch.max_rows();
  p.pad_total = pad;
  p.run_offsets = batch.d_run_offsets().get();
  p.pad_offsets = batch.d_pad_offsets().get();
  p.neighbors = m.nbr16.get();
  p.scores = batch.d_scores().get();
  p.run_key = m.run_key.get();
  p.perm_bits = m.perm_bits.get();
  p.cur = m.cur.get();
  p.curpred = m.curpred.get();
  p.best = m.best.get();
  p.bestpred = m.bestpred.get();
  p.state = m.state.get();
  p.flows = m.flows.get();
  p.history = m.history.get();
  check(hipStreamSynchronize(s), "setup sync");
  impl_ = impl.release();
}

DeviceEvolution::~DeviceEvolution() { delete impl_; }

void DeviceEvolution::run(int cycles) {
  Impl& m = *impl_;
  if (cycles < 0 || m.done_cycles + cycles > m.cycles_cap) throw InvalidArgument("more cycles than configured.");
  const hipStream_t s = m.batch.stream();
  EvoParams p = m.prm;
  p.init = m.initialized ? 0 : 1;
  p.cycle_begin = m.done_cycles;
  p.cycle_count = cycles;
  EventTimer t;
  t.start(s);
  check(launch_evo_persistent(p, s), "evo_persistent");
  t.stop(s);
  kernel_seconds_ = t.ms() * 1e-3;
  m.initialized = true;
  m.done_cycles += cycles;
}

void DeviceEvolution::export_results(int* clbest, double* accbest, double* vect_acc) {
  Impl& m = *impl_;
  const hipStream_t s = m.batch.stream();
  std::vector<unsigned short> best(static_cast<size_t>(m.batch.pad_total()));
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
