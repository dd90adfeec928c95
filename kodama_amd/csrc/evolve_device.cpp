// Host side of the device-resident evolution: state allocation, the 16-bit
// neighbour table, one persistent launch per run() (all runs, all requested
// cycles; the first launch also evaluates the starting labels) and result
// export.
#include "evolve_device.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <map>
#include <mutex>
#include <queue>

#include "evolve_launch.h"
#include "launch.h"
#include "mlp_launch.h"

namespace kodama_hip {

namespace {
std::atomic<int> g_split_max{0};
std::atomic<int> g_cooperative{1};
std::atomic<int> g_lds_flows{1};
constexpr int kMinSliceRows = 512;  // a split run's members vote at least this many rows each
// Time per row of the replayed phases relative to the vote's.
constexpr double kReplayShare = 0.33;
constexpr int kSpinLimit = 1 << 20;

// Measured cost of each run in earlier evolutions of a batch shape (key: run
// count, k, classes, layout and every run's rows): wall-clock ticks per
// evaluation divided by (kReplayShare + 1 / members), so that it
// extrapolates to other member counts. The member assignment of the next evolution of
// that shape starts from them: the rows alone miss how many of a run's rows
// need the full vote.
std::mutex g_cost_mutex;
std::map<std::vector<int>, std::vector<double>> g_run_costs;
constexpr size_t kMaxCostShapes = 64;

uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
}  // namespace

// All device state in one allocation (one hipMalloc instead of ten), laid
// out so that the host-written part (starting labels, permutation widths, run
// keys) is one upload and the exported part (best labels, state, history) one
// download.
struct DeviceEvolution::Impl {
  KnnBatch& batch;
  std::vector<IVec> values;
  int runs = 0, classes = 0, cycles_cap = 0;
  uint64_t seed = 0;
  PoolBuffer arena;
  size_t up_begin = 0, up_bytes = 0, down_begin = 0, down_bytes = 0;
  size_t off_best = 0, off_state = 0, off_history = 0, off_params = 0, off_packed = 0;
  size_t off_cur = 0, off_bits = 0, off_key = 0;
  EvoParams prm;
  int done_cycles = 0;
  bool initialized = false;
  bool large = false;
  bool learned = false;
  int hash_bits = 0;  // LDS flow table (0: the per-workgroup HBM table)
  std::vector<int> shape_key;  // the batch shape of the measured run costs
  IVec group;                  // members per run
  size_t off_time = 0;         // per-workgroup elapsed ticks
  EventTimer timer;
  bool timed = false;
  // Split runs: the exchange region (xbuf, xmatch, error) cleared before
  // every launch, and the error word read back at export.
  int workgroups = 0;
  size_t off_exchange = 0, exchange_words = 0, off_error = 0;
  bool exchange_clear = false;  // cleared since the last launch
  PinnedBuffer<int> h_error;

  explicit Impl(KnnBatch& b) : batch(b) {}
};

namespace {

struct ArenaLayout {
  size_t total = 0;
  size_t take(size_t bytes) {
    const size_t at = (total + 255) / 256 * 256;
    total = at + bytes;
    return at;
  }
};

}  // namespace

DeviceEvolution::DeviceEvolution(KnnBatch& batch, const std::vector<IVec>& values,
                                 const kodama_hip_core_options& options, const float* params) {
  if (options.evolutionary_search == 0) {
    throw InvalidArgument("the device-resident evolution implements the evolutionary KODAMA.matrix loop.");
  }
  auto impl = std::make_unique<Impl>(batch);
  Impl& m = *impl;
  m.values = values;
  m.runs = batch.runs();
  m.classes = batch.classes();
  m.cycles_cap = std::max(1, options.cycles);
  m.seed = options.seed;
  m.learned = params != nullptr;
  if (m.classes > 1024) throw InvalidArgument("device-resident evolution supports up to 1024 classes per run.");
  // Runs whose row queue and flow list do not fit LDS beside their labels take
  // the large-run layout (those two in global scratch), up to 16383 rows.
  m.large = batch.max_rows() > kEvoMaxRows || evo_lds_bytes(m.classes, batch.max_rows()) > kEvoMaxLds;
  if (batch.max_rows() > kEvoMaxRowsLarge) {
    throw InvalidArgument("device-resident evolution supports up to 16383 rows per run.");
  }
  if (evo_lds_bytes(m.classes, batch.max_rows(), false, m.large) > kEvoMaxLds) {
    throw InvalidArgument("device-resident evolution: the run state does not fit the LDS of one compute unit.");
  }
  if (m.learned && evo_lds_bytes(m.classes, batch.max_rows(), true, m.large) > kEvoMaxLds) {
    throw InvalidArgument("the learned proposal's network does not fit the LDS of one compute unit beside the run state.");
  }
  // Flow counts in an LDS table when it fits beside the run state (the
  // per-workgroup HBM table otherwise, and for the large-run layout).
  if (!m.large && g_lds_flows.load() != 0) {
    const int bits = evo_hash_bits(batch.max_rows());
    if (evo_lds_bytes(m.classes, batch.max_rows(), m.learned, false, bits) <= kEvoMaxLds) m.hash_bits = bits;
  }
  // Split the largest runs over up to kEvoMaxGroup workgroups while the
  // workgroups stay within what the device holds at once: repeatedly one more
  // member for the run with the longest estimated cycle, as long as each
  // member keeps at least kMinSliceRows rows. Every member replays the
  // proposal, absorption and acceptance over all of the run's rows and votes
  // its slice, so a cycle takes about rows * (kReplayShare + 1 / members)
  // vote-row times (kReplayShare measured on the 100-run batch, k = 30).
  // After an evolution of the same batch shape the rows give way to each
  // run's measured time in that model (results do not depend on it).
  int device = 0;
  check(hipGetDevice(&device), "hipGetDevice");
  m.shape_key = {m.runs, batch.k(), m.classes, m.learned ? 1 : 0, m.large ? 1 : 0, m.hash_bits};
  for (int r = 0; r < m.runs; ++r) m.shape_key.push_back(batch.run_rows(r));
  std::vector<double> measured;
  {
    std::lock_guard<std::mutex> lock(g_cost_mutex);
    const auto it = g_run_costs.find(m.shape_key);
    if (it != g_run_costs.end()) measured = it->second;
  }
  IVec group(static_cast<size_t>(m.runs), 1);
  const int max_group = std::min(kEvoMaxGroup, g_split_max.load() == 0 ? kEvoMaxGroup : g_split_max.load());
  if (max_group > 1) {
    int cus = 0;
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
    const int capacity =
      cus * evo_resident_per_cu(batch.k(), m.classes, batch.max_rows(), m.learned, m.large, m.hash_bits);
    if (capacity >= m.runs) {
      std::priority_queue<std::pair<double, int>> pq;
      const auto cycle_cost = [&](int r) {
        const double unit = measured.empty() ? static_cast<double>(batch.run_rows(r)) : measured[r];
        return unit * (kReplayShare + 1.0 / group[r]);
      };
      for (int r = 0; r < m.runs; ++r) pq.emplace(cycle_cost(r), r);
      int total = m.runs;
      while (total < capacity && !pq.empty()) {
        const int r = pq.top().second;
        pq.pop();
        const int rows = batch.run_rows(r);
        if (group[r] >= max_group || rows / (group[r] + 1) < kMinSliceRows) continue;
        ++group[r];
        ++total;
        pq.emplace(cycle_cost(r), r);
      }
    }
  }
  int workgroups = 0;
  for (int g : group) workgroups += g;
  const bool split = workgroups > m.runs;
  m.workgroups = split ? workgroups : m.runs;
  const size_t flow_items = m.hash_bits > 0 ? 0 : static_cast<size_t>(m.workgroups) * m.classes * m.classes;
  const int pad = batch.pad_total();
  const size_t label_bytes = sizeof(unsigned short) * static_cast<size_t>(pad);
  const hipStream_t s = batch.stream();

  ArenaLayout lay;
  const size_t off_nbr = lay.take(sizeof(unsigned short) * static_cast<size_t>(batch.k()) * pad);
  const size_t off_flows = lay.take(sizeof(int) * flow_items);
  const size_t off_curpred = lay.take(label_bytes);
  m.off_cur = lay.take(label_bytes);  // upload: cur .. run_key
  m.off_bits = lay.take(sizeof(int) * m.runs);
  m.off_key = lay.take(sizeof(uint64_t) * m.runs);
  const size_t off_bestpred = lay.take(label_bytes);
  m.off_best = lay.take(label_bytes);  // download: best .. history
  m.off_state = lay.take(sizeof(double) * m.runs * kEvoStateStride);
  m.off_history = lay.take(sizeof(double) * static_cast<size_t>(m.runs) * m.cycles_cap);
  m.off_time = lay.take(sizeof(unsigned long long) * static_cast<size_t>(m.workgroups));
  // The learned proposal's network: fp32 parameters and their tiled form.
  m.off_params = lay.take(sizeof(float) * kMlpParams);
  m.off_packed = lay.take(sizeof(uint32_t) * kMlpTiled);
  // Large-run layout: each workgroup's flow list and row queue.
  const size_t off_large =
    m.large ? lay.take(sizeof(int) * 2 * static_cast<size_t>(m.workgroups) * batch.max_rows()) : 0;
  // Split runs: workgroup tables and the exchange region.
  const size_t off_wg = split ? lay.take(sizeof(int) * (2 * static_cast<size_t>(m.workgroups) + m.runs)) : 0;
  if (split) {
    // Two halves (alternate steps) of the predictions and match counts, then
    // the error word.
    m.exchange_words = 2 * (static_cast<size_t>(pad) + static_cast<size_t>(m.runs) * kEvoMaxGroup) + 1;
    m.off_exchange = lay.take(sizeof(unsigned) * m.exchange_words);
    m.off_error = m.off_exchange + sizeof(unsigned) * (m.exchange_words - 1);
  }
  m.up_begin = m.off_cur;
  m.up_bytes = m.off_key + sizeof(uint64_t) * m.runs - m.off_cur;
  m.down_begin = m.off_best;
  m.down_bytes = m.off_time + sizeof(unsigned long long) * static_cast<size_t>(m.workgroups) - m.off_best;
  // Rounded up to 2 MiB so that evolutions of similar size reuse one pool block.
  m.arena.allocate((lay.total + (2u << 20) - 1) / (2u << 20) * (2u << 20), device, s);
  unsigned char* base = m.arena.get();
  check(launch_evo_zero(reinterpret_cast<int*>(base + off_flows), flow_items, s), "evo_zero");
  if (split) {  // the first launch's exchange region, cleared with the set-up
    check(launch_evo_zero(reinterpret_cast<int*>(base + m.off_exchange), m.exchange_words, s), "exchange clear");
    m.exchange_clear = true;
  }
  check(launch_evo_zero(reinterpret_cast<int*>(base + m.off_time), 2 * static_cast<size_t>(m.workgroups), s),
        "evo_zero");
  check(launch_evo_narrow(batch.d_neighbors().get(), batch.d_run_offsets().get(), batch.d_pad_offsets().get(),
                          batch.d_block_run().get(), reinterpret_cast<unsigned short*>(base + off_nbr),
                          static_cast<size_t>(batch.k()) * pad, pad, s),
        "evo_narrow");
  if (split) {
    IVec tables;
    tables.reserve(2 * static_cast<size_t>(m.workgroups) + m.runs);
    for (int r = 0; r < m.runs; ++r)
      for (int g = 0; g < group[r]; ++g) tables.push_back(r);
    for (int r = 0; r < m.runs; ++r)
      for (int g = 0; g < group[r]; ++g) tables.push_back(g);
    tables.insert(tables.end(), group.begin(), group.end());
    check(hipMemcpyAsync(base + off_wg, tables.data(), sizeof(int) * tables.size(), hipMemcpyHostToDevice, s),
          "upload workgroup tables");
    // The copy reads a pageable vector: complete it before the vector goes.
    check(hipStreamSynchronize(s), "workgroup tables sync");
    m.h_error.resize(1);
  }
  if (m.learned) {
    float* d_params = reinterpret_cast<float*>(base + m.off_params);
    check(hipMemcpyAsync(d_params, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, s), "upload params");
    check(launch_mlp_pack(d_params, reinterpret_cast<uint32_t*>(base + m.off_packed), s), "mlp_pack");
  }

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
  p.neighbors = reinterpret_cast<const unsigned short*>(base + off_nbr);
  p.scores = batch.d_scores().get();
  p.run_key = reinterpret_cast<const uint64_t*>(base + m.off_key);
  p.perm_bits = reinterpret_cast<const int*>(base + m.off_bits);
  p.cur = reinterpret_cast<unsigned short*>(base + m.off_cur);
  p.curpred = reinterpret_cast<unsigned short*>(base + off_curpred);
  p.best = reinterpret_cast<unsigned short*>(base + m.off_best);
  p.bestpred = reinterpret_cast<unsigned short*>(base + off_bestpred);
  p.state = reinterpret_cast<double*>(base + m.off_state);
  p.flows = m.hash_bits > 0 ? nullptr : reinterpret_cast<int*>(base + off_flows);
  p.hash_bits = m.hash_bits;
  p.history = reinterpret_cast<double*>(base + m.off_history);
  p.wg_time = reinterpret_cast<unsigned long long*>(base + m.off_time);
  p.large_scratch = m.large ? reinterpret_cast<int*>(base + off_large) : nullptr;
  p.learn_packed = m.learned ? reinterpret_cast<const uint32_t*>(base + m.off_packed) : nullptr;
  if (split) {
    const int* wg = reinterpret_cast<const int*>(base + off_wg);
    p.wg_run = wg;
    p.wg_member = wg + m.workgroups;
    p.run_group = wg + 2 * m.workgroups;
    p.xbuf = reinterpret_cast<unsigned*>(base + m.off_exchange);
    p.xmatch = p.xbuf + 2 * static_cast<size_t>(pad);
    p.error = reinterpret_cast<int*>(base + m.off_error);
    p.spin_limit = kSpinLimit;
    p.workgroups = m.workgroups;
    p.cooperative = g_cooperative.load();
  }
  m.group = group;
  // The set-up is complete before the timed evolution starts.
  check(hipStreamSynchronize(s), "setup sync");
  impl_ = impl.release();
}

DeviceEvolution::~DeviceEvolution() { delete impl_; }

void DeviceEvolution::start(const std::vector<IVec>& codes) {
  Impl& m = *impl_;
  if (static_cast<int>(codes.size()) != m.runs) throw InvalidArgument("one label vector per run.");
  const hipStream_t s = m.batch.stream();
  // Host image of the upload region: starting labels (the kernel's first
  // evaluation writes their predictions), permutation widths, run keys. Built
  // in the batch's pinned staging when it fits (an asynchronous copy the
  // launch follows in stream order), else in pageable memory.
  std::vector<unsigned char> pageable;
  unsigned char* img = m.batch.stage_up(m.up_bytes);
  if (img == nullptr) {
    pageable.assign(m.up_bytes, 0);
    img = pageable.data();
  }
  auto* h_init = reinterpret_cast<unsigned short*>(img);
  auto* h_bits = reinterpret_cast<int*>(img + (m.off_bits - m.off_cur));
  auto* h_key = reinterpret_cast<uint64_t*>(img + (m.off_key - m.off_cur));
  // The runs' rows are disjoint slices of the image: filled by host threads.
#pragma omp parallel for schedule(static)
  for (int r = 0; r < m.runs; ++r) {
    for (size_t i = 0; i < codes[r].size(); ++i) {
      h_init[m.batch.pad_offset(r) + i] = static_cast<unsigned short>(codes[r][i]);
    }
    const int n = static_cast<int>(codes[r].size());
    int bits = 1;
    while ((1 << bits) < n) ++bits;
    h_bits[r] = bits;
    h_key[r] = splitmix(m.seed + static_cast<uint64_t>(r) * 0xD1B54A32D192ED03ull);
  }
  check(hipMemcpyAsync(m.arena.get() + m.up_begin, img, m.up_bytes, hipMemcpyHostToDevice, s), "upload");
  check(hipStreamSynchronize(s), "upload sync");
}

void DeviceEvolution::wait() {
  if (impl_->timed) impl_->timer.wait();
}

namespace {

void launch_cycles(DeviceEvolution::Impl& m, int cycles, bool learned, float fallback) {
  if (cycles < 0 || m.done_cycles + cycles > m.cycles_cap) throw InvalidArgument("more cycles than configured.");
  const hipStream_t s = m.batch.stream();
  EvoParams p = m.prm;
  p.init = m.initialized ? 0 : 1;
  p.cycle_begin = m.done_cycles;
  p.cycle_count = cycles;
  if (!learned) p.learn_packed = nullptr;
  p.learn_fallback = fallback;
  if (p.wg_run != nullptr && !m.exchange_clear) {
    check(launch_evo_zero(reinterpret_cast<int*>(m.arena.get() + m.off_exchange), m.exchange_words, s),
          "exchange clear");
  }
  m.exchange_clear = false;  // the launch leaves tagged words behind
  if (!m.timed) m.timer.start(s);  // kernel time spans the first launch to the last
  check(launch_evo_persistent(p, s), "evo_persistent");
  m.timer.stop(s);
  m.timed = true;
  m.initialized = true;
  m.done_cycles += cycles;
}

}  // namespace

void DeviceEvolution::run(int cycles) { launch_cycles(*impl_, cycles, false, 0.0f); }

int DeviceEvolution::workgroups() const { return impl_->workgroups; }
bool DeviceEvolution::lds_flow_table() const { return impl_->hash_bits > 0; }

void set_evolution_split(int max_group) { g_split_max.store(max_group); }
void set_cooperative_launch(bool on) { g_cooperative.store(on ? 1 : 0); }
void set_lds_flow_table(bool on) { g_lds_flows.store(on ? 1 : 0); }
int evolution_split() { return g_split_max.load(); }

void DeviceEvolution::run_learned(int cycles, float fallback) {
  if (!impl_->learned) throw InvalidArgument("the evolution was set up without a network.");
  launch_cycles(*impl_, cycles, true, fallback);
}

double DeviceEvolution::kernel_seconds() const { return impl_->timed ? impl_->timer.ms() * 1e-3 : 0.0; }

void DeviceEvolution::export_results(int* clbest, double* accbest, double* vect_acc) {
  Impl& m = *impl_;
  const hipStream_t s = m.batch.stream();
  std::vector<unsigned char> pageable;
  unsigned char* down = m.batch.stage_down(m.down_bytes);
  if (down == nullptr) {
    pageable.resize(m.down_bytes);
    down = pageable.data();
  }
  check(hipMemcpyAsync(down, m.arena.get() + m.down_begin, m.down_bytes, hipMemcpyDeviceToHost, s), "download");
  if (m.h_error.size() > 0) {
    check(hipMemcpyAsync(m.h_error.get(), m.arena.get() + m.off_error, sizeof(int), hipMemcpyDeviceToHost, s),
          "download error word");
  }
  check(hipStreamSynchronize(s), "export sync");
  if (m.h_error.size() > 0 && m.h_error[0] != 0) {
    throw DeviceError("device-resident evolution: a split run's exchange wait ran out.");
  }
  const auto* best = reinterpret_cast<const unsigned short*>(down + (m.off_best - m.down_begin));
  const auto* state = reinterpret_cast<const double*>(down + (m.off_state - m.down_begin));
  const auto* hist = reinterpret_cast<const double*>(down + (m.off_history - m.down_begin));
  // Each run's cost: its slowest member's ticks per evaluation, per unit of
  // the assignment's model; averaged with an earlier measurement of the shape.
  if (m.done_cycles > 0) {
    const auto* ticks = reinterpret_cast<const unsigned long long*>(down + (m.off_time - m.down_begin));
    std::vector<double> cost(static_cast<size_t>(m.runs));
    for (int r = 0, w = 0; r < m.runs; ++r) {
      unsigned long long t = 0;
      for (int g = 0; g < m.group[r]; ++g, ++w) t = std::max(t, ticks[w]);
      cost[r] = static_cast<double>(t) / (m.done_cycles + 1) / (kReplayShare + 1.0 / m.group[r]);
    }
    std::lock_guard<std::mutex> lock(g_cost_mutex);
    auto it = g_run_costs.find(m.shape_key);
    if (it != g_run_costs.end()) {
      for (int r = 0; r < m.runs; ++r) it->second[r] = 0.5 * (it->second[r] + cost[r]);
    } else {
      if (g_run_costs.size() >= kMaxCostShapes) g_run_costs.clear();
      g_run_costs.emplace(m.shape_key, std::move(cost));
    }
  }
  const IVec& off = m.batch.run_offsets();
#pragma omp parallel for schedule(static)
  for (int r = 0; r < m.runs; ++r) {
    const int L = off[r + 1] - off[r];
    const IVec& v = m.values[r];
    const unsigned short* b = best + m.batch.pad_offset(r);
    int* out = clbest + off[r];
    for (int i = 0; i < L; ++i) out[i] = v[b[i]];
    accbest[r] = state[static_cast<size_t>(r) * kEvoStateStride + 2];
    if (vect_acc) {
      for (int c = 0; c < m.done_cycles; ++c) vect_acc[static_cast<size_t>(r) * m.done_cycles + c] =
        hist[static_cast<size_t>(r) * m.cycles_cap + c];
    }
  }
}

}  // namespace kodama_hip
