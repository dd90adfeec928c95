// C ABI of libkodama_hip.so, graph half: the exact self-kNN graph
// (KODAMAKNNGraph_CUDA, graph_cluster.cpp:790-837) and the KODAMA dissimilarity
// rewrite (KODAMADissimilarityInPlace, kodama_matrix.cpp:2685-2776, 4383-4463).
#include <algorithm>
#include <memory>
#include <mutex>
#include <numeric>

#include "../../include/kodama_hip.h"
#include "capi_common.h"
#include "engine.h"
#include "graph_launch.h"
#include "knn_mfma.h"
#include "launch.h"

using namespace kodama_hip;

namespace {

void require(bool ok, const char* msg) {
  if (!ok) throw InvalidArgument(msg);
}

}  // namespace

extern "C" {

int kodama_hip_knn_graph(const float* x, int32_t n, int32_t p, int32_t k, int32_t metric, int32_t device,
                         int32_t* indices_one_based, float* distances) {
  return guarded([&] {
    require(x != nullptr && indices_one_based != nullptr && distances != nullptr, "null pointer.");
    require(n >= 2 && p >= 1, "KODAMAKNNGraph requires at least two rows.");
    require(metric >= KODAMA_HIP_COSINE && metric <= KODAMA_HIP_EUCLIDEAN, "unknown metric.");
    const int kk = std::max(1, std::min(k, n - 1));
    require(kk <= (p <= 128 ? kMaxScanK : kMaxK),
            "graph neighbours above 128 (above 64 beyond 128 feature columns) are not supported by the HIP kNN kernel.");
    require(p <= kMaxDims, "the HIP kNN kernel supports 1..256 feature columns.");
    // Cosine rows are normalised as the reference's exact search expects
    // (native_knn.cpp:1005-1035).
    const std::vector<float> xs = prepare_rows(x, n, p, metric, 1);
    select_device(device);
    Stream st(device);
    const hipStream_t s = st.get();
    DeviceBuffer<float> d_x(static_cast<size_t>(n) * p);
    DeviceBuffer<int> d_ex(static_cast<size_t>(n));
    DeviceBuffer<int> d_idx(static_cast<size_t>(n) * kk);
    DeviceBuffer<float> d_dist(static_cast<size_t>(n) * kk);
    std::vector<int> ex(static_cast<size_t>(n));
    std::iota(ex.begin(), ex.end(), 0);
    d_x.upload(xs.data(), d_x.size(), s);
    d_ex.upload(ex.data(), ex.size(), s);
    device_exact_knn(d_x.get(), n, d_x.get(), n, p, kk, metric, d_ex.get(), d_idx.get(), d_dist.get(),
                     max_row_norm2(xs.data(), n, p), s, &last_knn_fallbacks());
    check(launch_graph_publish(d_idx.get(), d_dist.get(), d_idx.size(), metric, s), "graph_publish launch");
    d_idx.download(indices_one_based, d_idx.size(), s);
    d_dist.download(distances, d_dist.size(), s);
    st.sync();
  });
}

int kodama_hip_dissimilarity_inplace(int32_t* indices_one_based, float* distances, int32_t samples,
                                     int32_t neighbors, const int32_t* run_labels, int32_t runs, int32_t device) {
  return guarded([&] {
    if (runs <= 0 || samples <= 0 || neighbors <= 0) return;  // the reference leaves the graph as is
    require(indices_one_based != nullptr && distances != nullptr && run_labels != nullptr, "null pointer.");
    require(dissimilarity_lds_bytes(neighbors) <= 64 * 1024, "dissimilarity supports up to 1365 neighbours per row.");
    select_device(device);
    Stream st(device);
    const hipStream_t s = st.get();
    const int stride = (runs + 3) / 4 * 4;
    const size_t edges = static_cast<size_t>(samples) * neighbors;
    DeviceBuffer<int> d_idx(edges), d_res(static_cast<size_t>(runs) * samples),
      d_lab(static_cast<size_t>(samples) * stride);
    DeviceBuffer<float> d_dist(edges);
    d_idx.upload(indices_one_based, edges, s);
    d_dist.upload(distances, edges, s);
    d_res.upload(run_labels, d_res.size(), s);
    check(launch_labels_transpose(d_res.get(), runs, samples, stride, d_lab.get(), s), "labels_transpose launch");
    check(launch_dissimilarity(d_idx.get(), d_dist.get(), samples, neighbors, d_lab.get(), stride, true, true, s),
          "dissimilarity launch");
    d_idx.download(indices_one_based, edges, s);
    d_dist.download(distances, edges, s);
    st.sync();
  });
}

// The allocation-free twin of kodama_hip_dissimilarity_inplace: the caller's
// device buffers, the caller's stream, and the transposed labels in the
// library's static scratch (no hipMalloc, hipFree or pool call).
int kodama_hip_dissimilarity_inplace_noalloc(int32_t* d_indices_one_based, float* d_distances, int32_t samples,
                                             int32_t neighbors, const int32_t* d_run_labels, int32_t runs,
                                             int32_t device, void* stream) {
  return guarded([&] {
    if (runs <= 0 || samples <= 0 || neighbors <= 0) return;  // the reference leaves the graph as is
    require(d_indices_one_based != nullptr && d_distances != nullptr && d_run_labels != nullptr, "null pointer.");
    require(dissimilarity_lds_bytes(neighbors) <= 64 * 1024, "dissimilarity supports up to 1365 neighbours per row.");
    const int stride = (runs + 3) / 4 * 4;
    select_device(device);
    int* labels = nullptr;
    size_t bytes = 0;
    check(noalloc_scratch(device, &labels, &bytes), "scratch address");
    require(static_cast<size_t>(samples) * stride * sizeof(int) <= bytes,
            "samples x runs (rounded up to 4) exceeds the 64 MiB scratch of the allocation-free dissimilarity.");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // The transpose writes every entry the rewrite reads (padding included),
    // so the scratch needs no clearing between calls.
    check(launch_labels_transpose(d_run_labels, runs, samples, stride, labels, s), "labels_transpose launch");
    check(launch_dissimilarity(d_indices_one_based, d_distances, samples, neighbors, labels, stride, true, true, s),
          "dissimilarity launch");
  });
}

// ---- device-resident graph handle --------------------------------------------
// KODAMAGraphHandle / KODAMAGraphMaterialize (kodama.hpp:481-530,
// kodama_matrix.cpp:2422-2490, 4364-4381): the exact graph stays on the device
// in its public form (1-based ids, public distances) until the caller
// materialises it; the KODAMA dissimilarity rewrites it in place there.
struct kodama_hip_graph {
  int device = 0;
  int samples = 0;
  int neighbors = 0;
  DeviceBuffer<int> idx;
  DeviceBuffer<float> dist;
  std::mutex lock;  // one caller at a time per handle
};

int kodama_hip_graph_build(const float* x, int32_t n, int32_t p, int32_t k, int32_t metric, int32_t device,
                           kodama_hip_graph** out) {
  return guarded([&] {
    require(out != nullptr, "null pointer.");
    *out = nullptr;
    require(x != nullptr, "null pointer.");
    require(n >= 2 && p >= 1, "KODAMAKNNGraph requires at least two rows.");
    require(metric >= KODAMA_HIP_COSINE && metric <= KODAMA_HIP_EUCLIDEAN, "unknown metric.");
    const int kk = std::max(1, std::min(k, n - 1));
    require(kk <= (p <= 128 ? kMaxScanK : kMaxK),
            "graph neighbours above 128 (above 64 beyond 128 feature columns) are not supported by the HIP kNN kernel.");
    require(p <= kMaxDims, "the HIP kNN kernel supports 1..256 feature columns.");
    const std::vector<float> xs = prepare_rows(x, n, p, metric, 1);
    select_device(device);
    std::unique_ptr<kodama_hip_graph> g(new kodama_hip_graph);
    g->device = device;
    g->samples = n;
    g->neighbors = kk;
    g->idx.resize(static_cast<size_t>(n) * kk);
    g->dist.resize(static_cast<size_t>(n) * kk);
    Stream st(device);
    const hipStream_t s = st.get();
    DeviceBuffer<float> d_x(static_cast<size_t>(n) * p);
    DeviceBuffer<int> d_ex(static_cast<size_t>(n));
    std::vector<int> ex(static_cast<size_t>(n));
    std::iota(ex.begin(), ex.end(), 0);
    d_x.upload(xs.data(), d_x.size(), s);
    d_ex.upload(ex.data(), ex.size(), s);
    device_exact_knn(d_x.get(), n, d_x.get(), n, p, kk, metric, d_ex.get(), g->idx.get(), g->dist.get(),
                     max_row_norm2(xs.data(), n, p), s, &last_knn_fallbacks());
    check(launch_graph_publish(g->idx.get(), g->dist.get(), g->idx.size(), metric, s), "graph_publish launch");
    st.sync();
    *out = g.release();
  });
}

int kodama_hip_graph_info(const kodama_hip_graph* graph, int32_t* samples, int32_t* neighbors, int32_t* device) {
  return guarded([&] {
    require(graph != nullptr, "null graph handle.");
    if (samples) *samples = graph->samples;
    if (neighbors) *neighbors = graph->neighbors;
    if (device) *device = graph->device;
  });
}

int kodama_hip_graph_materialize(kodama_hip_graph* graph, int32_t* indices_one_based, float* distances) {
  return guarded([&] {
    require(graph != nullptr, "null graph handle.");
    require(indices_one_based != nullptr && distances != nullptr, "null pointer.");
    std::lock_guard<std::mutex> hold(graph->lock);
    select_device(graph->device);
    Stream st(graph->device);
    graph->idx.download(indices_one_based, graph->idx.size(), st.get());
    graph->dist.download(distances, graph->dist.size(), st.get());
    st.sync();
  });
}

int kodama_hip_graph_dissimilarity(kodama_hip_graph* graph, const int32_t* run_labels, int32_t runs) {
  return guarded([&] {
    require(graph != nullptr, "null graph handle.");
    if (runs <= 0) return;  // the reference leaves the graph as is
    require(run_labels != nullptr, "null pointer.");
    require(dissimilarity_lds_bytes(graph->neighbors) <= 64 * 1024,
            "dissimilarity supports up to 1365 neighbours per row.");
    std::lock_guard<std::mutex> hold(graph->lock);
    select_device(graph->device);
    Stream st(graph->device);
    const hipStream_t s = st.get();
    const int n = graph->samples;
    const int stride = (runs + 3) / 4 * 4;
    DeviceBuffer<int> d_res(static_cast<size_t>(runs) * n), d_lab(static_cast<size_t>(n) * stride);
    d_res.upload(run_labels, d_res.size(), s);
    check(launch_labels_transpose(d_res.get(), runs, n, stride, d_lab.get(), s), "labels_transpose launch");
    check(launch_dissimilarity(graph->idx.get(), graph->dist.get(), n, graph->neighbors, d_lab.get(), stride, true,
                               true, s),
          "dissimilarity launch");
    st.sync();
  });
}

void kodama_hip_graph_release(kodama_hip_graph* graph) {
  if (graph == nullptr) return;
  try {
    select_device(graph->device);
  } catch (...) {
  }
  delete graph;
}

}  // extern "C"
