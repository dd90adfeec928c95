// C-ABI driver over the reference's KODAMA.matrix building blocks, used only by
// the test suite to pin the oracle's restatement. TEST INFRASTRUCTURE ONLY.
//
// kmeans_labels and quota_sample_landmarks live in an unnamed namespace of the
// reference's kodama_matrix.cpp, so this translation unit compiles that source
// file where it lies (oracle/Makefile adds its directory to the include path)
// and calls them from the inside. The library it goes into
// (oracle/_ref/libkodama_ref_matrix.so) links every other reference object and
// leaves out kodama_matrix.o, whose definitions this unit provides instead.

#include "kodama_matrix.cpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>

namespace {

thread_local std::string g_matrix_error;

template <class F>
int guarded_call(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_matrix_error = e.what();
    return 1;
  } catch (...) {
    g_matrix_error = "unknown exception";
    return 1;
  }
}

}  // namespace

extern "C" {

const char* ref_matrix_last_error(void) { return g_matrix_error.c_str(); }

// kmeans_labels (kodama_matrix.cpp:664-799), CPU backend, one thread.
int ref_kmeans_labels(const float* x, int32_t n, int32_t p, int32_t k, uint64_t rng_seed, int32_t discard,
                      int32_t max_iter, int32_t* labels) {
  return guarded_call([&] {
    const std::vector<float> xv(x, x + static_cast<size_t>(n) * p);
    std::mt19937_64 rng(rng_seed);
    for (int i = 0; i < discard; ++i) (void)rng();
    const std::vector<int> l = kodama::kmeans_labels(xv, n, p, k, rng, max_iter, 1, kodama::Backend::CPU, 0);
    std::memcpy(labels, l.data(), sizeof(int32_t) * l.size());
  });
}

// index_strata + quota_sample_landmarks (kodama_matrix.cpp:800-944), shuffled strata.
int ref_quota_sample(const int32_t* strata_labels, int32_t n, int32_t strata, int32_t target, uint64_t rng_seed,
                     int32_t discard, int32_t* rows_out, int32_t* count_out) {
  return guarded_call([&] {
    const std::vector<int> lab(strata_labels, strata_labels + n);
    const kodama::IndexedStrata s = kodama::index_strata(lab, strata);
    std::mt19937_64 rng(rng_seed);
    for (int i = 0; i < discard; ++i) (void)rng();
    const kodama::LandmarkSample out = kodama::quota_sample_landmarks(s.offsets, s.rows, target, rng, false);
    std::memcpy(rows_out, out.rows.data(), sizeof(int32_t) * out.rows.size());
    *count_out = static_cast<int32_t>(out.rows.size());
  });
}

// normalize_external_graph (kodama_matrix.cpp:1292-1330); index_base -1 is Auto.
int ref_normalize_graph(const int32_t* graph_indices, const float* graph_distances, int32_t n,
                        int32_t graph_neighbors, int32_t index_base, int32_t width, int32_t* indices_out,
                        float* distances_out) {
  return guarded_call([&] {
    kodama::NeighborGraph g;
    const size_t total = static_cast<size_t>(n) * graph_neighbors;
    g.indices.assign(graph_indices, graph_indices + total);
    g.distances.assign(graph_distances, graph_distances + total);
    g.neighbors = graph_neighbors;
    g.index_base = index_base == 0 ? kodama::GraphIndexBase::Zero
                   : index_base == 1 ? kodama::GraphIndexBase::One : kodama::GraphIndexBase::Auto;
    const kodama::NeighborGraph out = kodama::normalize_external_graph(g, n, width);
    if (out.neighbors != width) throw std::runtime_error("unexpected normalised width");
    std::memcpy(indices_out, out.indices.data(), sizeof(int32_t) * out.indices.size());
    std::memcpy(distances_out, out.distances.data(), sizeof(float) * out.distances.size());
  });
}

}  // extern "C"
