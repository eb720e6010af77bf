// C++ drop-in for the reference's graph types (src/graph.cuh:26-239), over
// the C ABI of libhcamd.so.  Header-only; link with -lhcamd.
//
// Same names, members and behaviour as the reference, with two deliberate
// differences:
//   * weight_t is int32_t (SSSP here uses integer weights; the reference's
//     float weight_t and numeric_limits<int>::infinity() == 0 make an int
//     build of it wrong, src/graph.cuh:27,32).  INF_WEIGHT is HC_INF_DIST.
//   * GenericCSRGraph is move-only (the reference's destructor frees arrays a
//     copy would share).
#ifndef HC_CPP_GRAPH_H
#define HC_CPP_GRAPH_H

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <istream>
#include <random>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../hc_capi.h"

using nid_t = std::int32_t;     // vertex id
using weight_t = std::int32_t;  // edge weight / distance
using offset_t = std::int64_t;  // edge index

const weight_t INF_WEIGHT = HC_INF_DIST;
const nid_t INVALID_NODE = -1;

struct wnode_t {
    nid_t v;
    weight_t w;
};

namespace hc {

// d + w saturated at INF_WEIGHT (an unreached d stays unreached): the
// relaxation of the library's integer SSSP (weights >= 0, d <= INF_WEIGHT).
inline weight_t relax_host(weight_t d, weight_t w) {
    const uint32_t c = (uint32_t)d + (uint32_t)w;
    return c < (uint32_t)INF_WEIGHT ? (weight_t)c : INF_WEIGHT;
}

// Error behaviour of the reference (print and exit, src/cuda.cuh:14-21,
// src/util.h:16-22) for the drop-in layer.
inline void check_or_exit(int rc, const char *what) {
    if (rc != HC_OK) {
        std::fprintf(stderr, "%s failed: %s\n", what, hc_last_error());
        std::exit(EXIT_FAILURE);
    }
}

}  // namespace hc

template <typename NeighborT>
class GenericCSRGraph {
    class Neighborhood {
    public:
        Neighborhood(nid_t nid_, offset_t *index_, NeighborT *neighbors_)
            : nid(nid_), index(index_), neighbors(neighbors_) {}
        using iterator = NeighborT *;
        iterator begin() { return &neighbors[index[nid]]; }
        iterator end() { return &neighbors[index[nid + 1]]; }

    private:
        nid_t nid;
        offset_t *index;
        NeighborT *neighbors;
    };

public:
    offset_t *index = nullptr;
    NeighborT *neighbors = nullptr;
    nid_t num_nodes = 0;
    offset_t num_edges = 0;

    GenericCSRGraph() = default;
    GenericCSRGraph(const GenericCSRGraph &) = delete;
    GenericCSRGraph &operator=(const GenericCSRGraph &) = delete;
    GenericCSRGraph(GenericCSRGraph &&o) noexcept { *this = std::move(o); }
    GenericCSRGraph &operator=(GenericCSRGraph &&o) noexcept {
        if (this != &o) {
            delete[] index;
            delete[] neighbors;
            index = o.index;
            neighbors = o.neighbors;
            num_nodes = o.num_nodes;
            num_edges = o.num_edges;
            o.index = nullptr;
            o.neighbors = nullptr;
            o.num_nodes = 0;
            o.num_edges = 0;
        }
        return *this;
    }
    ~GenericCSRGraph() {
        delete[] index;
        delete[] neighbors;
    }

    offset_t get_degree(nid_t nid) const { return index[nid + 1] - index[nid]; }
    Neighborhood get_neighbors(nid_t nid) const { return Neighborhood(nid, index, neighbors); }
    bool has_self_cycle(nid_t nid) const;
};

template <>
inline bool GenericCSRGraph<nid_t>::has_self_cycle(nid_t nid) const {
    for (nid_t nei : get_neighbors(nid))
        if (nei == nid) return true;
    return false;
}

template <>
inline bool GenericCSRGraph<wnode_t>::has_self_cycle(nid_t nid) const {
    for (wnode_t nei : get_neighbors(nid))
        if (nei.v == nid) return true;
    return false;
}

using CSRWGraph = GenericCSRGraph<wnode_t>;
using CSRUWGraph = GenericCSRGraph<nid_t>;

namespace hc {

// Structure-of-arrays view of a reference graph for the C ABI: the index is
// shared, the AoS neighbour list is split (weighted graphs only).
template <typename NeighborT>
struct HostView {
    hc_host_graph_t c{};
    std::vector<nid_t> ids;
    std::vector<weight_t> weights;
    explicit HostView(const GenericCSRGraph<NeighborT> &g) {
        c.num_nodes = g.num_nodes;
        c.num_edges = g.num_edges;
        c.index = g.index;
        if constexpr (std::is_same<NeighborT, wnode_t>::value) {
            ids.resize(g.num_edges);
            weights.resize(g.num_edges);
            for (offset_t i = 0; i < g.num_edges; i++) {
                ids[i] = g.neighbors[i].v;
                weights[i] = g.neighbors[i].w;
            }
            c.neighbors = ids.data();
            c.weights = weights.data();
        } else {
            c.neighbors = g.neighbors;
            c.weights = nullptr;
        }
    }
};

// Takes ownership of a library host graph's arrays as a reference graph.
template <typename NeighborT>
GenericCSRGraph<NeighborT> adopt(hc_host_graph_t &h) {
    GenericCSRGraph<NeighborT> g;
    g.num_nodes = h.num_nodes;
    g.num_edges = h.num_edges;
    g.index = new offset_t[(size_t)h.num_nodes + 1];
    std::memcpy(g.index, h.index, sizeof(offset_t) * ((size_t)h.num_nodes + 1));
    g.neighbors = new NeighborT[(size_t)h.num_edges];
    for (offset_t i = 0; i < h.num_edges; i++) {
        if constexpr (std::is_same<NeighborT, wnode_t>::value)
            g.neighbors[i] = wnode_t{h.neighbors[i], h.weights ? h.weights[i] : 1};
        else
            g.neighbors[i] = h.neighbors[i];
    }
    hc_host_graph_free(&h);
    return g;
}

}  // namespace hc

// SourcePicker (src/graph.cuh:106-141): mt19937(37525), uniform ids,
// rejects degree-0 vertices and lone self-loops.
template <typename GraphT>
class SourcePicker {
public:
    SourcePicker(const GraphT *const g_, nid_t const_source_ = -1)
        : g(g_), const_source(const_source_), udist(0, g_->num_nodes - 1) {
        rng = std::mt19937(rand_seed);
    }
    nid_t next_vertex() {
        if (const_source >= 0) return const_source;
        nid_t source;
        do {
            source = udist(rng);
        } while (g->get_degree(source) == 0 ||
                 (g->get_degree(source) == 1 && g->has_self_cycle(source)));
        return source;
    }
    void reset() { rng = std::mt19937(rand_seed); }

private:
    const GraphT *const g;
    nid_t const_source;
    std::mt19937 rng;
    std::uniform_int_distribution<nid_t> udist;
    const int rand_seed = 37525;
};

// compute_equal_edge_ranges (src/graph.cuh:156-201): num_segments + 1
// boundaries, new[]-allocated (the caller delete[]s them).
template <typename NeighborT>
nid_t *compute_equal_edge_ranges(const GenericCSRGraph<NeighborT> &g, const nid_t num_segments,
                                 nid_t chunk_size = 1) {
    hc::HostView<NeighborT> v(g);
    nid_t *seg = new nid_t[num_segments + 1];
    hc::check_or_exit(hc_equal_edge_ranges(&v.c, num_segments, chunk_size, seg),
                      "compute_equal_edge_ranges");
    return seg;
}

// Binary graph input (src/graph.cuh:207-239): int32 N, int64 m, int64
// index[N+1], neighbours ({int32 v; float w} per edge in .wsg files; int32
// ids in .sg files).  Validated by the library.
template <typename GraphT>
inline GraphT load_graph_from_file(const char *filename) {
    constexpr bool weighted = std::is_same<GraphT, CSRWGraph>::value;
    hc_host_graph_t h{};
    hc::check_or_exit(hc_host_graph_read(filename, weighted ? 1 : 0, &h), "load_graph_from_file");
    return hc::adopt<typename std::remove_pointer<decltype(GraphT::neighbors)>::type>(h);
}

// gapbs-style synthetic graph (the generator behind the reference's graph
// files): R-MAT (kron) or uniform, symmetrized, relabelled by degree.
template <typename GraphT>
inline GraphT generate_graph(bool uniform, int scale, int degree, int64_t seed = 27491095) {
    constexpr bool weighted = std::is_same<GraphT, CSRWGraph>::value;
    hc_host_graph_t h{};
    hc::check_or_exit(hc_host_graph_generate(uniform ? 1 : 0, scale, degree, weighted ? 1 : 0,
                                             seed, &h),
                      "generate_graph");
    return hc::adopt<typename std::remove_pointer<decltype(GraphT::neighbors)>::type>(h);
}

#endif  // HC_CPP_GRAPH_H
