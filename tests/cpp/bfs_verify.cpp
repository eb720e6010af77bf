// BFS verifier in the shape of the reference's tests/bfs/bfs_verify.cu:16-117:
// every BFS driver is called by its reference name on the graph file given
// as argument, from the first SourcePicker source, and its parents are
// checked against depths from a FIFO BFS.  Built against include/hc only
// (plus the oracle's BFS, as a second check of those depths).  Unlike the
// reference's verifier it exits 1 when a check fails; without a GPU the GPU
// drivers are reported as skipped.
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <functional>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/hc/kernels.h"
#include "../../oracle/oracle.h"

namespace {

int failures = 0;

// Depth of every vertex from source_id (INVALID_NODE: unreachable).
std::vector<nid_t> proper_depths(const CSRUWGraph &g, nid_t source_id) {
    std::vector<nid_t> depth(g.num_nodes, INVALID_NODE);
    depth[source_id] = 0;
    std::deque<nid_t> fifo{source_id};
    while (!fifo.empty()) {
        const nid_t u = fifo.front();
        fifo.pop_front();
        for (nid_t v : g.get_neighbors(u))
            if (depth[v] == INVALID_NODE) {
                depth[v] = depth[u] + 1;
                fifo.push_back(v);
            }
    }
    return depth;
}

// A parent array is right when exactly the reachable vertices have one, the
// source is its own parent and every other parent is one level closer.
bool parents_ok(const std::vector<nid_t> &depth, const nid_t *parents, nid_t source_id) {
    long bad = 0;
    for (nid_t u = 0; u < (nid_t)depth.size(); u++) {
        const nid_t p = parents[u];
        bool ok;
        if (p == INVALID_NODE)
            ok = depth[u] == INVALID_NODE;
        else if (u == source_id)
            ok = p == source_id;
        else
            ok = p >= 0 && p < (nid_t)depth.size() && depth[p] != INVALID_NODE &&
                 depth[u] == depth[p] + 1;
        if (!ok && bad++ < 8)
            std::cout << " > " << u << ": parent " << p << " depth " << depth[u] << std::endl;
    }
    if (bad) std::cout << " > " << bad << " wrong parent(s)" << std::endl;
    return bad == 0;
}

void verify(const std::vector<nid_t> &depth, nid_t source_id, const std::string &name,
            const std::function<nid_t *()> &get_parents) {
    nid_t *parents = get_parents();
    const bool ok = parents_ok(depth, parents, source_id);
    std::cout << "Verifying " << name << " kernel ... " << (ok ? "PASS" : "FAIL") << std::endl;
    failures += !ok;
    delete[] parents;
}

}  // namespace

int main(int argc, char *argv[]) {
    if (argc != 2) {
        std::cerr << "Usage: " << argv[0] << " [graph.sg]" << std::endl;
        return EXIT_FAILURE;
    }
    CSRUWGraph g = load_graph_from_file<CSRUWGraph>(argv[1]);
    std::cout << "Loaded graph." << std::endl;
    SourcePicker<CSRUWGraph> sp(&g);
    const nid_t source_id = sp.next_vertex();

    const std::vector<nid_t> depth = proper_depths(g, source_id);
    std::vector<int32_t> odepth(g.num_nodes);
    or_bfs_depth(g.num_nodes, g.index, g.neighbors, source_id, odepth.data());
    const bool oracle_ok = odepth == depth;
    std::cout << "Oracle depths agree ... " << (oracle_ok ? "PASS" : "FAIL") << std::endl;
    failures += !oracle_ok;

    verify(depth, source_id, "BFS CPU push by node", [&]() {
        nid_t *parents = nullptr;
        bfs_push_cpu<epoch_bfs_push_cpu_by_node>(g, source_id, &parents);
        return parents;
    });
    verify(depth, source_id, "BFS CPU push by edge", [&]() {
        nid_t *parents = nullptr;
        bfs_push_cpu<epoch_bfs_push_cpu_by_edge>(g, source_id, &parents);
        return parents;
    });
    verify(depth, source_id, "BFS CPU pull", [&]() {
        nid_t *parents = nullptr;
        bfs_pull_cpu(g, source_id, &parents);
        return parents;
    });
    verify(depth, source_id, "BFS CPU DO", [&]() {
        nid_t *parents = nullptr;
        bfs_do_cpu(g, source_id, &parents);
        return parents;
    });

    if (hc_init(0) != HC_OK) {
        std::cout << "SKIP GPU kernels: " << hc_last_error() << std::endl;
    } else {
        verify(depth, source_id, "BFS GPU one-to-one", [&]() {
            nid_t *parents = nullptr;
            bfs_gpu(g, source_id, &parents, epoch_bfs_pull_gpu_one_to_one);
            return parents;
        });
        verify(depth, source_id, "BFS GPU warp", [&]() {
            nid_t *parents = nullptr;
            bfs_gpu(g, source_id, &parents, epoch_bfs_pull_gpu_warp);
            return parents;
        });
        verify(depth, source_id, "BFS GPU sync warp", [&]() {
            nid_t *parents = nullptr;
            bfs_gpu(g, source_id, &parents, epoch_bfs_sync_pull_gpu_warp);
            return parents;
        });
        verify(depth, source_id, "BFS heterogeneous", [&]() {
            nid_t *parents = nullptr;
            bfs_do_heterogeneous(g, source_id, &parents);
            return parents;
        });
    }
    std::cout << failures << " failure(s)" << std::endl;
    return failures ? EXIT_FAILURE : EXIT_SUCCESS;
}
