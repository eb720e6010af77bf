// SSSP verifier in the shape of the reference's tests/sssp/sssp_verify.cu:
// 16-110: the serial pull oracle from the first SourcePicker source, then
// every SSSP driver called by its reference name and compared with it vertex
// by vertex (integer distances: exact).  Also checks the serial oracle
// against Dijkstra (both the reference-named one and the oracle's) and a
// .skd schedule loaded into the reference's schedule types.  Built against
// include/hc only (plus the oracle).  Exits 1 when a check fails; without a
// GPU the GPU drivers are reported as skipped.
//
// usage: sssp_verify graph.wsg [schedule.skd]
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/hc/schedule.h"
#include "../../oracle/oracle.h"

namespace {

int failures = 0;

bool same_dist(const weight_t *oracle_dist, const weight_t *dist, nid_t num_nodes) {
    long bad = 0;
    for (nid_t u = 0; u < num_nodes; u++)
        if (oracle_dist[u] != dist[u] && bad++ < 8)
            std::cout << " > " << u << ": " << dist[u] << " != " << oracle_dist[u] << std::endl;
    if (bad) std::cout << " > " << bad << " wrong distance(s)" << std::endl;
    return bad == 0;
}

void report(const std::string &what, bool ok) {
    std::cout << "Verifying " << what << " ... " << (ok ? "PASS" : "FAIL") << std::endl;
    failures += !ok;
}

}  // namespace

int main(int argc, char *argv[]) {
    if (argc != 2 && argc != 3) {
        std::cerr << "Usage: " << argv[0] << " [graph.wsg] [schedule.skd]" << std::endl;
        return EXIT_FAILURE;
    }
    CSRWGraph g = load_graph_from_file<CSRWGraph>(argv[1]);
    std::cout << "Loaded graph." << std::endl;
    SourcePicker<CSRWGraph> sp(&g);
    const nid_t source_id = sp.next_vertex();

    weight_t *oracle_dist = nullptr;
    sssp_pull_cpu_serial(g, source_id, &oracle_dist);
    std::cout << "Computed oracle distance." << std::endl;
    {
        weight_t *dj = nullptr;
        sssp_dijkstras_cpu_serial(g, source_id, &dj);
        report("sssp_dijkstras_cpu_serial vs sssp_pull_cpu_serial",
               same_dist(oracle_dist, dj, g.num_nodes));
        delete[] dj;
        std::vector<int32_t> ids(g.num_edges), wts(g.num_edges), od(g.num_nodes);
        for (offset_t i = 0; i < g.num_edges; i++) {
            ids[i] = g.neighbors[i].v;
            wts[i] = g.neighbors[i].w;
        }
        or_sssp_dijkstra(g.num_nodes, g.index, ids.data(), wts.data(), source_id, od.data());
        report("oracle Dijkstra vs sssp_pull_cpu_serial",
               same_dist(oracle_dist, od.data(), g.num_nodes));
    }

    weight_t *init_dist = new weight_t[g.num_nodes];
    for (nid_t i = 0; i < g.num_nodes; i++) init_dist[i] = INF_WEIGHT;
    init_dist[source_id] = 0;

    {
        weight_t *dist = nullptr;
        sssp_pull_cpu(g, epoch_sssp_pull_cpu_one_to_one, init_dist, &dist);
        report("SSSP CPU kernel", same_dist(oracle_dist, dist, g.num_nodes));
        delete[] dist;
    }

    // A schedule file in the reference's .skd layout, read into its types.
    if (argc == 3) {
        SSSPHeteroSchedule schedule = load_sssp_schedule(argv[2]);
        std::ostringstream os;
        os << schedule;
        std::cout << os.str();
        report("load_sssp_schedule", schedule.num_segments > 0 && !schedule.blocks.empty() &&
                                         os.str().find("SSSP Heterogeneous Schedule") == 0);
        bool cpu_only = true;
        for (const SSSPScheduleBlock &b : schedule.blocks)
            cpu_only &= std::holds_alternative<SSSPCPUScheduleBlock>(b);
        if (cpu_only || hc_init(0) == HC_OK) {
            weight_t *dist = nullptr;
            sssp_pull_heterogeneous(g, schedule, init_dist, &dist);
            report("SSSP heterogeneous kernel (loaded schedule)",
                   same_dist(oracle_dist, dist, g.num_nodes));
            delete[] dist;
        } else {
            std::cout << "SKIP loaded schedule: it has GPU blocks and there is no GPU" << std::endl;
        }
    }

    if (hc_init(0) != HC_OK) {
        std::cout << "SKIP GPU kernels: " << hc_last_error() << std::endl;
    } else {
        {
            weight_t *dist = nullptr;
            sssp_pull_gpu(g, epoch_sssp_pull_gpu_one_to_one, init_dist, &dist);
            report("SSSP GPU naive kernel", same_dist(oracle_dist, dist, g.num_nodes));
            delete[] dist;
        }
        {
            weight_t *dist = nullptr;
            sssp_pull_gpu(g, epoch_sssp_pull_gpu_warp_red, init_dist, &dist);
            report("SSSP GPU warp red kernel", same_dist(oracle_dist, dist, g.num_nodes));
            delete[] dist;
        }
        {
            weight_t *dist = nullptr;
            sssp_pull_gpu(g, epoch_sssp_pull_gpu_block_red, init_dist, &dist, 64, 256);
            report("SSSP GPU block red kernel", same_dist(oracle_dist, dist, g.num_nodes));
            delete[] dist;
        }
        {
            enable_all_peer_access();
            weight_t *dist = nullptr;
            sssp_pull_heterogeneous(g, init_dist, &dist);
            report("SSSP heterogeneous kernel", same_dist(oracle_dist, dist, g.num_nodes));
            delete[] dist;
        }
    }
    delete[] init_dist;
    delete[] oracle_dist;
    std::cout << failures << " failure(s)" << std::endl;
    return failures ? EXIT_FAILURE : EXIT_SUCCESS;
}
