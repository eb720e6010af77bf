// PageRank verifier in the shape of the reference's tests/pr/pr_verify.cu:
// 16-110: pr_pull_cpu with epoch_pr_pull_cpu_one_to_one gives the reference
// scores; every GPU epoch kernel through pr_pull_gpu (and the heterogeneous
// driver) must match them within 5e-4 per vertex, the reference's tolerance.
// The CPU scores are also compared with the oracle's restatement of the same
// Gauss-Seidel iteration.  Scores are float (the reference keeps them in its
// float weight_t).  Exits 1 when a check fails; without a GPU the GPU
// drivers are reported as skipped.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/hc/kernels.h"
#include "../../oracle/oracle.h"

namespace {

int failures = 0;

bool close_scores(const float *oracle_score, const float *score, nid_t num_nodes, float eps) {
    long bad = 0;
    for (nid_t u = 0; u < num_nodes; u++)
        if (!(std::fabs(oracle_score[u] - score[u]) <= eps) && bad++ < 8)
            std::cout << " > " << u << ": " << score[u] << " != " << oracle_score[u] << std::endl;
    if (bad) std::cout << " > " << bad << " score(s) off by more than " << eps << std::endl;
    return bad == 0;
}

void report(const std::string &what, bool ok) {
    std::cout << "Verifying " << what << " ... " << (ok ? "PASS" : "FAIL") << std::endl;
    failures += !ok;
}

}  // namespace

int main(int argc, char *argv[]) {
    if (argc != 2) {
        std::cerr << "Usage: " << argv[0] << " [graph.wsg]" << std::endl;
        return EXIT_FAILURE;
    }
    CSRWGraph g = load_graph_from_file<CSRWGraph>(argv[1]);
    std::cout << "Loaded graph." << std::endl;
    SourcePicker<CSRWGraph> sp(&g);
    const nid_t source_id = sp.next_vertex();

    float *init_score = new float[g.num_nodes];
    for (nid_t i = 0; i < g.num_nodes; i++) init_score[i] = 1.0f / g.num_nodes;
    init_score[source_id] = 0.0f;
    const float eps = 0.0005f;

    float *oracle_score = nullptr;
    pr_pull_cpu(g, epoch_pr_pull_cpu_one_to_one, init_score, &oracle_score);
    std::cout << "Computed oracle score." << std::endl;
    {
        // The same in-place iteration restated serially by the oracle: the
        // threaded run visits vertices in another order, so the two agree
        // to the stopping threshold's scale, not bit for bit.
        std::vector<int32_t> ids(g.num_edges);
        for (offset_t i = 0; i < g.num_edges; i++) ids[i] = g.neighbors[i].v;
        std::vector<float> gs(g.num_nodes);
        or_pr_gauss_seidel_f32(g.num_nodes, g.index, ids.data(), 1000, gs.data());
        report("PR CPU kernel vs the oracle's Gauss-Seidel",
               close_scores(gs.data(), oracle_score, g.num_nodes, 1e-5f));
    }
    {
        float *score = nullptr;
        pr_pull_cpu(g, epoch_pr_pull_cpu_one_to_one, init_score, &score);
        report("PR CPU kernel", close_scores(oracle_score, score, g.num_nodes, eps));
        delete[] score;
    }

    if (hc_init(0) != HC_OK) {
        std::cout << "SKIP GPU kernels: " << hc_last_error() << std::endl;
    } else {
        {
            float *score = nullptr;
            pr_pull_gpu(g, epoch_pr_pull_gpu_one_to_one, init_score, &score);
            report("PR GPU naive kernel", close_scores(oracle_score, score, g.num_nodes, eps));
            delete[] score;
        }
        {
            float *score = nullptr;
            pr_pull_gpu(g, epoch_pr_pull_gpu_warp_red, init_score, &score);
            report("PR GPU warp red kernel", close_scores(oracle_score, score, g.num_nodes, eps));
            delete[] score;
        }
        {
            float *score = nullptr;
            pr_pull_gpu(g, epoch_pr_pull_gpu_block_red, init_score, &score, 64, 256);
            report("PR GPU block red kernel", close_scores(oracle_score, score, g.num_nodes, eps));
            delete[] score;
        }
        {
            enable_all_peer_access_pr();
            float *score = nullptr;
            pr_pull_heterogeneous(g, init_score, &score);
            report("PR heterogeneous kernel", close_scores(oracle_score, score, g.num_nodes, eps));
            delete[] score;
        }
    }
    delete[] init_score;
    delete[] oracle_score;
    std::cout << failures << " failure(s)" << std::endl;
    return failures ? EXIT_FAILURE : EXIT_SUCCESS;
}
