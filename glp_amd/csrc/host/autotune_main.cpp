// glp_autotune: tunes tier thresholds and LDS sizes on a graph and writes a
// key=value file the library reads through GLP_TUNING or GPULP tuning=.
//   glp_autotune graph.bin out.tuning [method] [steps] [ifWeighted]
#include <glp.h>

#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: glp_autotune graph.bin out.tuning [method] [steps] [ifWeighted]\n");
        return 0;
    }
    const int method = argc > 3 ? std::atoi(argv[3]) : GLP_CLASSIC;
    const int steps = argc > 4 ? std::atoi(argv[4]) : 5;
    const int weighted = argc > 5 ? std::atoi(argv[5]) : 0;
    glp_graph* g = nullptr;
    if (glp_graph_load_bin(argv[1], 0, weighted, &g) != GLP_OK) {
        std::fprintf(stderr, "glp_autotune: %s\n", glp_last_error());
        return 1;
    }
    glp_params p;
    glp_params_default(&p);
    p.method = method;
    p.weighted = weighted;
    if (method == GLP_LAYERED) {
        p.gamma_num = 1;
        p.gamma_den = 4;
    }
    double best = 0;
    int rc = glp_autotune(g, &p, steps, argv[2], &best);
    if (rc != GLP_OK) std::fprintf(stderr, "glp_autotune: %s\n", glp_last_error());
    else std::printf("best %.3f ms for %d steps; written to %s\n", best, steps, argv[2]);
    glp_graph_free(g);
    return rc == GLP_OK ? 0 : 1;
}
