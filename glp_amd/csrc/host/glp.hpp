// C++ host layer over libglp.so's C ABI, shaped like the reference's host
// classes so a caller of the reference finds the same objects:
//
//   CSRGraph(path, startvalue, direct)    header/graph.h:62-198
//   Test_result                           header/test_result.h:5-23
//   classicLprop(G).run(niter)            prop/classicLprop.cuh:12-21,454-462
//   MultiLprop(G).run(niter)              prop/MultiLprop.cuh (layered)
//   SLPAprop(G).run(niter)                prop/SLPAprop.cuh
//   UserdefinedLprop(G).run(niter)        prop/UserdefinedLprop.cuh (method 4)
//
// As in the reference loop (prop/classicLprop.cuh:361), run(niter) executes
// niter-1 propagation steps; run_steps(k) executes exactly k.  Errors throw
// glp::Error carrying the library's status code and message.
#pragma once

#include <glp.h>

#include <algorithm>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace glp {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
    int code;
};

inline void check(int rc, const char* what) {
    if (rc != GLP_OK) throw Error(rc, std::string(what) + ": " + glp_last_error());
}

// header/test_result.h:5-23; times in seconds like the reference Timer.
struct Test_result {
    double load_graph_to_cpu = 0;
    double load_graph_to_gpu = 0;
    double other_proprocess = 0;
    double label_load = 0;  // no label copy exists here (buffers swap): always 0
    double lp = 0;
    double tiny_node_process = 0;    // packed tier (deg <= pack_max_deg)
    double small_node_process = 0;   // folded into the packed tier: 0
    double medium_node_process = 0;  // wave tier
    double big_node_process = 0;     // block tier
    double huge_node_process = 0;    // hub tier
    double tiny_distribution = 0;
    double small_distribution = 0;
    double medium_distribution = 0;
    double big_distribution = 0;
    double huge_distribution = 0;
};

class CSRGraph {
public:
    // startvalue is subtracted from every id as the reference does (graph.h:236)
    CSRGraph(const std::string& path, int startvalue, bool direct, bool weighted = false) {
        if (startvalue != 0) throw Error(GLP_E_UNSUPPORTED, "startvalue other than 0");
        check(glp_graph_load_bin(path.c_str(), direct ? 1 : 0, weighted ? 1 : 0, &h), "load");
        int32_t nn, md, w;
        int64_t mm;
        check(glp_graph_info(h, &nn, &mm, &md, &w), "info");
        n = nn;
        m = mm;
    }
    explicit CSRGraph(glp_graph* handle) : h(handle) {
        int32_t nn, md, w;
        int64_t mm;
        check(glp_graph_info(h, &nn, &mm, &md, &w), "info");
        n = nn;
        m = mm;
    }
    CSRGraph(const CSRGraph&) = delete;
    CSRGraph& operator=(const CSRGraph&) = delete;
    ~CSRGraph() { glp_graph_free(h); }
    glp_graph* handle() const { return h; }
    int32_t n = 0;
    int64_t m = 0;

private:
    glp_graph* h = nullptr;
};

// A device-resident LP job: the common part of the propagators.
class Propagator {
public:
    Propagator(CSRGraph* G, const glp_params& p) : G(G), params(p) {}
    virtual ~Propagator() {
        if (s) glp_session_free(s);
    }
    // Executes exactly k steps from L_0[v] = v.
    void run_steps(int32_t k) {
        open();
        check(glp_session_step(s, k), "step");
        check(glp_session_sync(s), "sync");
    }
    std::vector<int32_t> labels() {
        std::vector<int32_t> out(G->n);
        check(glp_session_labels(s, out.data()), "labels");
        return out;
    }
    std::vector<int64_t> changes() {
        std::vector<int64_t> out(4096);
        int32_t k = 0;
        check(glp_session_changes(s, out.data(), (int32_t)out.size(), &k), "changes");
        out.resize(k);
        return out;
    }
    glp_timing timing() {
        glp_timing t;
        check(glp_session_timing(s, &t), "timing");
        return t;
    }
    Test_result result() {
        glp_timing t = timing();
        Test_result r;
        r.load_graph_to_gpu = t.upload_ms / 1e3;
        r.other_proprocess = t.plan_ms / 1e3;
        r.lp = t.steps_ms / 1e3;
        r.tiny_node_process = t.tier_ms[GLP_TIER_PACKED] / 1e3;
        r.medium_node_process = t.tier_ms[GLP_TIER_WAVE] / 1e3;
        r.big_node_process = t.tier_ms[GLP_TIER_BLOCK] / 1e3;
        r.huge_node_process = (t.tier_ms[GLP_TIER_HUB] + t.tier_ms[GLP_TIER_AUX]) / 1e3;
        const double n = G->n > 0 ? (double)G->n : 1.0;
        r.tiny_distribution = t.tier_vertices[GLP_TIER_PACKED] / n;
        r.medium_distribution = t.tier_vertices[GLP_TIER_WAVE] / n;
        r.big_distribution = t.tier_vertices[GLP_TIER_BLOCK] / n;
        r.huge_distribution = t.tier_vertices[GLP_TIER_HUB] / n;
        return r;
    }
    glp_session* session() { return s; }
    // Seeded LP: sets the seed mask (n bytes) and resets the labels.
    void set_seeds(const std::vector<uint8_t>& seeds) {
        open();
        check(glp_session_set_seeds(s, seeds.data(), (int32_t)seeds.size()), "seeds");
    }

protected:
    void open() {
        if (!s) check(glp_session_create(G->handle(), &params, &s), "session");
    }
    CSRGraph* G;
    glp_params params;
    glp_session* s = nullptr;
};

inline glp_params make_params(int32_t method) {
    glp_params p;
    glp_params_default(&p);
    p.method = method;
    p.profile = 1;
    return p;
}

class classicLprop : public Propagator {
public:
    explicit classicLprop(CSRGraph* G, bool weighted = false)
        : Propagator(G, [&] { auto p = make_params(GLP_CLASSIC); p.weighted = weighted; return p; }()) {}
    Test_result* run(int niter) {
        run_steps(niter > 1 ? niter - 1 : 0);
        return new Test_result(result());
    }
};

// Layered LP as MultiLprop::run drives it (prop/MultiLprop.cuh:308-373):
// ten rounds s = 0..9 of niter steps with delta = 2^(s+1), score
// (delta+1)*C - V, i.e. gamma = 1/delta; labels and volumes carry over
// between rounds (the round re-initialisation is overwritten by the label
// sync in the reference, D13).  Returns seconds.
class MultiLprop : public Propagator {
public:
    explicit MultiLprop(CSRGraph* G, bool weighted = false, int rounds = 10)
        : Propagator(G, [&] { auto p = make_params(GLP_LAYERED); p.gamma_num = 1; p.gamma_den = 2;
                              p.weighted = weighted; return p; }()),
          rounds(rounds) {}
    double run(int niter) {
        open();
        for (int r = 0; r < rounds; ++r) {
            check(glp_session_set_gamma(s, 1, 1 << (r + 1)), "gamma");
            check(glp_session_step(s, niter), "step");
        }
        check(glp_session_sync(s), "sync");
        return timing().steps_ms / 1e3;
    }
    int rounds;
};

// SLPA: niter-1 listener steps (prop/SLPAprop.cuh:337), memory of niter
// rows, post-processing threshold r.  Returns seconds.
class SLPAprop : public Propagator {
public:
    explicit SLPAprop(CSRGraph* G, uint64_t seed = 1, int r_num = 1, int r_den = 10)
        : Propagator(G, [&] { auto p = make_params(GLP_SLPA); p.seed = seed; p.slpa_r_num = r_num;
                              p.slpa_r_den = r_den; return p; }()) {}
    double run(int niter) {
        params.slpa_T = niter > 1 ? niter - 1 : 1;
        run_steps(niter > 1 ? niter - 1 : 0);
        return timing().steps_ms / 1e3;
    }
    // dominant label per vertex and the thresholded label sets
    void communities(std::vector<int32_t>& dominant, std::vector<int32_t>& nsets,
                     std::vector<int32_t>& sets, int32_t kmax) {
        dominant.resize(G->n);
        nsets.resize(G->n);
        sets.resize((size_t)G->n * kmax);
        check(glp_session_slpa_post(s, kmax, dominant.data(), nsets.data(), sets.data()), "post");
    }
};

// Seed file as layeredLprop::assign_seeds_for_device reads it
// (prop/layeredLprop.cuh:161-180): one vertex id per line.  Ids outside
// [0, n) are rejected (the reference writes out of bounds).
inline std::vector<uint8_t> read_seed_file(const std::string& path, int32_t n) {
    std::ifstream ifs(path);
    if (!ifs) throw Error(GLP_E_IO, "cannot read seed file '" + path + "'");
    std::vector<uint8_t> seeds((size_t)std::max(n, 0), 0);
    std::string line;
    while (std::getline(ifs, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        const long id = std::stol(line);
        if (id < 0 || id >= n) throw Error(GLP_E_ARG, "seed id " + line + " outside the graph");
        seeds[(size_t)id] = 1;
    }
    return seeds;
}

// Seeded LP as prop/layeredLprop.cuh runs it (despite its name that class is
// the seed-constrained propagator; main.cu includes it without a method
// number): seeds from a seed file, L_0 = v on seeds and -1 elsewhere
// (header/kernel.cuh:453-466), run(niter) executes niter-1 steps like the
// loop at prop/layeredLprop.cuh:388.  Returns seconds.
class layeredLprop : public Propagator {
public:
    layeredLprop(CSRGraph* G, const std::string& seedfile, bool weighted = false)
        : Propagator(G, [&] { auto p = make_params(GLP_SEEDED); p.weighted = weighted; return p; }()),
          seeds(read_seed_file(seedfile, G->n)) {}
    double run(int niter) {
        set_seeds(seeds);
        check(glp_session_step(s, niter > 1 ? niter - 1 : 0), "step");
        check(glp_session_sync(s), "sync");
        return timing().steps_ms / 1e3;
    }
    std::vector<uint8_t> seeds;
};

// Method 4: the compiled-in user hooks (glp_amd/csrc/user_hooks.h).
class UserdefinedLprop : public Propagator {
public:
    explicit UserdefinedLprop(CSRGraph* G) : Propagator(G, make_params(GLP_USER)) {}
    Test_result* run(int niter) {
        run_steps(niter > 1 ? niter - 1 : 0);
        return new Test_result(result());
    }
};

}  // namespace glp
