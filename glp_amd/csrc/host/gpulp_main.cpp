// GPULP: the command-line driver, argument- and log-compatible with the
// reference's main.cu.
//
//   GPULP method niter input.bin output.log ifDirected [ifWeighted] [key=value ...]
//   GPULP method=1 iter=20 input=x.bin output=out.log ifdirect=0 ifweighted=0
//
// method: 1 classic, 2 layered, 3 SLPA, 4 user hooks (main.cu:63-101), and
// this library's counting methods 5 count-min, 6 frequent, 7 buffer, and
// 8 seeded LP (prop/layeredLprop.cuh; needs seeds=PATH).  As in
// the reference, niter iterations execute niter-1 propagation steps
// (prop/classicLprop.cuh:361) and one result line is appended to the output
// file in main.cu:106-121's format.  Extra keys (extensions):
//   steps=K        execute exactly K steps
//   gamma=p/q      layered with a fixed gamma instead of the 10-round schedule
//   seed=S r=p/q   SLPA stream seed and threshold
//   cm=d,W[,a]     count-min rows, width, aliased rows (method 5)
//   cap=C          frequent / buffer table capacity (methods 6, 7)
//   seeds=PATH     seed file, one vertex id per line (method 8)
//   gather=1       separate gather pass;  replay=1 HIP-graph step replay
//   labels=PATH    write the final labels as little-endian int32
//   tuning=PATH    key=value tuning file;  autotune=1 tune on this graph first
//   json=1         also print a JSON summary line (GTEPS per step)
#include "glp.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>

static void usage() {
    std::printf(
        "usage: GPULP method niter input.bin output.log ifDirected [ifWeighted] [key=value ...]\n"
        "       GPULP method=1 iter=20 input=g.bin output=out.log ifdirect=0 ifweighted=0\n"
        "methods: 1 classic, 2 layered, 3 SLPA, 4 user hooks, 5 count-min, 6 frequent, 7 buffer,\n"
        "         8 seeded (seeds=PATH)\n");
}

static bool parse_ratio(const std::string& s, int32_t& a, int32_t& b) {
    return std::sscanf(s.c_str(), "%d/%d", &a, &b) == 2;
}

int main(int argc, char** argv) {
    std::map<std::string, std::string> kv;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto eq = a.find('=');
        if (eq == std::string::npos) pos.push_back(a);
        else kv[a.substr(0, eq)] = a.substr(eq + 1);
    }
    if (argc == 1) {
        usage();
        return 0;
    }
    auto take = [&](std::initializer_list<const char*> keys, size_t idx, const char* dflt) -> std::string {
        for (const char* k : keys)
            if (kv.count(k)) return kv[k];
        if (idx < pos.size()) return pos[idx];
        return dflt ? dflt : "";
    };
    const std::string method_s = take({"method"}, 0, nullptr);
    const std::string niter_s = take({"iter", "iterations", "niter"}, 1, nullptr);
    const std::string input = take({"input"}, 2, nullptr);
    const std::string output = take({"output"}, 3, "");
    const std::string direct_s = take({"ifdirect", "ifDirected", "directed"}, 4, "0");
    const std::string weighted_s = take({"ifweighted", "ifWeighted", "weighted"}, 5, "0");
    if (method_s.empty() || niter_s.empty() || input.empty()) {
        // main.cu:53-54 prints this and exits 0
        std::printf("need 6 argc, input %d argc \n", argc);
        usage();
        return 0;
    }
    const int method = std::atoi(method_s.c_str());
    const int niter = std::atoi(niter_s.c_str());
    const bool directed = std::atoi(direct_s.c_str()) != 0;
    const bool weighted = std::atoi(weighted_s.c_str()) != 0;
    try {
        if (kv.count("tuning")) {
            std::ifstream in(kv["tuning"]);
            std::string line;
            while (std::getline(in, line)) {
                auto eq = line.find('=');
                if (eq != std::string::npos)
                    glp::check(glp_tuning_set(line.substr(0, eq).c_str(), std::atoi(line.c_str() + eq + 1)),
                               "tuning");
            }
        }
        glp::CSRGraph graph(input, 0, directed, weighted);
        std::cout << "n = " << graph.n << " m = " << graph.m << std::endl << std::endl;
        const int32_t steps = kv.count("steps") ? std::atoi(kv["steps"].c_str()) : (niter > 1 ? niter - 1 : 0);
        glp_params p = glp::make_params(method);
        p.weighted = weighted ? 1 : 0;
        p.gather = kv.count("gather") ? std::atoi(kv["gather"].c_str()) : 0;
        if (kv.count("seed")) p.seed = std::strtoull(kv["seed"].c_str(), nullptr, 10);
        if (kv.count("r") && !parse_ratio(kv["r"], p.slpa_r_num, p.slpa_r_den))
            throw glp::Error(GLP_E_ARG, "r must be num/den");
        if (kv.count("cm")) {
            int d = 3, w = 200, al = 0;
            std::sscanf(kv["cm"].c_str(), "%d,%d,%d", &d, &w, &al);
            p.cm_rows = d;
            p.cm_width = w;
            p.cm_aliased = al;
        }
        if (kv.count("cap")) p.table_cap = std::atoi(kv["cap"].c_str());
        if (kv.count("autotune") && std::atoi(kv["autotune"].c_str())) {
            double best = 0;
            glp::check(glp_autotune(graph.handle(), &p, std::max(steps, 1), nullptr, &best), "autotune");
            std::cout << "autotune: best " << best << " ms" << std::endl;
        }
        const bool replay = kv.count("replay") && std::atoi(kv["replay"].c_str());
        double running_time = 0;
        std::unique_ptr<glp::Test_result> myresult;
        std::unique_ptr<glp::Propagator> prop;
        int32_t executed = steps;
        if (method == GLP_LAYERED && !kv.count("gamma")) {
            auto* mp = new glp::MultiLprop(&graph, weighted);
            prop.reset(mp);
            running_time = mp->run(niter);
            executed = niter * mp->rounds;
        } else {
            if (method == GLP_LAYERED && !parse_ratio(kv["gamma"], p.gamma_num, p.gamma_den))
                throw glp::Error(GLP_E_ARG, "gamma must be num/den");
            if (method == GLP_SLPA) p.slpa_T = std::max(steps, 1);
            prop.reset(new glp::Propagator(&graph, p));
            if (method == GLP_SEEDED) {
                if (!kv.count("seeds")) throw glp::Error(GLP_E_ARG, "method 8 needs seeds=PATH");
                prop->set_seeds(glp::read_seed_file(kv["seeds"], graph.n));
            }
            if (replay) {
                prop->run_steps(0);
                glp::check(glp_session_step_graph(prop->session(), steps), "replay");
                glp::check(glp_session_sync(prop->session()), "sync");
            } else {
                prop->run_steps(steps);
            }
            if (method == GLP_CLASSIC || method == GLP_USER)
                myresult.reset(new glp::Test_result(prop->result()));
            else
                running_time = prop->timing().steps_ms / 1e3;
        }
        if (kv.count("labels")) {
            std::vector<int32_t> labels;
            if (method == GLP_SLPA) {
                std::vector<int32_t> dom(graph.n), ns(graph.n), sets;
                glp::check(glp_session_slpa_post(prop->session(), 0, dom.data(), ns.data(), nullptr), "post");
                labels = dom;
            } else {
                labels = prop->labels();
            }
            std::ofstream lf(kv["labels"], std::ios::binary | std::ios::trunc);
            lf.write(reinterpret_cast<const char*>(labels.data()), labels.size() * sizeof(int32_t));
        }
        std::cout << "success run";
        if (!output.empty()) {
            // main.cu:106-121
            std::ofstream ofs(output, std::ofstream::out | std::ios::app);
            if (!myresult) {
                ofs << "graph: " << input << " method: " << method << " iteration: " << niter
                    << " time: " << running_time;
                ofs << std::endl;
            } else {
                const glp::Test_result* r = myresult.get();
                ofs << "graph: " << input << " method: " << method << " iteration: " << niter
                    << " time: " << " load graph to GPU: " << r->load_graph_to_gpu;
                ofs << " other proprocess: " << r->other_proprocess << " label_load: " << r->label_load
                    << " LP: " << r->lp;
                ofs << " tiny_speed: " << r->tiny_node_process << " small_speed: " << r->small_node_process
                    << " medium_speed: " << r->medium_node_process << " big_speed: " << r->big_node_process
                    << " huge_speed: " << r->huge_node_process;
                ofs << " tiny_distrbution: " << r->tiny_distribution
                    << " small_distribution: " << r->small_distribution
                    << " medium_distribution: " << r->medium_distribution
                    << " big_distribution: " << r->big_distribution
                    << " huge_distribution: " << r->huge_distribution;
                ofs << std::endl;
            }
        }
        if (kv.count("json") && std::atoi(kv["json"].c_str())) {
            glp_timing t = prop->timing();
            const double ms = t.steps > 0 ? t.steps_ms / t.steps : 0.0;
            std::printf("\n{\"method\": %d, \"n\": %d, \"m\": %lld, \"steps\": %d, \"ms_per_step\": %.4f, "
                        "\"gteps\": %.4f}\n",
                        method, graph.n, (long long)graph.m, executed, ms,
                        ms > 0 ? graph.m / (ms / 1e3) / 1e9 : 0.0);
        }
    } catch (const glp::Error& e) {
        std::fprintf(stderr, "GPULP: %s\n", e.what());
        return 1;
    }
    return 0;
}
