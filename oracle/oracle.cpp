// CPU restatement of the reference's label-propagation semantics.
//
// TEST INFRASTRUCTURE ONLY: the checker for the GPU library and the timed
// CPU baseline.  See oracle.h for the reference line each function follows.
//
// Pinning status (what independently checks each function; the tests under
// tests/ hold the checks):
//   orc_load_bin         against the reference's own header/graph.h compiled
//                        unchanged into oracle/_ref/ref_loadbin (byte-equal
//                        CSR on fixture files)
//   orc_classic_step     closed forms (clique, star, path, bridge, isolated),
//                        brute force over all labelings of small graphs,
//                        multigraph multiplicity and weighted cases
//   orc_layered_step     gamma=0 reduces to classic; hand-derived scores
//   orc_seeded_step      closed forms (paths, star, competing seeds, isolated
//                        and unreachable vertices), all-seeds identity,
//                        invariants (seeds fixed, labels are seed ids, the
//                        labeled set only grows and advances one hop a step)
//   orc_slpa / _post     T=1 reduces to one classic step; stream KATs; a
//                        hand-traced three-step case whose speakers come
//                        from the stream's known answers; hand-derived sets
//   orc_countmin_step    no-collision widths reduce to classic; width-4 star
//                        cases argued by hand from the row hashes' cells
//                        (row 1/2 collisions decide, aliased != proper for
//                        d = 3 and 4); hash KATs from the constants
//   orc_table_step       capacity >= degree reduces to classic; hand traces
//   orc_cpulp_run        mode 1 equals repeated orc_classic_step
//   orc_classic_subset   equals orc_classic_step on every listed vertex
//                        (random graphs, weighted and unweighted)
//   orc_cpulp_layered_sample  timing only (the reference loop's live reads
//                        race); its labels are checked to be neighbour labels

#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <vector>

namespace {

struct Loaded {
    int32_t n = 0;
    std::vector<int32_t> off, nbr;
    std::vector<uint32_t> w;
};

// header/graph.h:227-258 (load_edgelist_bin) then :103-130 (adjacency list
// to CSR), with vector<vector> push_back exactly as the reference builds it.
int load(const char* path, int directed, int weighted, Loaded& out) {
    std::ifstream ifs(path, std::ios::binary);
    if (!ifs) return -1;
    const int words = weighted ? 3 : 2;
    std::vector<std::vector<int32_t>> adj;
    std::vector<std::vector<uint32_t>> adjw;
    int32_t max_u_v = 0;
    bool any = false;
    int32_t buf[3];
    while (ifs.read(reinterpret_cast<char*>(buf), sizeof(int32_t) * words)) {
        int32_t u = buf[0], v = buf[1];
        uint32_t wt = weighted ? (uint32_t)buf[2] : 1u;
        if (u < 0 || v < 0) return -1;
        any = true;
        if (max_u_v < u || max_u_v < v) max_u_v = std::max(u, v);
        if (u != v) {
            if ((int64_t)adj.size() <= u) { adj.resize((size_t)u + 1); adjw.resize((size_t)u + 1); }
            adj[u].push_back(v);
            adjw[u].push_back(wt);
            if (!directed) {
                if ((int64_t)adj.size() <= v) { adj.resize((size_t)v + 1); adjw.resize((size_t)v + 1); }
                adj[v].push_back(u);
                adjw[v].push_back(wt);
            }
        }
    }
    // n: adjacency-list size when undirected (:87); max id + 1 when directed
    // (the reference's :89 takes max id, one short).
    int32_t n = directed ? (any ? max_u_v + 1 : 0) : (int32_t)adj.size();
    out.n = n;
    out.off.assign((size_t)n + 1, 0);
    int64_t cur = 0;
    for (int32_t u = 0; u < n; ++u) {
        if ((size_t)u < adj.size()) {
            for (size_t k = 0; k < adj[u].size(); ++k) {
                out.nbr.push_back(adj[u][k]);
                if (weighted) out.w.push_back(adjw[u][k]);
                ++cur;
            }
        }
        out.off[u + 1] = (int32_t)cur;
    }
    return 0;
}

// argmax over (score, label) with the RS tie rule: highest score, then
// highest label.
template <class Score>
struct Best {
    bool has = false;
    Score s{};
    int32_t l = 0;
    void offer(Score sc, int32_t lab) {
        if (!has || sc > s || (sc == s && lab > l)) {
            has = true;
            s = sc;
            l = lab;
        }
    }
};

uint64_t splitmix64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" int orc_load_bin_size(const char* path, int directed, int weighted, int32_t* n, int64_t* m) {
    Loaded L;
    if (load(path, directed, weighted, L) != 0) return -1;
    *n = L.n;
    *m = (int64_t)L.nbr.size();
    return 0;
}

extern "C" int orc_load_bin(const char* path, int directed, int weighted, int32_t* offsets,
                            int32_t* neighbors, uint32_t* weights) {
    Loaded L;
    if (load(path, directed, weighted, L) != 0) return -1;
    std::copy(L.off.begin(), L.off.end(), offsets);
    std::copy(L.nbr.begin(), L.nbr.end(), neighbors);
    if (weights && weighted) std::copy(L.w.begin(), L.w.end(), weights);
    return 0;
}

extern "C" void orc_classic_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                                 const int32_t* prev, int32_t* next) {
    for (int32_t v = 0; v < n; ++v) {
        if (off[v] == off[v + 1]) {
            next[v] = prev[v];
            continue;
        }
        std::map<int32_t, uint64_t> count;
        for (int32_t j = off[v]; j < off[v + 1]; ++j) count[prev[nbr[j]]] += w ? w[j] : 1u;
        Best<uint64_t> b;
        for (auto& kv : count) b.offer(kv.second, kv.first);
        next[v] = b.l;
    }
}

extern "C" void orc_seeded_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                                const uint8_t* seeds, const int32_t* prev, int32_t* next) {
    for (int32_t v = 0; v < n; ++v) {
        if (seeds[v]) {
            next[v] = v;
            continue;
        }
        std::map<int32_t, uint64_t> count;
        for (int32_t j = off[v]; j < off[v + 1]; ++j)
            if (prev[nbr[j]] >= 0) count[prev[nbr[j]]] += w ? w[j] : 1u;
        if (count.empty()) {
            next[v] = prev[v];
            continue;
        }
        Best<uint64_t> b;
        for (auto& kv : count) b.offer(kv.second, kv.first);
        next[v] = b.l;
    }
}

extern "C" void orc_layered_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                                 int32_t p, int32_t q, const int32_t* prev, const int32_t* vprev,
                                 int32_t* next, int32_t* vnext) {
    for (int32_t v = 0; v < n; ++v) {
        if (off[v] == off[v + 1]) {
            next[v] = prev[v];
            continue;
        }
        std::map<int32_t, int64_t> count;
        for (int32_t j = off[v]; j < off[v + 1]; ++j) count[prev[nbr[j]]] += w ? w[j] : 1u;
        Best<int64_t> b;
        for (auto& kv : count)
            b.offer((int64_t)(q + p) * kv.second - (int64_t)p * vprev[kv.first], kv.first);
        next[v] = b.l;
    }
    for (int32_t l = 0; l < n; ++l) vnext[l] = 0;
    for (int32_t v = 0; v < n; ++v) vnext[next[v]]++;
}

extern "C" uint32_t orc_stream_u32(uint64_t seed, int32_t t, int32_t u) {
    uint64_t x = (seed << 40) ^ ((uint64_t)(uint32_t)t << 32) ^ (uint64_t)(uint32_t)u;
    return (uint32_t)splitmix64_mix(x + 0x9e3779b97f4a7c15ULL);
}

extern "C" void orc_slpa(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                         uint64_t seed, int32_t T, int32_t* mem) {
    for (int32_t v = 0; v < n; ++v) mem[v] = v;  // M_v[0] = v
    std::vector<int32_t> spoken(n);
    for (int32_t t = 1; t <= T; ++t) {
        for (int32_t u = 0; u < n; ++u) {
            uint32_t row = orc_stream_u32(seed, t, u) % (uint32_t)t;
            spoken[u] = mem[(size_t)row * n + u];
        }
        int32_t* cur = mem + (size_t)t * n;
        const int32_t* last = mem + (size_t)(t - 1) * n;
        for (int32_t v = 0; v < n; ++v) {
            if (off[v] == off[v + 1]) {
                cur[v] = last[v];
                continue;
            }
            std::map<int32_t, uint64_t> count;
            for (int32_t j = off[v]; j < off[v + 1]; ++j) count[spoken[nbr[j]]] += w ? w[j] : 1u;
            Best<uint64_t> b;
            for (auto& kv : count) b.offer(kv.second, kv.first);
            cur[v] = b.l;
        }
    }
}

extern "C" void orc_slpa_post(int32_t n, int32_t rows, const int32_t* mem, int32_t r_num, int32_t r_den,
                              int32_t kmax, int32_t* dominant, int32_t* nsets, int32_t* sets) {
    for (int32_t v = 0; v < n; ++v) {
        std::map<int32_t, int64_t> count;
        for (int32_t i = 0; i < rows; ++i) count[mem[(size_t)i * n + v]]++;
        Best<int64_t> b;
        int32_t kept = 0;
        for (auto& kv : count) {  // ascending label order
            b.offer(kv.second, kv.first);
            if (kv.second * r_den >= (int64_t)r_num * rows) {
                if (kept < kmax) sets[(size_t)v * kmax + kept] = kv.first;
                ++kept;
            }
        }
        for (int32_t k = kept; k < kmax; ++k) sets[(size_t)v * kmax + k] = -1;
        if (dominant) dominant[v] = b.l;
        if (nsets) nsets[v] = kept;
    }
}

extern "C" uint32_t orc_cm_hash(int32_t row, uint32_t k) {
    switch (row) {
        case 0: k ^= k >> 16; k *= 0x12ebcb3b; k ^= k >> 13; k *= 0xc123ae53; k ^= k >> 16; return k;
        case 1: k ^= k >> 16; k *= 0x21eb21bb; k ^= k >> 13; k *= 0x2cb2ea35; k ^= k >> 16; return k;
        case 2: k ^= k >> 16; k *= 0x8febce61; k ^= k >> 13; k *= 0xa2b2be25; k ^= k >> 16; return k;
        case 3: k ^= k >> 12; k *= 0x13ebdb3a; k ^= k >> 12; k *= 0xc123ae53; k ^= k >> 14; return k;
        default: k ^= k >> 13; k *= 0x312b521b; k ^= k >> 5; k *= 0x1cb3ea22; k ^= k >> 17; return k;
    }
}

extern "C" void orc_countmin_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                                  int32_t rows, int32_t width, int32_t aliased, const int32_t* prev,
                                  int32_t* next) {
    // (row, column) cells one insertion of `key` increments
    auto cells = [&](uint32_t key) {
        std::vector<std::pair<int, uint32_t>> c;
        if (aliased && rows == 3) {        // countmin.cuh:185-191
            c = {{0, orc_cm_hash(0, key) % width}, {1, orc_cm_hash(1, key) % width},
                 {1, orc_cm_hash(2, key) % width}};
        } else if (aliased && rows == 4) { // countmin.cuh:197-221
            c = {{0, orc_cm_hash(0, key) % width}, {1, orc_cm_hash(1, key) % width},
                 {1, orc_cm_hash(2, key) % width}, {1, orc_cm_hash(2, key) % width}};
        } else {
            for (int r = 0; r < rows; ++r) c.push_back({r, orc_cm_hash(r, key) % width});
        }
        return c;
    };
    for (int32_t v = 0; v < n; ++v) {
        if (off[v] == off[v + 1]) {
            next[v] = prev[v];
            continue;
        }
        std::vector<std::vector<uint64_t>> sketch(rows, std::vector<uint64_t>(width, 0));
        for (int32_t j = off[v]; j < off[v + 1]; ++j) {
            uint32_t key = (uint32_t)prev[nbr[j]] + 1u;
            for (auto& c : cells(key)) sketch[c.first][c.second] += w ? w[j] : 1u;
        }
        Best<uint64_t> b;
        for (int32_t j = off[v]; j < off[v + 1]; ++j) {
            uint32_t key = (uint32_t)prev[nbr[j]] + 1u;
            uint64_t est = UINT64_MAX;
            for (auto& c : cells(key)) est = std::min(est, sketch[c.first][c.second]);
            b.offer(est, prev[nbr[j]]);
        }
        next[v] = b.l;
    }
}

extern "C" void orc_table_step(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                               int32_t cap, int32_t buffer, const int32_t* prev, int32_t* next) {
    for (int32_t v = 0; v < n; ++v) {
        std::map<int32_t, uint64_t> table;  // label -> counter, at most cap entries
        for (int32_t j = off[v]; j < off[v + 1]; ++j) {
            const int32_t l = prev[nbr[j]];
            const uint64_t f = w ? w[j] : 1u;
            auto it = table.find(l);
            if (it != table.end()) {
                it->second += f;
            } else if ((int32_t)table.size() < cap) {
                table[l] = f;
            } else if (!buffer) {
                // Misra-Gries: f unit steps
                uint64_t units = f;
                while (units > 0) {
                    auto hit = table.find(l);
                    if (hit != table.end()) { hit->second += units; break; }
                    if ((int32_t)table.size() < cap) { table[l] = units; break; }
                    for (auto t = table.begin(); t != table.end();) {
                        if (--t->second == 0) t = table.erase(t);
                        else ++t;
                    }
                    --units;
                }
            } else {
                for (auto t = table.begin(); t != table.end();) {
                    t->second = t->second > f ? t->second - f : 0;
                    if (t->second == 0) t = table.erase(t);
                    else ++t;
                }
            }
        }
        Best<uint64_t> b;
        for (auto& kv : table) b.offer(kv.second, kv.first);
        next[v] = b.has ? b.l : prev[v];
    }
}

extern "C" double orc_cpulp_run(int32_t n, const int32_t* off, const int32_t* nbr, int32_t niter,
                                int32_t threads, int32_t mode, int32_t* labels_out) {
    std::vector<int32_t> labels(n), prelabels(n);
    for (int32_t i = 0; i < n; ++i) labels[i] = prelabels[i] = i;
    double total = 0;
    for (int32_t it = 0; it < niter; ++it) {
        auto t0 = std::chrono::system_clock::now();
        for (int32_t k = 0; k < n; ++k) prelabels[k] = labels[k];
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int32_t u = 0; u < n; ++u) {
            std::map<int32_t, int> label_count;
            int32_t max_label = prelabels[u];
            int max_score = 0;
            const int32_t b = off[u], e = off[u + 1];
            if (mode == 0) {
                // tutorial/CPUlp.cuh:53-78: live labels[], first label to
                // reach the running maximum wins
                for (int32_t j = b; j < e; ++j) {
                    int32_t label = labels[nbr[j]];
                    int c = ++label_count[label];
                    if (c > max_score) {
                        max_score = c;
                        max_label = label;
                    }
                }
            } else {
                for (int32_t j = b; j < e; ++j) ++label_count[prelabels[nbr[j]]];
                for (auto& kv : label_count)
                    if (kv.second > max_score || (kv.second == max_score && kv.first > max_label)) {
                        max_score = kv.second;
                        max_label = kv.first;
                    }
            }
            if (prelabels[u] != max_label) labels[u] = max_label;
        }
        auto t1 = std::chrono::system_clock::now();
        total += std::chrono::duration<double>(t1 - t0).count();
    }
    if (labels_out) std::copy(labels.begin(), labels.end(), labels_out);
    return total;
}

extern "C" double orc_cpulp_sample(int32_t n, const int32_t* off, const int32_t* nbr, int32_t* labels,
                                   const int32_t* verts, int64_t nverts, int32_t threads,
                                   int32_t balanced) {
    std::vector<int32_t> prelabels(labels, labels + n);
    // balanced = 0: the reference's `omp parallel for` with the default
    // (static) schedule over vertices in id order (tutorial/CPUlp.cuh:50-51),
    // so with degree-sorted ids the first thread gets the hubs, as in a whole
    // step; balanced = 1: dynamic chunks of 64 vertices
    if (balanced) omp_set_schedule(omp_sched_dynamic, 64);
    else omp_set_schedule(omp_sched_static, 0);
    auto t0 = std::chrono::system_clock::now();
#pragma omp parallel for num_threads(threads) schedule(runtime)
    for (int64_t k = 0; k < nverts; ++k) {
        const int32_t u = verts[k];
        std::map<int32_t, int> label_count;
        int32_t max_label = prelabels[u];
        int max_score = 0;
        for (int32_t j = off[u]; j < off[u + 1]; ++j) {
            int32_t label = labels[nbr[j]];
            int c = ++label_count[label];
            if (c > max_score) {
                max_score = c;
                max_label = label;
            }
        }
        if (prelabels[u] != max_label) labels[u] = max_label;
    }
    auto t1 = std::chrono::system_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

extern "C" void orc_classic_subset(int32_t n, const int32_t* off, const int32_t* nbr, const uint32_t* w,
                                   const int32_t* prev, const int32_t* verts, int64_t nverts, int32_t threads,
                                   int32_t* next) {
    (void)n;
#pragma omp parallel num_threads(threads)
    {
        std::vector<std::pair<int32_t, uint32_t>> seen;
#pragma omp for schedule(dynamic, 16)
        for (int64_t k = 0; k < nverts; ++k) {
            const int32_t v = verts[k];
            const int32_t b = off[v], e = off[v + 1];
            if (b == e) {
                next[k] = prev[v];
                continue;
            }
            seen.clear();
            for (int32_t j = b; j < e; ++j) seen.emplace_back(prev[nbr[j]], w ? w[j] : 1u);
            std::sort(seen.begin(), seen.end());
            Best<uint64_t> best;
            for (size_t i = 0; i < seen.size();) {
                size_t j = i;
                uint64_t f = 0;
                while (j < seen.size() && seen[j].first == seen[i].first) f += seen[j++].second;
                best.offer(f, seen[i].first);
                i = j;
            }
            next[k] = best.l;
        }
    }
}

extern "C" double orc_cpulp_layered_sample(int32_t n, const int32_t* off, const int32_t* nbr, int32_t* labels,
                                           const int32_t* Vr, int32_t* Vw, const int32_t* verts, int64_t nverts,
                                           int32_t threads, int32_t p, int32_t q) {
    std::vector<int32_t> prelabels(labels, labels + n);
    omp_set_schedule(omp_sched_static, 0);
    auto t0 = std::chrono::system_clock::now();
#pragma omp parallel for num_threads(threads) schedule(runtime)
    for (int64_t k = 0; k < nverts; ++k) {
        const int32_t u = verts[k];
        std::map<int32_t, int> label_count;
        int32_t max_label = prelabels[u];
        int64_t max_score = INT64_MIN;
        for (int32_t j = off[u]; j < off[u + 1]; ++j) {
            const int32_t label = labels[nbr[j]];
            const int c = ++label_count[label];
            const int64_t s = (int64_t)(q + p) * c - (int64_t)p * Vr[label];
            if (s > max_score) {
                max_score = s;
                max_label = label;
            }
        }
        __atomic_fetch_add(&Vw[max_label], 1, __ATOMIC_RELAXED);
        if (prelabels[u] != max_label) labels[u] = max_label;
    }
    auto t1 = std::chrono::system_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}
