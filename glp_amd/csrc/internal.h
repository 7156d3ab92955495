// Internal host-side declarations shared by the libglp.so translation units.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <vector>
#include <memory>

#include "../../include/glp.h"
#include "../../include/glp_tools.h"

namespace glp {

// Thread-local error message; set_error returns the code so call sites can
// `return set_error(GLP_E_ARG, "...")`.
int set_error(int code, const char* fmt, ...);
void clear_error();

// Host CSR.  offsets has n+1 entries, neighbors m, weights m or empty.
struct HostCSR {
    int32_t n = 0;
    int64_t m = 0;
    std::vector<int32_t> offsets;
    std::vector<int32_t> neighbors;
    std::vector<uint32_t> weights;
    int32_t max_degree = 0;
    // 1: every edge u->v has an edge v->u (the frontier marks a changed
    // vertex's out-neighbours, so it needs this); 0: not; -1: not yet known
    int8_t symmetric = -1;
    bool weighted() const { return !weights.empty(); }
    void finish();  // computes max_degree
};

struct DeviceGraph;  // defined in device.h (owns HBM copies and the tier plan)

}  // namespace glp

struct glp_graph {
    glp::HostCSR csr;
    glp::DeviceGraph* dev = nullptr;  // created lazily by the first session
    ~glp_graph();
};

namespace glp {

// SplitMix64 finaliser (Steele, Lea, Flood 2014 "Fast splittable PRNGs").
inline uint64_t splitmix64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
inline uint32_t stream_u32(uint64_t seed, uint32_t t, uint32_t u) {
    uint64_t x = (seed << 40) ^ (uint64_t(t) << 32) ^ uint64_t(u);
    return uint32_t(splitmix64_mix(x + 0x9e3779b97f4a7c15ULL));
}

// Sequential SplitMix64 generator for host-side generators.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() { s += 0x9e3779b97f4a7c15ULL; return splitmix64_mix(s); }
    // uniform double in [0,1) with 53 random bits
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    // uniform integer in [0, k)
    uint64_t below(uint64_t k) { return k ? next() % k : 0; }
};

int host_threads(int requested);

// Builds a symmetric, deduplicated CSR (neighbours ascending) from directed
// pairs (u,v) packed as (u<<32|v); self-loops must already be absent.
// pairs is consumed (sorted in place).
int csr_from_pairs(int32_t n, std::vector<uint64_t>& pairs, int threads, HostCSR& out);
// Whether every edge u->v has an edge v->u (multiplicities may differ);
// computed once (transpose by counting, then each in-list looked up in the
// sorted out-list) and kept in g.symmetric.
bool csr_is_symmetric(HostCSR& g);

}  // namespace glp
