// Seeded synthetic graph generators for the benchmark workloads.  Every
// output is a pure function of the arguments (independent of thread count):
// per-edge or per-block random streams are derived from the seed with the
// SplitMix64 finaliser, never from a shared sequential state.

#include "internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include <omp.h>

namespace glp {
namespace {

inline uint64_t derive(uint64_t seed, uint64_t a, uint64_t b = 0) {
    return splitmix64_mix(seed ^ splitmix64_mix(a * 0x9e3779b97f4a7c15ULL + b + 1));
}

// R-MAT endpoint pair of draw i.  Each level consumes 32 random bits and
// picks a quadrant with integer thresholds floor(p * 2^32).
struct RmatDraw {
    int scale;
    uint64_t seed;
    uint64_t tA, tAB, tABC;
    inline void operator()(uint64_t i, uint32_t& u, uint32_t& v) const {
        uint64_t st = derive(seed, i);
        uint32_t uu = 0, vv = 0;
        uint64_t word = 0;
        for (int lvl = 0; lvl < scale; ++lvl) {
            uint32_t r;
            if ((lvl & 1) == 0) {
                st += 0x9e3779b97f4a7c15ULL;
                word = splitmix64_mix(st);
                r = uint32_t(word);
            } else {
                r = uint32_t(word >> 32);
            }
            uint32_t ub, vb;
            if (r < tA) { ub = 0; vb = 0; }
            else if (r < tAB) { ub = 0; vb = 1; }
            else if (r < tABC) { ub = 1; vb = 0; }
            else { ub = 1; vb = 1; }
            uu = (uu << 1) | ub;
            vv = (vv << 1) | vb;
        }
        u = uu;
        v = vv;
    }
};

}  // namespace
}  // namespace glp

using namespace glp;

extern "C" uint32_t glp_stream_u32(uint64_t seed, int32_t t, int32_t u) {
    return glp::stream_u32(seed, (uint32_t)t, (uint32_t)u);
}

extern "C" int glp_stream_fill(uint64_t seed, int32_t t, int32_t u0, int32_t count, uint32_t* out) {
    if (!out || count < 0) return set_error(GLP_E_ARG, "bad stream_fill arguments");
    for (int32_t i = 0; i < count; ++i) out[i] = glp::stream_u32(seed, (uint32_t)t, (uint32_t)(u0 + i));
    return GLP_OK;
}

extern "C" int glp_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                            uint64_t seed, int32_t degree_sort, int32_t wmax, int32_t threads,
                            glp_graph** out) {
    clear_error();
    if (!out || scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 ||
        a + b + c > 1.0 || wmax < 0)
        return set_error(GLP_E_ARG, "bad rmat arguments");
    *out = nullptr;
    threads = host_threads(threads);
    const int32_t n = int32_t(1) << scale;
    const uint64_t draws = (uint64_t)edge_factor << scale;
    RmatDraw gen;
    gen.scale = scale;
    gen.seed = seed;
    const double two32 = 4294967296.0;
    gen.tA = (uint64_t)std::floor(a * two32);
    gen.tAB = (uint64_t)std::floor((a + b) * two32);
    gen.tABC = (uint64_t)std::floor((a + b + c) * two32);
    // pass 1: count both directions per source to size the pair array
    std::vector<uint64_t> pairs;
    {
        std::vector<int64_t> cnt(threads, 0);
#pragma omp parallel num_threads(threads)
        {
            int tid = omp_get_thread_num();
            int64_t local = 0;
#pragma omp for schedule(static)
            for (int64_t i = 0; i < (int64_t)draws; ++i) {
                uint32_t u, v;
                gen((uint64_t)i, u, v);
                if (u != v) local += 2;
            }
            cnt[tid] = local;
        }
        int64_t total = std::accumulate(cnt.begin(), cnt.end(), int64_t(0));
        pairs.resize(total);
    }
    // pass 2: write (u,v) and (v,u); positions from a per-chunk prefix so the
    // array content is deterministic
    {
        const int64_t chunk = 1 << 20;
        const int64_t nchunks = ((int64_t)draws + chunk - 1) / chunk;
        std::vector<int64_t> cpos(nchunks + 1, 0);
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t k = 0; k < nchunks; ++k) {
            int64_t cnt = 0;
            for (int64_t i = k * chunk; i < std::min<int64_t>((k + 1) * chunk, draws); ++i) {
                uint32_t u, v;
                gen((uint64_t)i, u, v);
                if (u != v) cnt += 2;
            }
            cpos[k + 1] = cnt;
        }
        for (int64_t k = 0; k < nchunks; ++k) cpos[k + 1] += cpos[k];
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t k = 0; k < nchunks; ++k) {
            int64_t p = cpos[k];
            for (int64_t i = k * chunk; i < std::min<int64_t>((k + 1) * chunk, draws); ++i) {
                uint32_t u, v;
                gen((uint64_t)i, u, v);
                if (u == v) continue;
                pairs[p++] = (uint64_t(u) << 32) | v;
                pairs[p++] = (uint64_t(v) << 32) | u;
            }
        }
    }
    auto g = std::make_unique<glp_graph>();
    int rc = csr_from_pairs(n, pairs, threads, g->csr);
    if (rc != GLP_OK) return rc;
    g->csr.symmetric = 1;  // every pair was added in both directions
    if (wmax > 0) {
        // symmetric weights: w(u,v) = w(v,u), a hash of the unordered pair
        HostCSR& cs = g->csr;
        cs.weights.resize(cs.m);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096)
        for (int32_t u = 0; u < cs.n; ++u) {
            for (int32_t j = cs.offsets[u]; j < cs.offsets[u + 1]; ++j) {
                uint32_t v = (uint32_t)cs.neighbors[j];
                uint64_t lo = std::min<uint32_t>(u, v), hi = std::max<uint32_t>(u, v);
                uint64_t h = derive(seed ^ 0x5745494748545321ULL, (hi << 32) | lo);
                cs.weights[j] = 1u + (uint32_t)(h % (uint64_t)wmax);
            }
        }
    }
    if (degree_sort) {
        glp_graph* sorted = nullptr;
        rc = glp_graph_degree_sort(g.get(), &sorted, nullptr);
        if (rc != GLP_OK) return rc;
        *out = sorted;
        return GLP_OK;
    }
    *out = g.release();
    return GLP_OK;
}

namespace {

// Geometric skipping over `total` Bernoulli(p) trials; calls emit(idx) for
// every success in increasing idx order.
template <class F>
void bernoulli_skip(uint64_t total, double p, SplitMix64& rng, F emit) {
    if (p <= 0 || total == 0) return;
    if (p >= 1) {
        for (uint64_t i = 0; i < total; ++i) emit(i);
        return;
    }
    const double lq = std::log1p(-p);
    int64_t idx = -1;
    for (;;) {
        double u = rng.uniform();
        double skip = std::floor(std::log1p(-u) / lq);  // number of failures
        if (!(skip < 1.8e19)) return;
        idx += 1 + (int64_t)skip;
        if (idx < 0 || (uint64_t)idx >= total) return;
        emit((uint64_t)idx);
    }
}

}  // namespace

extern "C" int glp_gen_planted(int32_t n, int32_t k, double deg_in, double deg_out, uint64_t seed,
                               int32_t threads, int32_t* community_out, glp_graph** out) {
    clear_error();
    if (!out || n < 1 || k < 1 || k > n || deg_in < 0 || deg_out < 0)
        return set_error(GLP_E_ARG, "bad planted-partition arguments");
    *out = nullptr;
    threads = host_threads(threads);
    // community c holds ids [start[c], start[c+1]); sizes differ by at most one
    std::vector<int64_t> start(k + 1, 0);
    for (int32_t c = 0; c < k; ++c) start[c + 1] = start[c] + n / k + (c < n % k ? 1 : 0);
    // blocks (c1 <= c2) in a fixed order, each with its own stream
    std::vector<std::pair<int32_t, int32_t>> blocks;
    for (int32_t c1 = 0; c1 < k; ++c1)
        for (int32_t c2 = c1; c2 < k; ++c2) blocks.emplace_back(c1, c2);
    std::vector<std::vector<uint64_t>> found(blocks.size());
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
    for (int64_t bi = 0; bi < (int64_t)blocks.size(); ++bi) {
        int32_t c1 = blocks[bi].first, c2 = blocks[bi].second;
        int64_t s1 = start[c1 + 1] - start[c1], s2 = start[c2 + 1] - start[c2];
        SplitMix64 rng(derive(seed, (uint64_t)c1 * 1000003ULL + (uint64_t)c2, 7));
        std::vector<uint64_t>& f = found[bi];
        if (c1 == c2) {
            if (s1 < 2) continue;
            double p = deg_in / double(s1 - 1);
            uint64_t total = (uint64_t)s1 * (s1 - 1) / 2;
            // row-major upper triangle: row i holds pairs (i, i+1..s1-1)
            int64_t row = 0, row_start = 0;
            bernoulli_skip(total, p, rng, [&](uint64_t idx) {
                while ((int64_t)idx >= row_start + (s1 - 1 - row)) {
                    row_start += s1 - 1 - row;
                    ++row;
                }
                int64_t col = row + 1 + ((int64_t)idx - row_start);
                uint32_t u = (uint32_t)(start[c1] + row), v = (uint32_t)(start[c1] + col);
                f.push_back((uint64_t(u) << 32) | v);
            });
        } else {
            double p = (n - s1) > 0 ? deg_out / double(n - s1) : 0.0;
            uint64_t total = (uint64_t)s1 * (uint64_t)s2;
            bernoulli_skip(total, p, rng, [&](uint64_t idx) {
                uint32_t u = (uint32_t)(start[c1] + (int64_t)(idx / s2));
                uint32_t v = (uint32_t)(start[c2] + (int64_t)(idx % s2));
                f.push_back((uint64_t(u) << 32) | v);
            });
        }
    }
    size_t total = 0;
    for (auto& f : found) total += 2 * f.size();
    std::vector<uint64_t> pairs;
    pairs.reserve(total);
    for (auto& f : found) {
        for (uint64_t e : f) {
            pairs.push_back(e);
            pairs.push_back((e << 32) | (e >> 32));
        }
        std::vector<uint64_t>().swap(f);
    }
    auto g = std::make_unique<glp_graph>();
    int rc = csr_from_pairs(n, pairs, threads, g->csr);
    if (rc != GLP_OK) return rc;
    g->csr.symmetric = 1;  // every pair was added in both directions
    if (community_out)
        for (int32_t c = 0; c < k; ++c)
            for (int64_t v = start[c]; v < start[c + 1]; ++v) community_out[v] = c;
    *out = g.release();
    return GLP_OK;
}

namespace {
// inverse-CDF sample of a continuous power law x^-tau on [lo, hi]
double powerlaw(SplitMix64& rng, double tau, double lo, double hi) {
    double u = rng.uniform();
    if (std::fabs(tau - 1.0) < 1e-12) return lo * std::pow(hi / lo, u);
    double e = 1.0 - tau;
    double a = std::pow(lo, e), b = std::pow(hi, e);
    return std::pow(a + u * (b - a), 1.0 / e);
}
template <class T>
void shuffle(std::vector<T>& v, SplitMix64& rng) {
    for (size_t i = v.size(); i > 1; --i) std::swap(v[i - 1], v[rng.below(i)]);
}
}  // namespace

extern "C" int glp_gen_lfr(int32_t n, double tau1, int32_t kmin, int32_t kmax, double tau2,
                           int32_t cmin, int32_t cmax, double mu, uint64_t seed,
                           int32_t* community_out, glp_graph** out) {
    clear_error();
    if (!out || n < 2 || kmin < 1 || kmax < kmin || cmin < 2 || cmax < cmin || mu < 0 || mu > 1)
        return set_error(GLP_E_ARG, "bad LFR arguments");
    *out = nullptr;
    SplitMix64 rng(derive(seed, 0x4c4652));
    // degrees
    std::vector<int32_t> deg(n);
    for (int32_t v = 0; v < n; ++v) {
        double x = powerlaw(rng, tau1, kmin, kmax + 1.0);
        deg[v] = std::min<int32_t>(kmax, (int32_t)std::floor(x));
    }
    // community sizes until they cover n
    std::vector<int32_t> sizes;
    int64_t covered = 0;
    while (covered < n) {
        int32_t s = std::min<int32_t>(cmax, (int32_t)std::floor(powerlaw(rng, tau2, cmin, cmax + 1.0)));
        if (covered + s > n) s = (int32_t)(n - covered);
        if (s < cmin && !sizes.empty()) {
            sizes.back() += s;
        } else {
            sizes.push_back(s);
        }
        covered += s;
    }
    const int32_t k = (int32_t)sizes.size();
    // membership: shuffled vertex order filled community by community
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    shuffle(order, rng);
    std::vector<int32_t> comm(n);
    std::vector<std::vector<int32_t>> members(k);
    {
        int64_t p = 0;
        for (int32_t c = 0; c < k; ++c)
            for (int32_t i = 0; i < sizes[c]; ++i) {
                comm[order[p]] = c;
                members[c].push_back(order[p]);
                ++p;
            }
    }
    std::vector<uint64_t> pairs;
    pairs.reserve((size_t)std::accumulate(deg.begin(), deg.end(), int64_t(0)) + 16);
    // intra-community configuration model
    std::vector<int32_t> kout(n);
    for (int32_t c = 0; c < k; ++c) {
        std::vector<int32_t> stubs;
        const int32_t cs = (int32_t)members[c].size();
        for (int32_t v : members[c]) {
            int32_t kin = (int32_t)std::lround((1.0 - mu) * deg[v]);
            kin = std::min(kin, cs - 1);
            kout[v] = deg[v] - kin;
            for (int32_t i = 0; i < kin; ++i) stubs.push_back(v);
        }
        shuffle(stubs, rng);
        for (size_t i = 0; i + 1 < stubs.size(); i += 2) {
            uint32_t u = stubs[i], v = stubs[i + 1];
            if (u == v) continue;
            pairs.push_back((uint64_t(u) << 32) | v);
            pairs.push_back((uint64_t(v) << 32) | u);
        }
    }
    // inter-community configuration model
    {
        std::vector<int32_t> stubs;
        for (int32_t v = 0; v < n; ++v)
            for (int32_t i = 0; i < kout[v]; ++i) stubs.push_back(v);
        shuffle(stubs, rng);
        for (size_t i = 0; i + 1 < stubs.size(); i += 2) {
            uint32_t u = stubs[i], v = stubs[i + 1];
            if (u == v || comm[u] == comm[v]) continue;
            pairs.push_back((uint64_t(u) << 32) | v);
            pairs.push_back((uint64_t(v) << 32) | u);
        }
    }
    auto g = std::make_unique<glp_graph>();
    int rc = csr_from_pairs(n, pairs, 0, g->csr);
    if (rc != GLP_OK) return rc;
    g->csr.symmetric = 1;  // every pair was added in both directions
    if (community_out) std::memcpy(community_out, comm.data(), sizeof(int32_t) * n);
    *out = g.release();
    return GLP_OK;
}

extern "C" int glp_gen_mix_params(uint64_t seed, int32_t i, int32_t nmin, int32_t nmax, double dmin,
                                  double dmax, double fmin, double fmax, int32_t* n, int32_t* k,
                                  double* deg_in, double* deg_out, uint64_t* graph_seed) {
    if (!n || !k || !deg_in || !deg_out || !graph_seed || nmin < 2 || nmax < nmin || i < 0)
        return set_error(GLP_E_ARG, "bad mix arguments");
    SplitMix64 rng(derive(seed, (uint64_t)i, 0x4d4958));
    double ln = std::log((double)nmin) + rng.uniform() * (std::log((double)nmax) - std::log((double)nmin));
    int32_t nn = (int32_t)std::min<double>(nmax, std::max<double>(nmin, std::floor(std::exp(ln))));
    double d = dmin + rng.uniform() * (dmax - dmin);
    double f = fmin + rng.uniform() * (fmax - fmin);
    *n = nn;
    *k = std::max<int32_t>(1, (nn + 999) / 1000);
    *deg_in = d * (1.0 - f);
    *deg_out = d * f;
    *graph_seed = rng.next();
    return GLP_OK;
}
