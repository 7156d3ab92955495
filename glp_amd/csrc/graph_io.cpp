// Host graph handling: the .bin edge-list reader and CSR builder, CSR import
// and export, the degree relabeling, and the edgelist2bin / normalize tools.
//
// Reader semantics follow header/graph.h:227-258 (load_edgelist_bin) and the
// adjacency-list -> CSR walk of header/graph.h:103-130: pairs are consumed in
// file order, self-loops dropped, each pair appended to adj[u] (and adj[v]
// when undirected), duplicates kept, neighbour order = insertion order.  The
// adjacency lists are built here by a two-pass counting placement instead of
// vector<vector>, which yields the same per-vertex order.

#include "internal.h"

#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <thread>

#include <omp.h>

namespace glp {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
void clear_error() { g_err.clear(); }
const char* last_error() { return g_err.c_str(); }

void HostCSR::finish() {
    max_degree = 0;
    for (int32_t v = 0; v < n; ++v)
        max_degree = std::max(max_degree, offsets[v + 1] - offsets[v]);
}

int host_threads(int requested) {
    if (requested > 0) return requested;
    int hw = (int)std::thread::hardware_concurrency();
    const char* env = getenv("OMP_NUM_THREADS");
    if (env && atoi(env) > 0) hw = atoi(env);
    return hw > 0 ? hw : 1;
}

static bool read_file(const char* path, std::vector<char>& buf) {
    std::ifstream ifs(path, std::ios::binary | std::ios::ate);
    if (!ifs) return false;
    std::streamsize size = ifs.tellg();
    if (size < 0) return false;
    ifs.seekg(0, std::ios::beg);
    buf.resize((size_t)size);
    if (size > 0 && !ifs.read(buf.data(), size)) return false;
    return true;
}

// Two-pass CSR placement from records in file order.  rec_words = 2 (u,v) or
// 3 (u,v,w).  Follows header/graph.h:235-253 and :103-130.
static int build_from_records(const int32_t* rec, size_t nrec, int rec_words,
                              bool directed, HostCSR& out) {
    int64_t max_u_v = 0;     // header/graph.h:238-240
    int64_t adj_size = 0;    // adj_list.size() after the resizes at :242-250
    bool any = false;
    for (size_t i = 0; i < nrec; ++i) {
        int32_t u = rec[i * rec_words], v = rec[i * rec_words + 1];
        if (u < 0 || v < 0)
            return set_error(GLP_E_FORMAT, "negative vertex id in record %zu (%d,%d)", i, u, v);
        any = true;
        max_u_v = std::max<int64_t>(max_u_v, std::max(u, v));
        if (u != v) {
            adj_size = std::max<int64_t>(adj_size, (int64_t)u + 1);
            if (!directed) adj_size = std::max<int64_t>(adj_size, (int64_t)v + 1);
        }
    }
    // Undirected: n = adj_list.size() (:87).  Directed: the reference sets
    // n = max id (:89), one short of the id range; n = max id + 1 here.
    int64_t n = directed ? (any ? max_u_v + 1 : 0) : adj_size;
    if (n > INT32_MAX - 1) return set_error(GLP_E_RANGE, "vertex count %lld exceeds int32", (long long)n);
    std::vector<int64_t> deg(n + 1, 0);
    int64_t m = 0;
    for (size_t i = 0; i < nrec; ++i) {
        int32_t u = rec[i * rec_words], v = rec[i * rec_words + 1];
        if (u == v) continue;
        deg[u]++; m++;
        if (!directed) { deg[v]++; m++; }
    }
    if (m > INT32_MAX) return set_error(GLP_E_RANGE, "edge count %lld exceeds int32 CSR offsets", (long long)m);
    out.n = (int32_t)n;
    out.m = m;
    out.offsets.assign(n + 1, 0);
    for (int64_t v = 0; v < n; ++v) out.offsets[v + 1] = out.offsets[v] + (int32_t)deg[v];
    out.neighbors.assign(m, 0);
    if (rec_words == 3) out.weights.assign(m, 0); else out.weights.clear();
    std::vector<int32_t> cur(out.offsets.begin(), out.offsets.end() - 1);
    for (size_t i = 0; i < nrec; ++i) {
        int32_t u = rec[i * rec_words], v = rec[i * rec_words + 1];
        if (u == v) continue;
        int32_t p = cur[u]++;
        out.neighbors[p] = v;
        if (rec_words == 3) out.weights[p] = (uint32_t)rec[i * 3 + 2];
        if (!directed) {
            int32_t q = cur[v]++;
            out.neighbors[q] = u;
            if (rec_words == 3) out.weights[q] = (uint32_t)rec[i * 3 + 2];
        }
    }
    out.finish();
    out.symmetric = directed ? -1 : 1;
    return GLP_OK;
}

bool csr_is_symmetric(HostCSR& g) {
    if (g.symmetric >= 0) return g.symmetric == 1;
    const int32_t n = g.n;
    std::vector<int64_t> toff((size_t)n + 1, 0);
    for (int64_t j = 0; j < g.m; ++j) toff[(size_t)g.neighbors[j] + 1]++;
    for (int32_t v = 0; v < n; ++v) toff[v + 1] += toff[v];
    std::vector<int32_t> tin((size_t)g.m);
    {
        std::vector<int64_t> cur(toff.begin(), toff.end() - 1);
        for (int32_t u = 0; u < n; ++u)
            for (int32_t j = g.offsets[u]; j < g.offsets[u + 1]; ++j) tin[cur[g.neighbors[j]]++] = u;
    }
    int ok = 1;
    const int threads = host_threads(0);
#pragma omp parallel num_threads(threads) reduction(&& : ok)
    {
        std::vector<int32_t> out;
#pragma omp for schedule(dynamic, 4096)
        for (int32_t v = 0; v < n; ++v) {
            if (!ok) continue;
            out.assign(g.neighbors.begin() + g.offsets[v], g.neighbors.begin() + g.offsets[v + 1]);
            std::sort(out.begin(), out.end());
            for (int64_t k = toff[v]; k < toff[v + 1]; ++k)
                if (!std::binary_search(out.begin(), out.end(), tin[k])) {
                    ok = 0;
                    break;
                }
        }
    }
    g.symmetric = ok ? 1 : 0;
    return ok != 0;
}

int csr_from_pairs(int32_t n, std::vector<uint64_t>& pairs, int threads, HostCSR& out) {
    threads = host_threads(threads);
    // counting placement by source, then per-vertex sort + dedup
    std::vector<int64_t> deg(n + 1, 0);
    const int64_t np = (int64_t)pairs.size();
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < np; ++i) {
        uint32_t u = uint32_t(pairs[i] >> 32);
        __atomic_fetch_add(&deg[u], 1, __ATOMIC_RELAXED);
    }
    std::vector<int64_t> off(n + 1, 0);
    for (int32_t v = 0; v < n; ++v) off[v + 1] = off[v] + deg[v];
    std::vector<int32_t> tmp(np);
    std::vector<int64_t> cur(off.begin(), off.end() - 1);
#pragma omp parallel for num_threads(threads) schedule(static)
    for (int64_t i = 0; i < np; ++i) {
        uint32_t u = uint32_t(pairs[i] >> 32);
        int64_t p = __atomic_fetch_add(&cur[u], 1, __ATOMIC_RELAXED);
        tmp[p] = int32_t(pairs[i] & 0xffffffffu);
    }
    std::vector<uint64_t>().swap(pairs);
    std::vector<int64_t> ndeg(n, 0);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096)
    for (int32_t v = 0; v < n; ++v) {
        int32_t* b = tmp.data() + off[v];
        int32_t* e = tmp.data() + off[v + 1];
        std::sort(b, e);
        ndeg[v] = std::unique(b, e) - b;
    }
    int64_t m = 0;
    out.n = n;
    out.offsets.assign((size_t)n + 1, 0);
    for (int32_t v = 0; v < n; ++v) {
        m += ndeg[v];
        if (m > INT32_MAX) return set_error(GLP_E_RANGE, "edge count exceeds int32 CSR offsets");
        out.offsets[v + 1] = (int32_t)m;
    }
    out.m = m;
    out.neighbors.resize(m);
    out.weights.clear();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096)
    for (int32_t v = 0; v < n; ++v)
        std::memcpy(out.neighbors.data() + out.offsets[v], tmp.data() + off[v],
                    sizeof(int32_t) * ndeg[v]);
    out.finish();
    return GLP_OK;
}

}  // namespace glp

using namespace glp;

glp_graph::~glp_graph() {}

extern "C" const char* glp_last_error(void) { return glp::last_error(); }
extern "C" const char* glp_version(void) { return "glp-mi355x 0.1 (gfx950)"; }

extern "C" int glp_graph_load_bin(const char* path, int32_t directed, int32_t weighted,
                                  glp_graph** out) {
    clear_error();
    if (!path || !out) return set_error(GLP_E_ARG, "null argument");
    *out = nullptr;
    std::vector<char> buf;
    if (!read_file(path, buf)) return set_error(GLP_E_IO, "cannot read '%s'", path);
    const int words = weighted ? 3 : 2;
    // The reference reads 8 bytes at a time and drops a trailing partial pair
    // (header/graph.h:235); the same truncation applies here.
    size_t nrec = buf.size() / (sizeof(int32_t) * words);
    auto g = std::make_unique<glp_graph>();
    int rc = build_from_records(reinterpret_cast<const int32_t*>(buf.data()), nrec, words,
                                directed != 0, g->csr);
    if (rc != GLP_OK) return rc;
    if (weighted) {
        for (uint32_t w : g->csr.weights)
            if (w == 0) return set_error(GLP_E_FORMAT, "zero edge weight");
    }
    *out = g.release();
    return GLP_OK;
}

extern "C" int glp_graph_from_csr(int32_t n, int64_t m, const int32_t* offsets,
                                  const int32_t* neighbors, const uint32_t* weights,
                                  glp_graph** out) {
    clear_error();
    if (!out || n < 0 || m < 0 || !offsets || (m > 0 && !neighbors))
        return set_error(GLP_E_ARG, "bad CSR arguments");
    *out = nullptr;
    if (m > INT32_MAX) return set_error(GLP_E_RANGE, "m=%lld exceeds int32 offsets", (long long)m);
    if (offsets[0] != 0 || (int64_t)offsets[n] != m)
        return set_error(GLP_E_FORMAT, "offsets must start at 0 and end at m");
    for (int32_t v = 0; v < n; ++v)
        if (offsets[v + 1] < offsets[v]) return set_error(GLP_E_FORMAT, "offsets not monotone at %d", v);
    for (int64_t j = 0; j < m; ++j)
        if (neighbors[j] < 0 || neighbors[j] >= n)
            return set_error(GLP_E_FORMAT, "neighbor %d at %lld out of range", neighbors[j], (long long)j);
    auto g = std::make_unique<glp_graph>();
    g->csr.n = n;
    g->csr.m = m;
    g->csr.offsets.assign(offsets, offsets + n + 1);
    g->csr.neighbors.assign(neighbors, neighbors + m);
    if (weights) {
        g->csr.weights.assign(weights, weights + m);
        for (uint32_t w : g->csr.weights)
            if (w == 0) return set_error(GLP_E_FORMAT, "zero edge weight");
    }
    g->csr.finish();
    *out = g.release();
    return GLP_OK;
}

extern "C" int glp_graph_symmetric(glp_graph* g, int32_t* symmetric) {
    clear_error();
    if (!g || !symmetric) return set_error(GLP_E_ARG, "null argument");
    *symmetric = csr_is_symmetric(g->csr) ? 1 : 0;
    return GLP_OK;
}

extern "C" int glp_graph_info(const glp_graph* g, int32_t* n, int64_t* m,
                              int32_t* max_degree, int32_t* weighted) {
    if (!g) return set_error(GLP_E_ARG, "null graph");
    if (n) *n = g->csr.n;
    if (m) *m = g->csr.m;
    if (max_degree) *max_degree = g->csr.max_degree;
    if (weighted) *weighted = g->csr.weighted() ? 1 : 0;
    return GLP_OK;
}

extern "C" int glp_graph_csr(const glp_graph* g, int32_t* offsets, int32_t* neighbors,
                             uint32_t* weights) {
    if (!g) return set_error(GLP_E_ARG, "null graph");
    if (offsets) std::memcpy(offsets, g->csr.offsets.data(), sizeof(int32_t) * (g->csr.n + 1));
    if (neighbors && g->csr.m)
        std::memcpy(neighbors, g->csr.neighbors.data(), sizeof(int32_t) * g->csr.m);
    if (weights && g->csr.weighted())
        std::memcpy(weights, g->csr.weights.data(), sizeof(uint32_t) * g->csr.m);
    return GLP_OK;
}

extern "C" int glp_graph_save_bin(const glp_graph* g, const char* path) {
    if (!g || !path) return set_error(GLP_E_ARG, "null argument");
    std::ofstream ofs(path, std::ios::binary | std::ios::trunc);
    if (!ofs) return set_error(GLP_E_IO, "cannot write '%s'", path);
    const HostCSR& c = g->csr;
    std::vector<int32_t> out;
    out.reserve(1 << 20);
    for (int32_t u = 0; u < c.n; ++u) {
        for (int32_t j = c.offsets[u]; j < c.offsets[u + 1]; ++j) {
            int32_t v = c.neighbors[j];
            if (v <= u) continue;
            out.push_back(u);
            out.push_back(v);
            if (c.weighted()) out.push_back((int32_t)c.weights[j]);
        }
        if (out.size() > (1u << 20)) {
            ofs.write(reinterpret_cast<const char*>(out.data()), out.size() * 4);
            out.clear();
        }
    }
    ofs.write(reinterpret_cast<const char*>(out.data()), out.size() * 4);
    if (!ofs) return set_error(GLP_E_IO, "write failed on '%s'", path);
    return GLP_OK;
}

extern "C" int glp_graph_degree_sort(const glp_graph* g, glp_graph** out, int32_t* perm_out) {
    if (!g || !out) return set_error(GLP_E_ARG, "null argument");
    const HostCSR& c = g->csr;
    const int32_t n = c.n;
    std::vector<int32_t> perm(n);  // new -> old
    for (int32_t v = 0; v < n; ++v) perm[v] = v;
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
        return (c.offsets[a + 1] - c.offsets[a]) > (c.offsets[b + 1] - c.offsets[b]);
    });
    std::vector<int32_t> inv(n);  // old -> new
    for (int32_t v = 0; v < n; ++v) inv[perm[v]] = v;
    auto ng = std::make_unique<glp_graph>();
    HostCSR& d = ng->csr;
    d.n = n;
    d.m = c.m;
    d.offsets.assign((size_t)n + 1, 0);
    for (int32_t v = 0; v < n; ++v)
        d.offsets[v + 1] = d.offsets[v] + (c.offsets[perm[v] + 1] - c.offsets[perm[v]]);
    d.neighbors.resize(c.m);
    if (c.weighted()) d.weights.resize(c.m);
    const int threads = host_threads(0);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 4096)
    for (int32_t v = 0; v < n; ++v) {
        int32_t o = perm[v];
        int32_t b = c.offsets[o], e = c.offsets[o + 1], p = d.offsets[v];
        // neighbour lists stay ascending in the new ids when they were ascending
        // in the old ids only if the relabeling is monotone; keep them sorted
        // by new id so generated graphs keep the "neighbours ascending" form.
        std::vector<std::pair<int32_t, uint32_t>> tmp;
        tmp.reserve(e - b);
        for (int32_t j = b; j < e; ++j)
            tmp.emplace_back(inv[c.neighbors[j]], c.weighted() ? c.weights[j] : 0u);
        std::stable_sort(tmp.begin(), tmp.end(),
                         [](const std::pair<int32_t, uint32_t>& x,
                            const std::pair<int32_t, uint32_t>& y) { return x.first < y.first; });
        for (size_t k = 0; k < tmp.size(); ++k) {
            d.neighbors[p + k] = tmp[k].first;
            if (c.weighted()) d.weights[p + k] = tmp[k].second;
        }
    }
    d.finish();
    d.symmetric = c.symmetric;  // a relabelling keeps (a)symmetry
    if (perm_out) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * n);
    *out = ng.release();
    return GLP_OK;
}

// ---- tools ---------------------------------------------------------------

// tools/edgelist2bin.cpp:13-73.  Lines whose first character is '#' are
// skipped (:23); a line that does not hold two integers is skipped (the
// reference would reuse indeterminate values there).
extern "C" int glp_tool_edgelist2bin(const char* txt_path, const char* out_path) {
    if (!txt_path || !out_path) return set_error(GLP_E_ARG, "null argument");
    std::ifstream ifs(txt_path);
    if (!ifs) return set_error(GLP_E_IO, "cannot read '%s'", txt_path);
    std::vector<std::vector<int>> adj;
    std::set<int> vset;
    std::string s;
    while (std::getline(ifs, s)) {
        if (!s.empty() && s[0] == '#') continue;
        int u, v;
        if (sscanf(s.c_str(), "%d %d", &u, &v) != 2) continue;
        if (u == v) continue;
        if (u > v) std::swap(u, v);
        if (u < 0) return set_error(GLP_E_FORMAT, "negative id in '%s'", txt_path);
        vset.insert(u);
        vset.insert(v);
        if ((size_t)v >= adj.size()) adj.resize((size_t)v + 1);
        adj[u].push_back(v);
        adj[v].push_back(u);
    }
    std::vector<int> vlist(vset.begin(), vset.end());  // ascending
    std::map<int, int> map;
    for (size_t i = 0; i < vlist.size(); ++i) map[vlist[i]] = (int)i;
    std::ofstream ofs(out_path, std::ios::binary | std::ios::trunc);
    if (!ofs) return set_error(GLP_E_IO, "cannot write '%s'", out_path);
    for (size_t i = 0; i < vlist.size(); ++i) {
        int u = vlist[i];
        int x = map[u];
        for (int yv : adj[u]) {
            int y = map[yv];
            if (x >= y) continue;
            int32_t w[2] = {x, y};
            ofs.write(reinterpret_cast<const char*>(w), sizeof w);
        }
    }
    return ofs ? GLP_OK : set_error(GLP_E_IO, "write failed on '%s'", out_path);
}

namespace {
struct NodeDeg { int node; int degree; };
struct ByDegreeDesc {
    bool operator()(const NodeDeg& i, const NodeDeg& j) const { return i.degree > j.degree; }
};
}  // namespace

// tools/normalize.cpp:262-348.  Equal-degree vertices come out in the order
// std::sort leaves them with this comparator on the ascending-id input
// (:307-314); the same libstdc++ algorithm runs here.
extern "C" int glp_tool_normalize(const char* bin_path, const char* out_path) {
    if (!bin_path || !out_path) return set_error(GLP_E_ARG, "null argument");
    std::vector<char> buf;
    if (!read_file(bin_path, buf)) return set_error(GLP_E_IO, "cannot read '%s'", bin_path);
    const int32_t* rec = reinterpret_cast<const int32_t*>(buf.data());
    size_t nrec = buf.size() / 8;
    std::vector<std::vector<int>> adj;
    std::set<int> vset;
    int max_id = 0;
    for (size_t i = 0; i < nrec; ++i) {
        if (rec[2 * i] == '#') continue;  // :274
        if (rec[2 * i] == INT32_MIN || rec[2 * i + 1] == INT32_MIN)
            return set_error(GLP_E_FORMAT, "id INT32_MIN has no absolute value");
        int u = std::abs(rec[2 * i]), v = std::abs(rec[2 * i + 1]);
        vset.insert(u);
        vset.insert(v);
        max_id = std::max(max_id, std::max(u, v));
        if (u == v) continue;
        if (u > v) std::swap(u, v);
        if ((size_t)v >= adj.size()) adj.resize((size_t)v + 1);
        adj[u].push_back(v);
        adj[v].push_back(u);
    }
    // a vertex seen only in self-loops may lie beyond adj's end; it has degree 0
    if (!vset.empty() && (size_t)max_id >= adj.size()) adj.resize((size_t)max_id + 1);
    std::vector<int> vlist(vset.begin(), vset.end());
    std::vector<NodeDeg> nd(vlist.size());
    for (size_t i = 0; i < vlist.size(); ++i) {
        nd[i].node = vlist[i];
        nd[i].degree = (int)adj[vlist[i]].size();
    }
    std::sort(nd.begin(), nd.end(), ByDegreeDesc());
    for (size_t i = 0; i < nd.size(); ++i) vlist[i] = nd[i].node;
    std::map<int, int> map;
    for (size_t i = 0; i < vlist.size(); ++i) map[vlist[i]] = (int)i;
    std::ofstream ofs(out_path, std::ios::binary | std::ios::trunc);
    if (!ofs) return set_error(GLP_E_IO, "cannot write '%s'", out_path);
    for (size_t i = 0; i < vlist.size(); ++i) {
        int u = vlist[i];
        int x = map[u];
        for (int yv : adj[u]) {
            int y = map[yv];
            if (x >= y) continue;
            int32_t w[2] = {x, y};
            ofs.write(reinterpret_cast<const char*>(w), sizeof w);
        }
    }
    return ofs ? GLP_OK : set_error(GLP_E_IO, "write failed on '%s'", out_path);
}

// tools/direct_edgelist2bin.cpp:113-205 (buildfrombin, the path its main
// runs): reads .bin pairs (u,v), drops self-loops, keeps an edge only when
// u or v is below id_cutoff (the reference hard-codes 5013066 at :137;
// id_cutoff <= 0 keeps every edge), keeps u -> v only (directed, duplicates
// kept), orders the vertices seen by non-increasing out-degree with the same
// unstable std::sort on the ascending-id list, and writes each vertex's
// pairs (x, y) with its neighbours in ascending original id.
extern "C" int glp_tool_direct_edgelist2bin(const char* bin_path, const char* out_path,
                                           int32_t id_cutoff) {
    if (!bin_path || !out_path) return set_error(GLP_E_ARG, "null argument");
    std::vector<char> buf;
    if (!read_file(bin_path, buf)) return set_error(GLP_E_IO, "cannot read '%s'", bin_path);
    const int32_t* rec = reinterpret_cast<const int32_t*>(buf.data());
    const size_t nrec = buf.size() / 8;
    std::vector<std::vector<int>> adj;
    std::set<int> vset;
    for (size_t i = 0; i < nrec; ++i) {
        const int u = rec[2 * i], v = rec[2 * i + 1];
        if (u == v) continue;
        if (id_cutoff > 0 && !(u < id_cutoff || v < id_cutoff)) continue;
        if (u < 0 || v < 0) return set_error(GLP_E_FORMAT, "negative vertex id in '%s'", bin_path);
        vset.insert(u);
        vset.insert(v);
        const size_t top = (size_t)std::max(u, v);
        if (top >= adj.size()) adj.resize(top + 1);
        adj[u].push_back(v);
    }
    std::vector<int> vlist(vset.begin(), vset.end());
    std::vector<NodeDeg> nd(vlist.size());
    for (size_t i = 0; i < vlist.size(); ++i) {
        nd[i].node = vlist[i];
        nd[i].degree = (int)adj[vlist[i]].size();
    }
    std::sort(nd.begin(), nd.end(), ByDegreeDesc());
    for (size_t i = 0; i < nd.size(); ++i) vlist[i] = nd[i].node;
    std::map<int, int> map;
    for (size_t i = 0; i < vlist.size(); ++i) map[vlist[i]] = (int)i;
    std::ofstream ofs(out_path, std::ios::binary | std::ios::trunc);
    if (!ofs) return set_error(GLP_E_IO, "cannot write '%s'", out_path);
    for (size_t i = 0; i < vlist.size(); ++i) {
        std::vector<int>& nb = adj[vlist[i]];
        std::sort(nb.begin(), nb.end());
        const int x = map[vlist[i]];
        for (int yv : nb) {
            const int32_t w[2] = {x, map[yv]};
            ofs.write(reinterpret_cast<const char*>(w), sizeof w);
        }
    }
    return ofs ? GLP_OK : set_error(GLP_E_IO, "write failed on '%s'", out_path);
}
