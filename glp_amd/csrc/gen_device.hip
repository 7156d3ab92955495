// Device-side R-MAT generator: the graph glp_gen_rmat builds on the host
// (degree_sort = 0), built directly in HBM.
//
// Draw i of the host generator (gen.cpp, RmatDraw) is a pure function of
// (seed, i), so a thread per draw reproduces it exactly.  Each draw with
// u != v yields the directed pairs (u, v) and (v, u); a self-loop yields two
// sentinel keys instead.  The 2E keys (u << scale | v) are radix-sorted, the
// duplicates dropped, and the sorted unique keys are the CSR: neighbours
// ascending per vertex, exactly the host's sort-and-unique per vertex.
// Weights are the host's hash of the unordered pair, so w(u,v) = w(v,u).
// The offsets come from the positions where the sorted keys' source
// changes.  The sort and the unique are rocPRIM device primitives (bench and
// serving infrastructure; the LP step itself uses none).

#include "device.h"

#include <cmath>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

namespace glp {
namespace {

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
// gen.cpp derive(seed, a, b)
__device__ __forceinline__ uint64_t derive_dev(uint64_t seed, uint64_t a, uint64_t b) {
    return mix64(seed ^ mix64(a * kGolden + b + 1));
}

struct RmatArgs {
    uint64_t seed, tA, tAB, tABC, draws;
    int32_t scale;
};

// Draw i: per level one 32-bit word (two levels per SplitMix64 output) picks
// the quadrant by integer thresholds (gen.cpp RmatDraw).
__global__ __launch_bounds__(256) void k_rmat_draw(RmatArgs r, unsigned long long* __restrict__ keys) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.draws) return;
    uint64_t st = derive_dev(r.seed, i, 0);
    uint32_t uu = 0, vv = 0;
    uint64_t word = 0;
    for (int lvl = 0; lvl < r.scale; ++lvl) {
        uint32_t x;
        if ((lvl & 1) == 0) {
            st += kGolden;
            word = mix64(st);
            x = (uint32_t)word;
        } else {
            x = (uint32_t)(word >> 32);
        }
        uint32_t ub, vb;
        if (x < r.tA) { ub = 0; vb = 0; }
        else if (x < r.tAB) { ub = 0; vb = 1; }
        else if (x < r.tABC) { ub = 1; vb = 0; }
        else { ub = 1; vb = 1; }
        uu = (uu << 1) | ub;
        vv = (vv << 1) | vb;
    }
    const unsigned long long sentinel = 1ull << (2 * r.scale);
    if (uu == vv) {
        keys[2 * i] = sentinel;
        keys[2 * i + 1] = sentinel;
    } else {
        keys[2 * i] = ((unsigned long long)uu << r.scale) | vv;
        keys[2 * i + 1] = ((unsigned long long)vv << r.scale) | uu;
    }
}

// Degree of each source over the m unique keys (the sentinel sorts last and
// is excluded from m).
// Offsets straight from the sorted keys: off[v] is the first position whose
// source is >= v, so position i (0 <= i <= m) writes off[v] for the
// sources src(i-1) < v <= src(i), with src(-1) = -1 and src(m) = n.  Every
// entry is written once, and a hub costs nothing beyond its own positions
// (a degree histogram's atomics serialise on a hub's counter).
__global__ __launch_bounds__(256) void k_key_offsets(const unsigned long long* __restrict__ keys, int64_t m,
                                                     int32_t scale, int32_t n, int32_t* __restrict__ off) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > m) return;
    const int64_t hi = i < m ? (int64_t)(keys[i] >> scale) : (int64_t)n;
    const int64_t lo = i > 0 ? (int64_t)(keys[i - 1] >> scale) : -1;
    for (int64_t v = lo + 1; v <= hi; ++v) off[v] = (int32_t)i;
}

// Neighbour ids, and weights 1 + H(min, max) % wmax (gen.cpp).
__global__ __launch_bounds__(256) void k_key_csr(const unsigned long long* __restrict__ keys, int64_t m,
                                                 int32_t scale, uint64_t wseed, uint32_t wmax,
                                                 int32_t* __restrict__ nbr, uint32_t* __restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const unsigned long long k = keys[i];
    const uint32_t u = (uint32_t)(k >> scale);
    const uint32_t v = (uint32_t)(k & ((1ull << scale) - 1));
    nbr[i] = (int32_t)v;
    if (w) {
        const uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
        w[i] = 1u + (uint32_t)(derive_dev(wseed, (hi << 32) | lo, 0) % (uint64_t)wmax);
    }
}

__global__ __launch_bounds__(256) void k_max_degree(const int32_t* __restrict__ off, int32_t n,
                                                    int32_t* __restrict__ out) {
    int32_t best = 0;
    for (int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
        best = max(best, off[v + 1] - off[v]);
    best = max(best, __shfl_xor(best, 32));
    best = max(best, __shfl_xor(best, 16));
    best = max(best, __shfl_xor(best, 8));
    best = max(best, __shfl_xor(best, 4));
    best = max(best, __shfl_xor(best, 2));
    best = max(best, __shfl_xor(best, 1));
    if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

unsigned grid_of(int64_t count) { return (unsigned)((count + 255) / 256); }

// Device buffers released on every exit path.
struct Scratch {
    void* p[4] = {nullptr, nullptr, nullptr, nullptr};
    ~Scratch() {
        for (void* q : p) (void)hipFree(q);
    }
};

}  // namespace

int gen_rmat_device(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                    int32_t wmax, hipStream_t s, DeviceCSR* out) {
    *out = DeviceCSR();
    if (scale < 1 || scale > 30 || edge_factor < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0 || wmax < 0)
        return set_error(GLP_E_ARG, "bad rmat arguments");
    const int32_t n = int32_t(1) << scale;
    RmatArgs r;
    r.seed = seed;
    r.scale = scale;
    r.draws = (uint64_t)edge_factor << scale;
    const double two32 = 4294967296.0;
    r.tA = (uint64_t)std::floor(a * two32);
    r.tAB = (uint64_t)std::floor((a + b) * two32);
    r.tABC = (uint64_t)std::floor((a + b + c) * two32);
    const int64_t nkeys = 2 * (int64_t)r.draws;
    if (nkeys > (int64_t)UINT32_MAX) return set_error(GLP_E_RANGE, "too many draws for a device build");
    Scratch sc;
    unsigned long long *keys = nullptr, *sorted = nullptr;
    GLP_HIP(hipMalloc(&sc.p[0], (size_t)nkeys * 8));
    GLP_HIP(hipMalloc(&sc.p[1], (size_t)nkeys * 8));
    keys = static_cast<unsigned long long*>(sc.p[0]);
    sorted = static_cast<unsigned long long*>(sc.p[1]);
    hipLaunchKernelGGL(k_rmat_draw, dim3(grid_of((int64_t)r.draws)), dim3(256), 0, s, r, keys);
    GLP_HIP(hipGetLastError());
    // sort the used bits only: keys < 2^(2 scale), the sentinel is bit 2 scale
    const unsigned end_bit = (unsigned)(2 * scale + 1);
    size_t tmp_bytes = 0, tmp2 = 0;
    GLP_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, keys, sorted, (size_t)nkeys, 0u, end_bit, s));
    uint32_t* d_count = nullptr;
    GLP_HIP(rocprim::unique(nullptr, tmp2, sorted, keys, d_count, (size_t)nkeys,
                            rocprim::equal_to<unsigned long long>(), s));
    tmp_bytes = (std::max(tmp_bytes, tmp2) + 255) & ~(size_t)255;
    GLP_HIP(hipMalloc(&sc.p[2], tmp_bytes + 256));
    d_count = reinterpret_cast<uint32_t*>(static_cast<char*>(sc.p[2]) + tmp_bytes);
    GLP_HIP(rocprim::radix_sort_keys(sc.p[2], tmp_bytes, keys, sorted, (size_t)nkeys, 0u, end_bit, s));
    size_t t2 = tmp_bytes;
    GLP_HIP(rocprim::unique(sc.p[2], t2, sorted, keys, d_count, (size_t)nkeys,
                            rocprim::equal_to<unsigned long long>(), s));
    uint32_t nuniq = 0;
    unsigned long long last = 0;
    GLP_HIP(hipMemcpyAsync(&nuniq, d_count, 4, hipMemcpyDeviceToHost, s));
    GLP_HIP(hipStreamSynchronize(s));
    if (nuniq > 0) {
        GLP_HIP(hipMemcpyAsync(&last, keys + (nuniq - 1), 8, hipMemcpyDeviceToHost, s));
        GLP_HIP(hipStreamSynchronize(s));
    }
    const int64_t m = (int64_t)nuniq - (nuniq > 0 && last == (1ull << (2 * scale)) ? 1 : 0);
    if (m > INT32_MAX) return set_error(GLP_E_RANGE, "edge count exceeds int32 CSR offsets");
    DeviceCSR o;
    o.n = n;
    o.m = m;
    GLP_HIP(hipMalloc(&o.off, ((size_t)n + 1) * 4));
    DeviceCSR* guard = &o;
    struct Release {
        DeviceCSR** g;
        ~Release() {
            if (*g) {
                (void)hipFree((*g)->off);
                (void)hipFree((*g)->nbr);
                (void)hipFree((*g)->w);
            }
        }
    } rel{&guard};
    hipLaunchKernelGGL(k_key_offsets, dim3(grid_of(m + 1)), dim3(256), 0, s, keys, m, scale, n, o.off);
    GLP_HIP(hipGetLastError());
    GLP_HIP(hipMalloc(&o.nbr, std::max<size_t>((size_t)m, 1) * 4));
    if (wmax > 0) GLP_HIP(hipMalloc(&o.w, std::max<size_t>((size_t)m, 1) * 4));
    if (m > 0) {
        hipLaunchKernelGGL(k_key_csr, dim3(grid_of(m)), dim3(256), 0, s, keys, m, scale,
                           seed ^ 0x5745494748545321ULL, (uint32_t)std::max(wmax, 1), o.nbr, o.w);
        GLP_HIP(hipGetLastError());
    }
    int32_t* d_max = reinterpret_cast<int32_t*>(d_count);  // the unique count is read: reuse its word
    GLP_HIP(hipMemsetAsync(d_max, 0, 4, s));
    hipLaunchKernelGGL(k_max_degree, dim3(std::min<unsigned>(grid_of(n), 1024)), dim3(256), 0, s, o.off, n,
                       d_max);
    GLP_HIP(hipGetLastError());
    GLP_HIP(hipMemcpyAsync(&o.max_degree, d_max, 4, hipMemcpyDeviceToHost, s));
    GLP_HIP(hipStreamSynchronize(s));
    o.wmax = (uint32_t)wmax;
    *out = o;
    guard = nullptr;
    return GLP_OK;
}

}  // namespace glp
