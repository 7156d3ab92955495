// glp_run_noalloc: glp_run for callers that cannot allocate device memory.
//
// Everything the run needs on the device -- the CSR, the tier plan, both
// label buffers, the layered volumes, the hub tier's scratch and the change
// counters -- is laid out in `scratch`, a file-scope device array of
// kNoAllocBytes that each GPU holds from the library's load until the
// process ends.  Host code never uses the array's name as a device address:
// the address is taken once per device with hipGetSymbolAddress and every
// launch and copy receives that pointer.  The call makes no hipMalloc,
// hipFree, hipMemset or stream / event creation: copies go through the null
// stream, and everything that must start at zero is cleared by fill kernels
// (the array keeps the previous call's contents).  Calls in this process
// share each device's array, so they are serialised by one lock.

#include "device.h"
#include "session.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace glp {
namespace {

constexpr size_t kNoAllocBytes = (size_t)GLP_NOALLOC_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_noalloc_mu;              // serialises calls (they share the arrays)
void* g_scratch_ptr[kMaxDevices] = {};  // per device: the address of `scratch`

// Bump layout inside the array; with base == nullptr it only sizes.
struct Arena {
    char* base = nullptr;
    size_t at = 0;
    template <class T>
    T* take(size_t count) {
        const size_t here = (at + 255) & ~(size_t)255;
        at = here + count * sizeof(T);
        return base ? reinterpret_cast<T*>(base + here) : nullptr;
    }
};

struct Layout {
    DeviceGraph* dg;  // CSR and plan arrays set
    int32_t* L[2];
    int32_t* V[2];
    HubScratch hs;
    unsigned long long* changes;
};

// Places every device array of the run; returns the bytes used.
size_t place(Arena& ar, const HostCSR& g, const Plan& p, bool layered, Layout* lay) {
    DeviceGraph* dg = lay->dg;
    const size_t n = std::max<size_t>((size_t)g.n, 1);
    dg->off = ar.take<int32_t>((size_t)g.n + 1);
    dg->nbr = ar.take<int32_t>(std::max<size_t>((size_t)g.m, 1));
    dg->w = g.weighted() ? ar.take<uint32_t>(std::max<size_t>((size_t)g.m, 1)) : nullptr;
    dg->d_groups = ar.take<int2>(p.groups.size());
    dg->d_half = ar.take<int4>(p.half.size());
    dg->d_wave = ar.take<int4>(p.wave.size());
    for (int c = 0; c < kMaxBlockClasses; ++c) dg->d_block[c] = ar.take<int32_t>(p.block[c].size());
    dg->d_hubs = ar.take<int32_t>(p.hubs.size());
    dg->d_hub_chunks = ar.take<int2>(p.hub_chunks.size());
    dg->d_hub_chunk0 = ar.take<int32_t>(p.hub_chunk0.size());
    dg->d_hub_base = ar.take<int64_t>(p.hub_base.size());
    dg->d_hub_parts = ar.take<int2>(p.hub_parts.size());
    dg->d_hub_hist_base = ar.take<int64_t>(p.hub_hist_base.size());
    lay->L[0] = ar.take<int32_t>(n);
    lay->L[1] = ar.take<int32_t>(n);
    lay->V[0] = layered ? ar.take<int32_t>(n) : nullptr;
    lay->V[1] = layered ? ar.take<int32_t>(n) : nullptr;
    lay->hs = HubScratch();
    if (!p.hubs.empty()) {
        lay->hs.pairs = ar.take<unsigned long long>((size_t)p.pair_slots);
        lay->hs.hist = ar.take<uint32_t>((size_t)p.hist_slots);
        lay->hs.lstart = ar.take<uint32_t>((size_t)p.hist_slots);
        lay->hs.best = ar.take<unsigned long long>(p.hubs.size());
    }
    lay->hs.work = ar.take<uint32_t>(1);
    lay->changes = ar.take<unsigned long long>(kChangeSlots);
    return ar.at;
}

template <class T>
int put(T* dst, const std::vector<T>& src, hipStream_t s) {
    if (src.empty()) return GLP_OK;
    GLP_HIP(hipMemcpyAsync(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, s));
    return GLP_OK;
}

int run_noalloc(const glp_graph* g, const glp_params* prm, int32_t steps, int32_t* labels_out,
                int32_t* step_dump) {
    const HostCSR& c = g->csr;
    int rc = check_params(c, prm);
    if (rc != GLP_OK) return rc;
    if (prm->method != GLP_CLASSIC && prm->method != GLP_LAYERED)
        return set_error(GLP_E_UNSUPPORTED, "the allocation-free run takes classic and layered LP");
    if (prm->gather || prm->frontier)
        return set_error(GLP_E_UNSUPPORTED, "the allocation-free run fuses the gather and has no frontier");
    const bool layered = prm->method == GLP_LAYERED;
    Plan plan;
    if ((rc = build_plan(c, default_tuning(), plan)) != GLP_OK) return rc;
    DeviceGraph dg;
    dg.owns_csr = false;
    dg.owns_plan = false;
    dg.n = c.n;
    dg.m = c.m;
    Layout lay{};
    lay.dg = &dg;
    Arena sizing;
    const size_t need = place(sizing, c, plan, layered, &lay);
    if (need > kNoAllocBytes)
        return set_error(GLP_E_RANGE, "graph, plan and state need %zu bytes; the allocation-free run holds %zu",
                         need, kNoAllocBytes);
    int dev = 0;
    GLP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(GLP_E_UNSUPPORTED, "device %d", dev);
    if (!g_scratch_ptr[dev]) {
        // once per device, with that device current: the array's address
        GLP_HIP(hipGetSymbolAddress(&g_scratch_ptr[dev], HIP_SYMBOL(scratch)));
    }
    Arena ar;
    ar.base = static_cast<char*>(g_scratch_ptr[dev]);
    place(ar, c, plan, layered, &lay);
    dg.plan = std::move(plan);
    const Plan& p = dg.plan;
    const hipStream_t s = nullptr;  // the null stream: no stream is created
    if ((rc = put(dg.off, c.offsets, s)) || (rc = put(dg.nbr, c.neighbors, s)) ||
        (dg.w && (rc = put(dg.w, c.weights, s))) || (rc = put(dg.d_groups, p.groups, s)) ||
        (rc = put(dg.d_half, p.half, s)) || (rc = put(dg.d_wave, p.wave, s)) || (rc = put(dg.d_hubs, p.hubs, s)) ||
        (rc = put(dg.d_hub_chunks, p.hub_chunks, s)) || (rc = put(dg.d_hub_chunk0, p.hub_chunk0, s)) ||
        (rc = put(dg.d_hub_base, p.hub_base, s)) || (rc = put(dg.d_hub_parts, p.hub_parts, s)) ||
        (rc = put(dg.d_hub_hist_base, p.hub_hist_base, s)))
        return rc;
    for (int k = 0; k < p.nclasses; ++k)
        if ((rc = put(dg.d_block[k], p.block[k], s))) return rc;
    // the array holds the previous call's data: clear what starts at zero
    const int64_t n = c.n;
    if ((rc = launch_iota(lay.L[0], n, s)) || (rc = launch_iota(lay.L[1], n, s)) ||
        (rc = launch_zero_bytes(lay.changes, 8 * kChangeSlots, s)))
        return rc;
    if (lay.hs.best && (rc = launch_zero_bytes(lay.hs.best, (int64_t)p.hubs.size() * 8, s))) return rc;
    if (layered && (rc = launch_fill_i32(lay.V[0], 1, n, s))) return rc;  // V_0(l) = 1
    const Method method = layered ? Method::Layered : Method::Classic;
    int cur = 0;
    for (int32_t t = 0; t < steps; ++t) {
        StepArgs a;
        std::memset(&a, 0, sizeof a);
        a.off = dg.off;
        a.nbr = dg.nbr;
        a.w = dg.w;
        a.n = c.n;
        a.t = t + 1;
        a.changes = lay.changes;
        a.L = lay.L[cur];
        a.keep = lay.L[cur];
        a.Lw = lay.L[cur ^ 1];
        if (layered) {
            a.Vr = lay.V[cur];
            a.Vw = lay.V[cur ^ 1];
            a.score_mul = prm->gamma_den + prm->gamma_num;
            a.score_vmul = prm->gamma_num;
            if ((rc = launch_fill_i32(a.Vw, 0, n, s))) return rc;
        }
        if ((rc = launch_step(method, prm->weighted != 0, false, dg, lay.hs, a, s, nullptr))) return rc;
        if (layered && (rc = launch_volume(a.Lw, a.Vw, c.n, s))) return rc;
        cur ^= 1;
        if (step_dump && n)
            GLP_HIP(hipMemcpyAsync(step_dump + (size_t)t * n, lay.L[cur], sizeof(int32_t) * n,
                                   hipMemcpyDeviceToHost, s));
    }
    if (labels_out && n)
        GLP_HIP(hipMemcpyAsync(labels_out, lay.L[cur], sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
    GLP_HIP(hipStreamSynchronize(s));
    return GLP_OK;
}

}  // namespace
}  // namespace glp

using namespace glp;

extern "C" int glp_run_noalloc(const glp_graph* g, const glp_params* p, int32_t steps, int32_t* labels_out,
                               int32_t* step_dump) {
    clear_error();
    if (!g || !p || steps < 0 || (g->csr.n > 0 && !labels_out))
        return set_error(GLP_E_ARG, "bad arguments");
    int rc = ensure_device();
    if (rc != GLP_OK) return rc;
    init_kernels();
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return run_noalloc(g, p, steps, labels_out, step_dump);
}
