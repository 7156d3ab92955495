// Host tier planner.
//
// Replaces the reference's planning in prop/classicLprop.cuh:196-332:
// binary_search tier boundaries (:129-158, which assume ids sorted by
// non-increasing degree), the tiny-vertex warp packing (:264-308) and the hub
// block assignment (compute_num_blocks / exclusive_scan / assign_blocks,
// header/kernel.cuh:6011-6046).  Here every vertex is classified by its own
// degree, so no id order is assumed, and low-degree vertices are packed into
// wavefront groups along runs of consecutive ids (their CSR edges are then
// contiguous).

#include "device.h"
#include "device_prims.h"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace glp {

namespace {
struct TuningField {
    const char* name;
    int32_t Tuning::*field;
};
const TuningField kFields[] = {
    {"pack_max_deg", &Tuning::pack_max_deg}, {"half_max_deg", &Tuning::half_max_deg},
    {"half_lanes", &Tuning::half_lanes},
    {"wave_max_deg", &Tuning::wave_max_deg},
    {"block_max_deg", &Tuning::block_max_deg}, {"hub_chunk", &Tuning::hub_chunk},
    {"part_keys", &Tuning::part_keys},       {"part_limit", &Tuning::part_limit},
    {"block_threads", &Tuning::block_threads}, {"fused_small", &Tuning::fused_small},
};
}  // namespace

int tuning_set(Tuning& t, const char* key, int32_t value) {
    for (const auto& f : kFields)
        if (key && std::string(key) == f.name) {
            t.*(f.field) = value;
            return GLP_OK;
        }
    return set_error(GLP_E_ARG, "unknown tuning key '%s'", key ? key : "(null)");
}

const char* tuning_key(int i) {
    return i >= 0 && i < (int)(sizeof(kFields) / sizeof(kFields[0])) ? kFields[i].name : nullptr;
}

int tuning_get(const Tuning& t, const char* key, int32_t* value) {
    for (const auto& f : kFields)
        if (key && std::string(key) == f.name) {
            *value = t.*(f.field);
            return GLP_OK;
        }
    return set_error(GLP_E_ARG, "unknown tuning key '%s'", key ? key : "(null)");
}

Tuning& default_tuning() {
    static Tuning t;
    static bool loaded = false;
    if (!loaded) {
        loaded = true;
        // key=value lines written by the autotuner
        const char* path = getenv("GLP_TUNING");
        if (path) {
            std::ifstream in(path);
            std::string line;
            while (std::getline(in, line)) {
                std::istringstream ss(line);
                std::string k;
                int v;
                if (std::getline(ss, k, '=') && (ss >> v)) (void)tuning_set(t, k.c_str(), v);
            }
        }
    }
    return t;
}

static int check_tuning(const Tuning& t) {
    if (t.pack_max_deg < 0 || t.pack_max_deg > 64) return set_error(GLP_E_ARG, "pack_max_deg must be in [0,64]");
    if (t.half_max_deg < 0 || t.half_max_deg > kHalfMaxDeg)
        return set_error(GLP_E_ARG, "half_max_deg must be in [0, %d]", kHalfMaxDeg);
    if (t.half_lanes != 0 && t.half_lanes != 16 && t.half_lanes != 32)
        return set_error(GLP_E_ARG, "half_lanes must be 0 (chosen per launch), 16 or 32");
    if (t.wave_max_deg < t.pack_max_deg || t.wave_max_deg > kWaveCap / 2)
        return set_error(GLP_E_ARG, "wave_max_deg must be in [pack_max_deg, %d]", kWaveCap / 2);
    // the largest LDS table (2*deg slots of 8 bytes) must fit a CU's 160 KB
    if (t.block_max_deg < t.wave_max_deg || t.block_max_deg > 8192)
        return set_error(GLP_E_ARG, "block_max_deg must be in [wave_max_deg, 8192]");
    // LDS tables: next_pow2(2*hub_chunk) and next_pow2(2*part_keys) slots of 8 bytes
    if (t.hub_chunk < 256 || t.hub_chunk > 8192 || (t.hub_chunk & 255))
        return set_error(GLP_E_ARG, "hub_chunk must be a multiple of 256 in [256, 8192]");
    if (t.part_keys < 256 || t.part_keys > 4096)
        return set_error(GLP_E_ARG, "part_keys must be in [256, 4096]");
    // a pass stops after `limit` distinct labels (+ at most a workgroup's
    // worth in flight): limit <= part_keys keeps that inside the table
    if (t.part_limit < 0 || t.part_limit > t.part_keys)
        return set_error(GLP_E_ARG, "part_limit must be in [0, part_keys]");
    if (t.block_threads != 256 && t.block_threads != 512 && t.block_threads != 1024)
        return set_error(GLP_E_ARG, "block_threads must be 256, 512 or 1024");
    if (t.fused_small != 0 && t.fused_small != 1) return set_error(GLP_E_ARG, "fused_small must be 0 or 1");
    return GLP_OK;
}

int build_plan(const HostCSR& g, const Tuning& t, Plan& p) {
    int rc = check_tuning(t);
    if (rc != GLP_OK) return rc;
    p = Plan();
    p.tune = t;
    // block classes: LDS table capacity 2*deg rounded up to a power of two
    {
        uint32_t lo = dev::next_pow2(2u * (uint32_t)std::max(t.wave_max_deg + 1, 1));
        uint32_t hi = dev::next_pow2(2u * (uint32_t)std::max(t.block_max_deg, 1));
        lo = std::max(lo, 64u);
        int c = 0;
        // classes double in size; the last absorbs everything up to hi
        uint32_t cap = lo;
        while (c < kMaxBlockClasses) {
            p.class_cap[c++] = (int32_t)cap;
            if (cap >= hi) break;
            cap = (c == kMaxBlockClasses - 1) ? hi : cap * 2;
        }
        p.nclasses = t.block_max_deg > t.wave_max_deg ? c : 0;
    }
    const int32_t n = g.n;
    int32_t gs = -1, gc = 0, ge = 0;  // open packed group: start, vertex count, edges
    auto close_group = [&]() {
        if (gs >= 0) p.groups.push_back(make_int2(gs, gc));
        gs = -1;
        gc = 0;
        ge = 0;
    };
    int64_t hub_chunks = 0;
    for (int32_t v = 0; v < n; ++v) {
        const int32_t d = g.offsets[v + 1] - g.offsets[v];
        if (d <= t.pack_max_deg) {
            if (gs >= 0 && (gc == 64 || ge + d > 64)) close_group();
            if (gs < 0) gs = v;
            ++gc;
            ge += d;
            p.edges[0] += d;
            p.verts[0]++;
            continue;
        }
        close_group();
        if (d <= t.wave_max_deg) {
            (d <= t.half_max_deg ? p.half : p.wave)
                .push_back(make_int4(v, g.offsets[v], g.offsets[v + 1], 0));
            p.edges[1] += d;
            p.verts[1]++;
        } else if (d <= t.block_max_deg) {
            const uint32_t need = std::max(dev::next_pow2(2u * (uint32_t)d), 64u);
            int c = 0;
            while (c < p.nclasses - 1 && (uint32_t)p.class_cap[c] < need) ++c;
            p.block[c].push_back(v);
            p.edges[2] += d;
            p.verts[2]++;
        } else {
            const int32_t h = (int32_t)p.hubs.size();
            p.hubs.push_back(v);
            const int32_t nch = (d + t.hub_chunk - 1) / t.hub_chunk;
            p.hub_chunk0.push_back((int32_t)p.hub_chunks.size());
            for (int32_t c = 0; c < nch; ++c) p.hub_chunks.push_back(make_int2(h, c));
            hub_chunks += nch;
            p.hub_base.push_back(p.pair_slots);
            p.pair_slots += (int64_t)nch * t.hub_chunk;
            // one counting workgroup per bucket; empty buckets exit at once
            const int32_t nb = (int32_t)hub_buckets(d, t.part_keys);
            for (int32_t q = 0; q < nb; ++q) p.hub_parts.push_back(make_int2(h, q));
            p.hub_max_buckets = std::max(p.hub_max_buckets, nb);
            p.hub_hist_base.push_back(p.hist_slots);
            p.hist_slots += (int64_t)nb * nch;
            p.edges[3] += d;
            p.verts[3]++;
        }
    }
    close_group();
    // groups holding only isolated vertices go last: classic LP skips them
    // (such a vertex keeps its label and both label buffers already hold it)
    {
        std::vector<int2> with_edges, isolated;
        for (const int2& gr : p.groups) {
            const bool any = g.offsets[gr.x + gr.y] > g.offsets[gr.x];
            (any ? with_edges : isolated).push_back(gr);
        }
        p.groups_with_edges = (int64_t)with_edges.size();
        p.groups = std::move(with_edges);
        p.groups.insert(p.groups.end(), isolated.begin(), isolated.end());
    }
    // half-tier vertices of degree <= 32 first: with 16 lanes per vertex
    // they load two edges per lane instead of four
    {
        std::vector<int4> shrt, rest;
        for (const int4& ent : p.half) (ent.z - ent.y <= 32 ? shrt : rest).push_back(ent);
        p.half_short = (int64_t)shrt.size();
        p.half = std::move(shrt);
        p.half.insert(p.half.end(), rest.begin(), rest.end());
    }
    // wave-tier vertices of degree <= 128 first: their waves load two edges
    // per lane instead of four
    {
        std::vector<int4> shrt, rest;
        for (const int4& ent : p.wave) (ent.z - ent.y <= 128 ? shrt : rest).push_back(ent);
        p.wave_short = (int64_t)shrt.size();
        p.wave = std::move(shrt);
        p.wave.insert(p.wave.end(), rest.begin(), rest.end());
    }
    if (hub_chunks > INT32_MAX) return set_error(GLP_E_RANGE, "too many hub chunks");
    return GLP_OK;
}

}  // namespace glp

extern "C" int glp_tuning_set(const char* key, int32_t value) {
    glp::clear_error();
    return glp::tuning_set(glp::default_tuning(), key, value);
}

extern "C" int glp_tuning_get(const char* key, int32_t* value) {
    glp::clear_error();
    if (!value) return glp::set_error(GLP_E_ARG, "null value");
    return glp::tuning_get(glp::default_tuning(), key, value);
}
