// Batches of uneven images: many gaussian_blur_f32, filter_clamped (u8 image,
// f32 kernel, u8 output) and match_template (GrayImage) jobs planned together
// and replayed as one hipGraph. The reference has no batch API (a batch there
// is a loop of per-image calls); each job computes exactly what the
// corresponding single call computes.
//
// Plan: the blur jobs become device job tables of the persistent fixed-point
// blur kernel (one launch per kernel variant, each walking the tiles of all
// its images), the filter jobs one table of the ragged 2-D filter kernel (one
// launch), the match jobs device tables of the fused int8 match kernel (one
// launch per K-block count and method, each walking the tiles of all its
// pairs); pairs that kernel does not take (templates above 128 px, images
// beyond its 32-bit offsets) run the direct kernel, one launch each. The
// blurs run on the caller's stream, the filters and the match tables on side
// streams forked from it (the remaining pairs spread over six more, largest
// first onto the least loaded) and joined back, captured once into a graph.
// The autotuner times the candidate launch shapes of each kind on the real
// jobs (filter tile shape, rows of 32 outputs per match workgroup for each
// match table, persistent blur workgroups per CU), keeps the fastest before
// capture, and remembers each match table's choice for its shape class.
#include "iph_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

namespace {

struct MatchSpec {
    const uint8_t *img;
    uint32_t w, h;
    const uint8_t *tpl;
    uint32_t tw, th;
    int32_t method;
    float *out;
    size_t ws_off, ws_bytes;
};

struct FallbackBlur {
    int spec;
    size_t ws_off, ws_bytes;
};

// Match jobs of one fused-kernel launch (one K-block count and method),
// consecutive in the job table.
struct MatchGroup {
    int kb, method, first, count, rt, rt_max;
    int waves;  // waves per CU the chunk height aims for (12 or 8)
    int tc;     // template rows per chunk, at most
    long long ntiles;
    size_t lds;
    uint32_t max_tw, max_th;
    double outputs;  // map pixels of all the group's matches
};

}  // namespace

struct iph_batch {
    std::vector<iph::BlurSpec> blur;
    std::vector<iph::FilterSpec> filt;
    std::vector<MatchSpec> match;
    std::vector<iph::MatchJob> mjobs;     // fused-kernel matches, grouped (host copy)
    std::vector<MatchGroup> mgroups;
    std::vector<int> match_fallback;      // matches on the direct path
    size_t mjob_table_off = 0;

    bool planned = false;
    int device = -1;
    char *mem = nullptr;
    size_t mem_bytes = 0;
    size_t blur_table_off = 0, filt_table_off = 0;
    std::vector<iph::BlurGroup> blur_groups;
    std::vector<FallbackBlur> blur_fallback;
    int filt_tiles = 0;
    uint32_t filt_max_kw = 1, filt_max_kh = 1;
    int blur_grid_pct = 100, filt_variant = 0;
    // side[0]: the filters; side[1]: the match tables; side[2 .. kSide - 1]: the
    // matches on the direct path
    static constexpr int kSide = 8;
    hipStream_t side[kSide] = {};
    hipStream_t capture = nullptr;  // the graph is captured here (the caller's may be the null stream)
    hipEvent_t fork = nullptr, join[kSide] = {};
    std::vector<int> match_lane;  // side stream (2 ..) of each direct-path match
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string tuning;
};

namespace iph {
namespace {

void unplan(iph_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (int i = 0; i < iph_batch::kSide; i++) {
        if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
        if (b->join[i]) (void)hipEventDestroy(b->join[i]);
        b->side[i] = nullptr;
        b->join[i] = nullptr;
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    b->capture = nullptr;
    if (b->mem) (void)hipFree(b->mem);
    b->exec = nullptr;
    b->graph = nullptr;
    b->fork = nullptr;
    b->mem = nullptr;
    b->mem_bytes = 0;
    b->planned = false;
}

iph_status enqueue_blur(iph_batch *b, hipStream_t s) {
    IPH_TRY(launch_blur_jobs(b->mem + b->blur_table_off, (int)b->blur.size(), b->blur_groups,
                             b->blur_grid_pct, s));
    for (const FallbackBlur &f : b->blur_fallback) {
        const BlurSpec &sp = b->blur[f.spec];
        IPH_TRY(launch_separable(sp.src, sp.dst, sp.w, sp.h, 1, (size_t)sp.w * sp.h, IPH_U8,
                                 sp.taps, sp.taps, sp.len, IPH_F32, b->mem + f.ws_off,
                                 f.ws_bytes, s));
    }
    return IPH_OK;
}

iph_status enqueue_filters(iph_batch *b, hipStream_t s) {
    return launch_filter_jobs(b->mem + b->filt_table_off, (int)b->filt.size(), b->filt_tiles,
                              b->filt_variant, b->filt_max_kw, b->filt_max_kh, s);
}

iph_status enqueue_match_group(iph_batch *b, const MatchGroup &g, hipStream_t s) {
    const MatchJob *table = reinterpret_cast<const MatchJob *>(b->mem + b->mjob_table_off);
    return launch_match_jobs(table + g.first, g.count, g.ntiles, g.kb, g.method, g.rt, g.tc,
                             g.lds, s);
}

// The match jobs: the fused-kernel tables, then the pairs the fused kernel
// does not take (templates above 128 px, or images beyond its offsets: the
// direct path); with `spread`, the tables on side[1] and each pair on its
// side stream, else all on s.
iph_status enqueue_matches(iph_batch *b, hipStream_t s, bool spread = false) {
    for (const MatchGroup &g : b->mgroups)
        IPH_TRY(enqueue_match_group(b, g, spread ? b->side[1] : s));
    for (size_t k = 0; k < b->match_fallback.size(); k++) {
        const MatchSpec &m = b->match[b->match_fallback[k]];
        hipStream_t ms = spread ? b->side[b->match_lane[k]] : s;
        IPH_TRY(launch_match(m.img, IPH_U8, m.w, m.h, m.tpl, m.tw, m.th, m.method,
                             resolve_path(IPH_PATH_AUTO, m.w, m.h, m.tw, m.th), m.out,
                             b->mem + m.ws_off, m.ws_bytes, ms));
    }
    return IPH_OK;
}

// Lays out and uploads the job-table slice of group g for g.rt waves per
// workgroup (its tile list, chunk height and LDS size).
iph_status layout_match_group(iph_batch *b, MatchGroup &g) {
    g.ntiles = match_jobs_layout(b->mjobs.data() + g.first, g.count, g.rt);
    g.tc = fused_match_tc(g.kb, g.max_th, g.rt, g.method, g.waves);
    g.lds = 0;
    for (int i = g.first; i < g.first + g.count; i++) {
        const MatchJob &j = b->mjobs[i];
        g.lds = std::max(g.lds, fused_match_lds(g.kb, (uint32_t)j.tw, (uint32_t)j.th, g.rt,
                                                g.method, g.tc));
    }
    IPH_TRY_HIP(hipMemcpy(b->mem + b->mjob_table_off + (size_t)g.first * sizeof(MatchJob),
                          b->mjobs.data() + g.first, (size_t)g.count * sizeof(MatchJob),
                          hipMemcpyHostToDevice));
    return IPH_OK;
}

// Waves per workgroup and per CU the autotuner chose for a shape class of
// match tables: (K blocks, method, tallest and widest template rounded up to
// 16, log2 of the map pixels).
using MatchClass = std::tuple<int, int, int, int, int>;
std::mutex g_match_cache_mu;
std::map<MatchClass, std::pair<int, int>> g_match_cache;

MatchClass match_class(const MatchGroup &g) {
    int lg = 0;
    while (std::ldexp(1.0, lg + 1) <= std::max(g.outputs, 1.0)) lg++;
    return MatchClass(g.kb, g.method, (int)((g.max_th + 15) / 16 * 16),
                      (int)((g.max_tw + 15) / 16 * 16), lg);
}

// All jobs: blurs on `s`, filters and matches on the side streams forked
// from `s` and joined back into it.
iph_status enqueue_all(iph_batch *b, hipStream_t s) {
    IPH_TRY_HIP(hipEventRecord(b->fork, s));
    for (int i = 0; i < iph_batch::kSide; i++)
        IPH_TRY_HIP(hipStreamWaitEvent(b->side[i], b->fork, 0));
    IPH_TRY(enqueue_blur(b, s));
    IPH_TRY(enqueue_filters(b, b->side[0]));
    IPH_TRY(enqueue_matches(b, s, true));
    for (int i = 0; i < iph_batch::kSide; i++) {
        IPH_TRY_HIP(hipEventRecord(b->join[i], b->side[i]));
        IPH_TRY_HIP(hipStreamWaitEvent(s, b->join[i], 0));
    }
    return IPH_OK;
}

iph_status upload_filter_table(iph_batch *b, int variant) {
    if (b->filt.empty()) return IPH_OK;
    std::vector<unsigned char> host(filter_jobs_table_bytes((int)b->filt.size()));
    IPH_TRY(filter_jobs_table(b->filt.data(), (int)b->filt.size(), variant, host.data(),
                              &b->filt_tiles));
    IPH_TRY_HIP(hipMemcpy(b->mem + b->filt_table_off, host.data(), host.size(),
                          hipMemcpyHostToDevice));
    b->filt_variant = variant;
    return IPH_OK;
}

// Milliseconds of the fastest of `reps` runs of fn on stream s.
template <typename F> iph_status time_min(hipStream_t s, int reps, F fn, float *best) {
    hipEvent_t e0, e1;
    IPH_TRY_HIP(hipEventCreate(&e0));
    IPH_TRY_HIP(hipEventCreate(&e1));
    *best = 1e30f;
    iph_status st = IPH_OK;
    for (int r = 0; r < reps && st == IPH_OK; r++) {
        (void)hipEventRecord(e0, s);
        st = fn();
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) break;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) *best = std::min(*best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return st;
}

iph_status autotune(iph_batch *b, hipStream_t s) {
    char line[160];
    b->tuning.clear();
    if (!b->filt.empty()) {
        int best_v = 0;
        float best_ms = 1e30f;
        for (int v = 0; v < filter_jobs_variants(); v++) {
            IPH_TRY(upload_filter_table(b, v));
            float ms;
            IPH_TRY(time_min(s, 3, [&] { return enqueue_filters(b, s); }, &ms));
            std::snprintf(line, sizeof line, "filter tile %s: %.3f ms\n",
                          filter_jobs_variant_name(v), ms);
            b->tuning += line;
            if (ms < best_ms) {
                best_ms = ms;
                best_v = v;
            }
        }
        IPH_TRY(upload_filter_table(b, best_v));
    }
    for (MatchGroup &g : b->mgroups) {
        int best_r = g.rt_max, best_w = 12;
        float best_ms = 1e30f;
        for (int rt : {8, 4, 2, 1}) {
            if (rt > g.rt_max) continue;
            for (int waves : {12, 8}) {
                // 8 only where it gives taller chunks than 12
                if (waves == 8 && fused_match_tc(g.kb, g.max_th, rt, g.method, 8) ==
                                      fused_match_tc(g.kb, g.max_th, rt, g.method, 12))
                    continue;
                g.rt = rt;
                g.waves = waves;
                IPH_TRY(layout_match_group(b, g));
                float ms;
                IPH_TRY(time_min(s, 3, [&] { return enqueue_match_group(b, g, s); }, &ms));
                std::snprintf(line, sizeof line,
                              "match table (kb %d, %d pairs, templates <= %ux%u) rows/workgroup "
                              "%d, chunk %d: %.3f ms\n",
                              g.kb, g.count, g.max_tw, g.max_th, 32 * rt, g.tc, ms);
                b->tuning += line;
                if (ms < best_ms) {
                    best_ms = ms;
                    best_r = rt;
                    best_w = waves;
                }
            }
        }
        g.rt = best_r;
        g.waves = best_w;
        IPH_TRY(layout_match_group(b, g));
        std::lock_guard<std::mutex> lock(g_match_cache_mu);
        g_match_cache[match_class(g)] = {best_r, best_w};
    }
    // the tile kernel's persistent grid (the per-wave kernel's grid is fixed)
    if (std::any_of(b->blur_groups.begin(), b->blur_groups.end(),
                    [](const BlurGroup &g) { return !g.wave; })) {
        int best_p = 100;
        float best_ms = 1e30f;
        for (int pct : {100, 67, 34}) {
            b->blur_grid_pct = pct;
            float ms;
            IPH_TRY(time_min(s, 3, [&] { return enqueue_blur(b, s); }, &ms));
            std::snprintf(line, sizeof line, "blur persistent workgroups %d%%: %.3f ms\n", pct, ms);
            b->tuning += line;
            if (ms < best_ms) {
                best_ms = ms;
                best_p = pct;
            }
        }
        b->blur_grid_pct = best_p;
    }
    std::string rows;
    for (const MatchGroup &g : b->mgroups) rows += " " + std::to_string(32 * g.rt);
    std::snprintf(line, sizeof line, "chosen: filter %s, match tables rows%s, blur %d%%\n",
                  filter_jobs_variant_name(b->filt_variant), rows.empty() ? " -" : rows.c_str(),
                  b->blur_grid_pct);
    b->tuning += line;
    return IPH_OK;
}

iph_status plan(iph_batch *b, int32_t tune, hipStream_t s) {
    unplan(b);
    IPH_TRY(require_device());
    IPH_TRY_HIP(hipGetDevice(&b->device));
    // layout of the one device allocation
    WorkspacePlan wp;
    std::vector<unsigned char> blur_host(blur_jobs_table_bytes((int)b->blur.size()));
    std::vector<int> fallback;
    IPH_TRY(blur_jobs_table(b->blur.data(), (int)b->blur.size(), blur_host.data(),
                            b->blur_groups, fallback));
    b->blur_table_off = 0;
    wp.take<unsigned char>(blur_host.size());
    b->blur_fallback.clear();
    for (int i : fallback) {
        const BlurSpec &sp = b->blur[i];
        size_t n = separable_workspace(sp.w, sp.h, 1, IPH_U8, sp.len, IPH_F32);
        b->blur_fallback.push_back({i, align_up(wp.used, 256), n});
        wp.take<unsigned char>(std::max<size_t>(n, 1));
    }
    b->filt_table_off = align_up(wp.used, 256);
    wp.take<unsigned char>(filter_jobs_table_bytes((int)b->filt.size()));
    b->filt_max_kw = b->filt_max_kh = 1;
    for (const FilterSpec &f : b->filt) {
        b->filt_max_kw = std::max(b->filt_max_kw, f.kw);
        b->filt_max_kh = std::max(b->filt_max_kh, f.kh);
    }
    // matches: the fused kernel's tables, grouped by (K blocks, method); the
    // rest on the direct path with a workspace slice each
    b->mjobs.clear();
    b->mgroups.clear();
    b->match_fallback.clear();
    {
        struct Entry {
            int key, idx, rt_max;
        };
        std::vector<Entry> entries;
        for (size_t i = 0; i < b->match.size(); i++) {
            MatchSpec &m = b->match[i];
            const int rt_max =
                mfma_match_fits(m.w, m.h) ? fused_match_rt_max(m.tw, m.th, m.method) : 0;
            if (rt_max == 0) {
                b->match_fallback.push_back((int)i);
                m.ws_bytes = match_workspace(IPH_U8, m.w, m.h, m.tw, m.th, m.method,
                                             resolve_path(IPH_PATH_AUTO, m.w, m.h, m.tw, m.th));
                m.ws_off = align_up(wp.used, 256);
                wp.take<unsigned char>(std::max<size_t>(m.ws_bytes, 1));
                continue;
            }
            entries.push_back({fused_match_kb(m.tw) * 4 + m.method, (int)i, rt_max});
        }
        std::stable_sort(entries.begin(), entries.end(),
                         [](const Entry &a, const Entry &c) { return a.key < c.key; });
        for (const Entry &e : entries) {
            const MatchSpec &m = b->match[e.idx];
            if (b->mgroups.empty() || b->mgroups.back().kb * 4 + b->mgroups.back().method != e.key)
                b->mgroups.push_back({e.key / 4, e.key % 4, (int)b->mjobs.size(), 0, 8, 8, 12,
                                      32, 0, 0, 0u, 0u, 0.0});
            MatchGroup &g = b->mgroups.back();
            MatchJob j{};
            j.img = m.img;
            j.tpl = m.tpl;
            j.out = m.out;
            j.W = (int)m.w;
            j.H = (int)m.h;
            j.tw = (int)m.tw;
            j.th = (int)m.th;
            j.ow = (int)(m.w - m.tw + 1);
            j.oh = (int)(m.h - m.th + 1);
            b->mjobs.push_back(j);
            g.count++;
            g.rt_max = std::min(g.rt_max, e.rt_max);
            g.max_tw = std::max(g.max_tw, m.tw);
            g.max_th = std::max(g.max_th, m.th);
            g.outputs += (double)j.ow * j.oh;
        }
    }
    b->mjob_table_off = align_up(wp.used, 256);
    wp.take<unsigned char>(std::max<size_t>(b->mjobs.size() * sizeof(MatchJob), 1));
    b->mem_bytes = std::max<size_t>(wp.used, 256);
    if (hipMalloc(reinterpret_cast<void **>(&b->mem), b->mem_bytes) != hipSuccess) {
        b->mem = nullptr;
        set_error("cannot allocate %zu bytes of batch memory", b->mem_bytes);
        return IPH_ENOMEM;
    }
    IPH_TRY_HIP(hipMemcpy(b->mem + b->blur_table_off, blur_host.data(), blur_host.size(),
                          hipMemcpyHostToDevice));
    IPH_TRY(upload_filter_table(b, 1));  // 64x32 tiles until the autotuner says otherwise
    // match tables: the cached choice for their shape class, else the shape a
    // single call of the group's largest template would take
    for (MatchGroup &g : b->mgroups) {
        g.rt = std::min(g.rt_max, fused_match_rt_default(g.max_tw, g.max_th, g.method));
        g.waves = 12;
        g.ntiles = match_jobs_layout(b->mjobs.data() + g.first, g.count, g.rt);
        std::lock_guard<std::mutex> lock(g_match_cache_mu);
        auto it = g_match_cache.find(match_class(g));
        if (it != g_match_cache.end() && it->second.first <= g.rt_max) {
            g.rt = it->second.first;
            g.waves = it->second.second;
        }
    }
    for (MatchGroup &g : b->mgroups) IPH_TRY(layout_match_group(b, g));
    // direct-path pairs to their streams: largest first, each to the least loaded
    {
        const int nl = iph_batch::kSide - 2;
        const size_t nf = b->match_fallback.size();
        std::vector<size_t> order(nf);
        std::vector<double> cost(nf), load(nl, 0.0);
        for (size_t k = 0; k < nf; k++) {
            const MatchSpec &m = b->match[b->match_fallback[k]];
            order[k] = k;
            cost[k] = (double)(m.w - m.tw + 1) * (m.h - m.th + 1) * ((double)m.tw * m.th + 64.0);
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t c) { return cost[a] > cost[c]; });
        b->match_lane.assign(nf, 2);
        for (size_t k : order) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            b->match_lane[k] = 2 + l;
            load[l] += cost[k];
        }
    }
    for (int i = 0; i < iph_batch::kSide; i++) {
        IPH_TRY_HIP(hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking));
        IPH_TRY_HIP(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming));
    }
    IPH_TRY_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    IPH_TRY_HIP(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
    b->blur_grid_pct = 100;
    b->tuning = "not tuned\n";
    if (tune) IPH_TRY(autotune(b, s));
    IPH_TRY_HIP(hipStreamSynchronize(s));
    // one graph for the whole batch, captured on the batch's own stream
    IPH_TRY_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
    iph_status st = enqueue_all(b, b->capture);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->capture, &g);
    if (st != IPH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) {
        set_error("graph capture of the batch failed: %s", hipGetErrorString(e));
        return IPH_EDEVICE;
    }
    b->graph = g;
    IPH_TRY_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->planned = true;
    return IPH_OK;
}

iph_status check_batch(iph_batch *b) {
    if (!b) {
        set_error("null batch");
        return IPH_EINVAL;
    }
    return IPH_OK;
}

}  // namespace
}  // namespace iph

using namespace iph;

extern "C" {

IPH_API iph_status iph_batch_create(iph_batch **out) {
    clear_error();
    if (!out) {
        set_error("null output pointer");
        return IPH_EINVAL;
    }
    *out = new (std::nothrow) iph_batch();
    if (!*out) {
        set_error("out of host memory");
        return IPH_ENOMEM;
    }
    return IPH_OK;
}

IPH_API void iph_batch_destroy(iph_batch *b) {
    if (!b) return;
    unplan(b);
    delete b;
}

IPH_API iph_status iph_batch_add_gaussian_blur(iph_batch *b, const uint8_t *src, uint8_t *dst,
                                               uint32_t w, uint32_t h, float sigma) {
    clear_error();
    IPH_TRY(check_batch(b));
    // gaussian_blur_f32, filter/mod.rs:156-164
    if (!(sigma > 0.0f)) {
        set_error("sigma must be > 0.0");
        return IPH_EINVAL;
    }
    IPH_TRY(check_dims(w, h));
    std::vector<float> k;
    const int32_t len = gaussian_kernel(sigma, k);
    if (len < 0 || len > kBlurSpecMaxTaps) {
        set_error("batched gaussian_blur_f32 takes sigma up to 8 (at most %d taps)",
                  kBlurSpecMaxTaps);
        return IPH_EUNSUPPORTED;
    }
    if (w == 0 || h == 0) return IPH_OK;
    BlurSpec sp{};
    sp.src = src;
    sp.dst = dst;
    sp.w = w;
    sp.h = h;
    sp.len = (uint32_t)len;
    for (int32_t i = 0; i < len; i++) sp.taps[i] = k[i];
    unplan(b);
    b->blur.push_back(sp);
    return IPH_OK;
}

IPH_API iph_status iph_batch_add_filter_clamped(iph_batch *b, const uint8_t *src, uint8_t *dst,
                                                uint32_t w, uint32_t h, const float *kernel,
                                                uint32_t kw, uint32_t kh) {
    clear_error();
    IPH_TRY(check_batch(b));
    // Kernel::new, kernel.rs:24-35
    if (kw == 0 || kh == 0) {
        set_error("width and height must be non-zero");
        return IPH_EINVAL;
    }
    if (kw * kh > (uint32_t)kF2MaxJobTaps) {
        set_error("batched filter_clamped takes kernels of at most %d taps", kF2MaxJobTaps);
        return IPH_EUNSUPPORTED;
    }
    IPH_TRY(check_dims(w, h));
    if (w == 0 || h == 0) return IPH_OK;
    FilterSpec f{};
    f.src = src;
    f.dst = dst;
    f.w = w;
    f.h = h;
    f.kw = kw;
    f.kh = kh;
    std::memcpy(f.k, kernel, sizeof(float) * kw * kh);
    unplan(b);
    b->filt.push_back(f);
    return IPH_OK;
}

IPH_API iph_status iph_batch_add_match_template(iph_batch *b, const uint8_t *image, uint32_t w,
                                                uint32_t h, const uint8_t *tpl, uint32_t tw,
                                                uint32_t th, int32_t method, float *out) {
    clear_error();
    IPH_TRY(check_batch(b));
    IPH_TRY(check_match(IPH_U8, w, h, tw, th, method, IPH_PATH_AUTO));
    unplan(b);
    b->match.push_back({image, w, h, tpl, tw, th, method, out, 0, 0});
    return IPH_OK;
}

IPH_API iph_status iph_batch_plan(iph_batch *b, int32_t autotune, void *stream) {
    clear_error();
    IPH_TRY(check_batch(b));
    iph_status st = plan(b, autotune, static_cast<hipStream_t>(stream));
    if (st != IPH_OK) unplan(b);
    return st;
}

IPH_API iph_status iph_batch_run(iph_batch *b, void *stream) {
    clear_error();
    IPH_TRY(check_batch(b));
    if (!b->planned) {
        set_error("batch is not planned (iph_batch_plan)");
        return IPH_EINVAL;
    }
    IPH_TRY_HIP(hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream)));
    return IPH_OK;
}

IPH_API iph_status iph_batch_run_eager(iph_batch *b, void *stream) {
    clear_error();
    IPH_TRY(check_batch(b));
    if (!b->planned) {
        set_error("batch is not planned (iph_batch_plan)");
        return IPH_EINVAL;
    }
    return enqueue_all(b, static_cast<hipStream_t>(stream));
}

IPH_API iph_status iph_batch_tuning(iph_batch *b, char *buf, size_t cap) {
    clear_error();
    IPH_TRY(check_batch(b));
    if (!buf || cap == 0) return IPH_OK;
    std::snprintf(buf, cap, "%s", b->tuning.c_str());
    return IPH_OK;
}

}  // extern "C"
