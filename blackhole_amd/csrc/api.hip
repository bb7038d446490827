// api.hip - the C ABI of include/bhx.h: context, uploads, launches.
#include "context.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace bhx;

namespace {

int fail(bhx_context* ctx, int code, const std::string& msg) { return bhx_fail(ctx, code, msg); }

int check_scene(bhx_context* ctx, const bhx_scene* s) {
    if (!s) return fail(ctx, BHX_E_INVALID, "scene is NULL");
    if (s->width <= 0 || s->height <= 0) return fail(ctx, BHX_E_INVALID, "width/height must be > 0");
    const int S = std::max(1, s->sqrt_spp);
    const double n = static_cast<double>(s->width) * s->height * S * S;
    if (n > 2.0e9) return fail(ctx, BHX_E_TOO_LARGE, "more than 2e9 samples in one frame");
    // rtol / atol are taken as given, as Stepper does (render.cpp:80): rtol = 0 is a
    // pure absolute tolerance; a scale of 0 makes the error NaN, the attempt is
    // rejected, and after 40 retries the ray ends as captured.
    return BHX_OK;
}

// Centre-first order of 8 x 8 tiles: the photon ring (longest rays) sits near
// the optical axis, so starting there shortens the queue's tail.
int ensure_tiles(bhx_context* ctx, int W, int H, int S, bool band, double roll) {
    if (!band) roll = 0.0;
    const auto key = std::make_tuple(W, H, S, band, roll);
    auto it = ctx->tiles.find(key);
    if (it != ctx->tiles.end()) {
        ctx->cur_tiles = &it->second;
        return BHX_OK;
    }
    if (ctx->tiles.size() >= 256) {   // bound the cache: drain the device, then start over
        HIPCHK(hipDeviceSynchronize());
        for (auto& kv : ctx->tiles) {
            kv.second.order.release();
            kv.second.start.release();
        }
        ctx->tiles.clear();
    }
    std::vector<int32_t> order, start;
    host::tile_queue(W, H, S, 0, order, start, band, roll);
    const int n = static_cast<int>(order.size());
    bhx_context::TileSet& t = ctx->tiles[key];
    HIPCHK(t.order.ensure(sizeof(int32_t) * n));
    HIPCHK(hipMemcpy(t.order.p, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(t.start.ensure(sizeof(int32_t) * (n + 1)));
    HIPCHK(hipMemcpy(t.start.p, start.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
    t.tiles_x = (W + 7) / 8;
    t.n_tiles = n;
    ctx->cur_tiles = &t;
    return BHX_OK;
}

int upload_catalogue(bhx_context* ctx, const host::Catalogue& cat) {
    std::vector<uint32_t> start;
    std::vector<double> rows;
    host::bin_catalogue(cat, start, rows);
    HIPCHK(ctx->cell_start.ensure(sizeof(uint32_t) * start.size()));
    HIPCHK(hipMemcpy(ctx->cell_start.p, start.data(), sizeof(uint32_t) * start.size(), hipMemcpyHostToDevice));
    HIPCHK(ctx->star_rows.ensure(std::max<size_t>(8, sizeof(double) * rows.size())));
    if (!rows.empty())
        HIPCHK(hipMemcpy(ctx->star_rows.p, rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice));
    return BHX_OK;
}

int ensure_catalogue(bhx_context* ctx, const bhx_scene* s) {
    if (!s->stars_enabled) return BHX_OK;
    if (ctx->custom_cat) {
        if (ctx->cat_valid && ctx->cat_is_custom) return BHX_OK;
        int rc = upload_catalogue(ctx, ctx->custom);
        if (rc) return rc;
        ctx->cat_valid = true;
        ctx->cat_is_custom = true;
        return BHX_OK;
    }
    if (ctx->cat_valid && !ctx->cat_is_custom && ctx->cat_count == s->star_count &&
        ctx->cat_seed == s->seed)
        return BHX_OK;
    host::Catalogue cat;
    host::generate_catalogue(std::max(0, s->star_count), s->seed, cat);
    int rc = upload_catalogue(ctx, cat);
    if (rc) return rc;
    ctx->cat_valid = true;
    ctx->cat_is_custom = false;
    ctx->cat_count = s->star_count;
    ctx->cat_seed = s->seed;
    return BHX_OK;
}

int trace_grid(bhx_context* ctx, bool lean, int* grid, int* block) {
    *block = ctx->launch.block > 0 ? ctx->launch.block : 256;
    int per_cu = 0;
    HIPCHK(trace_occupancy(*block, lean, &per_cu));
    if (per_cu <= 0) return fail(ctx, BHX_E_HIP, "tracer kernel cannot be resident");
    if (ctx->launch.waves_per_simd > 0) {
        const int want = std::max(1, ctx->launch.waves_per_simd * 4 * 64 / *block);
        per_cu = std::min(per_cu, want);
    }
    *grid = per_cu * ctx->n_cu;
    return BHX_OK;
}

double ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms;
}

}  // namespace

extern "C" {

int bhx_abi_version(void) { return BHX_ABI_VERSION; }

const char* bhx_last_error(const bhx_context* ctx) {
    if (ctx) return ctx->err.c_str();
    const char* m = bhx_noalloc_last_error();   // the context-free entry point's
    return m[0] ? m : "no context";
}

int bhx_create(int device, bhx_context** out) {
    if (!out) return BHX_E_INVALID;
    *out = nullptr;
    bhx_context* ctx = new bhx_context();
    ctx->device = device;
    auto bail = [&](int rc) {
        bhx_destroy(ctx);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(BHX_E_HIP);
    if (hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return bail(BHX_E_HIP);
    const std::vector<double>& tab = host::bb_table();
    if (hipMalloc(&ctx->d_bb, sizeof(double) * tab.size()) != hipSuccess) return bail(BHX_E_NOMEM);
    if (hipMemcpy(ctx->d_bb, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(BHX_E_HIP);
    if (hipMalloc(&ctx->d_cnt, sizeof(Counters)) != hipSuccess) return bail(BHX_E_NOMEM);
    for (auto& e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(BHX_E_HIP);
    *out = ctx;
    return BHX_OK;
}

int bhx_destroy(bhx_context* ctx) {
    if (!ctx) return BHX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    if (ctx->d_bb) (void)hipFree(ctx->d_bb);
    if (ctx->d_cnt) (void)hipFree(ctx->d_cnt);
    for (auto& kv : ctx->tiles) {
        kv.second.order.release();
        kv.second.start.release();
    }
    for (DevBuf* b : {&ctx->cell_start, &ctx->star_rows,
                      &ctx->init, &ctx->q2ray, &ctx->evp, &ctx->flux_partial,
                      &ctx->exp_hist, &ctx->exp_state, &ctx->r_cls, &ctx->r_y, &ctx->r_jet, &ctx->r_xyz,
                      &ctx->pilot_init, &ctx->pilot_q2ray, &ctx->pilot_att, &ctx->pilot_cls, &ctx->order_keys,
                      &ctx->order_vals, &ctx->order_sizes, &ctx->cost_order, &ctx->cost_start, &ctx->order_tmp,
                      &ctx->png_in, &ctx->png_raw, &ctx->png_key, &ctx->png_pos, &ctx->png_prev, &ctx->png_match,
                      &ctx->png_tmp, &ctx->png_box, &ctx->png_sub})
        b->release();
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->order_free) (void)hipEventDestroy(ctx->order_free);
    if (ctx->order_stream) (void)hipStreamDestroy(ctx->order_stream);
    delete ctx;
    return BHX_OK;
}

int bhx_set_catalogue(bhx_context* ctx, int32_t n, const double* x, const double* y, const double* z,
                      const double* T, const double* irr) {
    if (!ctx) return BHX_E_INVALID;
    if (n < 0) return fail(ctx, BHX_E_INVALID, "negative catalogue size");
    ctx->cat_valid = false;
    if (n == 0) {
        ctx->custom_cat = false;
        ctx->custom = host::Catalogue{};
        return BHX_OK;
    }
    if (!x || !y || !z || !T || !irr) return fail(ctx, BHX_E_INVALID, "catalogue array is NULL");
    ctx->custom_cat = true;
    ctx->custom.x.assign(x, x + n);
    ctx->custom.y.assign(y, y + n);
    ctx->custom.z.assign(z, z + n);
    ctx->custom.T.assign(T, T + n);
    ctx->custom.irr.assign(irr, irr + n);
    return BHX_OK;
}

int bhx_set_launch(bhx_context* ctx, const bhx_launch* l) {
    if (!ctx || !l) return BHX_E_INVALID;
    if (l->block < 0 || (l->block % 64) != 0 || l->block > 1024)
        return fail(ctx, BHX_E_INVALID, "block must be a multiple of 64 up to 256");
    if (l->block > 256) return fail(ctx, BHX_E_INVALID, "block must be at most 256");
    if (l->refill < 0 || l->refill > 64) return fail(ctx, BHX_E_INVALID, "refill must be in [0, 64]");
    if (l->waves_per_simd < 0 || l->waves_per_simd > 8)
        return fail(ctx, BHX_E_INVALID, "waves_per_simd must be in [0, 8]");
    if (l->queue_order < 0 || l->queue_order > BHX_ORDER_COST)
        return fail(ctx, BHX_E_INVALID, "queue_order must be 0, BHX_ORDER_CENTRE or BHX_ORDER_COST");
    if (l->queue_order) ctx->launch.queue_order = l->queue_order;
    if (l->block) ctx->launch.block = l->block;
    if (l->waves_per_simd) ctx->launch.waves_per_simd = l->waves_per_simd;
    if (l->chunk) ctx->launch.chunk = l->chunk;
    if (l->refill) ctx->launch.refill = l->refill;
    return BHX_OK;
}

int bhx_get_launch(const bhx_context* ctx, bhx_launch* l) {
    if (!ctx || !l) return BHX_E_INVALID;
    *l = ctx->launch;
    return BHX_OK;
}

int bhx_trace_dev(bhx_context* ctx, const bhx_scene* s, const bhx_rays* rays, bhx_stats* stats,
                  void* stream) {
    if (!ctx) return BHX_E_INVALID;
    int rc = check_scene(ctx, s);
    if (rc) return rc;
    if (!rays) return fail(ctx, BHX_E_INVALID, "rays is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int Sspp = std::max(1, s->sqrt_spp);
    rc = ensure_tiles(ctx, s->width, s->height, Sspp, host::axis_band(*s), s->cam_roll);
    if (rc) return rc;
    const int64_t n_rays = static_cast<int64_t>(s->width) * s->height * Sspp * Sspp;
    HIPCHK(ctx->init.ensure(sizeof(double) * kInitRows * n_rays));
    HIPCHK(ctx->q2ray.ensure(sizeof(int32_t) * n_rays));
    RaygenArgs G = host::raygen_args(*s);
    G.tiles_x = ctx->cur_tiles->tiles_x;
    G.n_tiles = ctx->cur_tiles->n_tiles;
    G.tile_order = static_cast<const int32_t*>(ctx->cur_tiles->order.p);
    G.tile_start = static_cast<const int32_t*>(ctx->cur_tiles->start.p);
    G.init = static_cast<double*>(ctx->init.p);
    G.q2ray = static_cast<int32_t*>(ctx->q2ray.p);
    TraceArgs P = host::trace_args(*s);
    HIPCHK(ctx->evp.ensure(sizeof(EventParams)));
    P.ev = static_cast<const EventParams*>(ctx->evp.p);
    if (P.has_jet || P.has_sun) {
        // synchronous and only on change: the buffer may still be read by a
        // launch in flight, and a pageable source cannot be copied asynchronously
        const EventParams evh = host::event_params(*s);
        if (!ctx->ev_valid || std::memcmp(&evh, &ctx->ev_host, sizeof evh) != 0) {
            HIPCHK(hipStreamSynchronize(st));
            HIPCHK(hipMemcpy(ctx->evp.p, &evh, sizeof evh, hipMemcpyHostToDevice));
            ctx->ev_host = evh;
            ctx->ev_valid = true;
        }
    }
    P.refill = std::max(1, std::min(64, ctx->launch.refill));
    P.chunk = std::max(1, ctx->launch.chunk);
    P.init = G.init;
    P.q2ray = G.q2ray;
    P.counter = &ctx->d_cnt->tile_counter;
    P.cls = rays->cls;
    P.y = rays->y;
    P.jet = rays->jet;
    P.steps = rays->steps;
    P.attempts = rays->attempts;
    P.norm_err = rays->norm_err;
    P.carter = rays->carter;
    P.totals = ctx->d_cnt->totals;
    int grid = 0, block = 0;
    rc = trace_grid(ctx, trace_is_lean(P), &grid, &block);
    if (rc) return rc;
    // a frame with fewer rays than resident lanes: one-wave blocks, as many as
    // its rays fill, so every wave holds 64 rays and runs without a half-empty
    // partner sharing its SIMD's issue slots (the 256^2 frame: 2.37 -> 2.03 ms)
    if (n_rays < static_cast<int64_t>(grid) * block) {
        block = 64;
        grid = static_cast<int>((n_rays + 63) / 64);
    }
    // counters are cleared by fill kernels (never a memset node: the call may
    // be captured); the tracer's here, the pilot's with the order below
    Counters* cnt = ctx->d_cnt;
    HIPCHK(launch_zero_u64(cnt->totals, 4, st));
    HIPCHK(launch_zero_u64(cnt->classes, 8, st));
    HIPCHK(launch_zero_u32(&cnt->tile_counter, 1, st));
    // longest-first queue when every resident lane sees several rays (below
    // that the queue hardly drains before the end and the pilot is overhead)
    const bool by_cost = ctx->launch.queue_order == BHX_ORDER_COST && !P.has_jet &&
                         n_rays >= 4 * static_cast<int64_t>(grid) * block;
    bool order_side = false;
    if (by_cost) {
        // On the order stream unless the call is being captured: it waits only
        // for the last frame's ray generation to have read the previous order,
        // so back-to-back frames build it during the last frame's trace (whose
        // retiring waves free the CUs it needs) and shading.  Same order, same rays.
        hipStream_t os = st;
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(st, &cap));
        if (cap == hipStreamCaptureStatusNone) {
            if (!ctx->order_stream) {
                HIPCHK(hipStreamCreateWithFlags(&ctx->order_stream, hipStreamNonBlocking));
                HIPCHK(hipEventCreateWithFlags(&ctx->order_free, hipEventDisableTiming));
                HIPCHK(hipEventRecord(ctx->order_free, st));
            }
            os = ctx->order_stream;
            order_side = true;
            HIPCHK(hipStreamWaitEvent(os, ctx->order_free, 0));
        }
        HIPCHK(launch_zero_u32(&cnt->pilot_counter, 1, os));
        HIPCHK(launch_zero_u64(cnt->pilot_totals, 4, os));
        HIPCHK(hipEventRecord(ctx->ev[4], os));
        rc = cost_order(ctx, G, P, os);
        if (rc) return rc;
        HIPCHK(hipEventRecord(ctx->ev[5], os));
        if (order_side) HIPCHK(hipStreamWaitEvent(st, ctx->ev[5], 0));
        G.tile_order = static_cast<const int32_t*>(ctx->cost_order.p);
        G.tile_start = static_cast<const int32_t*>(ctx->cost_start.p);
    }
    HIPCHK(launch_raygen(G, st));
    if (order_side) HIPCHK(hipEventRecord(ctx->order_free, st));
    HIPCHK(hipEventRecord(ctx->ev[0], st));
    HIPCHK(launch_trace(P, grid, block, st));
    HIPCHK(hipEventRecord(ctx->ev[1], st));
    if (stats) {
        HIPCHK(hipStreamSynchronize(st));
        Counters c;
        HIPCHK(hipMemcpy(&c, ctx->d_cnt, sizeof c, hipMemcpyDeviceToHost));
        double mn, mc;
        std::memcpy(&mn, &c.totals[2], 8);
        std::memcpy(&mc, &c.totals[3], 8);
        stats->rays += n_rays;
        stats->steps += static_cast<int64_t>(c.totals[0]);
        stats->attempts += static_cast<int64_t>(c.totals[1]);
        stats->rhs_evals += static_cast<int64_t>(c.totals[1]) * 6 + n_rays;
        stats->max_norm_error = std::max(stats->max_norm_error, mn);
        stats->max_carter_drift = std::max(stats->max_carter_drift, mc);
        stats->trace_ms = ms_between(ctx->ev[0], ctx->ev[1]);
        stats->order_ms = by_cost ? ms_between(ctx->ev[4], ctx->ev[5]) : 0.0;
    }
    return BHX_OK;
}

int bhx_shade_dev(bhx_context* ctx, const bhx_scene* s, const bhx_rays* rays, const bhx_shaded* out,
                  bhx_stats* stats, void* stream) {
    if (!ctx) return BHX_E_INVALID;
    int rc = check_scene(ctx, s);
    if (rc) return rc;
    if (!rays || !rays->cls || !rays->y) return fail(ctx, BHX_E_INVALID, "rays.cls / rays.y required");
    if (!out) return fail(ctx, BHX_E_INVALID, "out is NULL");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    rc = ensure_catalogue(ctx, s);
    if (rc) return rc;
    ShadeArgs P = host::shade_args(*s);
    P.bb = ctx->d_bb;
    P.cell_start = static_cast<const uint32_t*>(ctx->cell_start.p);
    P.star = static_cast<const double*>(ctx->star_rows.p);
    if (s->stars_enabled && !P.cell_start) return fail(ctx, BHX_E_INVALID, "no star catalogue");
    P.cls = rays->cls;
    P.y = rays->y;
    P.jet = rays->jet;
    P.xyz = out->xyz;
    P.ray_xyz = out->ray_xyz;
    P.g = out->g;
    P.dir = out->dir;
    const int blocks = (P.n_pix + 255) / 256;
    HIPCHK(ctx->flux_partial.ensure(sizeof(double) * blocks));
    P.flux_partial = static_cast<double*>(ctx->flux_partial.p);
    P.class_counts = ctx->d_cnt->classes;
    HIPCHK(launch_zero_u64(ctx->d_cnt->classes, 8, st));
    HIPCHK(hipEventRecord(ctx->ev[2], st));
    HIPCHK(launch_shade(P, blocks, st));
    HIPCHK(hipEventRecord(ctx->ev[3], st));
    if (stats) {
        HIPCHK(hipStreamSynchronize(st));
        Counters c;
        HIPCHK(hipMemcpy(&c, ctx->d_cnt, sizeof c, hipMemcpyDeviceToHost));
        std::vector<double> fp(blocks);
        HIPCHK(hipMemcpy(fp.data(), P.flux_partial, sizeof(double) * blocks, hipMemcpyDeviceToHost));
        double flux = 0.0;
        for (double v : fp) flux += v;
        stats->captured += static_cast<int64_t>(c.classes[BHX_HORIZON]);
        stats->disc_hits += static_cast<int64_t>(c.classes[BHX_DISC]);
        stats->sun_hits += static_cast<int64_t>(c.classes[BHX_SUN]);
        stats->star_hits += static_cast<int64_t>(c.classes[BHX_STAR]);
        stats->escaped += static_cast<int64_t>(c.classes[BHX_SKY]);
        stats->exhausted += static_cast<int64_t>(c.classes[BHX_EXHAUSTED]);
        stats->jet_rays += static_cast<int64_t>(c.classes[6]);
        stats->total_flux += flux;
        stats->shade_ms = ms_between(ctx->ev[2], ctx->ev[3]);
    }
    return BHX_OK;
}

int bhx_render(bhx_context* ctx, const bhx_scene* s, double* xyz, bhx_stats* stats) {
    if (!ctx) return BHX_E_INVALID;
    int rc = check_scene(ctx, s);
    if (rc) return rc;
    if (!xyz) return fail(ctx, BHX_E_INVALID, "xyz is NULL");
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(ctx->device));
    const int S = std::max(1, s->sqrt_spp);
    const size_t n_rays = static_cast<size_t>(s->width) * s->height * S * S;
    const size_t n_pix = static_cast<size_t>(s->width) * s->height;
    HIPCHK(ctx->r_cls.ensure(n_rays));
    HIPCHK(ctx->r_y.ensure(8 * sizeof(double) * n_rays));
    HIPCHK(ctx->r_jet.ensure(sizeof(double) * n_rays));
    HIPCHK(ctx->r_xyz.ensure(3 * sizeof(double) * n_pix));
    bhx_rays rays{};
    rays.cls = static_cast<uint8_t*>(ctx->r_cls.p);
    rays.y = static_cast<double*>(ctx->r_y.p);
    rays.jet = static_cast<double*>(ctx->r_jet.p);
    bhx_shaded sh{};
    sh.xyz = static_cast<double*>(ctx->r_xyz.p);
    bhx_stats local;
    std::memset(&local, 0, sizeof local);
    rc = bhx_trace_dev(ctx, s, &rays, &local, nullptr);
    if (rc) return rc;
    rc = bhx_shade_dev(ctx, s, &rays, &sh, &local, nullptr);
    if (rc) return rc;
    HIPCHK(hipMemcpy(xyz, sh.xyz, 3 * sizeof(double) * n_pix, hipMemcpyDeviceToHost));
    if (stats) {
        stats->rays += local.rays;
        stats->steps += local.steps;
        stats->captured += local.captured;
        stats->disc_hits += local.disc_hits;
        stats->sun_hits += local.sun_hits;
        stats->star_hits += local.star_hits;
        stats->jet_rays += local.jet_rays;
        stats->escaped += local.escaped;
        stats->exhausted += local.exhausted;
        stats->total_flux += local.total_flux;
        stats->max_norm_error = std::max(stats->max_norm_error, local.max_norm_error);
        stats->max_carter_drift = std::max(stats->max_carter_drift, local.max_carter_drift);
        stats->attempts += local.attempts;
        stats->rhs_evals += local.rhs_evals;
        stats->trace_ms = local.trace_ms;
        stats->shade_ms = local.shade_ms;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return BHX_OK;
}

int bhx_auto_exposure_dev(bhx_context* ctx, const double* xyz, int64_t n, double key, double* exposure,
                          void* stream) {
    if (!ctx) return BHX_E_INVALID;
    if (!xyz || !exposure || n < 0) return fail(ctx, BHX_E_INVALID, "bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIPCHK(ctx->exp_hist.ensure(sizeof(unsigned int) * 2 * 4096));
    HIPCHK(ctx->exp_state.ensure(sizeof(ExposureState)));
    unsigned int* hist = static_cast<unsigned int*>(ctx->exp_hist.p);
    ExposureState* state = static_cast<ExposureState*>(ctx->exp_state.p);
    // both order statistics (image.cpp:29-36) selected on the device, then
    // one copy of the selection's state to the host
    HIPCHK(launch_zero_u32(hist, 2 * 4096, st));
    HIPCHK(launch_zero_u64(reinterpret_cast<unsigned long long*>(state), sizeof(ExposureState) / 8, st));
    if (n > 0) HIPCHK(launch_exposure(xyz, n, state, hist, st));
    ExposureState h{};
    HIPCHK(hipMemcpyAsync(&h, state, sizeof h, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (h.n_pos == 0) {
        *exposure = 1.0;
        return BHX_OK;
    }
    double hi, peak;
    std::memcpy(&hi, &h.prefix[0], 8);
    std::memcpy(&peak, &h.prefix[1], 8);
    double e = 1.0;
    if (hi > 0.0) e = key / hi;
    if (peak > 0.0) e = std::min(e, 3.0 * key / peak);
    *exposure = e;
    return BHX_OK;
}

int bhx_develop_dev(bhx_context* ctx, const double* xyz, int64_t n, double exposure, int32_t tm,
                    double log_decades, uint8_t* rgb, void* stream) {
    if (!ctx) return BHX_E_INVALID;
    if (!xyz || !rgb || n < 0) return fail(ctx, BHX_E_INVALID, "bad argument");
    if (tm < BHX_TM_LINEAR || tm > BHX_TM_HDR) return fail(ctx, BHX_E_INVALID, "unknown tone map");
    HIPCHK(hipSetDevice(ctx->device));
    const double log_k = std::pow(10.0, std::clamp(log_decades, 1.0, 20.0));
    if (n == 0) return BHX_OK;
    HIPCHK(launch_develop(xyz, n, exposure, tm, log_k, rgb, static_cast<hipStream_t>(stream)));
    return BHX_OK;
}

int bhx_develop_auto_dev(bhx_context* ctx, const double* xyz, int64_t n, double key, int32_t tm,
                         double log_decades, uint8_t* rgb, double* exposure_dev, void* stream) {
    if (!ctx) return BHX_E_INVALID;
    if (!xyz || !rgb || n < 0) return fail(ctx, BHX_E_INVALID, "bad argument");
    if (tm < BHX_TM_LINEAR || tm > BHX_TM_HDR) return fail(ctx, BHX_E_INVALID, "unknown tone map");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    const double log_k = std::pow(10.0, std::clamp(log_decades, 1.0, 20.0));
    HIPCHK(ctx->exp_hist.ensure(sizeof(unsigned int) * 2 * 4096));
    HIPCHK(ctx->exp_state.ensure(sizeof(ExposureState)));
    unsigned int* hist = static_cast<unsigned int*>(ctx->exp_hist.p);
    ExposureState* state = static_cast<ExposureState*>(ctx->exp_state.p);
    // the selection of bhx_auto_exposure_dev, then develop reads its state
    // on the device: the stream never waits for the host
    HIPCHK(launch_zero_u32(hist, 2 * 4096, st));
    HIPCHK(launch_zero_u64(reinterpret_cast<unsigned long long*>(state), sizeof(ExposureState) / 8, st));
    if (n > 0) HIPCHK(launch_exposure(xyz, n, state, hist, st));
    HIPCHK(launch_develop_auto(xyz, n, state, key, exposure_dev, tm, log_k, rgb, st));
    return BHX_OK;
}

int bhx_blackbody_xyz_dev(bhx_context* ctx, const double* T, int64_t n, double* xyz, void* stream) {
    if (!ctx) return BHX_E_INVALID;
    if (!T || !xyz || n < 0) return fail(ctx, BHX_E_INVALID, "bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(launch_bb(ctx->d_bb, T, n, xyz, static_cast<hipStream_t>(stream)));
    return BHX_OK;
}

}  // extern "C"
