// batch.hip - many scenes, one persistent tracer launch, replayed as a hipGraph.
//
// The reference renders one RenderConfig per render() call (src/render.cpp:
// 396-497); a caller with many scenes (an animation, a parameter sweep, a
// survey of spins and inclinations) calls it once per scene.  On a GPU with
// 256 CUs that sequence leaves most of the chip idle: a 128 x 128 frame has
// 16k rays against ~130k resident lanes, and every frame ends in a tail where
// the last photon-ring rays run alone.  Here all scenes' rays share one work
// queue: ray generation per scene, ONE persistent tracer launch in which each
// lane reads its own scene's constants from an LDS table, then shading per
// scene - the whole sequence captured once into a hipGraph and replayed.
// Results are those of the single-scene path ray for ray (same device code).
#include "context.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

using namespace bhx;

namespace {

struct SceneCounters {
    unsigned long long classes[8];   // 6 classes + jet rays (k_shade)
    unsigned long long trace[4];     // accepted steps, attempts, max norm bits, max Carter bits
};
// Device counter block: per-scene counters, then the batch's own.
struct BatchCounters {
    unsigned long long totals[4];
    unsigned int tile_counter;
    unsigned int pad;
};

}  // namespace

struct bhx_batch {
    bhx_context* ctx = nullptr;
    std::vector<bhx_scene> scenes;
    int n = 0;
    int64_t n_rays = 0, n_pix = 0;
    std::vector<int64_t> ray_base, pix_base;
    std::vector<RaygenArgs> gen;
    std::vector<ShadeArgs> shd;
    std::vector<int> shade_blocks;
    std::vector<int> flux_off;
    int flux_total = 0;
    std::vector<int> queue_order;      // scene indices in queue order
    TraceArgs trace{};
    DevBuf tile_order, tile_start, init, q2ray, q2scene, table;
    DevBuf cls, y, jet, steps, attempts, norm, carter, xyz;
    DevBuf counters, flux, cell_start, star_rows;
    // one launch of ray generation / of shading for every scene: the scenes'
    // arguments and each block's scene and first block of that scene
    DevBuf gen_table, gen_map, shd_table, shd_map;
    int gen_blocks = 0;
    bhx_launch launch{};
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double last_ms = 0.0;
};

namespace {

SceneCounters* scene_counters(bhx_batch* b) { return static_cast<SceneCounters*>(b->counters.p); }
BatchCounters* batch_counters(bhx_batch* b) {
    return reinterpret_cast<BatchCounters*>(scene_counters(b) + b->n);
}
size_t counter_words(int n) { return (sizeof(SceneCounters) * n + sizeof(BatchCounters)) / 8; }

void drop_graph(bhx_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int batch_grid(bhx_batch* b, int* grid, int* block) {
    bhx_context* ctx = b->ctx;
    *block = b->launch.block > 0 ? b->launch.block : 256;
    int per_cu = 0;
    HIPCHK(trace_batch_occupancy(*block, b->n, trace_is_lean(b->trace), &per_cu));
    if (per_cu <= 0) return bhx_fail(ctx, BHX_E_HIP, "batched tracer cannot be resident");
    if (b->launch.waves_per_simd > 0)
        per_cu = std::min(per_cu, std::max(1, b->launch.waves_per_simd * 4 * 64 / *block));
    *grid = per_cu * ctx->n_cu;
    return BHX_OK;
}

// The whole batch on one stream: counters, ray generation, tracer, shading.
int enqueue(bhx_batch* b, hipStream_t st) {
    bhx_context* ctx = b->ctx;
    int grid = 0, block = 0;
    int rc = batch_grid(b, &grid, &block);
    if (rc) return rc;
    TraceArgs P = b->trace;
    P.refill = std::max(1, std::min(64, b->launch.refill));
    P.chunk = std::max(1, b->launch.chunk);
    HIPCHK(launch_zero_u64(static_cast<unsigned long long*>(b->counters.p),
                           static_cast<int>(counter_words(b->n)), st));
    const int32_t* gmap = static_cast<const int32_t*>(b->gen_map.p);
    HIPCHK(launch_raygen_batch(static_cast<const RaygenArgs*>(b->gen_table.p), gmap, gmap + b->gen_blocks,
                               b->gen_blocks, st));
    HIPCHK(launch_trace_batch(P, grid, block, st));
    const int32_t* smap = static_cast<const int32_t*>(b->shd_map.p);
    HIPCHK(launch_shade_batch(static_cast<const ShadeArgs*>(b->shd_table.p), smap, smap + b->flux_total,
                              b->flux_total, st));
    return BHX_OK;
}


}  // namespace

extern "C" {

int bhx_batch_create(bhx_context* ctx, const bhx_scene* scenes, int32_t n, bhx_batch** out) {
    if (!ctx || !out) return BHX_E_INVALID;
    *out = nullptr;
    if (!scenes || n <= 0) return bhx_fail(ctx, BHX_E_INVALID, "empty batch");
    if (n > kMaxBatchScenes) return bhx_fail(ctx, BHX_E_TOO_LARGE, "more than 128 scenes in one batch");
    int64_t total = 0, total_pix = 0;
    int star_scene = -1;
    for (int i = 0; i < n; ++i) {
        const bhx_scene& s = scenes[i];
        if (s.width <= 0 || s.height <= 0) return bhx_fail(ctx, BHX_E_INVALID, "width/height must be > 0");
        const int S = std::max(1, s.sqrt_spp);
        total += static_cast<int64_t>(s.width) * s.height * S * S;
        total_pix += static_cast<int64_t>(s.width) * s.height;
        if (s.stars_enabled) {
            if (star_scene >= 0 && (scenes[star_scene].star_count != s.star_count ||
                                    scenes[star_scene].seed != s.seed))
                return bhx_fail(ctx, BHX_E_INVALID,
                                "scenes with stars in one batch must share star_count and seed");
            star_scene = i;
        }
    }
    if (total > 2000000000LL) return bhx_fail(ctx, BHX_E_TOO_LARGE, "more than 2e9 samples in one batch");
    HIPCHK(hipSetDevice(ctx->device));

    bhx_batch* b = new bhx_batch();
    b->ctx = ctx;
    b->scenes.assign(scenes, scenes + n);
    b->n = n;
    b->n_rays = total;
    b->n_pix = total_pix;
    b->launch = ctx->launch;
    auto bail = [&](int rc) {
        bhx_batch_destroy(b);
        return rc;
    };
#define BCHK(call)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            return bail(bhx_fail(ctx, BHX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
    } while (0)

    // ray / pixel bases in the order given; queue order = most rays first
    b->ray_base.resize(n);
    b->pix_base.resize(n);
    int64_t rb = 0, pb = 0;
    std::vector<int64_t> rays_of(n);
    for (int i = 0; i < n; ++i) {
        const bhx_scene& s = scenes[i];
        const int S = std::max(1, s.sqrt_spp);
        rays_of[i] = static_cast<int64_t>(s.width) * s.height * S * S;
        b->ray_base[i] = rb;
        b->pix_base[i] = pb;
        rb += rays_of[i];
        pb += static_cast<int64_t>(s.width) * s.height;
    }
    b->queue_order.resize(n);
    std::iota(b->queue_order.begin(), b->queue_order.end(), 0);
    std::stable_sort(b->queue_order.begin(), b->queue_order.end(),
                     [&](int x, int y) { return rays_of[x] > rays_of[y]; });

    // tile queues of every scene (each centre first: the photon ring's long
    // rays, then the sky's short ones), merged by their position in their
    // own scene's queue: the batch's queue takes every scene's centre tiles
    // first and ends on short rays, with no scene's long rays left for the end
    std::vector<int32_t> order_all, start_all;
    std::vector<size_t> toff(n), soff(n);
    std::vector<int> ntiles(n);
    struct QTile {
        double pos;   // (t + 1/2) / tiles of the scene
        int rank, scene, t;
    };
    std::vector<QTile> merged;
    for (int r = 0; r < n; ++r) {
        const int i = b->queue_order[r];
        const bhx_scene& s = scenes[i];
        std::vector<int32_t> order, start;
        host::tile_queue(s.width, s.height, std::max(1, s.sqrt_spp), 0, order, start, host::axis_band(s),
                         s.cam_roll);
        toff[i] = order_all.size();
        soff[i] = start_all.size();
        ntiles[i] = static_cast<int>(order.size());
        order_all.insert(order_all.end(), order.begin(), order.end());
        start_all.insert(start_all.end(), start.begin(), start.end());   // counts for now
        for (int t = 0; t < ntiles[i]; ++t) merged.push_back({(t + 0.5) / ntiles[i], r, i, t});
    }
    std::stable_sort(merged.begin(), merged.end(), [](const QTile& x, const QTile& y) {
        return x.pos < y.pos || (x.pos == y.pos && x.rank < y.rank);
    });
    {
        std::vector<int32_t> count(start_all.size());
        for (int i = 0; i < n; ++i)
            for (int t = 0; t < ntiles[i]; ++t) count[soff[i] + t] = start_all[soff[i] + t + 1] - start_all[soff[i] + t];
        int32_t q = 0;
        for (const QTile& m : merged) {
            start_all[soff[m.scene] + m.t] = q;
            q += count[soff[m.scene] + m.t];
        }
        for (int i = 0; i < n; ++i) start_all[soff[i] + ntiles[i]] = -1;   // no tile past the last
    }

    // device buffers
    const int64_t N = total;
    BCHK(b->tile_order.ensure(sizeof(int32_t) * order_all.size()));
    BCHK(b->tile_start.ensure(sizeof(int32_t) * start_all.size()));
    BCHK(hipMemcpy(b->tile_order.p, order_all.data(), sizeof(int32_t) * order_all.size(), hipMemcpyHostToDevice));
    BCHK(hipMemcpy(b->tile_start.p, start_all.data(), sizeof(int32_t) * start_all.size(), hipMemcpyHostToDevice));
    BCHK(b->init.ensure(sizeof(double) * kInitRows * N));
    BCHK(b->q2ray.ensure(sizeof(int32_t) * N));
    BCHK(b->q2scene.ensure(N));
    BCHK(b->cls.ensure(N));
    BCHK(b->y.ensure(sizeof(double) * 8 * N));
    BCHK(b->jet.ensure(sizeof(double) * N));
    BCHK(b->steps.ensure(sizeof(int32_t) * N));
    BCHK(b->attempts.ensure(sizeof(int32_t) * N));
    BCHK(b->norm.ensure(sizeof(double) * N));
    BCHK(b->carter.ensure(sizeof(double) * N));
    BCHK(b->xyz.ensure(sizeof(double) * 3 * total_pix));
    BCHK(b->counters.ensure(8 * counter_words(n)));

    // star catalogue owned by the batch (the graph keeps its pointers)
    if (star_scene >= 0) {
        host::Catalogue cat;
        if (ctx->custom_cat) cat = ctx->custom;
        else host::generate_catalogue(std::max(0, scenes[star_scene].star_count), scenes[star_scene].seed, cat);
        std::vector<uint32_t> start;
        std::vector<double> rows;
        host::bin_catalogue(cat, start, rows);
        BCHK(b->cell_start.ensure(sizeof(uint32_t) * start.size()));
        BCHK(hipMemcpy(b->cell_start.p, start.data(), sizeof(uint32_t) * start.size(), hipMemcpyHostToDevice));
        BCHK(b->star_rows.ensure(std::max<size_t>(8, sizeof(double) * rows.size())));
        if (!rows.empty())
            BCHK(hipMemcpy(b->star_rows.p, rows.data(), sizeof(double) * rows.size(), hipMemcpyHostToDevice));
    }

    // per-scene constants of the tracer (LDS table) and the launch arguments
    std::vector<SceneK> table(n);
    b->gen.resize(n);
    b->shd.resize(n);
    b->shade_blocks.resize(n);
    b->flux_off.resize(n);
    int foff = 0;
    for (int i = 0; i < n; ++i) {
        const bhx_scene& s = scenes[i];
        const TraceArgs T = host::trace_args(s);
        SceneK& k = table[i];
        std::memset(&k, 0, sizeof k);
        k.a = T.a; k.a2 = T.a2; k.r_h = T.r_h; k.r_stop = T.r_stop; k.rtol = T.rtol; k.atol = T.atol;
        k.r_escape = T.r_escape; k.h0 = T.h0; k.r_in = T.r_in; k.r_out = T.r_out; k.ns_R = T.ns_R;
        k.has_disc = T.has_disc; k.has_ns = T.has_ns; k.has_jet = T.has_jet; k.has_sun = T.has_sun;
        k.max_steps = T.max_steps;
        k.ev = host::event_params(s);

        RaygenArgs G = host::raygen_args(s);
        G.tiles_x = (s.width + 7) / 8;
        G.n_tiles = ntiles[i];
        G.tile_order = static_cast<const int32_t*>(b->tile_order.p) + toff[i];
        G.tile_start = static_cast<const int32_t*>(b->tile_start.p) + soff[i];
        G.n_rays = static_cast<int32_t>(N);
        G.init = static_cast<double*>(b->init.p);
        G.q2ray = static_cast<int32_t*>(b->q2ray.p);
        G.ray_base = static_cast<int32_t>(b->ray_base[i]);
        G.scene_id = i;
        G.q2scene = static_cast<uint8_t*>(b->q2scene.p);
        b->gen[i] = G;

        ShadeArgs P = host::shade_args(s);
        P.n_rays = N;
        P.ray_base = b->ray_base[i];
        P.bb = ctx->d_bb;
        P.cell_start = static_cast<const uint32_t*>(b->cell_start.p);
        P.star = static_cast<const double*>(b->star_rows.p);
        P.cls = static_cast<const uint8_t*>(b->cls.p);
        P.y = static_cast<const double*>(b->y.p);
        P.jet = static_cast<const double*>(b->jet.p);
        P.xyz = static_cast<double*>(b->xyz.p) + 3 * b->pix_base[i];
        P.class_counts = scene_counters(b)[i].classes;
        P.steps_in = static_cast<const int32_t*>(b->steps.p);
        P.attempts_in = static_cast<const int32_t*>(b->attempts.p);
        P.norm_in = static_cast<const double*>(b->norm.p);
        P.carter_in = static_cast<const double*>(b->carter.p);
        P.trace_totals = scene_counters(b)[i].trace;
        b->shade_blocks[i] = (P.n_pix + 255) / 256;
        b->flux_off[i] = foff;
        foff += b->shade_blocks[i];
        b->shd[i] = P;
    }
    b->flux_total = foff;
    BCHK(b->flux.ensure(sizeof(double) * foff));
    for (int i = 0; i < n; ++i) b->shd[i].flux_partial = static_cast<double*>(b->flux.p) + b->flux_off[i];
    // the batched launches' tables: block -> scene maps followed by each
    // scene's first block (ray generation in queue order, shading by scene)
    {
        std::vector<int32_t> gmap, gfirst(n), smap, sfirst(n);
        for (int i : b->queue_order) {
            gfirst[i] = static_cast<int32_t>(gmap.size());
            gmap.insert(gmap.end(), ntiles[i], i);
        }
        for (int i = 0; i < n; ++i) {
            sfirst[i] = b->flux_off[i];
            smap.insert(smap.end(), b->shade_blocks[i], i);
        }
        b->gen_blocks = static_cast<int>(gmap.size());
        gmap.insert(gmap.end(), gfirst.begin(), gfirst.end());
        smap.insert(smap.end(), sfirst.begin(), sfirst.end());
        BCHK(b->gen_table.ensure(sizeof(RaygenArgs) * n));
        BCHK(b->gen_map.ensure(sizeof(int32_t) * gmap.size()));
        BCHK(b->shd_table.ensure(sizeof(ShadeArgs) * n));
        BCHK(b->shd_map.ensure(sizeof(int32_t) * smap.size()));
        BCHK(hipMemcpy(b->gen_table.p, b->gen.data(), sizeof(RaygenArgs) * n, hipMemcpyHostToDevice));
        BCHK(hipMemcpy(b->gen_map.p, gmap.data(), sizeof(int32_t) * gmap.size(), hipMemcpyHostToDevice));
        BCHK(hipMemcpy(b->shd_table.p, b->shd.data(), sizeof(ShadeArgs) * n, hipMemcpyHostToDevice));
        BCHK(hipMemcpy(b->shd_map.p, smap.data(), sizeof(int32_t) * smap.size(), hipMemcpyHostToDevice));
    }
    BCHK(b->table.ensure(sizeof(SceneK) * n));
    BCHK(hipMemcpy(b->table.p, table.data(), sizeof(SceneK) * n, hipMemcpyHostToDevice));

    TraceArgs& P = b->trace;
    // the uniform event flags say whether any scene has jets or the Sun (which
    // tracer kernel the batch needs); each lane reads its own scene's flags
    P.has_jet = P.has_sun = 0;
    for (const SceneK& k : table) {
        P.has_jet |= k.has_jet;
        P.has_sun |= k.has_sun;
    }
    P.n_rays = static_cast<int32_t>(N);
    P.init = static_cast<const double*>(b->init.p);
    P.q2ray = static_cast<const int32_t*>(b->q2ray.p);
    P.counter = &batch_counters(b)->tile_counter;
    P.scenes = static_cast<const SceneK*>(b->table.p);
    P.q2scene = static_cast<const uint8_t*>(b->q2scene.p);
    P.n_scenes = n;
    P.cls = static_cast<uint8_t*>(b->cls.p);
    P.y = static_cast<double*>(b->y.p);
    P.jet = static_cast<double*>(b->jet.p);
    P.steps = static_cast<int32_t*>(b->steps.p);
    P.attempts = static_cast<int32_t*>(b->attempts.p);
    P.norm_err = static_cast<double*>(b->norm.p);
    P.carter = static_cast<double*>(b->carter.p);
    P.totals = batch_counters(b)->totals;

    BCHK(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
    for (auto& e : b->ev) BCHK(hipEventCreate(&e));
#undef BCHK
    *out = b;
    return BHX_OK;
}

int bhx_batch_destroy(bhx_batch* b) {
    if (!b) return BHX_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    drop_graph(b);
    for (DevBuf* d : {&b->tile_order, &b->tile_start, &b->init, &b->q2ray, &b->q2scene, &b->table,
                      &b->cls, &b->y, &b->jet, &b->steps, &b->attempts, &b->norm, &b->carter, &b->xyz,
                      &b->counters, &b->flux, &b->cell_start, &b->star_rows, &b->gen_table, &b->gen_map,
                      &b->shd_table, &b->shd_map})
        d->release();
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    delete b;
    return BHX_OK;
}

int bhx_batch_run(bhx_batch* b, int32_t use_graph, void* stream) {
    if (!b) return BHX_E_INVALID;
    bhx_context* ctx = b->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (!use_graph) {
        HIPCHK(hipEventRecord(b->ev[0], st));
        int rc = enqueue(b, st);
        if (rc) return rc;
        HIPCHK(hipEventRecord(b->ev[1], st));
        return BHX_OK;
    }
    if (!b->exec) {
        // capture on the batch's private stream, then launch into the caller's
        HIPCHK(hipStreamBeginCapture(b->cap, hipStreamCaptureModeRelaxed));
        int rc = enqueue(b, b->cap);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->cap, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return bhx_fail(ctx, BHX_E_HIP, std::string("capture: ") + hipGetErrorString(e));
        b->graph = g;
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipEventRecord(b->ev[0], st));
    HIPCHK(hipGraphLaunch(b->exec, st));
    HIPCHK(hipEventRecord(b->ev[1], st));
    return BHX_OK;
}

int bhx_batch_scene(const bhx_batch* b, int32_t i, bhx_rays* rays, int64_t* ray_base, int64_t* stride,
                    double** xyz) {
    if (!b || i < 0 || i >= b->n) return BHX_E_INVALID;
    if (rays) {
        rays->cls = static_cast<uint8_t*>(b->cls.p);
        rays->y = static_cast<double*>(b->y.p);
        rays->jet = static_cast<double*>(b->jet.p);
        rays->steps = static_cast<int32_t*>(b->steps.p);
        rays->attempts = static_cast<int32_t*>(b->attempts.p);
        rays->norm_err = static_cast<double*>(b->norm.p);
        rays->carter = static_cast<double*>(b->carter.p);
    }
    if (ray_base) *ray_base = b->ray_base[i];
    if (stride) *stride = b->n_rays;
    if (xyz) *xyz = static_cast<double*>(b->xyz.p) + 3 * b->pix_base[i];
    return BHX_OK;
}

int bhx_batch_copy_scene(bhx_batch* b, int32_t i, uint8_t* cls, double* y, double* jet, int32_t* steps,
                         int32_t* attempts, double* xyz, void* stream) {
    if (!b || i < 0 || i >= b->n) return BHX_E_INVALID;
    bhx_context* ctx = b->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    const bhx_scene& s = b->scenes[i];
    const int S = std::max(1, s.sqrt_spp);
    const int64_t n = static_cast<int64_t>(s.width) * s.height * S * S;
    const int64_t r0 = b->ray_base[i], N = b->n_rays;
    if (cls) HIPCHK(hipMemcpy(cls, static_cast<uint8_t*>(b->cls.p) + r0, n, hipMemcpyDeviceToHost));
    if (y)
        for (int c = 0; c < 8; ++c)
            HIPCHK(hipMemcpy(y + c * n, static_cast<double*>(b->y.p) + c * N + r0, sizeof(double) * n,
                             hipMemcpyDeviceToHost));
    if (jet) HIPCHK(hipMemcpy(jet, static_cast<double*>(b->jet.p) + r0, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (steps)
        HIPCHK(hipMemcpy(steps, static_cast<int32_t*>(b->steps.p) + r0, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    if (attempts)
        HIPCHK(hipMemcpy(attempts, static_cast<int32_t*>(b->attempts.p) + r0, sizeof(int32_t) * n,
                         hipMemcpyDeviceToHost));
    if (xyz)
        HIPCHK(hipMemcpy(xyz, static_cast<double*>(b->xyz.p) + 3 * b->pix_base[i],
                         sizeof(double) * 3 * s.width * s.height, hipMemcpyDeviceToHost));
    return BHX_OK;
}

int bhx_batch_stats(bhx_batch* b, bhx_stats* stats, double* batch_ms, void* stream) {
    if (!b || !stats) return BHX_E_INVALID;
    bhx_context* ctx = b->ctx;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    std::vector<unsigned long long> words(counter_words(b->n));
    HIPCHK(hipMemcpy(words.data(), b->counters.p, 8 * words.size(), hipMemcpyDeviceToHost));
    std::vector<double> flux(b->flux_total);
    HIPCHK(hipMemcpy(flux.data(), b->flux.p, sizeof(double) * flux.size(), hipMemcpyDeviceToHost));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->last_ms = ms;
    if (batch_ms) *batch_ms = ms;
    const SceneCounters* sc = reinterpret_cast<const SceneCounters*>(words.data());
    for (int i = 0; i < b->n; ++i) {
        const bhx_scene& s = b->scenes[i];
        const int S = std::max(1, s.sqrt_spp);
        const int64_t rays = static_cast<int64_t>(s.width) * s.height * S * S;
        bhx_stats& o = stats[i];
        std::memset(&o, 0, sizeof o);
        o.rays = rays;
        o.steps = static_cast<int64_t>(sc[i].trace[0]);
        o.attempts = static_cast<int64_t>(sc[i].trace[1]);
        o.rhs_evals = o.attempts * 6 + rays;
        std::memcpy(&o.max_norm_error, &sc[i].trace[2], 8);
        std::memcpy(&o.max_carter_drift, &sc[i].trace[3], 8);
        o.captured = static_cast<int64_t>(sc[i].classes[BHX_HORIZON]);
        o.disc_hits = static_cast<int64_t>(sc[i].classes[BHX_DISC]);
        o.sun_hits = static_cast<int64_t>(sc[i].classes[BHX_SUN]);
        o.star_hits = static_cast<int64_t>(sc[i].classes[BHX_STAR]);
        o.escaped = static_cast<int64_t>(sc[i].classes[BHX_SKY]);
        o.exhausted = static_cast<int64_t>(sc[i].classes[BHX_EXHAUSTED]);
        o.jet_rays = static_cast<int64_t>(sc[i].classes[6]);
        double f = 0.0;
        for (int k = 0; k < b->shade_blocks[i]; ++k) f += flux[b->flux_off[i] + k];
        o.total_flux = f;
        o.trace_ms = ms;
        o.seconds = ms * 1e-3;
    }
    return BHX_OK;
}

int bhx_batch_set_launch(bhx_batch* b, const bhx_launch* l) {
    if (!b || !l) return BHX_E_INVALID;
    bhx_context* ctx = b->ctx;
    if (l->block < 0 || (l->block % 64) != 0 || l->block > 256)
        return bhx_fail(ctx, BHX_E_INVALID, "block must be a multiple of 64 up to 256");
    if (l->refill < 0 || l->refill > 64) return bhx_fail(ctx, BHX_E_INVALID, "refill must be in [0, 64]");
    if (l->waves_per_simd < 0 || l->waves_per_simd > 8)
        return bhx_fail(ctx, BHX_E_INVALID, "waves_per_simd must be in [0, 8]");
    if (l->chunk < 0) return bhx_fail(ctx, BHX_E_INVALID, "chunk must be >= 0");
    bhx_launch nl = b->launch;
    if (l->block) nl.block = l->block;
    if (l->waves_per_simd) nl.waves_per_simd = l->waves_per_simd;
    if (l->chunk) nl.chunk = l->chunk;
    if (l->refill) nl.refill = l->refill;
    if (std::memcmp(&nl, &b->launch, sizeof nl) != 0) {
        HIPCHK(hipDeviceSynchronize());   // a replay in flight still uses the old graph
        drop_graph(b);
        b->launch = nl;
    }
    return BHX_OK;
}

int bhx_batch_get_launch(const bhx_batch* b, bhx_launch* l) {
    if (!b || !l) return BHX_E_INVALID;
    *l = b->launch;
    return BHX_OK;
}

// Coordinate search over the launch shape on this batch: each candidate is
// captured, run once to warm up and timed over `reps` graph replays; two
// sweeps over (block, occupancy, chunk, refill) starting from the current
// shape.  Leaves the best shape set on the batch.
int bhx_batch_autotune(bhx_batch* b, int32_t reps, bhx_launch* best, double* best_ms, void* stream) {
    if (!b) return BHX_E_INVALID;
    bhx_context* ctx = b->ctx;
    hipStream_t st = static_cast<hipStream_t>(stream);
    reps = std::max(1, reps);
    auto time_shape = [&](const bhx_launch& l, double* ms) -> int {
        int rc = bhx_batch_set_launch(b, &l);
        if (rc) return rc;
        rc = bhx_batch_run(b, 1, st);
        if (rc) return rc;
        double tot = 0.0;
        for (int r = 0; r < reps; ++r) {
            rc = bhx_batch_run(b, 1, st);
            if (rc) return rc;
            HIPCHK(hipEventSynchronize(b->ev[1]));
            float t = 0.0f;
            HIPCHK(hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
            tot += t;
        }
        *ms = tot / reps;
        return BHX_OK;
    };
    bhx_launch cur = b->launch;
    if (cur.waves_per_simd == 0) cur.waves_per_simd = 2;
    double cur_ms = 0.0;
    int rc = time_shape(cur, &cur_ms);
    if (rc) return rc;
    const int blocks[] = {64, 128, 256};
    const int waves[] = {1, 2};
    const int chunks[] = {16, 32, 64, 128, 256};
    const int refills[] = {4, 8, 16, 32, 48};
    for (int sweep = 0; sweep < 2; ++sweep) {
        for (int dim = 0; dim < 4; ++dim) {
            const int* vals = dim == 0 ? blocks : dim == 1 ? waves : dim == 2 ? chunks : refills;
            const int nv = dim == 0 ? 3 : dim == 1 ? 2 : 5;
            for (int k = 0; k < nv; ++k) {
                bhx_launch cand = cur;
                int* field = dim == 0 ? &cand.block : dim == 1 ? &cand.waves_per_simd
                           : dim == 2 ? &cand.chunk : &cand.refill;
                if (*field == vals[k]) continue;
                *field = vals[k];
                double ms = 0.0;
                rc = time_shape(cand, &ms);
                if (rc) return rc;
                if (ms < cur_ms) {
                    cur = cand;
                    cur_ms = ms;
                }
            }
        }
    }
    rc = bhx_batch_set_launch(b, &cur);
    if (rc) return rc;
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return BHX_OK;
}

}  // extern "C"
